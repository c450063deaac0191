"""CPU checks that tie constants the product uses to the scripts that derive them:
the FP64 operation count per attempted DOP853 step that the integrator's
roofline is computed from (tests/flop_count.cpp, evaluating the library's own
Hamiltonian and flux code on counting scalars), and the economized series of
log_pos (tests/log_series_coefficients.py)."""
import json
import os
import re
import shutil
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def test_step_flops_matches_the_counter(tmp_path):
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path / "flop_count")
    subprocess.run([cxx, "-x", "c++", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-I" + os.path.join(ROOT, "eobgpu_amd", "csrc"),
                    "-O1", "-std=c++17", "-o", exe, os.path.join(HERE, "flop_count.cpp")], check=True, capture_output=True)
    out = json.loads(subprocess.run([exe], check=True, capture_output=True, text=True).stdout)
    sys.path.insert(0, ROOT)
    import bench
    assert out["step_flops"] == bench.STEP_FLOPS
    assert out["step_ops"] + out["step_trans"] == out["step_flops"]
    assert out["rhs_ops"] * 12 < out["step_ops"]


def test_log_pos_coefficients_are_the_economized_series():
    sys.path.insert(0, HERE)
    import log_series_coefficients as L
    coef, err = L.economized(6)
    assert err < 1e-17
    src = open(os.path.join(ROOT, "eobgpu_amd", "csrc", "eob_fastmath.hpp")).read()
    body = src[src.index("EOB_HD double log_pos(double x)"):]
    body = body[:body.index("const double hfsq")]
    lits = [float.fromhex(x) for x in re.findall(r"0x1\.[0-9a-f]+p-?\d+", body)]
    # Horner from the highest power down: the file lists z^6 .. z^0
    assert lits == list(reversed(coef)), (lits, coef)
