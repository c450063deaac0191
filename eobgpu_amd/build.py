"""Build libeobhip.so (HIP kernels for gfx950 + the C ABI) in-tree.

    python -m eobgpu_amd.build            # compile into eobgpu_amd/libeobhip.so

hipcc cross-compiles for gfx950 without a GPU.  The library is loaded from the
package directory by eobgpu_amd._lib.
"""
from __future__ import annotations

import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libeobhip.so")
SOURCES = ["eob_kernels.hip", "spline_kernels.hip", "likelihood_kernels.hip", "detector_kernels.hip", "imr_kernels.hip"]
ARCH = os.environ.get("EOBHIP_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-munsafe-fp-atomics",
    # packed FP32 instructions are kept out of every HIP build
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.sep not in c or os.path.exists(c)):
            return c
    raise RuntimeError("hipcc not found")


def build(verbose: bool = True) -> str:
    srcs = [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    tmp = OUT + ".tmp"
    cmd = [hipcc(), *FLAGS, "-o", tmp, *srcs, "-lhipfft"]
    t0 = time.time()
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + " ".join(cmd) + "\n" + res.stdout + res.stderr)
    os.replace(tmp, OUT)
    if verbose:
        print(f"built {OUT} in {time.time() - t0:.1f}s")
    return OUT


if __name__ == "__main__":
    build()
    sys.exit(0)
