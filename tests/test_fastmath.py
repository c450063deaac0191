"""The right-hand side's arithmetic helpers (eob_fastmath.hpp) on the CPU.

log_pos against correctly rounded logarithms from Python's decimal module:
within 1 ulp over the radii and frequencies the integrator sees and over the
whole exponent range; NaN outside the positive finite numbers (the header is
compiled with contraction off, which rounds as the device build does).
qdiv / qrcp / qsqrt, started from single-precision estimates, against IEEE
division and square root (Python's, correctly rounded): within 1 ulp, and
equal in all but a handful of cases.  cbrt_pos against cube roots from
decimal: within 1 ulp."""
import ctypes
import math
import os
import shutil
import subprocess
from decimal import Decimal, getcontext

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def probe(tmp_path_factory):
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("g++ not available")
    out = str(tmp_path_factory.mktemp("fastmath") / "probe.so")
    subprocess.run([cxx, "-O2", "-ffp-contract=off", "-fPIC", "-shared", "-o", out,
                    os.path.join(HERE, "fastmath_probe.cpp")], check=True)
    lib = ctypes.CDLL(out)
    for name in ("probe_log_pos", "probe_qrcp", "probe_qsqrt", "probe_cbrt_pos"):
        getattr(lib, name).argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
    lib.probe_qdiv.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_int]
    lib.probe_qsqrt_recip.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_int]

    def run(x, fn="probe_log_pos", b=None):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty_like(x)
        if fn == "probe_qsqrt_recip":
            inv = np.empty_like(x)
            lib.probe_qsqrt_recip(x.ctypes.data, y.ctypes.data, inv.ctypes.data, len(x))
            return y, inv
        if b is None:
            getattr(lib, fn)(x.ctypes.data, y.ctypes.data, len(x))
        else:
            b = np.ascontiguousarray(b, dtype=np.float64)
            getattr(lib, fn)(x.ctypes.data, b.ctypes.data, y.ctypes.data, len(x))
        return y
    return run


def _ulp_err(x, y):
    getcontext().prec = 40
    worst = 0.0
    for xi, yi in zip(x, y):
        ref = Decimal(float(xi)).ln()
        r = float(ref)
        u = math.ulp(r) if r != 0.0 else math.ulp(0.0)
        worst = max(worst, float(abs(Decimal(float(yi)) - ref) / Decimal(u)))
    return worst


def test_within_one_ulp(probe):
    rng = np.random.default_rng(11)
    x = np.concatenate([
        rng.uniform(1.0, 100.0, 8000),          # radii
        10.0 ** rng.uniform(-2.5, 0.0, 8000),   # v = Omega^(1/3)
        rng.uniform(0.70, 1.42, 4000),          # around the reduction's split points
        2.0 ** rng.uniform(-1070.0, 1023.0, 4000),
        np.array([1.0, 2.0, 0.5, math.sqrt(0.5), np.nextafter(math.sqrt(0.5), 1.0), np.nextafter(1.0, 0.0),
                  np.nextafter(1.0, 2.0), 5e-324, 1.7976931348623157e308]),
    ])
    y = probe(x)
    assert y[-9] == 0.0 and y[-8] == math.log(2.0)
    assert _ulp_err(x, y) <= 1.0


def test_non_positive_and_non_finite(probe):
    y = probe(np.array([0.0, -0.0, -1.0, -np.inf, np.inf, np.nan]))
    assert np.all(np.isnan(y))


def _ulps(got, ref):
    return np.abs(got - ref) / np.spacing(np.abs(ref))


def test_division_and_square_root(probe):
    rng = np.random.default_rng(12)
    a = rng.choice([-1.0, 1.0], 50000) * 10.0 ** rng.uniform(-8, 8, 50000)
    b = 10.0 ** rng.uniform(-8, 8, 50000)
    root, inv = probe(b, "probe_qsqrt_recip")
    np.testing.assert_array_equal(root, probe(b, "probe_qsqrt"))
    assert np.max(_ulps(inv, (1.0 / np.sqrt(np.float128(b))).astype(np.float64))) <= 1.0   # 80-bit reference
    for got, ref in ((probe(a, "probe_qdiv", b), a / b), (probe(b, "probe_qrcp"), 1.0 / b),
                     (probe(b, "probe_qsqrt"), np.sqrt(b))):
        u = _ulps(got, ref)
        assert u.max() <= 1.0
        assert np.count_nonzero(got != ref) <= 5



def test_cube_root(probe):
    from decimal import Decimal, getcontext
    getcontext().prec = 40
    rng = np.random.default_rng(14)
    x = np.concatenate([10.0 ** rng.uniform(-6, 8, 6000), rng.uniform(0.5, 8.0, 2000), np.array([1.0, 8.0, 27.0, 0.125])])
    y = probe(x, "probe_cbrt_pos")
    worst = 0.0
    for xi, yi in zip(x, y):
        ref = Decimal(float(xi)) ** (Decimal(1) / 3)
        worst = max(worst, float(abs(Decimal(float(yi)) - ref) / Decimal(math.ulp(float(ref)))))
    assert worst <= 1.0
    assert np.all(np.isnan(probe(np.array([0.0, np.nan]), "probe_cbrt_pos")))
