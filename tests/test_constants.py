"""The transcendental constants typed into the HIP sources equal their values.

A constant mistyped in a late digit passes every check made at small
momenta and surfaces only where its term dominates (ln 7 enters the pr^8
radial-potential coefficient, which matters only in the plunge), so each
literal is checked here against numpy/math, on the CPU.
"""
import math
import os
import re

import pytest

CSRC = os.path.join(os.path.dirname(__file__), "..", "eobgpu_amd", "csrc")

NAMED = {
    "kPi": math.pi,
    "kPi2": math.pi ** 2,
    "kEulerGamma": 0.57721566490153286060651209,
    "kLn2": math.log(2.0),
    "kLn3": math.log(3.0),
    "kLn5": math.log(5.0),
    "kLn7": math.log(7.0),
}

# literals annotated in the sources: (file, literal, value)
INLINE = [
    ("eob_device.hpp", "0.9189385332046728", 0.5 * math.log(2.0 * math.pi)),
    ("eob_device.hpp", "0.3862742020231896", math.sqrt(15.0 / (32.0 * math.pi))),
    ("eob_device.hpp", "1.6487212707001282", math.sqrt(math.e)),
]


def _read(name):
    with open(os.path.join(CSRC, name)) as f:
        return f.read()


@pytest.mark.parametrize("fname", ["eob_math.hpp", "spline_kernels.hip"])
def test_named_constants(fname):
    src = _read(fname)
    found = dict(re.findall(r"constexpr double (k\w+)\s*=\s*([-0-9.eE+]+)\s*;", src))
    checked = 0
    for name, val in found.items():
        if name in NAMED:
            assert float(val) == pytest.approx(NAMED[name], rel=2e-16, abs=0), name
            checked += 1
    assert checked >= 1


@pytest.mark.parametrize("fname,lit,val", INLINE)
def test_inline_literals(fname, lit, val):
    assert lit in _read(fname)
    assert float(lit) == pytest.approx(val, rel=2e-16, abs=0)
