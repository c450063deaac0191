"""Seeded synthetic workloads for the batched SEOBNR path.

Each function draws the binary parameters of one workload with
``numpy.random.default_rng(seed)``; masses are in solar masses, spins are the
aligned components, ``r0`` is the initial separation in units of total mass M.

  small_batch      16 binaries,   q ~ U[1,5],  M = 30,            chi ~ U[-0.5,0.5], r0 = 20
  uneven_batch     1024 binaries, q ~ U[1,10], log10 M ~ U[1,2],  chi ~ U[-0.9,0.9], r0 ~ U[15,40]
  served_requests  Poisson arrivals (200/s) of proposal batches of U{1..256} binaries drawn as small_batch
  capacity_batch   16384 binaries with r0 ~ U[30,60] (otherwise as uneven_batch)
"""
from __future__ import annotations

import numpy as np

MTSUN_SI = 4.925491025543575903411922162094833998e-6


def _masses(q, M):
    m1 = M * q / (1.0 + q)
    m2 = M / (1.0 + q)
    return m1, m2


def small_batch(n=16, seed=1):
    rng = np.random.default_rng(seed)
    q = rng.uniform(1.0, 5.0, n)
    M = np.full(n, 30.0)
    chi1 = rng.uniform(-0.5, 0.5, n)
    chi2 = rng.uniform(-0.5, 0.5, n)
    m1, m2 = _masses(q, M)
    return dict(m1=m1, m2=m2, chi1=chi1, chi2=chi2, r0=np.full(n, 20.0))


def uneven_batch(n=1024, seed=2, r0_range=(15.0, 40.0)):
    rng = np.random.default_rng(seed)
    q = rng.uniform(1.0, 10.0, n)
    chi1 = rng.uniform(-0.9, 0.9, n)
    chi2 = rng.uniform(-0.9, 0.9, n)
    M = 10.0 ** rng.uniform(1.0, 2.0, n)
    r0 = rng.uniform(r0_range[0], r0_range[1], n)
    m1, m2 = _masses(q, M)
    return dict(m1=m1, m2=m2, chi1=chi1, chi2=chi2, r0=r0)


def capacity_batch(n=16384, seed=5):
    return uneven_batch(n=n, seed=seed, r0_range=(30.0, 60.0))


def served_requests(duration_s=2.0, rate_hz=200.0, max_batch=256, seed=4):
    """List of (arrival_time_s, params dict) for Poisson arrivals."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    while True:
        t += rng.exponential(1.0 / rate_hz)
        if t > duration_s:
            break
        nb = int(rng.integers(1, max_batch + 1))
        sub_seed = int(rng.integers(0, 2**31 - 1))
        out.append((t, small_batch(n=nb, seed=sub_seed)))
    return out
