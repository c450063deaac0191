"""Seeded synthetic workloads for the batched SEOBNR path.

Each function draws the binary parameters of one workload with
``numpy.random.default_rng(seed)``; masses are in solar masses, spins are the
aligned components, ``r0`` is the initial separation in units of total mass M.

  small_batch      16 binaries,   q ~ U[1,5],  M = 30,            chi ~ U[-0.5,0.5], r0 = 20
  uneven_batch     1024 binaries, q ~ U[1,10], log10 M ~ U[1,2],  chi ~ U[-0.9,0.9], r0 ~ U[15,40]
  served_requests  Poisson arrivals (200/s) of proposal batches of U{1..256} binaries drawn as small_batch
  capacity_batch   16384 binaries with r0 ~ U[30,60] (otherwise as uneven_batch)
  chirp_modes      512 synthetic (2,2) modes on 5k-60k chirp-spaced knots (spline + likelihood workload)
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


def chirp_modes(n=512, seed=3, N=1 << 20, dt=1.0 / 4096, Msun=30.0, K=None):
    """Synthetic (2,2) modes on non-uniform knots: K_b ~ U{5000..60000} knots with
    chirp spacing t_k = T (1 - (1 - u_k)^(3/8)) (u_k sorted uniform, t in units of
    M), an orbital phase whose GW frequency is 20 Hz x^(-3/8) (x = 1 - t/T) and an
    amplitude 1e-21 x^(-1/4); T is 5-90 % of the data span N dt.  K (optional)
    fixes the knot counts.  Returns (t, phi, re, im, K): lists of arrays, knot
    series phi and Re, Im of h22 e^{2 i phi}."""
    rng = np.random.default_rng(seed)
    K = rng.integers(5000, 60001, n) if K is None else np.asarray(K, dtype=np.int64)
    span = N * dt / (Msun * MTSUN_SI)                 # the data span in units of M
    T = span * rng.uniform(0.05, 0.9, n)
    ts, phis, res, ims = [], [], [], []
    for i in range(n):
        u = np.sort(rng.uniform(0.0, 1.0, int(K[i])))
        u[0], u[-1] = 0.0, 1.0 - 1e-6
        t = T[i] * (1.0 - (1.0 - u) ** 0.375)
        x = 1.0 - t / T[i]
        f0 = 20.0 * Msun * MTSUN_SI                   # 20 Hz in units of 1/M
        ts.append(t)
        phis.append(np.pi * f0 * T[i] * 1.6 * (1.0 - x ** 0.625))
        res.append(1e-21 * x ** -0.25)
        ims.append(np.zeros(int(K[i])))
    return ts, phis, res, ims, K
