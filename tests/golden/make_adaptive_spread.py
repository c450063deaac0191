"""Writes adaptive_spread.json: how far the oracle's own adaptive DOP853 results
move when the initial state moves by one ulp.

The reference's controller (4th-power error norm, alpha = 1/8) decides accept
or reject on the err = 1 boundary at almost every step, so accepted-step counts
and states at matched times are not reproducible between two floating-point
implementations; the GPU parity tests bound the device's differences from the
oracle by a multiple of this spread.  Oracle only (CPU); run from the repository
root:  python tests/golden/make_adaptive_spread.py

  trajectory: the 16-binary workload (q ~ U[1,5], M = 30, chi ~ U[-0.5,0.5],
    r0 = 20, seed 1), one adaptive DOPR853.integrate pass at rtol = atol = 1e-10
    (per binary), initial state (r0, pr0, pphi0) moved by +-1 ulp one component
    at a time; recorded: the largest relative change of a binary's accepted-step
    count, and of r, phi, pphi at matched times (cubic splines of both
    trajectories on 1500 times up to 98 % of the shorter one, relative to the
    component's largest magnitude);
  uneven trajectories: the first 64 binaries of the 1024-binary workload
    (q ~ U[1,10], chi ~ U[-0.9,0.9], log10 M ~ U[1,2], r0 ~ U[15,40] M, seed 2),
    two-pass adaptive runs; recorded: which binaries keep their stop code, and
    the largest relative change of the stored points of those binaries;
  log L: the first 32 binaries of the 1024-binary workload (seed 2), the whole
    pass (two-pass trajectory at rtol = 1e-10, (2,2) knots, splines, h+ on 2^20
    samples at 1/4096 s, rfft, log L against binary 0 injected at SNR 20 plus
    noise with the analytic PSD, seed 7); recorded: the largest relative change
    of log L under the same perturbations.
"""
import json
import os
import sys

import numpy as np
from scipy.interpolate import CubicSpline

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from eobgpu_amd import configs  # noqa: E402


def matched_spread(ta, ya, tb, yb):
    tt = np.linspace(0.0, 0.98 * min(ta[-1], tb[-1]), 1500)
    fa = CubicSpline(ta, ya)(tt)
    fb = CubicSpline(tb, yb)(tt)
    return float(np.max(np.abs(fa - fb)) / np.abs(fa).max())


def perturbations(cond):
    out = []
    for comp in (0, 2, 3):
        for sgn in (+1, -1):
            c = cond.copy()
            c[comp] = np.nextafter(c[comp], sgn * np.inf)
            out.append(c)
    return out


def trajectory_spread():
    p = configs.small_batch(n=16, seed=1)
    add = O.make_additional_args(p["m1"], p["m2"], p["chi1"], p["chi2"])
    co = O.fac_coeffs(add)
    pphi0, pr0, om0, add = O.initial_conditions(p["r0"], add, co)
    cond = np.array([p["r0"], np.zeros(16), pr0, pphi0])
    base = O.integrate(cond, add, co, max_step=4096, rtol=1e-10, atol=1e-10)
    worst = {"count": 0.0, "r": 0.0, "phi": 0.0, "pphi": 0.0}
    for c in perturbations(cond):
        dL, dO, sn, si, _ = O.integrate(c, add, co, max_step=4096, rtol=1e-10, atol=1e-10)
        worst["count"] = max(worst["count"], float(np.max(np.abs(sn - base[2]) / base[2])))
        for b in range(16):
            ta, ya = base[0][:base[2][b] + 1, b], base[1][:base[2][b] + 1, :, b]
            tb, yb = dL[:sn[b] + 1, b], dO[:sn[b] + 1, :, b]
            for name, comp in (("r", 0), ("phi", 1), ("pphi", 3)):
                worst[name] = max(worst[name], matched_spread(ta, ya[:, comp], tb, yb[:, comp]))
    return worst


def uneven_trajectory_spread(nb=64):
    """Two-pass adaptive trajectories of the first nb binaries of the 1024-binary
    workload: the largest relative change of a binary's stored points under the
    one-ulp perturbations, and which binaries keep their stop code under all of
    them."""
    p = configs.uneven_batch(n=1024, seed=2)
    idx = np.arange(nb)
    add = O.make_additional_args(p["m1"][idx], p["m2"][idx], p["chi1"][idx], p["chi2"][idx])
    co = O.fac_coeffs(add)
    pphi0, pr0, om0, add = O.initial_conditions(p["r0"][idx], add, co)
    cond = np.array([p["r0"][idx], np.zeros(nb), pr0, pphi0])
    base = O.run_trajectory(cond, add, co, max_step=4096, rtol=1e-10, atol=1e-10)
    runs = [O.run_trajectory(c, add, co, max_step=4096, rtol=1e-10, atol=1e-10) for c in perturbations(cond)]
    stable = np.ones(nb, dtype=bool)
    for t, traj, ns, si in runs:
        stable &= si == base[3]
    # a changed stop code moves the end of pass 1 and with it the whole count; the
    # count spread is taken over the binaries whose code does not change
    worst = max(float(np.max(np.abs(ns - base[2])[stable] / base[2][stable])) for t, traj, ns, si in runs)
    return {"count": worst, "stable_codes": stable.astype(int).tolist()}


def logl_spread(nb=32):
    p = configs.uneven_batch(n=1024, seed=2)
    idx = np.arange(nb)
    dt, N = 1.0 / 4096, 1 << 20
    df = 1.0 / (N * dt)
    sub = {k: p[k][idx] for k in ("m1", "m2", "chi1", "chi2", "r0")}
    f = np.arange(N // 2 + 1) * df
    w = O.psd_weight(f)
    # data: binary 0 at SNR 20 plus noise (the bench's recipe)
    _, _, h = O.pipeline_logl(sub["m1"][:1], sub["m2"][:1], sub["chi1"][:1], sub["chi2"][:1], sub["r0"][:1],
                              np.zeros(N // 2 + 1, dtype=complex), dt, N)
    h0 = np.fft.rfft(h[0])
    amp = 20.0 / np.sqrt(4 * df * np.sum(np.abs(h0) ** 2 * w))
    rng = np.random.default_rng(7)
    S = np.where(f > 0, 1.0 / np.where(w > 0, w, 1.0), 0.0)
    data = amp * h0 + np.sqrt(S / (4 * df)) * (rng.normal(size=f.size) + 1j * rng.normal(size=f.size)) / np.sqrt(2)

    def run(r0_shift=None, comp=None, sgn=0):
        add = O.make_additional_args(sub["m1"], sub["m2"], sub["chi1"], sub["chi2"])
        co = O.fac_coeffs(add)
        pphi0, pr0, om0, add = O.initial_conditions(sub["r0"], add, co)
        cond = np.array([sub["r0"], np.zeros(nb), pr0, pphi0])
        if comp is not None:
            cond[comp] = np.nextafter(cond[comp], sgn * np.inf)
        t, traj, ns, si = O.run_trajectory(cond, add, co, max_step=4096, rtol=1e-10, atol=1e-10)
        hp = O.hplus_from_trajectory(t, traj, ns, add, co, sub["m1"] + sub["m2"], dt, N, amp)
        return O.fd_logl(hp, data, df)

    base = run()
    worst = 0.0
    for comp in (0, 2, 3):
        for sgn in (+1, -1):
            worst = max(worst, float(np.max(np.abs(run(comp=comp, sgn=sgn) - base) / np.abs(base))))
    return worst


if __name__ == "__main__":
    out = {"trajectory_16_binaries_rtol1e-10": trajectory_spread()}
    print(out, flush=True)
    out["trajectory_64_uneven_rtol1e-10"] = uneven_trajectory_spread()
    print(out, flush=True)
    out["logl_rel_32_binaries_uneven_N2^20"] = logl_spread()
    print(out, flush=True)
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "adaptive_spread.json"), "w") as fh:
        json.dump(out, fh, indent=1)
