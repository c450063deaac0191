"""GPU tests on the benchmark workloads themselves:

  * 1024-binary uneven workload (q ~ U[1,10], chi ~ U[-0.9,0.9], log10 M ~ U[1,2],
    r0 ~ U[15,40] M, seed 2): its first 64 binaries with both passes at fixed
    step against the oracle's committed trajectories (identical counts and stop
    codes, states 1e-9), persistent and plain; adaptive within the oracle's own
    one-ulp spread;
  * spline + likelihood workload: synthetic chirp modes with 259 .. 60000
    knots splined and resampled to 2^20 samples, against scipy's not-a-knot
    CubicSpline and numpy's rfft and sum;
  * served workload: a short Poisson stream of proposal batches (r0 = 20 M), each
    result equal to the batch pipeline's on the same binaries, bit for bit;
  * capacity workload: 16384 binaries with r0 ~ U[30,60] M through the
    integrator with dense output on the device, checked by properties.
"""
import json
import os

import numpy as np
import pytest
import torch
from scipy.interpolate import CubicSpline

from eobgpu_amd import configs
from eobgpu_amd import pipeline as P

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _fixture(name):
    return json.load(open(os.path.join(GOLDEN, name)))


@pytest.mark.parametrize("mode", ["persistent", "plain"])
def test_uneven_fixed_step_against_oracle(dev, mode):
    g = _fixture("config2_fixed_64.json")
    p = configs.uneven_batch(n=1024, seed=2)
    n = len(g["num_steps"])
    add = P.make_additional_args(p["m1"][:n], p["m2"][:n], p["chi1"][:n], p["chi2"][:n])
    b = P.EOBBatch(add, max_step=g["max_step"], device=dev)
    b.set_initial_state(np.array(g["cond"]))
    print(f"uneven workload, 64 binaries, fixed step, {mode}", flush=True)
    b.integrate(rtol=g["rtol"], atol=g["rtol"], hInit=g["hInit"], fix_step=True, mode=mode, group=64)
    torch.cuda.synchronize()
    ns = b.num_steps.cpu().numpy()
    np.testing.assert_array_equal(ns, g["num_steps"])
    np.testing.assert_array_equal(b.stop_info.cpu().numpy(), g["stop_info"])
    # The right-hand sides agree to 1e-13 .. 1e-10 relative (the plunge end the
    # loosest: the device contracts into fused multiply-adds and has its own
    # elementary functions); over up to 51 718 fixed steps the differences
    # accumulate to 1.5e-9 at the last rows of the longest binaries (measured),
    # so the bound here is 3e-9 (the 16-binary runs of a few thousand steps hold 1e-9).
    for i, s in enumerate(g["samples"]):
        rows = torch.as_tensor(s["rows"], device=dev)
        d = b.dense[i, rows].cpu().numpy()
        np.testing.assert_allclose(d[:, 0], s["t"], rtol=1e-11, atol=1e-9)
        np.testing.assert_allclose(d[:, 1:], np.array(s["state"]), rtol=3e-9, atol=1e-13)


def test_uneven_adaptive_within_oracle_spread(oracle, dev):
    """Adaptive two-pass runs of the same 64 binaries: accepted-step counts within
    3x the oracle's own one-ulp count spread on this workload, the stop codes
    that the oracle's spread leaves stable equal."""
    spread = _fixture("adaptive_spread.json")["trajectory_64_uneven_rtol1e-10"]
    g = _fixture("config2_fixed_64.json")
    p = configs.uneven_batch(n=1024, seed=2)
    n = len(g["num_steps"])
    add = oracle.make_additional_args(p["m1"][:n], p["m2"][:n], p["chi1"][:n], p["chi2"][:n])
    co = oracle.fac_coeffs(add)
    cond = np.array(g["cond"])
    t_r, traj_r, ns_r, si_r = oracle.run_trajectory(cond, add, co, max_step=4096, rtol=1e-10, atol=1e-10)
    b = P.EOBBatch(add, max_step=4096, device=dev)
    b.set_initial_state(cond)
    b.integrate(rtol=1e-10, atol=1e-10, group=64)
    torch.cuda.synchronize()
    ns = b.num_steps.cpu().numpy()
    p1 = b.pass1_steps.cpu().numpy()
    si = b.stop_info.cpu().numpy()
    stable = np.array(spread["stable_codes"], dtype=bool)
    assert stable.sum() >= 0.9 * n
    np.testing.assert_array_equal(si[stable], si_r[stable])
    d = np.abs(ns - ns_r)[stable]
    assert np.all(d <= 3 * spread["count"] * ns_r[stable] + 2), (ns[stable], ns_r[stable])
    assert np.all(p1 < 4095) and np.all(ns < 4095)


def _scipy_hplus(t, phi, re, im, M, dt, N):
    dT = dt / (M * P.MTSUN_SI)
    nsamp = min(N, int((t[-1] - t[0]) / dT) + 1)
    tj = t[0] + np.arange(nsamp) * dT
    s = [CubicSpline(t, y, bc_type="not-a-knot") for y in (phi, re, im)]
    ph = s[0](tj)
    h = np.zeros(N)
    h[:nsamp] = s[1](tj) * np.cos(2 * ph) + s[2](tj) * np.sin(2 * ph)
    return h, s


def test_spline_workload_against_scipy(oracle, dev):
    """Lengths just above twice the 96-row chunk minimum, not a multiple of 16,
    about 5k, just above 32768 (256 chunks) and 60k, max_step the largest.
    The three short chirps with 33k-60k knots put more than 1024 knots under
    58 of their 4096-sample spans, so the resample kernels' bisection over the
    whole knot array (the path taken when a span's knots overflow the LDS
    window) runs as well as the LDS one."""
    K = np.array([259, 300, 5003, 5000, 32771, 33000, 45013, 60000])
    N, dt, Msun = 1 << 20, 1.0 / 4096, 30.0
    ts, phis, res, ims, K = configs.chirp_modes(n=K.size, seed=31, N=N, dt=dt, Msun=Msun, K=K)
    n, kmax = K.size, int(K.max())
    b = P.EOBBatch(P.make_additional_args(np.full(n, 0.6), np.full(n, 0.4), np.zeros(n), np.zeros(n)), max_step=kmax, device=dev)
    dense = np.zeros((n, kmax, 5))
    knots = np.zeros((n, kmax, 3))
    for i in range(n):
        dense[i, :K[i], 0] = ts[i]
        knots[i, :K[i]] = np.stack([phis[i], res[i], ims[i]], axis=1)
    b.dense.copy_(torch.from_numpy(dense))
    b.knots = torch.from_numpy(knots).to(dev)
    b.num_steps.copy_(torch.from_numpy(K.astype(np.int32)))
    coef = b.build_splines().cpu().numpy()
    hp = b.resample_hplus(np.full(n, Msun), dt, N)
    ht = P.rfft_batched(hp)
    h_ref = np.zeros((n, N))
    for i in range(n):
        h_ref[i], spl = _scipy_hplus(ts[i], phis[i], res[i], ims[i], Msun, dt, N)
        dx = np.diff(ts[i])
        for q in range(3):
            c = coef[i, :K[i] - 1, 4 * q:4 * q + 4]
            ymax = np.abs(spl[q].c[3]).max()
            if ymax == 0.0:
                assert not c.any()
                continue
            # each coefficient's term over its interval, in units of max |y|
            for k in (1, 2, 3):
                err = np.max(np.abs(c[:, k] - spl[q].c[3 - k]) * dx ** k) / ymax
                assert err < 1e-11, (K[i], q, k, err)
    df = 1.0 / (N * dt)
    h0 = np.fft.rfft(h_ref[0])
    w = oracle.psd_weight(np.arange(N // 2 + 1) * df)
    rng = np.random.default_rng(8)
    noise = np.sqrt(np.where(w > 0, 1.0 / np.where(w > 0, w, 1.0), 0.0) / (4 * df)) * (rng.normal(size=w.size) + 1j * rng.normal(size=w.size))
    amp = 20.0 / np.sqrt(4 * df * np.sum(np.abs(h0) ** 2 * w))
    data = amp * h0 + noise / np.sqrt(2)
    logl_ref = oracle.fd_logl(amp * h_ref, data, df)
    d = torch.tensor(data, device=dev)
    logl = P.fd_likelihood(ht * amp, d, df).cpu().numpy()
    np.testing.assert_allclose(logl, logl_ref, rtol=1e-10)
    lik = P.SupportLikelihood(d, N, df, device=dev)
    logl_s = lik(b, np.full(n, Msun), dt, amp=amp).cpu().numpy()
    np.testing.assert_allclose(logl_s, logl_ref, rtol=1e-10)


def test_served_poisson_stream_equals_batch(dev):
    """A 50 ms Poisson stream (200 requests/s, U{1..256} binaries each, r0 = 20 M):
    every request's log L equals the batch pipeline's on its binaries."""
    from eobgpu_amd.serving import Request, Server
    N, dt = 1 << 12, 1.0 / 2048
    rng = np.random.default_rng(3)
    F = N // 2 + 1
    data = (rng.normal(size=F) + 1j * rng.normal(size=F)) * 1e-3
    reqs = [Request(i, t, p) for i, (t, p) in enumerate(configs.served_requests(duration_s=0.05, seed=44))]
    assert len(reqs) >= 3
    srv = Server(data, df=1.0 / (N * dt), dt=dt, N=N, capacity=2048, max_step=4096, streams=4)
    srv.warm()
    res = {r.rid: r for r in srv.run(reqs)}
    assert sorted(res) == [r.rid for r in reqs]
    d = torch.as_tensor(data, dtype=torch.complex128, device=dev)
    for r in reqs:
        p = r.params
        b = P.EOBBatch(P.make_additional_args(p["m1"], p["m2"], p["chi1"], p["chi2"]), max_step=4096, device=dev)
        b.initial_conditions(p["r0"])
        b.integrate(rtol=1e-10, atol=1e-10, two_pass=True)
        b.hlm22_knots()
        b.build_splines()
        hp = b.resample_hplus(p["m1"] + p["m2"], dt, N)
        ht = P.rfft_batched(hp)
        ref = P.fd_likelihood(ht, d, 1.0 / (N * dt)).cpu().numpy()
        np.testing.assert_array_equal(res[r.rid].logl, ref)


def test_capacity_workload_properties(dev):
    """16384 binaries, r0 ~ U[30,60] M, two-pass adaptive trajectories with dense
    output on the device (5.4 GB at 8192 rows): every binary ends with a stop code
    (3, 4, 5, or the non-finite guard's -1) inside its rows, its times strictly
    increase and its states are finite."""
    p = configs.capacity_batch()
    n, ms = p["m1"].size, 8192
    b = P.EOBBatch.from_params(p["m1"], p["m2"], p["chi1"], p["chi2"], max_step=ms, device=dev)
    b.initial_conditions(p["r0"])
    print("capacity workload: 16384 binaries", flush=True)
    b.integrate(rtol=1e-10, atol=1e-10, group=64)
    torch.cuda.synchronize()
    ns = b.num_steps
    si = b.stop_info.cpu().numpy()
    assert set(np.unique(si).tolist()) <= {3, 4, 5, -1}, np.unique(si, return_counts=True)
    assert int(ns.max()) < ms - 1 and int(ns.min()) > 10
    rows = torch.arange(ms, device=dev)[None, :]
    live = rows < ns[:, None].to(torch.int64)
    t = b.dense[:, :, 0]
    inc = (t[:, 1:] > t[:, :-1]) | ~live[:, 1:]
    assert bool(inc.all())
    ok = torch.isfinite(b.dense).all(dim=2) | ~live
    good = si != -1
    assert bool(ok[torch.as_tensor(good, device=dev)].all())


def test_plunge_edge_binaries(oracle, dev):
    """The three binaries of the 1024-binary workload that the oracle's adaptive
    run ends with -1 (the next stage leaves the Hamiltonian's domain before any
    stop condition holds; the GPU's adaptive run ends two of them with code 4,
    the adaptive step sequences of two FP64 builds parting at rounding level
    there).  With pass 1 at fixed step the sequence is deterministic: identical
    stored points and stop codes, states 3e-9 (the fixture's bound, see the
    test above); adaptive, the device's codes are among those the oracle's own
    one-ulp spread reaches on these binaries: perturbing one initial state
    component by one ulp moves the oracle's endings between -1 and 4 or 5."""
    g = _fixture("plunge_edge.json")
    idx = np.array(g["index"])
    p = configs.uneven_batch(n=1024, seed=2)
    add = P.make_additional_args(p["m1"][idx], p["m2"][idx], p["chi1"][idx], p["chi2"][idx])
    cond = np.array(g["cond"])
    b = P.EOBBatch(add, max_step=g["max_step"], device=dev)
    b.set_initial_state(cond)
    print("plunge-edge binaries, fixed step", flush=True)
    b.integrate(rtol=g["rtol"], atol=g["rtol"], hInit=g["hInit"], fix_step=True, group=64)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(b.num_steps.cpu().numpy(), g["num_steps"])
    np.testing.assert_array_equal(b.stop_info.cpu().numpy(), g["stop_info"])
    for i, s in enumerate(g["samples"]):
        d = b.dense[i, torch.as_tensor(s["rows"], device=dev)].cpu().numpy()
        np.testing.assert_allclose(d[:, 0], s["t"], rtol=1e-11, atol=1e-9)
        np.testing.assert_allclose(d[:, 1:], np.array(s["state"]), rtol=3e-9, atol=1e-13)
    b2 = P.EOBBatch(add, max_step=4096, device=dev)
    b2.set_initial_state(cond)
    b2.integrate(rtol=1e-10, atol=1e-10, group=64)
    torch.cuda.synchronize()
    si = b2.stop_info.cpu().numpy()
    assert set(si.tolist()) <= {-1, 4} | set(g["stop_info"]), si
