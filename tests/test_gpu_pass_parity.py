"""End-to-end parity of the whole pass on the 16-binary workload (q ~ U[1,5],
M = 30, chi ~ U[-0.5,0.5], r0 = 20, seed 1): trajectory (both passes at fixed
step, a deterministic step sequence), (2,2) knots, splines, h+ on 2^14 samples
at 1/2048 s, FFT and log L, against the oracle's chain on the same initial
state and the same data (oracle.pipeline_logl: get_hlms, interpolate_TD,
spline evaluation, numpy rfft and the get_ll sum, eob/seobnrv4phm.py:742-807,
2138-2372).  The bound on log L is 1e-10 relative."""
import numpy as np
import pytest
import torch

from eobgpu_amd import configs
from eobgpu_amd import pipeline as P

pytestmark = pytest.mark.gpu

N, DT = 1 << 14, 1.0 / 2048


def _oracle_pass(oracle):
    p = configs.small_batch(n=16, seed=1)
    add = oracle.make_additional_args(p["m1"], p["m2"], p["chi1"], p["chi2"])
    co = oracle.fac_coeffs(add)
    pphi0, pr0, om0, add = oracle.initial_conditions(p["r0"], add, co)
    cond = np.array([p["r0"], np.zeros(16), pr0, pphi0])
    t, traj, ns, si = oracle.run_trajectory(cond, add, co, max_step=8192, rtol=1e-10, atol=1e-10, hInit=5.0, fix_step1=True)
    M = p["m1"] + p["m2"]
    h = oracle.hplus_from_trajectory(t, traj, ns, add, co, M, DT, N)
    df = 1.0 / (N * DT)
    f = np.arange(N // 2 + 1) * df
    w = oracle.psd_weight(f)
    h0 = np.fft.rfft(h[0])
    snr = np.sqrt(4 * df * np.sum(np.abs(h0) ** 2 * w))
    amp = 20.0 / snr
    rng = np.random.default_rng(7)
    sig = np.sqrt(np.where(f > 0, 1.0 / np.where(w > 0, w, 1.0), 0.0) / (4 * df))
    data = amp * h0 + sig * (rng.normal(size=f.size) + 1j * rng.normal(size=f.size)) / np.sqrt(2)
    logl = oracle.fd_logl(amp * h, data, df)
    return p, add, cond, ns, si, amp, data, logl, df


def test_whole_pass_logl_matches_oracle(oracle, dev):
    p, add, cond, ns_r, si_r, amp, data, logl_r, df = _oracle_pass(oracle)
    M = p["m1"] + p["m2"]
    b = P.EOBBatch(P.make_additional_args(p["m1"], p["m2"], p["chi1"], p["chi2"]), max_step=8192, device=dev)
    b.set_initial_state(cond)
    b.integrate(rtol=1e-10, atol=1e-10, hInit=5.0, fix_step=True, mode="persistent", group=64)
    b.hlm22_knots()
    b.build_splines()
    hp = b.resample_hplus(M, DT, N, amp=amp)
    ht = P.rfft_batched(hp)
    d = torch.tensor(data, device=dev)
    logl = P.fd_likelihood(ht, d, df).cpu().numpy()
    assert np.array_equal(b.num_steps.cpu().numpy(), ns_r)
    rel = np.abs(logl - logl_r) / np.abs(logl_r)
    print("log L relative difference, full-length path:", rel.max(), flush=True)
    assert rel.max() <= 1e-10, rel
    # the support-length likelihood (the headline path) on the same batch
    lik = P.SupportLikelihood(d, N, df, device=dev)
    out = torch.empty(16, dtype=torch.float64, device=dev)
    lik(b, M, DT, amp=amp, out=out)
    rel2 = np.abs(out.cpu().numpy() - logl_r) / np.abs(logl_r)
    print("log L relative difference, support-length path:", rel2.max(), flush=True)
    assert rel2.max() <= 1e-10, rel2
