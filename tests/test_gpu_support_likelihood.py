"""GPU: the support-length likelihood equals the full-length path (resample to N
samples, N-point rfft, sum over N/2+1 bins) within 1e-10 relative, term by
term against numpy, with templates both shorter and longer than N/2."""
import numpy as np
import pytest
import torch

from eobgpu_amd import configs
from eobgpu_amd import pipeline as P

pytestmark = pytest.mark.gpu


def _batch(n, seed, r0, N, dt, dev):
    p = configs.small_batch(n=n, seed=seed)
    p["r0"] = np.full(n, r0)
    b = P.EOBBatch(P.make_additional_args(p["m1"], p["m2"], p["chi1"], p["chi2"]), max_step=4096, device=dev)
    b.initial_conditions(p["r0"])
    b.integrate(rtol=1e-10, atol=1e-10, two_pass=True)
    b.hlm22_knots()
    b.build_splines()
    return b, p


@pytest.mark.parametrize("N,dt,r0", [(1 << 14, 1.0 / 4096, 14.0), (1 << 9, 1.0 / 1024, 16.0)])
def test_support_likelihood_matches_full_length(dev, N, dt, r0):
    b, p = _batch(6, 41, r0, N, dt, dev)
    M = p["m1"] + p["m2"]
    F = N // 2 + 1
    df = 1.0 / (N * dt)
    amp = np.linspace(0.5, 2.0, b.n)
    rng = np.random.default_rng(5)
    hp = b.resample_hplus(M, dt, N, amp=amp)
    ht = P.rfft_batched(hp)
    data = ht[0].cpu().numpy() * 0.7 + (rng.normal(size=F) + 1j * rng.normal(size=F)) * 0.05 * np.abs(ht[0].cpu().numpy()).max()
    dt_t = torch.as_tensor(data, dtype=torch.complex128, device=dev)
    ref = P.fd_likelihood(ht, dt_t, df).cpu().numpy()
    lik = P.SupportLikelihood(data, N, df, m_min_log2=6)
    terms = torch.empty((b.n, 2), dtype=torch.float64, device=dev)
    got = lik(b, M, dt, amp=amp, terms=terms).cpu().numpy()
    np.testing.assert_array_equal(lik.nsamp.cpu().numpy(), b.nsamp.cpu().numpy())
    np.testing.assert_allclose(got, ref, rtol=1e-10, atol=0)
    # the two template terms against numpy on the full-length spectra
    f = np.arange(F) * df
    w = np.where(f > 0, 1.0 / P.analytic_psd(f), 0.0)
    H = ht.cpu().numpy()
    cross = np.sum((np.conj(data)[None] * H * w[None]).real, axis=1)
    hh = np.sum(np.abs(H) ** 2 * w[None], axis=1)
    tm = terms.cpu().numpy()
    np.testing.assert_allclose(tm[:, 0], cross, rtol=1e-10, atol=1e-12 * np.abs(cross).max())
    np.testing.assert_allclose(tm[:, 1], hh, rtol=1e-10)
    ml = lik.mlog2.cpu().numpy()
    ns = lik.nsamp.cpu().numpy()
    assert np.all((1 << ml) >= np.minimum(2 * ns, N)) and np.all((1 << ml) <= N)
