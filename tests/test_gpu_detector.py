"""GPU: detector projection (antenna patterns, delays, detector-frame signal,
eob/seobnrv4phm.py:597-634, 742-767) against the oracle's numpy restatement,
and the two-polarization resample (h+ + i hx = amp h22: the real and imaginary
parts of the template, as the reference transforms them) against the h+ resample
and scipy's splines."""
import numpy as np
import pytest
import torch
from scipy.interpolate import CubicSpline

from eobgpu_amd import configs, detector
from eobgpu_amd import pipeline as P

pytestmark = pytest.mark.gpu


def _detectors(rng, k=3):
    D, V = [], []
    for _ in range(k):
        q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        x, y = q[:, 0], q[:, 1]
        D.append(0.5 * (np.outer(x, x) - np.outer(y, y)))
        V.append(rng.normal(size=3) * 6.4e6)
    return np.array(D), np.array(V)


def test_detector_response_and_signal(oracle, dev):
    rng = np.random.default_rng(21)
    n = 300
    ra, dec, psi, gmst = rng.uniform(0, 2 * np.pi, n), rng.uniform(-1.5, 1.5, n), rng.uniform(0, np.pi, n), rng.uniform(0, 7, n)
    D, V = _detectors(rng)
    Fp, Fc, dt = detector.get_detector_information(ra, dec, gmst, psi, D, V, device=dev)
    Fp_r, Fc_r, dt_r = oracle.detector_information(ra, dec, gmst, psi, D, V)
    np.testing.assert_allclose(Fp.cpu().numpy(), Fp_r, rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(Fc.cpu().numpy(), Fc_r, rtol=1e-13, atol=1e-15)
    # the delay is a dot product (unit vector . vertex / c) that cancels for
    # directions across the vertex: bounded relative to the delays' scale
    np.testing.assert_allclose(dt.cpu().numpy(), dt_r, rtol=1e-13, atol=1e-13 * np.abs(dt_r).max())
    F, df, kmin = 4097, 0.25, 80
    nb = 7
    hp = rng.normal(size=(nb, F)) + 1j * rng.normal(size=(nb, F))
    hc = rng.normal(size=(nb, F)) + 1j * rng.normal(size=(nb, F))
    sig = detector.detector_signal(torch.tensor(hp, device=dev), torch.tensor(hc, device=dev), Fp[:, :nb].contiguous(),
                                   Fc[:, :nb].contiguous(), dt[:, :nb].contiguous(), df, k_min=kmin)
    ref = oracle.detector_signal(hp, hc, Fp_r[:, :nb], Fc_r[:, :nb], dt_r[:, :nb], df, k_min=kmin)
    np.testing.assert_allclose(sig.cpu().numpy(), ref, rtol=1e-12, atol=1e-13)


def test_two_polarization_resample(dev):
    N, dt, Msun = 1 << 16, 1.0 / 4096, 30.0
    K = np.array([300, 2000, 5003])
    ts, phis, res, ims, K = configs.chirp_modes(n=K.size, seed=9, N=N, dt=dt, Msun=Msun, K=K)
    n, kmax = K.size, int(K.max())
    b = P.EOBBatch(P.make_additional_args(np.full(n, 0.6), np.full(n, 0.4), np.zeros(n), np.zeros(n)), max_step=kmax, device=dev)
    dense = np.zeros((n, kmax, 5))
    knots = np.zeros((n, kmax, 3))
    for i in range(n):
        dense[i, :K[i], 0] = ts[i]
        knots[i, :K[i]] = np.stack([phis[i], res[i], 0.3 * res[i] * np.sin(1e-3 * ts[i])], axis=1)
    b.dense.copy_(torch.from_numpy(dense))
    b.knots = torch.from_numpy(knots).to(dev)
    b.num_steps.copy_(torch.from_numpy(K.astype(np.int32)))
    b.build_splines()
    hp1 = b.resample_hplus(np.full(n, Msun), dt, N, amp=2.0).clone()
    hp, hc = detector.resample_hpc(b, np.full(n, Msun), dt, N, amp=2.0)
    np.testing.assert_array_equal(hp.cpu().numpy(), hp1.cpu().numpy())
    dT = dt / (Msun * P.MTSUN_SI)
    for i in range(n):
        ns = min(N, int((ts[i][-1] - ts[i][0]) / dT) + 1)
        tj = ts[i][0] + np.arange(ns) * dT
        s = [CubicSpline(ts[i], knots[i, :K[i], q], bc_type="not-a-knot")(tj) for q in range(3)]
        h22 = 2.0 * (s[1] + 1j * s[2]) * np.exp(-2j * s[0])
        scale = np.abs(h22).max()
        assert np.max(np.abs(hp.cpu().numpy()[i, :ns] - h22.real)) < 1e-10 * scale
        assert np.max(np.abs(hc.cpu().numpy()[i, :ns] - h22.imag)) < 1e-10 * scale
        assert not hc[i, ns:].any()
