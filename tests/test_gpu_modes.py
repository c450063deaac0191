"""Higher modes: EOBGetSpinFactorizedWaveform for every (l, m) against the
oracle (src/EOB.cu:1236-1613), the batch's modes on its knots (get_hlms,
eob/seobnrv4phm.py:2138-2221), splines of the 2M + 1 series and the mode sum
on the uniform grid against scipy's not-a-knot CubicSpline, and the
TDInterp_wrap2 drop-in (src/interpolate.cu:441-549) against its restatement,
with the reference's stride and with the padded one."""
import numpy as np
import pytest
import torch
from scipy.interpolate import CubicSpline

from eobgpu_amd import configs
from eobgpu_amd import pipeline as P
from eobgpu_amd import pyEOB, pytdinterp

pytestmark = pytest.mark.gpu

ALL_MODES = [(l, m) for l in range(2, 9) for m in range(1, l + 1)]


def _T(x, dev):
    return torch.tensor(np.ascontiguousarray(x), device=dev)


def _knot_inputs(oracle, add, rng, Lk):
    """Knot inputs of EOBGetSpinFactorizedWaveform as get_hlms forms them from states."""
    nb = add.shape[1]
    m1, m2 = add[0], add[1]
    eta = m1 * m2 / (m1 + m2) ** 2
    r = np.sort(rng.uniform(2.5, 25.0, (nb, Lk)), axis=1)[:, ::-1].copy()
    L = np.sqrt(r) * 1.02
    pr = -0.01 * rng.uniform(0, 1, (nb, Lk))
    phi = rng.uniform(0, 60, (nb, Lk))
    v = np.empty_like(r); Hr = np.empty_like(r); vPhi = np.empty_like(r)
    for b in range(nb):
        a1 = add[:, b:b + 1].repeat(Lk, axis=1)
        st = np.array([r[b], phi[b], pr[b], L[b]])
        g = oracle.grad_H(st, a1)
        st0 = st.copy(); st0[2] = 0.0
        gc = oracle.grad_H(st0, a1)
        H = oracle.evaluate_H(r[b], phi[b], pr[b], L[b], a1[0], a1[1], a1[2], a1[3])
        v[b] = np.cbrt(g[3])
        Hr[b] = eta[b] * H
        vPhi[b] = 1.0 / (gc[3] ** 2 * r[b] ** 3)
    return eta, r, phi, pr, L, v, Hr, vPhi


@pytest.mark.parametrize("use_hm", [1, 0])
def test_spin_factorized_waveform_all_modes(oracle, dev, use_hm):
    """Every (l, m), l = 2..8, through the drop-in (update_information's tables,
    the per-call prefix and h22_calib) against the oracle; the batch includes an
    equal-mass binary (eta = 1/4: odd-m modes keep only f_lm)."""
    rng = np.random.default_rng(21)
    nb, Lk = 7, 64
    p = configs.small_batch(n=nb, seed=5)
    p["m1"][0] = p["m2"][0] = 15.0
    add = oracle.make_additional_args(p["m1"], p["m2"], p["chi1"], p["chi2"])
    co = oracle.fac_coeffs(add, use_hm=use_hm)
    eta, r, phi, pr, L, v, Hr, vPhi = _knot_inputs(oracle, add, rng, Lk)
    m1, m2, c1, c2 = add[0], add[1], add[2], add[3]
    chiS, chiA = 0.5 * (c1 + c2), 0.5 * (c1 - c2)
    a = (1 - 2 * eta) * chiS + (m1 - m2) / (m1 + m2) * chiA
    ode = pyEOB.SEOBNRv5Class()
    pref = _T(add[10:80].T.reshape(-1), dev)
    ode.update_information(_T(m1, dev), _T(m2, dev), _T(eta, dev), _T(a, dev), _T(chiS, dev), _T(chiA, dev), use_hm, nb, pref)
    idx = torch.arange(nb, dtype=torch.int32, device=dev)
    args = [_T(x, dev) for x in (r, phi, pr, L, v, Hr, eta, vPhi)]
    h22c = 0.25
    for q, (l, m) in enumerate(ALL_MODES):
        pf = add[10 + 2 * q] + 1j * add[11 + 2 * q]
        out = torch.zeros(2 * nb * Lk, dtype=torch.float64, device=dev)
        pfi = _T(np.stack([pf.real, pf.imag], axis=1).reshape(-1), dev)
        ode.EOBGetSpinFactorizedWaveform(out, *args[:8], l, m, pfi, h22c, nb, Lk, idx)
        o = out.cpu().numpy().reshape(nb, Lk, 2)
        h = o[..., 0] + 1j * o[..., 1]
        h_r = oracle.hlm(l, m, r, phi, L, v, Hr, eta, vPhi, pf, co, h22_calib=h22c)
        scale = np.abs(h_r).max(axis=1, keepdims=True)
        scale[scale == 0] = 1.0
        err = np.max(np.abs(h - h_r) / scale)
        assert err < 1e-11, ((l, m), err)
    with pytest.raises(Exception, match="outside"):
        ode.EOBGetSpinFactorizedWaveform(out, *args[:8], 9, 1, pfi, h22c, nb, Lk, idx)
    ode.deallocate_information()


def _batch_with_trajectories(oracle, dev, n=6, seed=3, r0=14.0):
    p = configs.small_batch(n=n, seed=seed)
    add = oracle.make_additional_args(p["m1"], p["m2"], p["chi1"], p["chi2"])
    b = P.EOBBatch(P.make_additional_args(p["m1"], p["m2"], p["chi1"], p["chi2"]), max_step=4096, device=dev)
    b.initial_conditions(np.full(n, r0))
    b.integrate(rtol=1e-10, atol=1e-10, two_pass=True)
    torch.cuda.synchronize()
    return p, add, b


def test_modes_on_knots_match_oracle(oracle, dev):
    """The reference's five modes on the knots of integrated trajectories: each
    series h_lm e^{i m phi} against the oracle's get_hlms on the same states.
    The knot inputs (H and omega from the Hamiltonian's derivatives) carry the
    right-hand side's tolerance, which is 5e-10 relative on plunge states (the
    device contracts into fused multiply-adds); through v = omega^(1/3) and the
    modes' powers of v that reaches h_lm on the last knots at 4e-11 of max |h|
    for (2,2) and 3.6e-10 for (5,5) (measured: its amplitude grows as
    vPhi^5 rho^5), so the bound is 1e-9.  With identical inputs the waveform
    agrees with the oracle to 1e-11 for every mode (the drop-in test above).
    The (2,2) series is also checked against the (2,2) pipeline's own knots,
    where both use the device's inputs."""
    p, add, b = _batch_with_trajectories(oracle, dev)
    modes = P.REFERENCE_MODES
    kn = b.hlm_modes_knots(modes).cpu().numpy()
    ns = b.num_steps.cpu().numpy()
    dense = b.dense.cpu().numpy()
    nm = len(modes)
    traj = np.transpose(dense[:, :, 1:5], (0, 2, 1)).copy()   # [n][4][len]
    co = oracle.fac_coeffs(b.add.cpu().numpy())
    for j, (l, m) in enumerate(modes):
        h_r = oracle.get_hlm(l, m, traj, ns, b.add.cpu().numpy(), co)
        for i in range(b.n):
            k = ns[i]
            rot = h_r[i, :k] * np.exp(1j * m * dense[i, :k, 2])
            got = kn[i, :k, j] + 1j * kn[i, :k, nm + j]
            scale = np.abs(rot).max()
            assert np.max(np.abs(got - rot)) < 1e-9 * scale, ((l, m), i)
    np.testing.assert_array_equal(kn[0, :ns[0], 2 * nm], dense[0, :ns[0], 2])
    # the (2,2) series equals the (2,2) pipeline's knots
    k22 = b.hlm22_knots().cpu().numpy()
    for i in range(b.n):
        k = ns[i]
        np.testing.assert_allclose(kn[i, :k, 0], k22[i, :k, 1], rtol=0, atol=1e-13 * np.abs(k22[i, :k, 1:]).max())
        np.testing.assert_allclose(kn[i, :k, nm], k22[i, :k, 2], rtol=0, atol=1e-13 * np.abs(k22[i, :k, 1:]).max())


def test_mode_splines_and_sum_match_scipy(oracle, dev):
    """Splines of the 11 series and the mode sum on the uniform grid: scipy's
    not-a-knot CubicSpline of each series, h = sum_j S_j e^{-i m_j phi}; with
    the reference's five modes (a build of the sum for five modes) and with
    three (the build for any count)."""
    p, add, b = _batch_with_trajectories(oracle, dev, n=5, seed=8, r0=12.0)
    Msun, dt, N = 30.0, 1.0 / 4096, 1 << 15
    dT = dt / (Msun * P.MTSUN_SI)
    for modes in (P.REFERENCE_MODES, ((2, 2), (3, 3), (2, 1))):
        ms = [m for _, m in modes]
        nm = len(modes)
        b.hlm_modes_knots(modes)
        b.build_mode_splines()
        hp, hc = b.resample_modes(np.full(b.n, Msun), dt, N, amp=1.5)
        hp, hc = hp.cpu().numpy(), hc.cpu().numpy()
        kn = b.mode_knots.cpu().numpy()
        dense = b.dense.cpu().numpy()
        ns = b.num_steps.cpu().numpy()
        for i in range(b.n):
            k = ns[i]
            t = dense[i, :k, 0]
            nsamp = min(N, int((t[-1] - t[0]) / dT) + 1)
            tj = t[0] + np.arange(nsamp) * dT
            S = [CubicSpline(t, kn[i, :k, q], bc_type="not-a-knot")(tj) for q in range(2 * nm + 1)]
            h = sum((S[j] + 1j * S[nm + j]) * np.exp(-1j * ms[j] * S[2 * nm]) for j in range(nm)) * 1.5
            scale = np.abs(h).max()
            assert np.max(np.abs(hp[i, :nsamp] - h.real)) < 1e-9 * scale, (nm, i)
            assert np.max(np.abs(hc[i, :nsamp] - h.imag)) < 1e-9 * scale, (nm, i)
            assert not hp[i, nsamp:].any() and not hc[i, nsamp:].any()
    # the (2,2)-only sum is the (2,2) pipeline's h+ and hx
    b.hlm_modes_knots(((2, 2),))
    b.build_mode_splines()
    hp2, hc2 = b.resample_modes(np.full(b.n, Msun), dt, N, amp=1.5)
    b.hlm22_knots()
    b.build_splines()
    from eobgpu_amd import detector
    hp22, hc22 = detector.resample_hpc(b, np.full(b.n, Msun), dt, N, amp=1.5)
    scale = np.abs(hp22.cpu().numpy()).max()
    assert np.max(np.abs(hp2.cpu().numpy() - hp22.cpu().numpy())) < 1e-10 * scale
    assert np.max(np.abs(hc2.cpu().numpy() - hc22.cpu().numpy())) < 1e-10 * scale


@pytest.mark.parametrize("legacy", [True, False])
def test_td_interp2_dropin(oracle, dev, legacy):
    """TDInterp_wrap2 on the reference's layout [bin][2M+1][old_length] with
    uneven binaries: the reference's stride (old_lengths[b]) and the padded one,
    against the restatement; with the padded stride, also against scipy."""
    rng = np.random.default_rng(31)
    nb, M, old_length = 4, 3, 400
    ms = np.array([2, 1, 3], dtype=np.int32)
    ls = np.array([2, 2, 3], dtype=np.int32)
    old_lengths = np.array([400, 317, 250, 399], dtype=np.int32)
    y = np.zeros((nb, 2 * M + 1, old_length)); c1 = np.zeros_like(y); c2 = np.zeros_like(y); c3 = np.zeros_like(y)
    ts = np.zeros((nb, old_length))
    splines = []
    for b in range(nb):
        k = old_lengths[b]
        t = np.cumsum(rng.uniform(0.5, 1.5, k))
        ts[b, :k] = t
        ser = []
        for s in range(2 * M + 1):
            vals = np.sin(t / (5 + s)) + 0.1 * s if s < 2 * M else 0.05 * t
            cs = CubicSpline(t, vals, bc_type="not-a-knot")
            ser.append(cs)
            y[b, s, :k - 1] = cs.c[3]; c1[b, s, :k - 1] = cs.c[2]; c2[b, s, :k - 1] = cs.c[1]; c3[b, s, :k - 1] = cs.c[0]
            y[b, s, k - 1] = vals[-1]
        splines.append(ser)
    data_length = 600
    dataTime = np.linspace(0.0, 240.0, data_length)
    lengths = np.array([np.searchsorted(dataTime, ts[b, old_lengths[b] - 1], "right") for b in range(nb)], dtype=np.int32)
    lengths = np.minimum(lengths, data_length).astype(np.int32)
    max_length = int(lengths.max())
    inds = np.zeros((nb, max_length), dtype=np.int32)
    for b in range(nb):
        k = old_lengths[b]
        ii = np.searchsorted(ts[b, :k], dataTime[:lengths[b]], "right") - 1
        inds[b, :lengths[b]] = np.clip(ii, 0, k - 2)
    ref = oracle.td_interp2(dataTime, ts.reshape(-1), y.reshape(-1), c1.reshape(-1), c2.reshape(-1), c3.reshape(-1), old_length,
                            old_lengths, data_length, ms, inds.reshape(-1), lengths, max_length, legacy_stride=legacy)
    out = torch.zeros((nb, data_length), dtype=torch.complex128, device=dev)
    pytdinterp.TDInterp_wrap2(out, _T(dataTime, dev), _T(ts, dev), _T(y, dev), _T(c1, dev), _T(c2, dev), _T(c3, dev), old_length,
                              _T(old_lengths, dev), data_length, nb, M, _T(ls, dev), _T(ms, dev), _T(inds, dev), _T(lengths, dev),
                              max_length, legacy_stride=legacy)
    got = out.cpu().numpy()
    scale = np.abs(ref).max()
    assert np.max(np.abs(got - ref)) < 1e-12 * scale
    if not legacy:
        for b in range(nb):
            n = lengths[b]
            S = [splines[b][s](dataTime[:n]) for s in range(2 * M + 1)]
            h = sum((S[j] + 1j * S[M + j]) * np.exp(-1j * ms[j] * S[2 * M]) for j in range(M))
            assert np.max(np.abs(got[b, :n] - h)) < 1e-10 * np.abs(h).max()
    else:
        # the reference's stride differs from the layout for the binaries shorter than old_length
        full = old_lengths == old_length
        padded = oracle.td_interp2(dataTime, ts.reshape(-1), y.reshape(-1), c1.reshape(-1), c2.reshape(-1), c3.reshape(-1), old_length,
                                   old_lengths, data_length, ms, inds.reshape(-1), lengths, max_length, legacy_stride=False)
        for b in range(nb):
            same = np.allclose(ref[b], padded[b], rtol=0, atol=1e-12 * scale)
            assert same == bool(full[b]), b
