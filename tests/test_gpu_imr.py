"""Merger-ringdown attachment on the device against the oracle's restatement
(oracle/oracle_imr.py): the fit drop-ins of eob/eobstuff.py and
eob/seobnrv4phm.py, the Kerr QNMs, the NQC and ringdown parameters of a batch
and its IMR template (compute_full_waveform, eob/seobnrv4phm.py:2308-2556).

The oracle is fed the device's own trajectories and mode knots (the adaptive
integration is chaotic at rounding level, and the knot series are pinned by
their own tests), so what is compared here is the attachment itself.  Its
windows are the reference's (the NQC fit from 7 samples before the NQC time to
the series' end), the device cuts them 31 rows past the evaluation interval
(influence 0.268^31); the NQC coefficients come out of 3x3 and 2x2 solves
whose matrices are close to singular (q1, q2 and q3 differ by powers of r), so
the coefficients agree to 1e-7 relative while the template agrees to 1e-9 of
its peak."""
import numpy as np
import pytest
import torch

from eobgpu_amd import configs
from eobgpu_amd import eobstuff as ES
from eobgpu_amd import pipeline as P

pytestmark = pytest.mark.gpu

MODES = P.REFERENCE_MODES


@pytest.fixture(scope="module")
def OI():
    from oracle import oracle_imr
    return oracle_imr


def test_fit_dropins_match_oracle(OI, dev):
    rng = np.random.default_rng(4)
    n = 48
    m1 = rng.uniform(5, 60, n)
    m2 = rng.uniform(5, 60, n)
    c1 = rng.uniform(-0.9, 0.9, n)
    c2 = rng.uniform(-0.9, 0.9, n)
    M = m1 + m2
    m1s, m2s = m1 / M, m2 / M
    eta = m1s * m2s
    chiS, chiA = 0.5 * (c1 + c2), 0.5 * (c1 - c2)
    chi = np.array([OI.fit_chi(*x) for x in zip(m1s, m2s, c1, c2)])
    got = ES.EOBGetNRSpinPeakDeltaTv4(2, 2, m1s, m2s, c1, c2)
    np.testing.assert_allclose(got, [OI.nr_delta_t22(*x) for x in zip(m1s, m2s, c1, c2)], rtol=1e-13)
    for l, m in MODES:
        for fn, ofn in ((ES.EOBGetNRSpinPeakAmplitudeV4, OI.nr_amp), (ES.EOBGetNRSpinPeakADotV4, OI.nr_adot),
                        (ES.EOBGetNRSpinPeakADDotV4, OI.nr_addot)):
            got = fn(l, m, m1s, m2s, chiS, chiA)
            want = np.array([ofn(l, m, *x) for x in zip(m1s, m2s, c1, c2)])
            np.testing.assert_allclose(got, want, rtol=1e-11, atol=1e-14 * np.abs(want).max(), err_msg=str((fn.__name__, l, m)))
        for fn, ofn in ((ES.EOBGetNRSpinPeakOmegaV4, OI.nr_omega), (ES.EOBGetNRSpinPeakOmegaDotV4, OI.nr_omegadot)):
            got = fn(l, m, eta, chi)
            np.testing.assert_allclose(got, [ofn(l, m, e, c) for e, c in zip(eta, chi)], rtol=1e-13, err_msg=str((fn.__name__, l, m)))
    ells = np.array([l for l, _ in MODES])
    mms = np.array([m for _, m in MODES])
    for fn, name in ((ES.EOBCalculateRDAmplitudeCoefficient1, "c1f"), (ES.EOBCalculateRDAmplitudeCoefficient2, "c2f"),
                     (ES.EOBCalculateRDPhaseCoefficient1, "d1f"), (ES.EOBCalculateRDPhaseCoefficient2, "d2f")):
        got = fn(ells, mms, eta, chi)
        want = np.array([[OI.rd_fit(name, l, m, e, c) for l, m in MODES] for e, c in zip(eta, chi)])
        np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-15, err_msg=name)
    np.testing.assert_allclose(ES.compute_final_mass_SEOBNRv4(m1s, m2s, c1, c2), OI.final_mass(m1s, m2s, c1, c2), rtol=1e-14)
    np.testing.assert_allclose(ES.bbh_final_spin_non_precessing_HBR2016(m1s, m2s, c1, c2), OI.final_spin(m1s, m2s, c1, c2), rtol=1e-13)


@pytest.mark.parametrize("l,m", MODES)
def test_kerr_qnm_matches_oracle(OI, dev, l, m):
    """Leaver's method on the device (secant refinement along the spin path)
    against the oracle's Newton route, and compute_QNM's convention
    (omega_r - i / tau in units of the total mass)."""
    af = np.array([-0.7, -0.2, 0.0, 0.35, 0.69, 0.9])
    mf = np.full(af.size, 0.95)
    got = ES.compute_QNM(l, m, 0, af, mf)
    want = np.array([OI.kerr_qnm(l, m, a) for a in af]) / mf
    np.testing.assert_allclose(got, want, rtol=2e-10)
    assert np.all(got.imag < 0)


def _batch(dev, n=3, seed=12, r0=11.0):
    p = configs.small_batch(n=n, seed=seed)
    b = P.EOBBatch(P.make_additional_args(p["m1"], p["m2"], p["chi1"], p["chi2"]), max_step=4096, device=dev)
    b.initial_conditions(np.full(n, r0))
    b.integrate(rtol=1e-10, atol=1e-10, two_pass=True)
    b.hlm_modes_knots(MODES)
    b.build_mode_splines()
    torch.cuda.synchronize()
    return p, b


@pytest.fixture(scope="module")
def imr_case(OI, dev):
    p, b = _batch(dev)
    M = p["m1"] + p["m2"]
    dTM = 0.5
    dt = dTM * float(M.min()) * P.MTSUN_SI       # grid spacing 0.5 M for the lightest binary
    N = 1 << 14
    hp, hc = b.imr_template(M, dt, N)
    prm = {k: (v.cpu().numpy() if isinstance(v, torch.Tensor) else v) for k, v in b.imr.items()}
    ns = b.num_steps.cpu().numpy()
    dense = b.dense.cpu().numpy()
    kn = b.mode_knots.cpu().numpy()
    nm = len(MODES)
    ref = []
    for i in range(b.n):
        L = int(ns[i])
        series = np.array([kn[i, :L, j] for j in range(2 * nm)])
        dT = float(prm["dT_M"][i])
        tot, d = OI.imr_template(dense[i, :L, 0], dense[i, :L, 1], dense[i, :L, 2], dense[i, :L, 3], series, MODES, p["m1"][i], p["m2"][i],
                                 p["chi1"][i], p["chi2"][i], dT, N, detail=True)
        ref.append((tot, d))
    return dict(p=p, b=b, hp=hp.cpu().numpy(), hc=hc.cpu().numpy(), prm=prm, ref=ref, nsamp=b.nsamp.cpu().numpy(), M=M, dt=dt, N=N)


def test_imr_parameters_match_oracle(imr_case):
    c = imr_case
    prm = c["prm"]
    for i, (tot, d) in enumerate(c["ref"]):
        npts, idx, nadd = prm["bini"][i]
        assert (npts, idx, nadd) == (d["npts"], d["idx"], d["num_add"]), i
        bf = prm["binf"][i]
        assert bf[0] == d["t_peak"]
        np.testing.assert_allclose(bf[3:5], [d["mf"], d["af"]], rtol=1e-14)
        np.testing.assert_allclose(prm["nqc"][i], d["nqc"], rtol=1e-7, err_msg=str(i))
        for q in range(len(MODES)):
            r = d["rd"][q]
            got = prm["rd"][i, q]
            np.testing.assert_allclose(got[[0, 1, 3]], [r["amp"], r["damp"], r["omega"]], rtol=1e-9, atol=1e-12,
                                       err_msg=str((i, MODES[q])))
            # the phase at the match is unwrapped over the device's window, the
            # oracle's over the whole series: equal modulo 2 pi
            dphi = (got[2] - r["phi"]) / (2 * np.pi)
            assert abs(dphi - round(dphi)) < 1e-9, (i, MODES[q], dphi)
            np.testing.assert_allclose(got[4:11], [r["c1c"], r["c2c"], r["c1f"], r["c2f"], r["d1c"], r["d1f"], r["d2f"]], rtol=1e-8,
                                       err_msg=str((i, MODES[q])))
            np.testing.assert_allclose(got[11] + 1j * got[12], r["qnm"], rtol=1e-10)   # omega_r + i / tau


def test_imr_template_matches_oracle(imr_case):
    c = imr_case
    for i, (tot, d) in enumerate(c["ref"]):
        h = c["hp"][i] + 1j * c["hc"][i]
        scale = np.abs(tot).max()
        assert np.max(np.abs(h - tot)) < 1e-9 * scale, i
        assert c["nsamp"][i] == min(c["N"], d["idx"] + d["num_add"])
        assert np.all(h[d["idx"] + d["num_add"]:] == 0.0)


def test_batch_ringdown_length(imr_case):
    """batch_ringdown pads every binary's ringdown to the batch's longest (the reference's rule)."""
    c = imr_case
    b = c["b"]
    b.imr_template(c["M"], c["dt"], c["N"], batch_ringdown=True)
    bini = b.imr["bini"].cpu().numpy()
    nmax = int(c["prm"]["bini"][:, 2].max())
    assert np.all(bini[:, 2] == nmax)
    np.testing.assert_array_equal(b.nsamp.cpu().numpy(), np.minimum(c["N"], bini[:, 1] + nmax))


def test_nqc_factor_and_ringdown_dropins(OI, imr_case):
    """EOBNonQCCorrection and compute_MR_mode_free as drop-ins, on the oracle's
    attachment values."""
    c = imr_case
    tot, d = c["ref"][0]
    p = c["p"]
    co = d["nqc"][0]
    coeffs = dict(a1=co[0], a2=co[1], a3=co[2], a3S=0.0, a4=0.0, a5=0.0, b1=co[3], b2=co[4], b3=0.0, b4=0.0)
    got = ES.EOBNonQCCorrection(d["r"], None, d["pr"], None, d["om"], coeffs)
    np.testing.assert_allclose(got, OI.nqc_factor(co, d["r"], d["pr"], d["om"]), rtol=1e-13)
    M = p["m1"][0] + p["m2"][0]
    m1s, m2s = p["m1"][0] / M, p["m2"][0] / M
    nadd = d["num_add"]
    dt = float(c["prm"]["binf"][0][2])
    t = (d["t_match"] + np.arange(nadd) * dt)[None, :]
    ells = np.array([l for l, _ in MODES])
    mms = np.array([m for _, m in MODES])
    ap = dict(amp=np.array([[r["amp"] for r in d["rd"]]]), damp=np.array([[r["damp"] for r in d["rd"]]]),
              omega=np.array([[r["omega"] for r in d["rd"]]]), final_mass=np.array([d["mf"]]), final_spin=np.array([d["af"]]))
    phi = np.array([[r["phi"] for r in d["rd"]]])
    hring, _ = ES.compute_MR_mode_free(t, np.array([m1s]), np.array([m2s]), np.array([p["chi1"][0]]), np.array([p["chi2"][0]]), ap, ells,
                                       mms, t_match=np.array([d["t_match"]]), phi_match=phi)
    np.testing.assert_allclose(hring[0], d["hring"], rtol=1e-9, atol=1e-12 * np.abs(d["hring"]).max())


def test_equal_mass_equal_spin_is_refused(dev):
    """The reference raises for equal masses with equal (or no) spins
    (eob/seobnrv4phm.py:2391-2395): the odd-m modes vanish there."""
    m = np.array([20.0, 30.0]); c = np.array([0.3, 0.1])
    b = P.EOBBatch(P.make_additional_args(m, m.copy(), c, c.copy()), max_step=1024, device=dev)
    b.initial_conditions(np.full(2, 10.0))
    b.integrate(rtol=1e-10, atol=1e-10, two_pass=True)
    b.hlm_modes_knots(MODES)
    b.build_mode_splines()
    with pytest.raises(NotImplementedError):
        b.imr_params(2 * m, 1.0 / 16384)


@pytest.mark.parametrize("modes", [((2, 2),), ((2, 2), (3, 3)), ((2, 2), (2, 1), (4, 4)), ((2, 2), (3, 3), (2, 1), (5, 5))])
def test_imr_template_mode_subsets(OI, dev, modes):
    """The attachment with one to four of the modes (the (2,2) mode alone is the
    reference's default SEOBNRv4 configuration), against the oracle on the
    device's own knots: the template kernel is built once per mode count."""
    p = configs.small_batch(n=2, seed=31)
    b = P.EOBBatch(P.make_additional_args(p["m1"], p["m2"], p["chi1"], p["chi2"]), max_step=4096, device=dev)
    b.initial_conditions(np.full(2, 10.5))
    b.integrate(rtol=1e-10, atol=1e-10, two_pass=True)
    b.hlm_modes_knots(modes)
    b.build_mode_splines()
    M = p["m1"] + p["m2"]
    dt = 1.0 / 16384
    N = 1 << 14
    hp, hc = b.imr_template(M, dt, N)
    h = (hp + 1j * hc).cpu().numpy()
    ns = b.num_steps.cpu().numpy()
    dense = b.dense.cpu().numpy()
    kn = b.mode_knots.cpu().numpy()
    dTM = b.imr["dT_M"].cpu().numpy()
    for i in range(2):
        L = int(ns[i])
        series = kn[i, :L, :2 * len(modes)].T.copy()
        tot = OI.imr_template(dense[i, :L, 0], dense[i, :L, 1], dense[i, :L, 2], dense[i, :L, 3], series, modes,
                              p["m1"][i], p["m2"][i], p["chi1"][i], p["chi2"][i], float(dTM[i]), N)
        assert np.max(np.abs(h[i] - tot)) < 1e-9 * np.abs(tot).max(), i


class _GpuArray:
    """A GPU array of another library: only __cuda_array_interface__ and shape."""

    def __init__(self, t):
        self._t = t
        self.__cuda_array_interface__ = t.__cuda_array_interface__
        self.shape = tuple(t.shape)


def test_dropins_take_other_gpu_arrays(dev):
    rng = np.random.default_rng(7)
    m1, m2 = rng.uniform(0.3, 0.7, 5), rng.uniform(0.3, 0.7, 5)
    c1, c2 = rng.uniform(-0.5, 0.5, 5), rng.uniform(-0.5, 0.5, 5)
    want = ES.compute_final_mass_SEOBNRv4(m1, m2, c1, c2)
    got = ES.compute_final_mass_SEOBNRv4(*[_GpuArray(torch.tensor(x, device=dev)) for x in (m1, m2, c1, c2)])
    assert isinstance(got, torch.Tensor)
    np.testing.assert_array_equal(got.cpu().numpy(), want)
    eta = m1 * m2 / (m1 + m2) ** 2
    got = ES.EOBCalculateRDAmplitudeCoefficient1(np.array([2, 3]), np.array([2, 3]), _GpuArray(torch.tensor(eta, device=dev)),
                                                 _GpuArray(torch.tensor(c1, device=dev)))
    np.testing.assert_array_equal(got.cpu().numpy(), ES.EOBCalculateRDAmplitudeCoefficient1(np.array([2, 3]), np.array([2, 3]), eta, c1))
