"""Pins of the oracle's merger-ringdown attachment (oracle/oracle_imr.py) by
what the mathematics fixes, on the CPU:

* Kerr QNMs (Leaver's method): the Schwarzschild fundamental frequencies to 10
  digits, Kerr values of the standard tables at 4 digits, the (a, m) -> (-a, -m)
  symmetry of the Teukolsky equation and the first-order shift of the angular
  separation constant, dA/dc = -2 s^2 m / (l (l + 1)) at c = 0;
* the Kerr ISCO (6 M, 1 M, 9 M at a = 0, 1, -1) and the remnant mass built on it;
* the HBR2016 remnant spin: the equal-mass non-spinning value 0.6865 and the
  test-particle limit a_f -> chi_1;
* the NR fits' test-particle / equal-mass combination (exactly the equal-mass
  fit at eta = 1/4);
* the NQC solve by its defining property: after the correction the mode's
  amplitude, its first two derivatives, its frequency and the frequency's
  derivative at the NQC time equal the NR values;
* the ringdown attachment by continuity: the ringdown's first sample is the
  corrected inspiral mode at the match time.

The ringdown fit tables, the NR fit tables and the pesummary / lalsimulation
numbers the reference takes are not pinned beyond these (no fixture of them
exists in the reference tree)."""
import math

import numpy as np
import pytest

from oracle import oracle_imr as I

MODES = ((2, 2), (2, 1), (3, 3), (4, 4), (5, 5))


@pytest.mark.parametrize("l,w", [(2, 0.3736716844 - 0.0889623157j), (3, 0.5994432884 - 0.0927030479j),
                                 (4, 0.8091783775 - 0.0941639610j)])
def test_schwarzschild_qnms(l, w):
    """Leaver (1985) / Berti, Cardoso & Starinets (2009): M omega_l0 of Schwarzschild."""
    got = I.kerr_qnm(l, l, 0.0)
    assert abs(got - w) < 2e-10


@pytest.mark.parametrize("a,w", [(0.5, 0.4641 - 0.0856j), (0.7, 0.5326 - 0.0808j), (0.9, 0.6716 - 0.0649j)])
def test_kerr_22_qnm_table(a, w):
    """The (2,2,0) Kerr frequencies of the standard tables (4 digits)."""
    got = I.kerr_qnm(2, 2, a)
    assert abs(got.real - w.real) < 6e-5 and abs(got.imag - w.imag) < 6e-5


def test_qnm_spin_mirror_symmetry():
    """The Teukolsky equation is invariant under (a, m) -> (-a, -m)."""
    for l, m in ((2, 2), (2, 1), (3, 3)):
        np.testing.assert_allclose(I.kerr_qnm(l, -m, 0.6), I.kerr_qnm(l, m, -0.6), rtol=1e-10)


@pytest.mark.parametrize("l,m", [(2, 2), (2, 1), (2, -1), (2, -2), (3, 3), (4, 4), (5, 5), (3, -3)])
def test_angular_constant_small_c(l, m):
    """A(c) = (l - s)(l + s + 1) - 2 c s^2 m / (l (l + 1)) + O(c^2) (first-order
    perturbation: <cos theta> of the spin-weighted harmonic is -m s / (l (l + 1))),
    for the lowest l of each m -- the modes the ringdown uses."""
    s, h = -2, 1e-5
    A0 = complex((l - s) * (l + s + 1))
    Ap = I._newton(lambda y: I._angular_cf(y, h, s, m), A0)
    Am = I._newton(lambda y: I._angular_cf(y, -h, s, m), A0)
    assert abs((Ap + Am).real / 2 - A0.real) < 1e-8
    assert abs((Ap - Am).real / (2 * h) - (-2 * s * s * m / (l * (l + 1)))) < 1e-7


def test_isco_and_remnant_mass():
    assert I.kerr_isco_radius(0.0) == pytest.approx(6.0, abs=1e-14)
    assert I.kerr_isco_radius(-1.0) == pytest.approx(9.0, abs=1e-12)
    assert I.kerr_isco_radius(1.0 - 1e-12) == pytest.approx(1.0, abs=1e-3)
    # equal masses without spin: a_tl = 0, r_isco = 6, E_isco = sqrt(8/9)
    e = math.sqrt(8.0 / 9.0)
    mf = 1.0 - ((1.0 - e) * 0.25 + 16.0 * 0.0625 * (0.00258 - 0.0773 / (0.0 - 1.6939) - 0.25 * (1.0 - e)))
    assert I.final_mass(0.5, 0.5, 0.0, 0.0) == pytest.approx(mf, rel=1e-14)
    # test-particle limit: the remnant is the big hole
    assert I.final_mass(1.0, 1e-9, 0.3, 0.0) == pytest.approx(1.0, abs=1e-8)


def test_hbr2016_remnant_spin():
    """0.6865 for equal non-spinning masses (HBR2016's calibration value 0.68646),
    chi_1 in the test-particle limit, symmetric in the bodies."""
    assert I.final_spin(0.5, 0.5, 0.0, 0.0) == pytest.approx(0.68646, abs=2e-4)
    assert I.final_spin(1.0, 1e-10, 0.5, -0.3) == pytest.approx(0.5, abs=1e-8)
    assert I.final_spin(0.7, 0.3, 0.2, -0.6) == pytest.approx(I.final_spin(0.3, 0.7, -0.6, 0.2), rel=1e-14)


def test_peak_fit_equal_mass_limit():
    """The test-particle / equal-mass combination returns the equal-mass fit at eta = 1/4
    (to 1e-10: the combination's coefficients are printed to 12-16 digits)."""
    for chi in (-0.5, 0.0, 0.7):
        fEQ = 1.577457498227 - 0.0076949474494639085 * chi + 0.02188705616693344 * chi ** 2 + 0.023268366492696667 * chi ** 3
        assert I.nr_amp(2, 2, 0.5, 0.5, chi, chi) == pytest.approx(0.25 * fEQ, rel=1e-10)
        A1 = 0.01574321112717377 + 0.02244178140869133 * chi
        fTPL = -0.011209791668428353 + (0.0040867958978563915 + 0.0006333925136134493 * chi) * math.log(
            68.47466578100956 - 58.301487557007206 * chi)
        assert I._combine(0.25, A1, 0.01128156666995859 + 0.0002869276768158971 * chi, fTPL) == pytest.approx(
            0.01128156666995859 + 0.0002869276768158971 * chi, rel=1e-10)


def _oracle_binary(oracle, m1, m2, chi1, chi2, r0=11.0):
    add = oracle.make_additional_args(np.array([m1]), np.array([m2]), np.array([chi1]), np.array([chi2]))
    co = oracle.fac_coeffs(add)
    pphi0, pr0, om0, add = oracle.initial_conditions(np.array([r0]), add, co)
    cond = np.array([[r0], [0.0], [pr0[0]], [pphi0[0]]])
    t, traj, ns, si = oracle.run_trajectory(cond, add, co, max_step=4096, rtol=1e-10, atol=1e-10)
    L = int(ns[0])
    series = []
    hs = [oracle.get_hlm(l, m, traj, ns, add, co)[0, :L] * np.exp(1j * m * traj[0, 1, :L]) for l, m in MODES]
    series = np.array([h.real for h in hs] + [h.imag for h in hs])
    return t[0, :L], traj[0, :, :L], series, add


@pytest.fixture(scope="module")
def attached(oracle):
    m1, m2, c1, c2 = 30.0, 18.0, 0.4, -0.2
    t, traj, series, add = _oracle_binary(oracle, m1, m2, c1, c2)
    dT = 0.5
    N = int((t[-1] - t[0]) / dT) + 2000
    tot, d = I.imr_template(t, traj[0], traj[1], traj[2], series, MODES, m1, m2, c1, c2, dT, N, detail=True)
    return dict(t=t, traj=traj, series=series, m=(m1, m2, c1, c2), dT=dT, tot=tot, d=d)


def test_nqc_corrected_modes_meet_the_nr_values(attached):
    """The defining property of the NQC solve, per mode: the corrected mode,
    written as (|h| mag) e^{i (phase + psi)} with the NQC factor mag e^{i psi}, has
    over the fit window a spline of its amplitude with the NR amplitude and
    derivatives at the NQC time, and a spline of its phase with the NR frequency
    and its derivative (the frequency falls with time: the phase decreases).  The
    smooth form is used because mag turns negative on the last samples of the
    plunge, where the corrected mode's angle jumps by pi."""
    A, d = attached, attached["d"]
    m1, m2, c1, c2 = A["m"]
    M = m1 + m2
    m1s, m2s = m1 / M, m2 / M
    t, traj = A["t"], A["traj"]
    npts = d["npts"]
    tg = np.minimum(np.arange(npts) * A["dT"] + t[0], t.max())
    eta = I._eta(m1s, m2s)
    chi = I.fit_chi(m1s, m2s, c1, c2)
    for q, (l, m) in enumerate(MODES):
        T = d["t_peak"] - I.nr_delta_t22(m1s, m2s, c1, c2) - (10.0 if (l, m) == (5, 5) else 0.0)
        idx = int(np.argmin(np.abs(tg - T)))
        lo = idx - 7
        sl = slice(lo, npts)
        h = d["modes_raw"][q][sl]
        c = d["nqc"][q]
        r, pr, om = d["r"][sl], d["pr"][sl], d["om"][sl]
        rO = r * om
        mag = 1.0 + (pr * pr / (rO * rO)) * (c[0] + c[1] / r + c[2] / (r * np.sqrt(r)))
        psi = c[3] * pr / rO + pr ** 3 / rO * c[4]
        np.testing.assert_allclose(h * mag * np.exp(1j * psi), d["modes_nqc"][q][sl], rtol=1e-13, atol=1e-15)
        sa = I.spline1(tg[sl], np.abs(h) * mag)
        sp = I.spline1(tg[sl], np.unwrap(np.angle(h)) + psi)
        nr = [abs(I.nr_amp(l, m, m1s, m2s, c1, c2)), I.nr_adot(l, m, m1s, m2s, c1, c2), I.nr_addot(l, m, m1s, m2s, c1, c2)]
        got = [I.seval(sa, T, k)[0] for k in range(3)]
        np.testing.assert_allclose(got[0], nr[0], rtol=1e-9)
        assert abs(got[1] - nr[1]) < 1e-9 * nr[0]
        assert abs(got[2] - nr[2]) < 1e-9 * nr[0]
        w, wd = I.seval(sp, T, 1)[0], I.seval(sp, T, 2)[0]
        assert w < 0.0
        np.testing.assert_allclose(-w, I.nr_omega(l, m, eta, chi), rtol=1e-9)
        np.testing.assert_allclose(-wd, I.nr_omegadot(l, m, eta, chi), rtol=1e-7, atol=1e-12)


def test_ringdown_starts_on_the_corrected_mode(attached):
    """The ringdown's first sample is the corrected mode at the match sample (the
    ansatz is constrained to its amplitude and phase there); (5,5) matches 10 M
    earlier (:2482) and is left out."""
    d = attached["d"]
    idx = d["idx"]
    for q, (l, m) in enumerate(MODES):
        if (l, m) == (5, 5):
            continue
        h0 = d["hring"][q][0]
        hi = d["modes_nqc"][q][idx]
        assert abs(h0 - hi) < 1e-10 * abs(hi), (l, m)
    # the ringdown decays: the (2,2) QNM damping time of the remnant
    w = d["rd"][0]["qnm"]
    assert w.imag > 0.0
    h22 = np.abs(d["hring"][0])
    assert h22[-1] < 1e-6 * h22.max()


def test_template_is_inspiral_then_ringdown(attached):
    d, tot = attached["d"], attached["tot"]
    idx, nadd = d["idx"], d["num_add"]
    np.testing.assert_allclose(tot[:idx + 1], d["modes_nqc"][:, :idx + 1].sum(axis=0), rtol=0, atol=1e-15)
    np.testing.assert_allclose(tot[idx + 1:idx + nadd], d["hring"].sum(axis=0)[1:nadd], rtol=0, atol=1e-15)
    assert np.all(tot[idx + nadd:] == 0.0)
