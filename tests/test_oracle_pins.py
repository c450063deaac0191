"""CPU tests that pin the oracle (oracle/oracle_eob.cpp) to things other than itself:

  - the reference's own src/DOPR853.cu, compiled unmodified into oracle/_ref
    (error norm and controller, bit for bit);
  - the published Hairer-Wanner DOP853 tableau (as shipped in scipy.integrate);
  - the Kerr test-mass limit of the Hamiltonian;
  - published post-Newtonian EOB potentials: A(u) through 5.5PN
    (Bini-Damour 2013; Damour-Jaranowski-Schaefer 2015: 4PN a5, the 5PN log and
    the 5.5PN u^(13/2) terms), Dbar(u) through 4PN, Q through 4PN (its u^2 pr^4,
    u^3 pr^4 and u^2 pr^6 terms), the NLO spin-orbit gyro-gravitomagnetic
    factors (Damour-Jaranowski-Schaefer 2008);
  - finite differences of H for the AD gradient and Hessian;
  - the PN energy flux (quadrupole normalization + 1PN + tail) for the factorized flux;
  - scipy CubicSpline (not-a-knot) for the spline build, numpy for the likelihood,
    scipy.special for Gamma(z) and Y_lm.
Coefficients beyond those orders (e.g. the N3LO spin-orbit tables, the 5PN Q
terms (u^3 pr^6, the pr^8 terms, u^4 pr^4), the rho_lm tables past the orders
checked by the flux expansion) have no independent check here: parity unpinned.
"""
import ctypes
import math
import os
import re

import numpy as np
import pytest
import scipy.special as sps
from scipy.interpolate import CubicSpline

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _nonspinning_add(oracle, q, chi1=0.0, chi2=0.0):
    m1, m2 = q / (1 + q), 1 / (1 + q)
    return oracle.make_additional_args(np.array([m1]), np.array([m2]), np.array([chi1]), np.array([chi2]))


def _Heff(oracle, add, r, pr, L):
    m1, m2 = add[0, 0], add[1, 0]
    nu = m1 * m2 / (m1 + m2) ** 2
    H = oracle.evaluate_H(np.atleast_1d(r), 0 * np.atleast_1d(r), np.atleast_1d(pr) + 0 * np.atleast_1d(r), np.atleast_1d(L) + 0 * np.atleast_1d(r),
                          m1, m2, add[2, 0], add[3, 0])
    return ((nu * H) ** 2 - 1) / (2 * nu) + 1, H, nu


# --------------------------------------------------------------- DOP853
def test_tableau_matches_scipy(oracle):
    import scipy.integrate._ivp.dop853_coefficients as d
    tab = oracle.tableau()
    c, a, b, e5, bhh = tab[:12], tab[12:156].reshape(12, 12), tab[156:168], tab[168:180], tab[180:]
    np.testing.assert_array_equal(c, d.C[:12])
    np.testing.assert_array_equal(a, d.A[:12, :12])
    np.testing.assert_array_equal(b, d.B[:12])
    np.testing.assert_array_equal(e5, d.E5[:12])
    # E3 = B - bhh on stages 1, 9, 12 (scipy subtracts the same published bhh values)
    e3 = d.B[:12].copy()
    e3[0] -= bhh[0]; e3[8] -= bhh[1]; e3[11] -= bhh[2]
    np.testing.assert_array_equal(e3, d.E3[:12])


def test_dop853_order_on_linear_ode():
    """The tableau integrates y' = y to 8th order: one step of size h has error ~ h^9."""
    import scipy.integrate._ivp.dop853_coefficients as d
    errs = []
    for h in (0.4, 0.2):
        k = np.zeros(12)
        for s in range(12):
            k[s] = 1.0 + h * np.dot(d.A[s, :s], k[:s])
        y1 = 1.0 + h * np.dot(d.B[:12], k)
        errs.append(abs(y1 - math.exp(h)))
    assert 7.5 < math.log2(errs[0] / errs[1]) < 10.5


def test_error_norm_and_controller_match_reference_build(oracle):
    if oracle.ref_lib() is None:
        pytest.skip("oracle/_ref was not built (reference tree absent when building)")
    rng = np.random.default_rng(21)
    n = 500
    solOld = rng.normal(size=(4, n)) * np.array([[20.0], [300.0], [1e-3], [4.0]])
    ks = {k: rng.normal(size=(4, n)) * 1e-4 for k in ("k1", "k2", "k3", "k6", "k7", "k8", "k9", "k10")}
    h = rng.uniform(0.01, 300, n)
    e_o, s_o = oracle.error(solOld, h, ks)
    e_r, s_r = oracle.ref_error(solOld, h, ks)
    np.testing.assert_array_equal(s_o, s_r)
    np.testing.assert_allclose(e_o, e_r, rtol=4e-16, atol=0)
    err = np.concatenate([e_r[:200], [0.0, 1.0, np.nextafter(1.0, 2.0), 1e-300, 1e300], rng.uniform(0.2, 3.0, n - 205)])
    errOld = rng.uniform(1e-5, 2, n)
    prev = rng.uniform(size=n) < 0.4
    a = oracle.controller(err, errOld, prev, h)
    b = oracle.ref_controller(err, errOld, prev, h)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)


# ----------------------------------------------------------- Hamiltonian
def test_kerr_test_mass_limit(oracle):
    """nu -> 0: Heven is the equatorial Kerr geodesic Hamiltonian; Hodd is Kerr frame
    dragging plus the reference's nu-independent SO calibration term (1701/512) a L / r^4."""
    q = 1e7
    for a in (0.0, 0.5, -0.8):
        add = _nonspinning_add(oracle, q, chi1=a)
        X1 = add[0, 0]
        aa = a * X1
        for r, pr, L in ((10.0, 0.01, 3.5), (6.0, -0.05, 3.2), (30.0, 0.0, 5.6)):
            Heff, H, nu = _Heff(oracle, add, r, pr, L)
            Delta = r * r - 2 * r + aa * aa
            Lam = (r * r + aa * aa) ** 2 - aa * aa * Delta
            alpha2 = Delta * r * r / Lam
            hk = 2 * aa * r / Lam * L + math.sqrt(alpha2 * (1 + Delta / r ** 2 * pr ** 2 + r ** 2 / Lam * L ** 2))
            hk += L * aa * (1701.0 / 512.0) / r ** 4 / (r ** 3 + aa * aa * (r + 2))
            assert abs(Heff[0] - hk) < 2e-6, (a, r, Heff[0], hk)


def _A_and_Dbar(oracle, add, r):
    """A = Heff^2 and Dbar = (d^2 Heff^2/dpr^2 / 2) / A^2 at pr = pphi = 0 (nonspinning)."""
    Heff, H, nu = _Heff(oracle, add, r, 0.0, 0.0)
    arg = np.array([[r], [0.0], [0.0], [0.0]])
    hs = oracle.hessian_H(arg, add)[:, :, 0]
    # H = sqrt(1 + 2 nu (Heff - 1)) / nu (M = 1): H_prpr = Heff_prpr / (nu H) at pr = 0
    Heff_prpr = hs[2, 2] * nu * H[0]
    A = Heff[0] ** 2
    # Heff^2 = A (1 + A Dbar pr^2 + ...) at L = 0  =>  2 Heff Heff_prpr = 2 A^2 Dbar
    Dbar = Heff[0] * Heff_prpr / (A * A)
    return A, Dbar


def test_A_potential_5p5PN(oracle):
    eg, pi2 = 0.5772156649015329, math.pi ** 2
    for q in (1.0, 3.0):
        add = _nonspinning_add(oracle, q)
        nu = q / (1 + q) ** 2
        a4 = (94.0 / 3 - 41 * pi2 / 32) * nu
        a5c = (-4237.0 / 60 + 2275 * pi2 / 512 + 256 * math.log(2) / 5 + 128 * eg / 5) * nu + (-221.0 / 6 + 41 * pi2 / 32) * nu ** 2
        a5l = 64.0 * nu / 5
        a6l = -7004.0 / 105 * nu - 144.0 / 5 * nu ** 2
        a65 = 13696.0 / 525 * math.pi * nu
        us = np.array([0.006, 0.01, 0.015, 0.02])
        R = []
        for u in us:
            A, _ = _A_and_Dbar(oracle, add, 1 / u)
            lu = math.log(u)
            Apub = 1 - 2 * u + 2 * nu * u ** 3 + a4 * u ** 4 + (a5c + a5l * lu) * u ** 5 + a6l * lu * u ** 6 + a65 * u ** 6.5
            R.append((A - Apub) / u ** 6)
        R = np.array(R)
        # what is left is the constant u^6 coefficient plus O(u ln u): nearly flat in u
        assert np.ptp(R) < 0.25 * max(1.0, np.abs(R).max()), R


def test_Dbar_potential_4PN(oracle):
    eg, pi2 = 0.5772156649015329, math.pi ** 2
    for q in (1.0, 4.0):
        add = _nonspinning_add(oracle, q)
        nu = q / (1 + q) ** 2
        d4 = (-533.0 / 45 - 23761 * pi2 / 1536 + 1184 * eg / 15 - 6496 * math.log(2) / 15 + 2916 * math.log(3) / 5) * nu \
            + (123 * pi2 / 16 - 260) * nu ** 2
        d4l = 592.0 * nu / 15
        res = []
        us = np.array([0.006, 0.018])
        for u in us:
            _, D = _A_and_Dbar(oracle, add, 1 / u)
            Dpub = 1 + 6 * nu * u ** 2 + (52 * nu - 6 * nu ** 2) * u ** 3 + (d4 + d4l * math.log(u)) * u ** 4
            res.append(D - Dpub)
        # the residual is the next (u^5 ln u) order: its local power law is ~5, a wrong
        # 4PN coefficient would make it ~4
        p = math.log(abs(res[1] / res[0])) / math.log(us[1] / us[0])
        assert 4.6 < p < 5.6, (p, res)


def test_Q_potential_4PN(oracle):
    """Q at 3PN and 4PN (Damour-Jaranowski-Schaefer 2000, 2015):
    Q = 2 nu (4 - 3 nu) u^2 pr^4
        + [(-5308/15 + 496256/45 ln2 - 33048/5 ln3) nu - 83 nu^2 + 10 nu^3] u^3 pr^4
        + [(-827/3 - 2358912/25 ln2 + 1399437/50 ln3 + 390625/18 ln5) nu - 27/5 nu^2 + 6 nu^3] u^2 pr^6 + ...
    At L = 0 (nonspinning) Heff^2 / A - 1 = A Dbar pr^2 + Q is a polynomial in pr^2;
    its pr^4 and pr^6 coefficients are fitted at several u and then expanded in u."""
    l2, l3, l5 = math.log(2), math.log(3), math.log(5)
    us = np.array([0.002, 0.004, 0.007, 0.01, 0.014, 0.02, 0.03])
    prs = np.array([0.2, 0.35, 0.5, 0.65, 0.8, 1.0])
    for q in (1.0, 3.0):
        add = _nonspinning_add(oracle, q)
        nu = q / (1 + q) ** 2
        c4, c6 = [], []
        for u in us:
            A = _Heff(oracle, add, 1 / u, 0.0, 0.0)[0][0] ** 2
            f = np.array([_Heff(oracle, add, 1 / u, p, 0.0)[0][0] ** 2 / A - 1 for p in prs])
            c = np.linalg.lstsq(np.stack([prs ** 2, prs ** 4, prs ** 6, prs ** 8], 1), f, rcond=None)[0]
            c4.append(c[1])
            c6.append(c[2])
        # pr^4: u^2 (q2 + q3 u + (q4 + q4l ln u) u^2 + q92 u^2.5); pr^6: u^2 (p2 + p3 u + p72 u^1.5)
        a4 = np.linalg.lstsq(np.stack([np.ones_like(us), us, us ** 2, us ** 2 * np.log(us), us ** 2.5], 1),
                             np.array(c4) / us ** 2, rcond=None)[0]
        a6 = np.linalg.lstsq(np.stack([np.ones_like(us), us, us ** 1.5], 1), np.array(c6) / us ** 2, rcond=None)[0]
        q42 = 2 * nu * (4 - 3 * nu)
        q43 = (-5308.0 / 15 + 496256.0 / 45 * l2 - 33048.0 / 5 * l3) * nu - 83 * nu ** 2 + 10 * nu ** 3
        q62 = (-827.0 / 3 - 2358912.0 / 25 * l2 + 1399437.0 / 50 * l3 + 390625.0 / 18 * l5) * nu - 27.0 / 5 * nu ** 2 + 6 * nu ** 3
        assert a4[0] == pytest.approx(q42, rel=1e-7)
        assert a4[1] == pytest.approx(q43, rel=1e-5)
        assert a6[0] == pytest.approx(q62, rel=1e-6)


def test_spin_orbit_NLO_gyro_factors(oracle):
    """Hodd = pphi u^3 [gS (chi1 X1^2 + chi2 X2^2) + gS* nu (chi1 + chi2)] / Dn at pr = 0 with
    gS = 2 - 5 nu u / 8 + O(u^2), gS* = 3/2 - (9/8 + 3 nu / 4) u + O(u^2)."""
    q, c1, c2 = 2.5, 0.6, -0.3
    add = _nonspinning_add(oracle, q, chi1=c1, chi2=c2)
    X1, X2 = add[0, 0], add[1, 0]
    nu = X1 * X2
    wS, wT = c1 * X1 ** 2 + c2 * X2 ** 2, nu * (c1 + c2)
    L = 1.0
    res = []
    for u in (0.004, 0.008):
        r = 1 / u
        Hp, _, _ = _Heff(oracle, add, r, 0.0, L)
        Hm, _, _ = _Heff(oracle, add, r, 0.0, -L)
        Hodd = 0.5 * (Hp[0] - Hm[0])
        ap2 = (c1 * X1 + c2 * X2) ** 2
        Dn = 1 + ap2 * (u * u + 2 * u ** 3)
        G = Hodd * Dn / (L * u ** 3)
        lead = 2 * wS + 1.5 * wT
        nlo = (-5 * nu / 8) * wS + (-(9.0 / 8) - 3 * nu / 4) * wT
        res.append((G - lead - nlo * u) / u ** 2)
    assert max(abs(x) for x in res) < 50, res


def test_gradient_and_hessian_vs_finite_differences(oracle):
    rng = np.random.default_rng(3)
    n = 20
    q = rng.uniform(1, 8, n)
    add = oracle.make_additional_args(q / (1 + q), 1 / (1 + q), rng.uniform(-0.9, 0.9, n), rng.uniform(-0.9, 0.9, n))
    arg = np.array([rng.uniform(4, 40, n), rng.uniform(0, 6, n), rng.uniform(-0.05, 0.05, n), np.sqrt(rng.uniform(4, 40, n))])
    g = oracle.grad_H(arg, add)
    hs = oracle.hessian_H(arg, add)
    H = lambda a: oracle.evaluate_H(a[0], a[1], a[2], a[3], add[0], add[1], add[2], add[3])
    for j in (0, 2, 3):
        e = 1e-5 * np.maximum(np.abs(arg[j]), 1e-2)
        ap, am = arg.copy(), arg.copy()
        ap[j] += e; am[j] -= e
        # central differences of H ~ 10 carry ~1e-15 / e of rounding: compare at that level
        np.testing.assert_allclose(g[j], (H(ap) - H(am)) / (2 * e), rtol=1e-6, atol=3e-8)
        gp, gm = oracle.grad_H(ap, add), oracle.grad_H(am, add)
        for i in (0, 2, 3):
            np.testing.assert_allclose(hs[i, j], (gp[i] - gm[i]) / (2 * e), rtol=1e-5, atol=3e-8)
    assert np.allclose(hs, hs.transpose(1, 0, 2), rtol=1e-12, atol=1e-15)
    assert np.all(g[1] == 0.0)


# ----------------------------------------------------------------- flux
def test_flux_matches_pn_expansion(oracle):
    """Circular nonspinning orbits: F / ((32/5) nu^2 v^10) = 1 + (-1247/336 - 35 nu/12) v^2 + 4 pi v^3 + O(v^4)."""
    for q in (1.5, 3.0):
        add = _nonspinning_add(oracle, q)
        co = oracle.fac_coeffs(add)
        nu = add[0, 0] * add[1, 0]
        res = []
        for r in (200.0, 400.0, 800.0):
            pphi, pr, om, add2 = oracle.initial_conditions(np.array([r]), add, co)
            arg = np.array([[r], [0.0], [0.0], [pphi[0]]])
            F = oracle.rr_force(arg, add2, co)
            flux = -F[1, 0] * nu * om[0]
            v = om[0] ** (1 / 3)
            ratio = flux / (32.0 / 5 * nu ** 2 * v ** 10)
            pn = 1 + (-1247.0 / 336 - 35 * nu / 12) * v ** 2 + 4 * math.pi * v ** 3
            res.append((ratio - pn) / v ** 4)
        assert max(abs(x) for x in res) < 60, res


def test_newtonian_prefix_22_closed_form(oracle):
    from eobgpu_amd.pipeline import newton_prefixes
    m1, m2 = np.array([0.7, 0.5]), np.array([0.3, 0.5])
    pf = newton_prefixes(m1, m2)
    np.testing.assert_allclose(pf, oracle.newton_prefixes(m1, m2), rtol=1e-14, atol=1e-16)
    nu = m1 * m2
    y22 = math.sqrt(15.0 / (32 * math.pi))
    # |prefix_22 Y_2^-2(pi/2)| = 8 sqrt(pi/5) nu  (leading-order h_22 amplitude per unit v^2)
    np.testing.assert_allclose(np.abs(pf[1]) * y22, 8 * math.sqrt(math.pi / 5) * nu, rtol=1e-14)


def _published_coefficients(nu):
    """Nonspinning factorized-resummation coefficients as published:
    Damour-Iyer-Nagar 2008 (PRD 79, 064004) Eqs. 20-21 for delta_22 and the
    rho_lm 1PN terms, Pan et al. 2011 (PRD 83, 064003) Eq. A8 and Fujita-Iyer
    2010 (test-mass v^8, v^10 terms of rho_22) with the logarithms in
    eulerlog_2(v) = gamma + ln(4 v)."""
    pi = math.pi
    return {
        "delta22vh3": 7.0 / 3.0,
        "delta22vh6": 428.0 * pi / 105.0,
        "delta22vh9": -2203.0 / 81.0 + 1712.0 * pi ** 2 / 315.0,
        "delta22v5": -24.0 * nu,
        "rho22v2": -43.0 / 42.0 + 55.0 * nu / 84.0,
        "rho22v3": 0.0,
        "rho22v4": -20555.0 / 10584.0 - 33025.0 * nu / 21168.0 + 19583.0 * nu ** 2 / 42336.0,
        "rho22v6": 1556919113.0 / 122245200.0 + (41.0 * pi ** 2 / 192.0 - 48993925.0 / 9779616.0) * nu
                   - 6292061.0 / 3259872.0 * nu ** 2 + 10620745.0 / 39118464.0 * nu ** 3,
        "rho22v6l": -428.0 / 105.0,
        "rho22v8": -387216563023.0 / 160190110080.0,
        "rho22v8l": 9202.0 / 2205.0,
        "rho22v10": -16094530514677.0 / 533967033600.0,
        "rho22v10l": 439877.0 / 55566.0,
        "rho21v2": -59.0 / 56.0 + 23.0 * nu / 84.0,
        "rho33v2": -7.0 / 6.0 + 2.0 * nu / 3.0,
        "rho31v2": -13.0 / 18.0 - 2.0 * nu / 9.0,
        "rho44v2": (1614.0 - 5870.0 * nu + 2625.0 * nu ** 2) / (1320.0 * (3.0 * nu - 1.0)),
        "rho42v2": (1146.0 - 3530.0 * nu + 285.0 * nu ** 2) / (1320.0 * (3.0 * nu - 1.0)),
        "rho66v2": (-106.0 + 602.0 * nu - 861.0 * nu ** 2 + 273.0 * nu ** 3) / (84.0 * (1.0 - 5.0 * nu + 5.0 * nu ** 2)),
    }


def test_factorized_waveform_coefficients_published_values(oracle):
    """The (2,2) phase and amplitude tables and the 1PN rho_lm of the other modes
    against the literature (nonspinning; odd-m modes for unequal masses only, the
    reference leaves them unset at equal mass)."""
    q = np.array([1.0, 1.7, 3.0, 8.0])
    m1, m2 = q / (1 + q), 1 / (1 + q)
    add = oracle.make_additional_args(m1, m2, np.zeros(q.size), np.zeros(q.size))
    rows = oracle.fac_coeffs_named(add)
    for i, c in enumerate(rows):
        nu = m1[i] * m2[i]
        for k, v in _published_coefficients(nu).items():
            if q[i] == 1.0 and k in ("rho21v2", "rho33v2", "rho31v2"):
                continue
            assert abs(c[k] - v) <= 1e-14 * max(1.0, abs(v)), (q[i], k, c[k], v)


def test_h22_on_knots_against_factorized_form(oracle):
    """h_22 = h^N S_eff T_22 e^{i delta_22} rho_22^2 (Damour-Iyer-Nagar 2008 Eq. 2)
    assembled here from independent pieces: T_22 = Gamma(3 - 2 i k) / 2 e^{pi k}
    e^{2 i k ln(8 Omega / sqrt(e))} with k = 2 H Omega and scipy's complex Gamma,
    delta_22 and rho_22 from the published series (nonspinning), S_eff = H_eff,
    h^N = prefix_22 v_Phi^2 Y_2^{-2}(pi/2, phi) with scipy's Y_lm."""
    for qq in (1.3, 4.0):
        m1, m2 = qq / (1 + qq), 1 / (1 + qq)
        add = oracle.make_additional_args(np.array([m1]), np.array([m2]), np.zeros(1), np.zeros(1))
        co = oracle.fac_coeffs(add)
        nu = m1 * m2
        pub = _published_coefficients(nu)
        pf = add[12, 0] + 1j * add[13, 0]
        rng = np.random.default_rng(31)
        L = 40
        v = rng.uniform(0.15, 0.45, L)
        Hr = 1.0 - nu * v ** 2 * rng.uniform(0.3, 0.6, L)
        r = v ** -2 * rng.uniform(0.95, 1.05, L)
        vPhiIn = rng.uniform(0.9, 1.1, L) / (v ** 6 * r ** 3)
        phi = rng.uniform(0.0, 40.0, L)
        h_o = oracle.hlm22(r[None], phi[None], v[None], Hr[None], np.array([nu]), vPhiIn[None], np.array([pf]), co)[0]
        Om = v ** 3
        k = 2.0 * Hr * Om
        T = sps.gamma(3.0 - 2j * k) / 2.0 * np.exp(math.pi * k) * np.exp(2j * k * np.log(8.0 * Om / math.sqrt(math.e)))
        y32 = Hr * Om
        delta = 7.0 / 3.0 * y32 + pub["delta22vh6"] * y32 ** 2 + pub["delta22vh9"] * y32 ** 3 + pub["delta22v5"] * v ** 5
        el = 0.5772156649015329 + np.log(4.0 * v)
        rho = (1 + pub["rho22v2"] * v ** 2 + pub["rho22v4"] * v ** 4 + (pub["rho22v6"] + pub["rho22v6l"] * el) * v ** 6
               + (pub["rho22v8"] + pub["rho22v8l"] * el) * v ** 8 + (pub["rho22v10"] + pub["rho22v10l"] * el) * v ** 10)
        Seff = (Hr ** 2 - 1.0) / (2 * nu) + 1.0
        vPhi = r * np.cbrt(vPhiIn) * Om
        ylm = sps.sph_harm_y(2, -2, math.pi / 2, phi) if hasattr(sps, "sph_harm_y") else sps.sph_harm(-2, 2, phi, math.pi / 2)
        hN = pf * vPhi ** 2 * ylm
        h = hN * Seff * T * np.exp(1j * delta) * rho ** 2
        np.testing.assert_allclose(h_o, h, rtol=1e-12)


def _mode_points(rng, nu, L=24):
    v = rng.uniform(0.15, 0.45, L)
    Hr = 1.0 - nu * v ** 2 * rng.uniform(0.3, 0.6, L)
    r = v ** -2 * rng.uniform(0.95, 1.05, L)
    vPhiIn = rng.uniform(0.9, 1.1, L) / (v ** 6 * r ** 3)
    pphi = np.sqrt(r) * rng.uniform(0.98, 1.1, L)
    return v, Hr, r, vPhiIn, pphi


ALL_MODES = [(l, m) for l in range(2, 9) for m in range(1, l + 1)]


def test_higher_mode_moduli_equal_the_flux_modes(oracle):
    """|h_lm| of the waveform (complex Gamma tail, nested rho and f, Y_lm with
    its phase) equals |h_lm| as the flux computes it (the product form of |T_lm|,
    the flux's own rho/f polynomials, |Y_lm(pi/2)|) for all 35 modes where the
    two polynomials coincide (use_hm = 0); with use_hm = 1 the waveform's extra
    terms of (2,1), (3,3), (4,4), (5,5) make those four differ and no other."""
    rng = np.random.default_rng(41)
    for m1, m2, c1, c2 in ((0.7, 0.3, 0.4, -0.2), (0.55, 0.45, -0.6, 0.5), (0.5, 0.5, 0.3, -0.1)):
        add = oracle.make_additional_args(np.array([m1]), np.array([m2]), np.array([c1]), np.array([c2]))
        nu = m1 * m2 / (m1 + m2) ** 2
        v, Hr, r, vPhiIn, pphi = _mode_points(rng, nu)
        phi = rng.uniform(0.0, 40.0, v.size)
        vPhi2 = (r * np.cbrt(vPhiIn) * v ** 3) ** 2
        for use_hm in (0, 1):
            co = oracle.fac_coeffs(add, use_hm=use_hm)
            for q, (l, m) in enumerate(ALL_MODES):
                pf = add[10 + 2 * q, 0] + 1j * add[11 + 2 * q, 0]
                h = oracle.hlm(l, m, r[None], phi[None], pphi[None], v[None], Hr[None], np.array([nu]), vPhiIn[None], np.array([pf]), co)[0]
                a2 = oracle.flux_mode_abs2(l, m, v, Hr, pphi, vPhi2, abs(pf), nu, co[0])
                close = np.allclose(np.abs(h) ** 2, a2, rtol=1e-12, atol=0)
                if use_hm == 0:
                    assert close, ((m1, m2), (l, m))
                else:
                    extra = (l, m) in ((2, 1), (3, 3), (4, 4), (5, 5))
                    assert close != extra, ((m1, m2), (l, m))


def test_higher_mode_tail_and_multipole_phase(oracle):
    """The phases the modes share: for the modes with no delta_lm and no f_lm
    ((7, even m) and l = 8), arg h_lm at phi = 0 is arg T_lm + arg prefix (mod pi)
    with arg T_lm = 2 k ln(4 m Omega / sqrt(e)) + Im lnGamma(l + 1 - 2 i k),
    k = H m Omega, from scipy's loggamma; and every mode turns as e^{-i m phi}."""
    rng = np.random.default_rng(43)
    add = oracle.make_additional_args(np.array([0.62]), np.array([0.38]), np.array([0.3]), np.array([0.1]))
    co = oracle.fac_coeffs(add)
    nu = 0.62 * 0.38
    v, Hr, r, vPhiIn, pphi = _mode_points(rng, nu)
    for q, (l, m) in enumerate(ALL_MODES):
        pf = add[10 + 2 * q, 0] + 1j * add[11 + 2 * q, 0]

        def h_at(phi):
            return oracle.hlm(l, m, r[None], np.full((1, v.size), phi), pphi[None], v[None], Hr[None], np.array([nu]), vPhiIn[None],
                              np.array([pf]), co)[0]
        h0 = h_at(0.0)
        phi1 = 0.7
        np.testing.assert_allclose(h_at(phi1), h0 * np.exp(-1j * m * phi1), rtol=1e-12)
        if (l == 7 and m % 2 == 0) or l == 8:
            k = Hr * m * v ** 3
            argT = 2 * k * np.log(4 * m * v ** 3 / math.sqrt(math.e)) + sps.loggamma(l + 1 - 2j * k).imag
            d = np.angle(h0 / pf) - argT
            d = (d + math.pi / 2) % math.pi - math.pi / 2     # mod pi (the sign of Y_lm and rho^l)
            assert np.max(np.abs(d)) < 1e-11, (l, m)


def test_higher_mode_leading_phases_and_22_consistency(oracle):
    """delta_lm's leading (H Omega) coefficients of the four higher modes against
    Damour-Iyer-Nagar 2009 (delta_21 = 2/3, delta_33 = 13/10, delta_44 = (112 +
    219 nu) / (120 (1 - 3 nu)), delta_55 = (96875 + 857528 nu) / (131250 (1 - 2 nu))),
    and the general waveform's (2,2) equals the (2,2) evaluation."""
    for nu_m1 in (0.7, 0.55):
        m1, m2 = nu_m1, 1 - nu_m1
        add = oracle.make_additional_args(np.array([m1]), np.array([m2]), np.array([0.2]), np.array([-0.3]))
        nu = m1 * m2
        c = oracle.fac_coeffs_named(add)[0]
        assert c["delta21vh3"] == pytest.approx(2.0 / 3.0, rel=1e-15)
        assert c["delta33vh3"] == pytest.approx(13.0 / 10.0, rel=1e-15)
        assert c["delta44vh3"] == pytest.approx((112 + 219 * nu) / (120 * (1 - 3 * nu)), rel=1e-14)
        assert c["delta55vh3"] == pytest.approx((96875 + 857528 * nu) / (131250 * (1 - 2 * nu)), rel=1e-14)
        co = oracle.fac_coeffs(add)
        rng = np.random.default_rng(47)
        v, Hr, r, vPhiIn, pphi = _mode_points(rng, nu)
        phi = rng.uniform(0.0, 40.0, v.size)
        pf = np.array([add[12, 0] + 1j * add[13, 0]])
        for h22c in (0.0, 0.3):
            a = oracle.hlm(2, 2, r[None], phi[None], pphi[None], v[None], Hr[None], np.array([nu]), vPhiIn[None], pf, co, h22_calib=h22c)
            b = oracle.hlm22(r[None], phi[None], v[None], Hr[None], np.array([nu]), vPhiIn[None], pf, co, h22c)
            np.testing.assert_allclose(a, b, rtol=1e-14)


# ------------------------------------------------------ initial conditions
def test_initial_condition_residuals(oracle):
    """The circular start solves dH/dr = 0 at pr = 0 for pphi0, and the
    dissipative condition of IC_diss (src/EOB.cu:2666-2688) for pr0:
    dr/dt = dH/dpr with dr/dt = F_phi / (dpphi/dr) and dpphi/dr = -H_rr / H_rpphi
    along circular orbits; omega0 = dH/dpphi.  The residuals are evaluated here
    with the oracle's separately pinned gradient, Hessian and radiation-reaction
    force."""
    rng = np.random.default_rng(41)
    n = 12
    q = rng.uniform(1, 8, n)
    add = oracle.make_additional_args(q / (1 + q), 1 / (1 + q), rng.uniform(-0.8, 0.8, n), rng.uniform(-0.8, 0.8, n))
    co = oracle.fac_coeffs(add)
    r0 = rng.uniform(12.0, 40.0, n)
    pphi0, pr0, om0, add2 = oracle.initial_conditions(r0, add, co)
    circ = np.array([r0, np.zeros(n), np.zeros(n), pphi0])
    g0 = oracle.grad_H(circ, add)
    assert np.all(np.abs(g0[0]) <= 1e-12 * np.abs(g0[3]) / r0), g0[0]
    st = np.array([r0, np.zeros(n), pr0, pphi0])
    g = oracle.grad_H(st, add)
    hs = oracle.hessian_H(st, add)
    aw = add.copy()
    aw[4] = g[3]
    F = oracle.rr_force(st, aw, co)
    dLdr = -hs[0, 0] / hs[3, 0]
    resid = F[1] / dLdr - g[2]
    # the false position stops at |residual| < err = 1e-12 (absolute)
    assert np.all(np.abs(resid) < 1e-12), resid
    assert np.all(np.abs(resid) <= 1e-9 * np.abs(g[2])), resid / g[2]
    np.testing.assert_allclose(om0, g0[3], rtol=1e-14)
    assert np.all(pr0 < 0) and np.all(pr0 > -3e-2)


def test_initial_condition_drop_in_solvers(oracle):
    """The reference's solvers restated in the oracle: root_find_all's (r, pphi)
    solve IC_cons (dH/dpphi = omega0 and dH/dr = 0 at pr = 0, src/EOB.cu:2840-3044)
    from a perturbed start; root_find_scalar_all's pr solves IC_diss's dissipative
    condition (src/EOB.cu:2691-2838) -- both checked with the separately pinned
    gradient, Hessian and radiation-reaction force."""
    rng = np.random.default_rng(45)
    n = 10
    q = rng.uniform(1, 8, n)
    add = oracle.make_additional_args(q / (1 + q), 1 / (1 + q), rng.uniform(-0.8, 0.8, n), rng.uniform(-0.8, 0.8, n))
    co = oracle.fac_coeffs(add)
    r0 = rng.uniform(12.0, 30.0, n)
    pphi0, pr0, om0, add2 = oracle.initial_conditions(r0, add, co)
    x0 = np.array([r0 * rng.uniform(0.97, 1.03, n), pphi0 * rng.uniform(0.98, 1.02, n)])
    x = oracle.root_find_all(x0, add2)
    st = np.array([x[0], np.zeros(n), np.zeros(n), x[1]])
    g = oracle.grad_H(st, add2)
    assert np.all(np.abs(g[3] - add2[4]) <= 1e-12 * add2[4]), g[3] - add2[4]
    assert np.all(np.abs(g[0]) <= 1e-12 * g[3] / x[0]), g[0]
    np.testing.assert_allclose(x[0], r0, rtol=1e-9)
    np.testing.assert_allclose(x[1], pphi0, rtol=1e-9)
    args = np.array([r0, np.zeros(n), np.zeros(n), pphi0])
    bounds = np.array([np.full(n, -3e-2), np.zeros(n)])
    pr = oracle.root_find_scalar_all(bounds, args, add2, co)
    s2 = np.array([r0, np.zeros(n), pr, pphi0])
    g2 = oracle.grad_H(s2, add2)
    hs = oracle.hessian_H(s2, add2)
    aw = add2.copy()
    aw[4] = g2[3]
    F = oracle.rr_force(s2, aw, co)
    resid = F[1] / (-hs[0, 0] / hs[3, 0]) - g2[2]
    assert np.all(np.abs(resid) < 1e-12), resid
    out = oracle.ic_diss(pr, args, add2, co)[0]
    assert np.all(np.abs(out) < 1e-12), out


def test_dop853_attempt_stages(oracle):
    """One attempted step of the oracle's stepper equals the stages formed here
    from scipy's DOP853 tableau and the oracle's (separately pinned) right-hand
    side: k_s = f(y + h sum_j a_sj k_j), y_new = y + h sum_s b_s k_s."""
    import scipy.integrate._ivp.dop853_coefficients as d
    rng = np.random.default_rng(49)
    n = 6
    q = rng.uniform(1, 6, n)
    add = oracle.make_additional_args(q / (1 + q), 1 / (1 + q), rng.uniform(-0.5, 0.5, n), rng.uniform(-0.5, 0.5, n))
    co = oracle.fac_coeffs(add)
    r0 = rng.uniform(8.0, 20.0, n)
    pphi0, pr0, om0, add = oracle.initial_conditions(r0, add, co)
    y = np.array([r0, rng.uniform(0, 10, n), pr0, pphi0])
    h = rng.uniform(0.5, 4.0, n)
    k_o, yn_o, err_o = oracle.dopr_attempt(y, h, add, co)
    k = np.zeros((12, 4, n))
    for s_ in range(12):
        acc = np.zeros((4, n))
        for j in range(s_):
            acc = acc + d.A[s_, j] * k[j]
        k[s_] = oracle.rhs(y + h * acc, add, co)
    yn = y + h * np.tensordot(d.B[:12], k, axes=1)
    np.testing.assert_allclose(k_o, k, rtol=1e-13, atol=1e-30)
    np.testing.assert_allclose(yn_o, yn, rtol=1e-14)
    e_r, _ = oracle.error(y, h, {"k1": k[0], "k2": k[10], "k3": k[11], "k6": k[5], "k7": k[6], "k8": k[7], "k9": k[8], "k10": k[9]})
    np.testing.assert_allclose(err_o, e_r, rtol=1e-12)


# ------------------------------------------------------ special functions
def test_complex_gamma_and_ylm(oracle):
    for z in (3 - 0.4j, 3 - 2.5j, 1.2 + 0.7j, 7.5 - 3j):
        np.testing.assert_allclose(oracle.cgamma(z), sps.gamma(z), rtol=2e-14)
    for l in range(2, 9):
        for m in range(-l, l + 1):
            phi = 0.37 * l + 0.11 * m
            ref = sps.sph_harm_y(l, m, math.pi / 2, phi) if hasattr(sps, "sph_harm_y") else sps.sph_harm(m, l, phi, math.pi / 2)
            np.testing.assert_allclose(oracle.ylm_equator(l, m, phi), ref, rtol=1e-12, atol=1e-15)


# ------------------------------------------------------------ spline, likelihood
def test_spline_build_matches_scipy_not_a_knot(oracle):
    rng = np.random.default_rng(9)
    nbin, nsubs, maxlen = 3, 2, 200
    lengths = np.array([50, 200, 123], dtype=np.int32)
    t = np.zeros((maxlen, nbin)); y = np.zeros((maxlen, nsubs * nbin))
    for b in range(nbin):
        tk = np.cumsum(rng.uniform(0.1, 2.0, lengths[b]))
        t[:lengths[b], b] = tk
        for s in range(nsubs):
            y[:lengths[b], s * nbin + b] = np.cos(0.3 * tk + s) * np.exp(-0.01 * tk)
    B, c1, c2, c3 = oracle.interpolate_TD(t, y, lengths, nbin, nsubs)
    for b in range(nbin):
        L = lengths[b]
        for s in range(nsubs):
            j = s * nbin + b
            cs = CubicSpline(t[:L, b], y[:L, j], bc_type="not-a-knot")
            np.testing.assert_allclose(B[:L, j], cs(t[:L, b], 1), rtol=1e-9, atol=1e-12)
            np.testing.assert_allclose(c1[:L - 1, j], cs.c[2], rtol=1e-9, atol=1e-12)
            np.testing.assert_allclose(c2[:L - 1, j], cs.c[1], rtol=1e-8, atol=1e-11)
            np.testing.assert_allclose(c3[:L - 1, j], cs.c[0], rtol=1e-8, atol=1e-11)
            tn = np.sort(rng.uniform(t[0, b], t[L - 1, b], 300))
            np.testing.assert_allclose(oracle.spline_eval(tn, t[:L, b], y[:L, j], c1[:L, j], c2[:L, j], c3[:L, j], L), cs(tn),
                                       rtol=1e-9, atol=1e-12)


def test_likelihood_partials_match_numpy(oracle):
    rng = np.random.default_rng(10)
    nbin, nchan, F = 2, 2, 2500
    cn = lambda *s: rng.normal(size=s) + 1j * rng.normal(size=s)
    hp, hc, data = cn(nbin, F), cn(nbin, F), cn(nchan, F)
    psd = rng.uniform(0.5, 2, (nchan, F))
    Fp, Fc, ts = rng.normal(size=(nchan, nbin)), rng.normal(size=(nchan, nbin)), rng.normal(size=(nchan, nbin)) * 1e-3
    df = 0.5
    parts = oracle.likelihood_partials(hp, hc, data, psd, Fp, Fc, ts, df)
    f = np.arange(F) * df
    for b in range(nbin):
        for c in range(nchan):
            h = (Fp[c, b] * hp[b] + Fc[c, b] * hc[b]) * np.exp(-2j * np.pi * f * ts[c, b])
            ref = np.sum(np.abs(data[c] - h) ** 2 / psd[c])
            np.testing.assert_allclose(parts[:, b, c].sum().real, ref, rtol=1e-12)


# ------------------------------------------------------------ C ABI
def test_c_abi_exports_every_declared_symbol():
    from eobgpu_amd import _lib
    header = open(os.path.join(ROOT, "include", "eobhip.h")).read()
    declared = set(re.findall(r"\b(eobhip_\w+)\s*\(", header))
    L = _lib.lib()
    missing = [s for s in declared if not hasattr(L, s)]
    assert not missing, missing
    assert set(_lib.exported_symbols()) >= declared


def test_c_abi_tableau_matches_oracle(oracle):
    from eobgpu_amd import _lib
    out = np.zeros(183)
    assert _lib.lib().eobhip_dop853_tableau(out.ctypes.data_as(ctypes.c_void_p)) == 0
    np.testing.assert_array_equal(out, oracle.tableau())


# -------------------------------------------------------- detector projection
def test_detector_response_closed_form(oracle):
    """An L-shaped detector with arms along x and y, D = (xx - yy)/2, at the
    origin: with this wave frame (m = -u sin psi - v cos psi, n = -u cos psi +
    v sin psi) the antenna patterns are F+ = -1/2 (1 + cos^2 th) cos 2ph cos 2psi -
    cos th sin 2ph sin 2psi and Fx = 1/2 (1 + cos^2 th) cos 2ph sin 2psi -
    cos th sin 2ph cos 2psi; a vertex at distance c along the source direction
    gives a delay of 1 s; F+^2 + Fx^2 does not depend on psi."""
    rng = np.random.default_rng(5)
    n = 50
    ra, dec, psi = rng.uniform(0, 2 * np.pi, n), rng.uniform(-1.5, 1.5, n), rng.uniform(0, np.pi, n)
    gmst = rng.uniform(0, 2 * np.pi, n)
    D = np.zeros((2, 3, 3))
    D[0, 0, 0], D[0, 1, 1] = 0.5, -0.5
    D[1] = D[0]
    c = 299792458.0
    V = np.array([[0.0, 0.0, 0.0], [0.0, 0.0, c]])
    Fp, Fc, dt = oracle.detector_information(ra, dec, gmst, psi, D, V)
    th, ph = np.pi / 2 - dec, ra - gmst
    fp = -0.5 * (1 + np.cos(th) ** 2) * np.cos(2 * ph) * np.cos(2 * psi) - np.cos(th) * np.sin(2 * ph) * np.sin(2 * psi)
    fc = 0.5 * (1 + np.cos(th) ** 2) * np.cos(2 * ph) * np.sin(2 * psi) - np.cos(th) * np.sin(2 * ph) * np.cos(2 * psi)
    np.testing.assert_allclose(Fp[0], fp, atol=1e-14)
    np.testing.assert_allclose(Fc[0], fc, atol=1e-14)
    np.testing.assert_allclose(dt[0], 0.0, atol=0)
    np.testing.assert_allclose(dt[1], np.cos(th), atol=1e-15)
    Fp2, Fc2, _ = oracle.detector_information(ra, dec, gmst, psi + 0.7, D, V)
    np.testing.assert_allclose(Fp2 ** 2 + Fc2 ** 2, Fp ** 2 + Fc ** 2, atol=1e-14)
