// The factorized-resummation waveform h_lm on trajectory knots for every
// (l, m), l = 2..8, m = 1..l, in one mode-uniform form.
//
// Reference: EOBGetSpinFactorizedWaveform (src/EOB.cu:1236-1613), the
// Newtonian multipole EOBFluxCalculateNewtonianMultipole (src/EOB.cu:1012-1027)
// and the coefficient tables EOBCalcSpinFacWaveformCoefficients
// (src/EOB.cu:1946-2462).
//
//   h_lm = T_lm e^{i delta_lm} S_lm (rho_lm^l + f_lm) h^N_lm
//   T_lm = e^{pi hk + 2 i hk ln(4 k / sqrt(e))} Gamma(l + 1 - 2 i hk) / l!,  k = m Omega, hk = H k
//   S_lm = H_eff (l + m even), v pphi (odd)
//   h^N  = prefix_lm vPhi^(l+eps) Y_{l-eps}^{-m}(pi/2, phi)
//   rho^l is replaced by f alone for odd m at eta = 1/4.
// Every mode's delta, rho and f are the reference's nested polynomials
// written as sums of powers: rho = 1 + sum_k (c_k + cl_k eulerlog_m) v^k (k <=
// 10) + c_12 v^12, f = sum_k f_k v^k + i fh6 (H Omega)^2, delta = sum_k dh_k
// vh^k (k = 3, 4, 6, 7, 9, vh = (H Omega)^(1/3)) + sum_k dv_k v^k (k = 5..8).
// The rho coefficients are the flux table's (eob_flux.hpp; the flux evaluates
// the same polynomials) plus the terms only the waveform reads.
#pragma once
#include "eob_device.hpp"

namespace eob {

struct WaveMode {
    double c[11];                    // rho: c[k-1] v^k for k = 1..10, c[10] v^12
    double cl6, cl7, cl8, cl10;      // rho: eulerlog_m(v) v^k
    double f[7];                     // Re f: f[k-1] v^k, k = 1..7
    double fh6;                      // Im f: fh6 (H Omega)^2
    double dh3, dh4, dh6, dh7, dh9;  // delta: vh^k
    double dv5, dv6, dv7, dv8;       // delta: v^k
    double ynorm;                    // Y_{l-eps}^{-m}(pi/2, 0), with its sign
    double log2m, inv_lfac;          // ln 2m, 1/l!
    int l, m, eps, pad;
};

// index of (l, m) in the flux table's order (l = 2..8, m = 1..l)
EOB_HD int flux_index(int l, int m) { return (l - 1) * l / 2 - 1 + m - 1; }

// Y_n^{mm}(pi/2, 0) for n >= |mm| with n + |mm| even (Condon-Shortley phase,
// scipy's sph_harm convention); the equator value of the associated Legendre
// function is (-1)^((n+|mm|)/2) (n+|mm|-1)!! / (n-|mm|)!!
EOB_HD double ylm_equator_real(int n, int mm) {
    const int am = mm < 0 ? -mm : mm;
    double v = ylm_abs_eq(n, am);
    if (((n + am) / 2) & 1) v = -v;
    if (mm < 0 && (am & 1)) v = -v;   // Y_n^{-|mm|} = (-1)^|mm| conj(Y_n^{|mm|})
    return v;
}

// The waveform table of mode (l, m) of one binary: the flux table's entry z of
// the same mode (built with the same use_hm), the terms only the waveform
// reads, and the phases.  m1, m2 as the coefficient routine takes them (M = 1),
// eta = m1 m2 / M^2, chiS, chiA.  Reference lines: src/EOB.cu:1975-1984 (delta22),
// 2019-2075 (2,1), 2080-2132 (3,3), 2134-2154 (3,2), (3,1), 2178-2245 (4,m),
// 2263-2341 (5,m), 2353-2424 (6,m), (7,m); aDelta = 0 as there.
EOB_HD void wave_mode_setup(WaveMode& w, const ModeCoef& z, int l, int m, double m1, double m2, double eta, double chiS, double chiA,
                            int use_hm, const WaveCoef* wc22) {
    const double PI = kPi, PI2 = kPi * kPi;
    const double eta2 = eta * eta;
    double dM2 = 1.0 - 4.0 * eta;
    if (dM2 < 0.0 && dM2 > -1e-4) dM2 = 0.0;
    double dM = sqrt(dM2);
    if (m1 < m2) dM = -dM;
    const bool uneq = (dM2 != 0.0);
    const double chiA2 = chiA * chiA, chiS2 = chiS * chiS, chiA3 = chiA2 * chiA, chiS3 = chiS2 * chiS;
    const double m3e = -1.0 + 3.0 * eta, n2e = -1.0 + 2.0 * eta;
    for (int k = 0; k < 11; ++k) w.c[k] = z.c[k];
    w.cl6 = z.cl6; w.cl7 = z.cl7; w.cl8 = z.cl8; w.cl10 = z.cl10;
    for (int k = 0; k < 7; ++k) w.f[k] = 0.0;
    w.f[0] = z.f1;
    w.f[2] = z.f3;
    w.fh6 = 0.0;
    w.dh3 = w.dh4 = w.dh6 = w.dh7 = w.dh9 = 0.0;
    w.dv5 = w.dv6 = w.dv7 = w.dv8 = 0.0;
    w.l = l; w.m = m; w.eps = (l + m) & 1; w.pad = 0;
    w.log2m = log(2.0 * m);
    double lf = 1.0;
    for (int q = 2; q <= l; ++q) lf *= q;
    w.inv_lfac = 1.0 / lf;
    w.ynorm = ylm_equator_real(l - w.eps, -m);
    switch (10 * l + m) {
    case 22:
        if (wc22) {
            w.dh3 = wc22->delta22vh3; w.dh6 = wc22->delta22vh6; w.dh9 = wc22->delta22vh9;
            w.dv5 = wc22->delta22v5; w.dv6 = wc22->delta22v6; w.dv8 = wc22->delta22v8;
        }
        break;
    case 21:
        w.dh3 = 2.0 / 3.0;
        w.dh6 = (107.0 * PI) / 105.0;
        w.dh9 = -272.0 / 81.0 + (214.0 * PI2) / 315.0;
        w.dv5 = -493.0 * eta / 42.0;
        if (use_hm) {   // f21v4..f21v6 (f21v7c = 0)
            if (uneq) {
                w.f[3] = (-3.0 - 2.0 * eta) * chiA2 + (-3.0 + eta / 2.0) * chiS2 + (-6.0 + 21.0 * eta / 2.0) * chiS * chiA / dM;
                w.f[4] = (3.0 / 4.0 - 3.0 * eta) * chiA3 / dM
                         + (-81.0 / 16.0 + 1709.0 * eta / 1008.0 + 613.0 * eta2 / 1008.0 + (9.0 / 4.0 - 3 * eta) * chiA2) * chiS + 3.0 / 4.0 * chiS3
                         + (-81.0 / 16.0 - 703.0 * eta2 / 112.0 + 8797.0 * eta / 1008.0 + (9.0 / 4.0 - 6.0 * eta) * chiS2) * chiA / dM;
                w.f[5] = (4163.0 / 252.0 - 9287.0 * eta / 1008.0 - 85.0 * eta2 / 112.0) * chiA2
                         + (4163.0 / 252.0 - 2633.0 * eta / 1008.0 + 461.0 * eta2 / 1008.0) * chiS2
                         + (4163.0 / 126.0 - 1636.0 * eta / 21.0 + 1088.0 * eta2 / 63.0) * chiS * chiA / dM;
            } else {
                w.f[3] = (-6 + 21 * eta / 2.0) * chiS * chiA;
                w.f[4] = (3.0 / 4.0 - 3.0 * eta) * chiA3
                         + (-81.0 / 16.0 - 703.0 * eta2 / 112.0 + 8797.0 * eta / 1008.0 + (9.0 / 4.0 - 6.0 * eta) * chiS2) * chiA;
                w.f[5] = (4163.0 / 126.0 - 1636.0 * eta / 21.0 + 1088.0 * eta2 / 63.0) * chiS * chiA;
            }
        }
        break;
    case 33:
        w.dh3 = 13.0 / 10.0;
        w.dh6 = (39.0 * PI) / 7.0;
        w.dh9 = -227827.0 / 3000.0 + (78.0 * PI2) / 7.0;
        w.dv5 = -80897.0 * eta / 2430.0;
        if (use_hm) {
            if (uneq) {
                w.c[9] = -903823148417327.0 / 30566888352000.0;   // rho33v10
                w.cl10 = 87347.0 / 13860.0;
                w.f[3] = (3.0 / 2.0 * chiS2 * dM + (3.0 - 12 * eta) * chiA * chiS + dM * (3.0 / 2.0 - 6.0 * eta) * chiA2) / dM;
                w.f[4] = (dM * (241.0 / 30.0 * eta2 + 11.0 / 20.0 * eta + 2.0 / 3.0) * chiS + (407.0 / 30.0 * eta2 - 593.0 / 60.0 * eta + 2.0 / 3.0) * chiA) / dM;
                w.f[5] = (dM * (6.0 * eta2 - 27.0 / 2.0 * eta - 7.0 / 4.0) * chiS2 + (44.0 * eta2 - 1.0 * eta - 7.0 / 2.0) * chiA * chiS
                          + dM * (-12 * eta2 + 11.0 / 2.0 * eta - 7.0 / 4.0) * chiA2) / dM;
                w.fh6 = (dM * (593.0 / 108.0 * eta - 81.0 / 20.0) * chiS + (7339.0 / 540.0 * eta - 81.0 / 20.0) * chiA) / dM;
            } else {
                w.f[3] = (3.0 - 12 * eta) * chiA * chiS;
                w.f[4] = (407.0 / 30.0 * eta2 - 593.0 / 60.0 * eta + 2.0 / 3.0) * chiA;
                w.f[5] = (44.0 * eta2 - 1.0 * eta - 7.0 / 2.0) * chiA * chiS;
                w.fh6 = (7339.0 / 540.0 * eta - 81.0 / 20.0) * chiA;
            }
        }
        break;
    case 32:
        w.dh3 = (10.0 + 33.0 * eta) / (-15.0 * m3e);
        w.dh6 = (52.0 * PI) / 21.0;
        w.dh9 = -9112.0 / 405.0 + (208.0 * PI2) / 63.0;
        break;
    case 31:
        w.dh3 = 13.0 / 30.0;
        w.dh6 = (13.0 * PI) / 21.0;
        w.dh9 = -227827.0 / 81000.0 + (26.0 * PI2) / 63.0;
        w.dv5 = -17.0 * eta / 10.0;
        break;
    case 44:
        w.dh3 = (112.0 + 219.0 * eta) / (-120.0 * m3e);
        w.dh6 = (25136.0 * PI) / 3465.0;
        if (use_hm) {
            w.dh9 = -55144.0 / 375.0 + 201088.0 * PI2 / 10395.0;
            w.c[7] = -172066910136202271.0 / 19426955708160000.0;   // rho44v8
            w.cl8 = 845198.0 / 190575.0;
            w.c[9] = -17154485653213713419357.0 / 568432724020761600000.0;   // rho44v10
            w.cl10 = 22324502267.0 / 3815311500.0;
        }
        break;
    case 43:
        w.dh3 = (486.0 + 4961.0 * eta) / (810.0 * (1.0 - 2.0 * eta));
        w.dh6 = 1571.0 * PI / 385.0;
        break;
    case 42:
        w.dh3 = (7.0 * (1.0 + 6.0 * eta)) / (-15.0 * m3e);
        w.dh6 = (6284.0 * PI) / 3465.0;
        break;
    case 41:
        w.dh3 = (2.0 + 507.0 * eta) / (10.0 * (1.0 - 2.0 * eta));
        w.dh6 = 1571.0 * PI / 3465.0;
        break;
    case 55:
        w.dh3 = (96875.0 + 857528.0 * eta) / (131250.0 * (1 - 2 * eta));
        if (use_hm) {
            w.dh6 = 3865.0 / 429.0 * PI;
            w.dh9 = (-7686949127.0 + 954500400.0 * PI2) / 31783752.0;
            if (uneq) {
                w.cl6 = -1546.0 / 429.0;   // rho55v6l (rho55v6 is in the flux table)
                w.c[7] = -1213641959949291437.0 / 118143853747920000.0;
                w.cl8 = 376451.0 / 83655.0;
                w.c[9] = -150082616449726042201261.0 / 4837990810977324000000.0;
                w.cl10 = 2592446431.0 / 456756300.0;
                w.f[2] = chiA / dM * (10.0 / (3.0 * n2e) - 70.0 * eta / (3.0 * n2e) + 110.0 * eta2 / (3.0 * n2e))
                         + chiS * (10.0 / (3.0 * n2e) - 10.0 * eta / n2e + 10 * eta2 / n2e);
                w.f[3] = chiS2 * (-5.0 / (2.0 * n2e) + 5.0 * eta / n2e) + chiA * chiS / dM * (-5.0 / n2e + 30.0 * eta / n2e - 40.0 * eta2 / n2e)
                         + chiA2 * (-5.0 / (2.0 * n2e) + 15.0 * eta / n2e - 20.0 * eta2 / n2e);
            } else {
                w.f[2] = chiA * (10.0 / (3.0 * n2e) - 70.0 * eta / (3.0 * n2e) + 110.0 * eta2 / (3.0 * n2e));
                w.f[3] = chiA * chiS * (-5.0 / n2e + 30.0 * eta / n2e - 40.0 * eta2 / n2e);
            }
        }
        break;
    case 54: w.dh3 = 8.0 / 15.0; break;
    case 53: w.dh3 = 31.0 / 70.0; break;
    case 52: w.dh3 = 4.0 / 15.0; break;
    case 51: w.dh3 = 31.0 / 210.0; break;
    case 66: w.dh3 = 43.0 / 70.0; break;
    case 65: w.dh3 = 10.0 / 21.0; break;
    case 64: w.dh3 = 43.0 / 105.0; break;
    case 63: w.dh3 = 2.0 / 7.0; break;
    case 62: w.dh3 = 43.0 / 210.0; break;
    case 61: w.dh3 = 2.0 / 21.0; break;
    case 77: w.dh3 = 19.0 / 36.0; break;
    case 75: w.dh3 = 95.0 / 252.0; break;
    case 73: w.dh3 = 19.0 / 84.0; break;
    case 71: w.dh3 = 19.0 / 252.0; break;
    default: break;   // (7, even m) and l = 8: no phase
    }
}

// h_lm at one knot.  pre: the mode's Newtonian prefix; v = Omega^(1/3), Hreal =
// nu H, vPhiIn = 1/(omega_circ^2 r^3); eta as the caller passes it (clamped to
// 1/4 within 1e-4 here, as the reference does).  With rotated, h_lm e^{i m phi}
// (the slowly varying series the reference splines, eob/seobnrv4phm.py:2338):
// the multipole's e^{-i m phi} is left out.
EOB_HD Cplx hlm_knot(const WaveMode& w, double pre_re, double pre_im, double eta, double r, double phi, double pphi, double v,
                     double Hreal, double vPhiIn, bool rotated = false) {
    if (eta > 0.25 && eta < 0.25 + 1e-4) eta = 0.25;
    const int l = w.l, m = w.m;
    const double v2 = v * v;
    const double Omega = v2 * v;
    const double vh3 = Hreal * Omega;
    const double vh = cbrt(vh3);
    const double el = kEulerGamma + w.log2m + log(v);
    double vPhi = r * cbrt(vPhiIn);
    vPhi *= Omega;
    const double vPhi2 = vPhi * vPhi;
    // Newtonian multipole: prefix vPhi2^((l+eps)/2) Y(pi/2, 0) e^{-i m phi}
    double vp = 1.0;
    for (int q = 0; q < l + w.eps; ++q) vp *= vPhi;
    const double yv = w.ynorm * vp;
    double sy = 0.0, cy = 1.0;
    if (!rotated) sincos(-(double)m * phi, &sy, &cy);
    Cplx hN = cmul({pre_re, pre_im}, {yv * cy, yv * sy});
    const double Slm = w.eps ? v * pphi : (Hreal * Hreal - 1.0) / (2.0 * eta) + 1.0;
    const double k = (double)m * Omega;
    const double hk = Hreal * k;
    const Cplx G = cgamma({(double)(l + 1), -2.0 * hk});
    Cplx T = cmul(cexp_(kPi * hk, 2.0 * hk * log(4.0 * k / 1.6487212707001282)), G);
    T.re *= w.inv_lfac; T.im *= w.inv_lfac;
    const double vh6 = vh3 * vh3;
    const double v5 = v2 * v2 * v;
    const double delta = w.dh3 * vh3 + w.dh4 * vh3 * vh + w.dh6 * vh6 + w.dh7 * vh6 * vh + w.dh9 * vh6 * vh3 + w.dv5 * v5 + w.dv6 * v5 * v
                         + w.dv7 * v5 * v2 + w.dv8 * v5 * v2 * v;
    double acc = w.c[10];
    acc = acc * v;
    acc = acc * v + (w.c[9] + w.cl10 * el);
    acc = acc * v + w.c[8];
    acc = acc * v + (w.c[7] + w.cl8 * el);
    acc = acc * v + (w.c[6] + w.cl7 * el);
    acc = acc * v + (w.c[5] + w.cl6 * el);
    acc = acc * v + w.c[4];
    acc = acc * v + w.c[3];
    acc = acc * v + w.c[2];
    acc = acc * v + w.c[1];
    acc = acc * v + w.c[0];
    const double rho = 1.0 + acc * v;
    double fr = w.f[6];
    for (int q = 5; q >= 0; --q) fr = fr * v + w.f[q];
    fr *= v;
    const double fi = w.fh6 * vh6;
    double rl = 1.0;
    for (int q = 0; q < l; ++q) rl *= rho;
    const bool fonly = (eta == 0.25) && (m & 1);
    const Cplx amp = fonly ? Cplx{fr, fi} : Cplx{rl + fr, fi};
    Cplx h = cmul(T, {cos(delta), sin(delta)});
    h = cmul(h, amp);
    h.re *= Slm; h.im *= Slm;
    return cmul(h, hN);
}

}  // namespace eob
