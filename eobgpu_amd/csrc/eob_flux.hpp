// Spin-factorized energy flux over the 35 (l,m) modes l = 2..8, m = 1..l, in a
// mode-uniform form so that one lane can evaluate any mode with the same
// instruction stream (mode-parallel flux across the lanes of a group).
//
// Reference: EOBSpinFactorizedFlux (src/EOB.cu:2506-2560), the flux waveform
// EOBFluxGetSpinFactorizedWaveform (src/EOB.cu:1706-1941) and its coefficient
// tables EOBCalcSpinFacWaveformCoefficients (src/EOB.cu:1946-2462).
//
// Per mode the reference computes |h_lm| = T_lm S_lm (rho_lm^l + f_lm) |h^N_lm|.
// Only |h_lm|^2 enters the flux, so this evaluates the square directly:
//   T_lm^2   = x/(1-exp(-x)) * prod_{i=1..l}(4 hk^2 + i^2) / (l!)^2,  x = 4 pi hk, hk = H m Omega
//   |h^N|^2  = |prefix_lm Y_{l-eps}^{-m}(pi/2)|^2 vPhi^(2(l+eps))
//   rho_lm   = 1 + sum_k (c_k + cl_k eulerlog_m(v)) v^k   (k = 1..10, 12)
// and exp(-4 pi H m Omega) is formed as exp(-4 pi H Omega)^m (one exp per RHS).
// The table values are the published factorized-resummation coefficients
// (Pan et al. 2011; Taracchini et al. 2012, 2014; Cotesta et al. 2018) and are
// identical to the reference's.
#pragma once
#include "eob_math.hpp"

namespace eob {

constexpr int kNumModes = 35;

struct ModeCoef {
    double c[11];                  // rho coefficient of v^(k+1), k = 0..9; c[10] multiplies v^12
    double cl6, cl7, cl8, cl10;    // coefficients of eulerlog v^k
    double f1, f3;                 // auxiliary f_lm: f1 v + f3 v^3
    double N2;                     // |prefix_lm|^2 |Y_{l-eps}^{-m}(pi/2,0)|^2
    double inv_lfac2;              // 1/(l!)^2
    double pp[9];                  // prod_{i=1..l} (h + i^2) = sum_j pp[j] h^j (integers; zero above j = l)
    double log2m;                  // ln(2m)
    double m2;                     // m^2
    int l, m, lpe, flags;          // lpe = l + eps; flags: bit0 even (l+m), bit1 f-only (eta = 1/4, odd m)
};

// (l, m) of flux mode index i in the reference's order (l = 2..8, m = 1..l)
EOB_HD void mode_lm(int i, int& l, int& m) {
    int l0 = 2, base = 0;
    while (i >= base + l0) { base += l0; ++l0; }
    l = l0; m = i - base + 1;
}

// |Y_l^m(pi/2, 0)|, closed form sqrt((2l+1)/(4pi) (l-m)!/(l+m)!) (l+m-1)!!/(l-m)!!
EOB_HD double ylm_abs_eq(int l, int m) {
    if ((l + m) & 1) return 0.0;
    double fl = 1.0, fh = 1.0, da = 1.0, db = 1.0;
    for (int i = 2; i <= l - m; ++i) fl *= i;
    for (int i = 2; i <= l + m; ++i) fh *= i;
    for (int i = l + m - 1; i > 1; i -= 2) da *= i;
    for (int i = l - m; i > 1; i -= 2) db *= i;
    return sqrt((2.0 * l + 1.0) / (4.0 * kPi) * fl / fh) * da / db;
}

// Per-binary (2,2) waveform phase coefficients and scalars used on trajectory knots.
struct WaveCoef {
    double delta22vh3, delta22vh6, delta22vh9, delta22v5, delta22v6, delta22v8;
    double pf22_re, pf22_im, h22_calib;
};

// Builds the 35-mode table.  m1, m2 scaled to M = 1 as the caller supplies them;
// prefixes: 70 doubles (Re, Im per mode) in flux-mode order.
// Inputs follow SEOBNRv5::update_information (src/EOB.cu:3355-3374): eta, the
// Kerr-like spin a (tplspin), chiS, chiA.  prefixes may be null (N2 left 0).
// The part of mode i's table that does not depend on the binary's PN
// coefficients: (l, m), 1/(l!)^2, the product polynomial, ln 2m, m^2,
// |prefix Y|^2, and the coefficients the tables below set, zeroed.
EOB_HD void flux_mode_init(ModeCoef& z, int i, double pre_re, double pre_im, bool have_prefix) {
    for (int k = 0; k < 11; ++k) z.c[k] = 0.0;
    z.cl6 = z.cl7 = z.cl8 = z.cl10 = 0.0;
    z.f1 = z.f3 = 0.0;
    int l, m;
    mode_lm(i, l, m);
    z.l = l; z.m = m;
    const int eps = (l + m) & 1;
    z.lpe = l + eps;
    double lf = 1.0;
    for (int q = 2; q <= l; ++q) lf *= q;
    z.inv_lfac2 = 1.0 / (lf * lf);
    // expand prod_{i=1..l} (h + i^2) one factor at a time (exact: the
    // coefficients stay below 2^53)
    for (int j = 0; j < 9; ++j) z.pp[j] = (j == 0) ? 1.0 : 0.0;
    for (int q = 1; q <= l; ++q) {
        for (int j = 8; j >= 1; --j) z.pp[j] = z.pp[j - 1] + (double)(q * q) * z.pp[j];
        z.pp[0] *= (double)(q * q);
    }
    z.log2m = log(2.0 * m);
    z.m2 = (double)(m * m);
    const double pa2 = have_prefix ? pre_re * pre_re + pre_im * pre_im : 0.0;
    const double y = ylm_abs_eq(l - eps, m);
    z.N2 = pa2 * y * y;
}

// tables_only: the caller has run flux_mode_init on every mode (one mode per
// lane); only the PN coefficient tables and the (2,2) waveform terms are set.
EOB_HD void flux_setup(ModeCoef* mc, WaveCoef& wc, double m1, double m2, double eta, double a, double chiS, double chiA,
                       const double* prefixes, double h22_calib, int use_hm, bool tables_only = false) {
    const double PI = kPi, PI2 = kPi * kPi;
    const double eta2 = eta * eta, eta3 = eta2 * eta, eta4 = eta2 * eta2;
    double dM2 = 1.0 - 4.0 * eta;
    if (dM2 < 0.0 && dM2 > -1e-4) dM2 = 0.0;
    double dM = sqrt(dM2);
    if (m1 < m2) dM = -dM;
    const bool uneq = (dM2 != 0.0);
    const double chiA2 = chiA * chiA, chiS2 = chiS * chiS, chiA3 = chiA2 * chiA, chiS3 = chiS2 * chiS;
    double a2 = a * a, a3 = a2 * a;
    const double m3e = -1.0 + 3.0 * eta, m3e2 = m3e * m3e;
    const double n2e = -1.0 + 2.0 * eta;
    const double p5 = 1.0 - 5.0 * eta + 5.0 * eta2;
    const double p7 = -1.0 + 7.0 * eta - 14.0 * eta2 + 7.0 * eta3;
    const double p6 = -1.0 + 6.0 * eta - 10.0 * eta2 + 4.0 * eta3;
    const double q3 = dM2 + 3.0 * eta2;

    if (!tables_only)
        for (int i = 0; i < kNumModes; ++i)
            flux_mode_init(mc[i], i, prefixes ? prefixes[2 * i] : 0.0, prefixes ? prefixes[2 * i + 1] : 0.0, prefixes != nullptr);
    // rho(v) = 1 + sum c[k-1] v^k; helpers name the reference's coefficient symbols
#define RHO(idx, k) mc[idx].c[(k) - 1]
    // (2,2) -- mode 1
    RHO(1, 2) = -43.0 / 42.0 + (55.0 * eta) / 84.0;
    RHO(1, 3) = (-2.0 * (chiS + chiA * dM - chiS * eta)) / 3.0;
    RHO(1, 4) = -20555.0 / 10584.0 + 0.5 * (chiS + chiA * dM) * (chiS + chiA * dM) - (33025.0 * eta) / 21168.0 + (19583.0 * eta2) / 42336.0;
    RHO(1, 5) = (-34.0 / 21.0 + 49.0 * eta / 18.0 + 209.0 * eta2 / 126.0) * chiS + (-34.0 / 21.0 - 19.0 * eta / 42.0) * dM * chiA;
    RHO(1, 6) = 1556919113.0 / 122245200.0 + (89.0 * a2) / 252.0 - (48993925.0 * eta) / 9779616.0 - (6292061.0 * eta2) / 3259872.0
                + (10620745.0 * eta3) / 39118464.0 + (41.0 * eta * PI2) / 192.0;
    mc[1].cl6 = -428.0 / 105.0;
    RHO(1, 7) = a3 / 3.0 + chiA * dM * (18733.0 / 15876.0 + (50140.0 * eta) / 3969.0 + (97865.0 * eta2) / 63504.0)
                + chiS * (18733.0 / 15876.0 + (74749.0 * eta) / 5292.0 - (245717.0 * eta2) / 63504.0 + (50803.0 * eta3) / 63504.0);
    RHO(1, 8) = -387216563023.0 / 160190110080.0 + (18353.0 * a2) / 21168.0 - a2 * a2 / 8.0;
    mc[1].cl8 = 9202.0 / 2205.0;
    RHO(1, 10) = -16094530514677.0 / 533967033600.0;
    mc[1].cl10 = 439877.0 / 55566.0;
    mc[1].c[10] = h22_calib;
    wc.delta22vh3 = 7.0 / 3.0;
    wc.delta22vh6 = -4.0 / 3.0 * (dM * chiA + chiS * (1 - 2 * eta)) + (428.0 * PI) / 105.0;
    wc.delta22vh9 = -2203.0 / 81.0 + (1712.0 * PI2) / 315.0;
    wc.delta22v5 = -24.0 * eta;
    wc.delta22v6 = 0.0;
    wc.delta22v8 = 0.0;
    wc.pf22_re = prefixes ? prefixes[2] : 0.0;
    wc.pf22_im = prefixes ? prefixes[3] : 0.0;
    wc.h22_calib = h22_calib;

    if (use_hm) { a = 0.0; a2 = 0.0; a3 = 0.0; }

    // (2,1) -- mode 0
    if (uneq) {
        RHO(0, 2) = -59.0 / 56 + (23.0 * eta) / 84.0;
        RHO(0, 4) = -47009.0 / 56448.0 - (865.0 * a2) / 1792.0 - (405.0 * a2 * a2) / 2048.0 - (10993.0 * eta) / 14112.0 + (617.0 * eta2) / 4704.0;
        RHO(0, 5) = (-98635.0 * a) / 75264.0 + (2031.0 * a * a2) / 7168.0 - (1701.0 * a2 * a3) / 8192.0;
        RHO(0, 6) = 7613184941.0 / 2607897600.0 + (9032393.0 * a2) / 1806336.0 + (3897.0 * a2 * a2) / 16384.0 - (15309.0 * a3 * a3) / 65536.0;
        mc[0].cl6 = -107.0 / 105.0;
        RHO(0, 7) = (-3859374457.0 * a) / 1159065600.0 - (55169.0 * a3) / 16384.0 + (18603.0 * a2 * a3) / 65536.0 - (72171.0 * a2 * a2 * a3) / 262144.0;
        mc[0].cl7 = 107.0 * a / 140.0;
        RHO(0, 8) = -1168617463883.0 / 911303737344.0;
        mc[0].cl8 = 6313.0 / 5880.0;
        RHO(0, 10) = -63735873771463.0 / 16569158860800.0;
        mc[0].cl10 = 5029963.0 / 5927040.0;
        mc[0].f1 = (-3.0 * (chiS + chiA / dM)) / 2.0;
        mc[0].f3 = (chiS * dM * (427.0 + 79.0 * eta) + chiA * (147.0 + 280.0 * dM * dM + 1251.0 * eta)) / 84.0 / dM;
    } else {
        mc[0].f1 = -3.0 * chiA / 2.0;
        mc[0].f3 = (chiS * dM * (427.0 + 79.0 * eta) + chiA * (147.0 + 280.0 * dM * dM + 1251.0 * eta)) / 84.0;
    }
    // (3,1) -- mode 2
    if (uneq) {
        RHO(2, 2) = -13.0 / 18.0 - (2.0 * eta) / 9.0;
        RHO(2, 4) = 101.0 / 7128.0 - (5.0 * a2) / 6.0 - (1685.0 * eta) / 1782.0 - (829.0 * eta2) / 1782.0;
        RHO(2, 5) = (4.0 * a) / 9.0;
        RHO(2, 6) = 11706720301.0 / 6129723600.0 - (49.0 * a2) / 108.0;
        mc[2].cl6 = -26.0 / 63.0;
        RHO(2, 7) = (-2579.0 * a) / 5346.0 + a * a2 / 9.0;
        RHO(2, 8) = 2606097992581.0 / 4854741091200.0;
        mc[2].cl8 = 169.0 / 567.0;
        mc[2].f3 = (chiA * (-4.0 + 11.0 * eta) + chiS * dM * (-4.0 + 13.0 * eta)) / (2.0 * dM);
    } else {
        mc[2].f3 = -chiA * 5.0 / 8.0;
    }
    // (3,2) -- mode 3
    RHO(3, 1) = (4.0 * chiS * eta) / (-3.0 * m3e);
    RHO(3, 2) = (328.0 - 1115.0 * eta + 320.0 * eta2) / (270.0 * m3e);
    RHO(3, 3) = 2.0 / 9.0 * a;
    RHO(3, 4) = a2 / 3.0 + (-1444528.0 + 8050045.0 * eta - 4725605.0 * eta2 - 20338960.0 * eta3 + 3085640.0 * eta4) / (1603800.0 * m3e2);
    RHO(3, 5) = (-2788.0 * a) / 1215.0;
    RHO(3, 6) = 5849948554.0 / 940355325.0 + (488.0 * a2) / 405.0;
    mc[3].cl6 = -104.0 / 63.0;
    RHO(3, 8) = -10607269449358.0 / 3072140846775.0;
    mc[3].cl8 = 17056.0 / 8505.0;
    // (3,3) -- mode 4
    if (uneq) {
        RHO(4, 2) = -7.0 / 6.0 + (2.0 * eta) / 3.0;
        RHO(4, 4) = -6719.0 / 3960.0 + a2 / 2.0 - (1861.0 * eta) / 990.0 + (149.0 * eta2) / 330.0;
        RHO(4, 5) = (-4.0 * a) / 3.0;
        RHO(4, 6) = use_hm ? (3203101567.0 / 227026800.0 + (5.0 * a2) / 36.0 + (-129509.0 / 25740.0 + 41.0 / 192.0 * PI2) * eta
                              - 274621.0 / 154440.0 * eta2 + 12011.0 / 46332.0 * eta3)
                           : (3203101567.0 / 227026800.0 + (5.0 * a2) / 36.0);
        mc[4].cl6 = -26.0 / 7.0;
        RHO(4, 7) = (5297.0 * a) / 2970.0 + a * a2 / 3.0;
        RHO(4, 8) = -57566572157.0 / 8562153600.0;
        mc[4].cl8 = 13.0 / 3.0;
        mc[4].f3 = (chiS * dM * (-4.0 + 5.0 * eta) + chiA * (-4.0 + 19.0 * eta)) / (2.0 * dM);
    } else {
        mc[4].f3 = chiA * 3.0 / 8.0;
    }
    // (4,1) -- mode 5
    if (uneq) {
        RHO(5, 2) = (602.0 - 1385.0 * eta + 288.0 * eta2) / (528.0 * n2e);
        RHO(5, 4) = -7775491.0 / 21141120.0 + (3.0 * a2) / 8.0;
        RHO(5, 5) = (-20033.0 * a) / 55440.0 - (5 * a * a2) / 6.0;
        RHO(5, 6) = 1227423222031.0 / 1758095539200.0;
        mc[5].cl6 = -1571.0 / 6930.0;
        mc[5].f1 = (5.0 * (chiA - chiS * dM) * eta) / (2.0 * dM * n2e);
    } else {
        mc[5].f1 = -5.0 * chiA / 4.0;
    }
    // (4,2) -- mode 6
    RHO(6, 2) = (1146.0 - 3530.0 * eta + 285.0 * eta2) / (1320.0 * m3e);
    RHO(6, 3) = (chiA * (10.0 - 21.0 * eta) * dM + chiS * (10.0 - 59.0 * eta + 78.0 * eta2)) / (15.0 * m3e);
    RHO(6, 4) = a2 / 2.0 + (-114859044.0 + 295834536.0 * eta + 1204388696.0 * eta2 - 3047981160.0 * eta3 - 379526805.0 * eta4) / (317116800.0 * m3e2);
    RHO(6, 5) = (-7.0 * a) / 110.0;
    RHO(6, 6) = 848238724511.0 / 219761942400.0 + (2323.0 * a2) / 3960.0;
    mc[6].cl6 = -3142.0 / 3465.0;
    // (4,3) -- mode 7
    if (uneq) {
        RHO(7, 2) = (222.0 - 547.0 * eta + 160.0 * eta2) / (176.0 * n2e);
        RHO(7, 4) = -6894273.0 / 7047040.0 + (3.0 * a2) / 8.0;
        RHO(7, 5) = (-12113.0 * a) / 6160.0;
        RHO(7, 6) = 1664224207351.0 / 195343948800.0;
        mc[7].cl6 = -1571.0 / 770.0;
        mc[7].f1 = (5.0 * (chiA - chiS * dM) * eta) / (2.0 * dM * n2e);
    } else {
        mc[7].f1 = -5.0 * chiA / 4.0;
    }
    // (4,4) -- mode 8
    RHO(8, 2) = (1614.0 - 5870.0 * eta + 2625.0 * eta2) / (1320.0 * m3e);
    RHO(8, 3) = (chiA * (10.0 - 39.0 * eta) * dM + chiS * (10.0 - 41.0 * eta + 42.0 * eta2)) / (15.0 * m3e);
    if (use_hm) {
        RHO(8, 4) = (-511573572.0 + 2338945704.0 * eta - 313857376.0 * eta2 - 6733146000.0 * eta3 + 1252563795.0 * eta4) / (317116800.0 * m3e2)
                    + chiS2 / 2.0 + dM * chiS * chiA + dM2 * chiA2 / 2.0;
        RHO(8, 5) = chiA * dM * (-8280.0 + 42716.0 * eta - 57990.0 * eta2 + 8955 * eta3) / (6600.0 * m3e2)
                    + chiS * (-8280.0 + 66284.0 * eta - 176418.0 * eta2 + 128085.0 * eta3 + 88650 * eta4) / (6600.0 * m3e2);
    } else {
        RHO(8, 4) = a2 / 2.0 + (-511573572.0 + 2338945704.0 * eta - 313857376.0 * eta2 - 6733146000.0 * eta3 + 1252563795.0 * eta4) / (317116800.0 * m3e2);
        RHO(8, 5) = (-69.0 * a) / 55.0;
    }
    RHO(8, 6) = 16600939332793.0 / 1098809712000.0 + (217.0 * a2) / 3960.0;
    mc[8].cl6 = -12568.0 / 3465.0;
    // (5,1)..(5,5) -- modes 9..13
    if (uneq) {
        RHO(9, 2) = (319.0 - 626.0 * eta + 8.0 * eta2) / (390.0 * n2e);
        RHO(9, 3) = (-2.0 * a) / 3.0;
        RHO(9, 4) = -31877.0 / 304200.0 + a2 / 2.0;
        RHO(9, 5) = 139.0 * a / 975.0;
    }
    RHO(10, 2) = (-15828.0 + 84679.0 * eta - 104930.0 * eta2 + 21980.0 * eta3) / (13650.0 * p5);
    RHO(10, 3) = (-2.0 * a) / 15.0;
    RHO(10, 4) = -7187914.0 / 15526875.0 + (2.0 * a2) / 5.0;
    if (uneq) {
        RHO(11, 2) = (375.0 - 850.0 * eta + 176.0 * eta2) / (390.0 * n2e);
        RHO(11, 3) = (-2.0 * a) / 3.0;
        RHO(11, 4) = -410833.0 / 709800.0 + a2 / 2.0;
        RHO(11, 5) = -103.0 * a / 325.0;
    }
    RHO(12, 2) = (-17448.0 + 96019.0 * eta - 127610.0 * eta2 + 33320.0 * eta3) / (13650.0 * p5);
    RHO(12, 3) = (-2.0 * a) / 15.0;
    RHO(12, 4) = -16213384.0 / 15526875.0 + (2.0 * a2) / 5.0;
    if (uneq) {
        RHO(13, 2) = (487.0 - 1298.0 * eta + 512.0 * eta2) / (390.0 * n2e);
        RHO(13, 3) = (-2.0 * a) / 3.0;
        RHO(13, 4) = -3353747.0 / 2129400.0 + a2 / 2.0;
        RHO(13, 5) = -241.0 * a / 195.0;
        if (use_hm) RHO(13, 6) = 190606537999247.0 / 11957879934000.0;
    }
    // (6,1)..(6,6) -- modes 14..19
    if (uneq) { RHO(14, 2) = (-161.0 + 694.0 * eta - 670.0 * eta2 + 124.0 * eta3) / (144.0 * q3); RHO(14, 3) = -2.0 * a / 9.0; }
    RHO(15, 2) = (-74.0 + 378.0 * eta - 413.0 * eta2 + 49.0 * eta3) / (84.0 * p5);
    RHO(15, 3) = (-2.0 * a) / 3.0;
    RHO(15, 4) = -817991.0 / 3298680.0 + a2 / 2.0;
    if (uneq) { RHO(16, 2) = (-169.0 + 742.0 * eta - 750.0 * eta2 + 156.0 * eta3) / (144.0 * q3); RHO(16, 3) = -2.0 * a / 9.0; }
    RHO(17, 2) = (-86.0 + 462.0 * eta - 581.0 * eta2 + 133.0 * eta3) / (84.0 * p5);
    RHO(17, 3) = (-2.0 * a) / 3.0;
    RHO(17, 4) = -476887.0 / 659736.0 + a2 / 2.0;
    if (uneq) { RHO(18, 2) = (-185.0 + 838.0 * eta - 910.0 * eta2 + 220.0 * eta3) / (144.0 * q3); RHO(18, 3) = -2.0 * a / 9.0; }
    RHO(19, 2) = (-106.0 + 602.0 * eta - 861.0 * eta2 + 273.0 * eta3) / (84.0 * p5);
    RHO(19, 3) = (-2.0 * a) / 3.0;
    RHO(19, 4) = -1025435.0 / 659736.0 + a2 / 2.0;
    // (7,1)..(7,7) -- modes 20..26
    if (uneq) { RHO(20, 2) = (-618.0 + 2518.0 * eta - 2083.0 * eta2 + 228.0 * eta3) / (714.0 * q3); RHO(20, 3) = -2.0 * a / 3.0; }
    RHO(21, 2) = (16832.0 - 123489.0 * eta + 273924.0 * eta2 - 190239.0 * eta3 + 32760.0 * eta4) / (14994.0 * p7);
    if (uneq) { RHO(22, 2) = (-666.0 + 2806.0 * eta - 2563.0 * eta2 + 420.0 * eta3) / (714.0 * q3); RHO(22, 3) = -2.0 * a / 3.0; }
    RHO(23, 2) = (17756.0 - 131805.0 * eta + 298872.0 * eta2 - 217959.0 * eta3 + 41076.0 * eta4) / (14994.0 * p7);
    if (uneq) { RHO(24, 2) = (-762.0 + 3382.0 * eta - 3523.0 * eta2 + 804.0 * eta3) / (714.0 * q3); RHO(24, 3) = -2.0 * a / 3.0; }
    RHO(25, 2) = (2144.0 - 16185.0 * eta + 37828.0 * eta2 - 29351.0 * eta3 + 6104.0 * eta4) / (1666.0 * p7);
    if (uneq) { RHO(26, 2) = (-906.0 + 4246.0 * eta - 4963.0 * eta2 + 1380.0 * eta3) / (714.0 * q3); RHO(26, 3) = -2.0 * a / 3.0; }
    // (8,1)..(8,8) -- modes 27..34
    if (uneq) RHO(27, 2) = (20022.0 - 126451.0 * eta + 236922.0 * eta2 - 138430.0 * eta3 + 21640.0 * eta4) / (18240.0 * p6);
    RHO(28, 2) = (2462.0 - 17598.0 * eta + 37119.0 * eta2 - 22845.0 * eta3 + 3063.0 * eta4) / (2736.0 * p7);
    if (uneq) RHO(29, 2) = (20598.0 - 131059.0 * eta + 249018.0 * eta2 - 149950.0 * eta3 + 24520.0 * eta4) / (18240.0 * p6);
    RHO(30, 2) = (2666.0 - 19434.0 * eta + 42627.0 * eta2 - 28965.0 * eta3 + 4899.0 * eta4) / (2736.0 * p7);
    if (uneq) RHO(31, 2) = (4350.0 - 28055.0 * eta + 54642.0 * eta2 - 34598.0 * eta3 + 6056.0 * eta4) / (3648.0 * p6);
    RHO(32, 2) = (1002.0 - 7498.0 * eta + 17269.0 * eta2 - 13055.0 * eta3 + 2653.0 * eta4) / (912.0 * p7);
    if (uneq) RHO(33, 2) = (23478.0 - 154099.0 * eta + 309498.0 * eta2 - 207550.0 * eta3 + 38920 * eta4) / (18240.0 * p6);
    RHO(34, 2) = (3482.0 - 26778.0 * eta + 64659.0 * eta2 - 53445.0 * eta3 + 12243.0 * eta4) / (2736.0 * p7);
#undef RHO
    (void)chiS3; (void)chiA3;

    // flux uses the mass ratio of the flux call; odd m at exactly eta = 1/4 keeps only f_lm
    double eta_f = eta;
    if (eta_f > 0.25 && eta_f < 0.25 + 1e-4) eta_f = 0.25;
    for (int i = 0; i < kNumModes; ++i) {
        ModeCoef& z = mc[i];
        int fl = ((z.l + z.m) & 1) ? 0 : 1;
        if (eta_f == 0.25 && (z.m & 1)) fl |= 2;
        z.flags = fl;
    }
}

// m^2 Omega^2 |h_lm|^2 for one mode (the flux summand before the 1/(8 pi) factor).
//   Heff = (H^2-1)/(2 eta)+1, vPhi2 = (vPhi)^2; the tail factor's exp(-4 pi m H Omega)
//   is evaluated per mode (the reference raises exp(-4 pi H Omega) to the m-th power,
//   the same number within rounding)
// (templated on the scalar so the FP64 operation count can be taken with a counting type)
template <class T>
EOB_HD T flux_mode(const ModeCoef& z, T v, T logv, T H, T Omega, T Heff, T pphi, T vPhi2) {
    const T hk = H * (double)z.m * Omega;
    const T x = 4.0 * kPi * hk;
    const T Em = exp(-x);
    const T hk4 = 4.0 * hk * hk;
    // prod_{i=1..l} (hk4 + i^2) as a polynomial in hk4 (the same degree-8
    // Horner on every lane; all terms are positive, so it loses nothing to
    // cancellation against the reference's running product)
    T prod = z.pp[8];
#pragma unroll
    for (int j = 7; j >= 0; --j) prod = prod * hk4 + z.pp[j];
    const T T2 = qdiv(x, T(1.0) - Em) * prod * z.inv_lfac2;
    const T el = kEulerGamma + z.log2m + logv;
    // Horner over v^12 .. v^1
    T acc = z.c[10];
    acc = acc * v;                                  // v^11 coefficient is zero
    acc = acc * v + (z.c[9] + z.cl10 * el);
    acc = acc * v + z.c[8];
    acc = acc * v + (z.c[7] + z.cl8 * el);
    acc = acc * v + (z.c[6] + z.cl7 * el);
    acc = acc * v + (z.c[5] + z.cl6 * el);
    acc = acc * v + z.c[4];
    acc = acc * v + z.c[3];
    acc = acc * v + z.c[2];
    acc = acc * v + z.c[1];
    acc = acc * v + z.c[0];
    const T rho = 1.0 + acc * v;
    const T aux = v * (z.f1 + v * v * z.f3);
    T rpl = ipow15(rho, z.l);   // l <= 8 and l + eps <= 9
    rpl = (z.flags & 2) ? aux : rpl + aux;
    const T S = (z.flags & 1) ? Heff : v * pphi;
    const T hN2 = z.N2 * ipow15(vPhi2, z.lpe);
    return z.m2 * Omega * Omega * T2 * S * S * rpl * rpl * hN2;
}

}  // namespace eob
