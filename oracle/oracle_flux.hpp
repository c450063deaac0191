// Oracle test infrastructure (CPU only; never shipped, never measured).
//
// Factorized-waveform PN coefficients, the spin-factorized energy flux and the
// radiation-reaction force, restated after the reference:
//   EOBCalcSpinFacWaveformCoefficients   src/EOB.cu:1946-2462 (struct include/EOB.hh:16-361)
//   EOBFluxGetSpinFactorizedWaveform     src/EOB.cu:1706-1941
//   fast_ylm / Newtonian multipole |.|   src/EOB.cu:1680-1701
//   EOBSpinFactorizedFlux                src/EOB.cu:2506-2560
//   RR_force                             src/EOB.cu:2564-2642
// The coefficient values are the published factorized-resummation tables
// (Pan et al. 2011, Taracchini et al. 2012/2014, Cotesta et al. 2018) and must
// equal the reference's.  Coefficients the reference leaves unset (its struct is
// uninitialised device memory) are zero here; the flux never reads them for
// unequal masses, and the (2,2) waveform reads only 22 entries.
#pragma once
#include <cmath>
#include <cstring>

namespace orc {

struct FacCoeffs {
    // (2,2) phase and amplitude
    double delta22vh3, delta22vh6, delta22vh9, delta22v5, delta22v6, delta22v8;
    double rho22v2, rho22v3, rho22v4, rho22v5, rho22v6, rho22v6l, rho22v7, rho22v8, rho22v8l, rho22v10, rho22v10l;
    // (2,1)
    double rho21v1, rho21v2, rho21v3, rho21v4, rho21v5, rho21v6, rho21v6l, rho21v7, rho21v7l, rho21v8, rho21v8l, rho21v10, rho21v10l;
    double f21v1, f21v3, f21v4, f21v5, f21v6, f21v7c;
    // (3,3)
    double rho33v2, rho33v3, rho33v4, rho33v5, rho33v6, rho33v6l, rho33v7, rho33v8, rho33v8l, rho33v10, rho33v10l;
    double f33v3, f33v4, f33v5, f33v6, f33vh6;
    // (3,2)
    double rho32v, rho32v2, rho32v3, rho32v4, rho32v5, rho32v6, rho32v6l, rho32v8, rho32v8l;
    // (3,1)
    double rho31v2, rho31v3, rho31v4, rho31v5, rho31v6, rho31v6l, rho31v7, rho31v8, rho31v8l, f31v3;
    // (4,m)
    double rho44v2, rho44v3, rho44v4, rho44v5, rho44v6, rho44v6l, rho44v8, rho44v8l, rho44v10, rho44v10l;
    double rho43v, rho43v2, rho43v4, rho43v5, rho43v6, rho43v6l, f43v;
    double rho42v2, rho42v3, rho42v4, rho42v5, rho42v6, rho42v6l;
    double rho41v, rho41v2, rho41v4, rho41v5, rho41v6, rho41v6l, f41v;
    // (5,m)
    double rho55v2, rho55v3, rho55v4, rho55v5, rho55v6, rho55v6l, rho55v8, rho55v8l, rho55v10, rho55v10l;
    double f55v3, f55v4, f55v5c;
    double rho54v2, rho54v3, rho54v4;
    double rho53v2, rho53v3, rho53v4, rho53v5;
    double rho52v2, rho52v3, rho52v4;
    double rho51v2, rho51v3, rho51v4, rho51v5;
    // (6,m)
    double rho66v2, rho66v3, rho66v4, rho65v2, rho65v3, rho64v2, rho64v3, rho64v4;
    double rho63v2, rho63v3, rho62v2, rho62v3, rho62v4, rho61v2, rho61v3;
    // (7,m)
    double rho77v2, rho77v3, rho76v2, rho75v2, rho75v3, rho74v2, rho73v2, rho73v3, rho72v2, rho71v2, rho71v3;
    // (8,m)
    double rho88v2, rho87v2, rho86v2, rho85v2, rho84v2, rho83v2, rho82v2, rho81v2;
    // phases delta_lm of the modes other than (2,2), as the waveform reads them
    double delta21vh3, delta21vh6, delta21vh7, delta21vh9, delta21v5, delta21v7;
    double delta33vh3, delta33vh6, delta33vh9, delta33v5;
    double delta32vh3, delta32vh4, delta32vh6, delta32vh9;
    double delta31vh3, delta31vh6, delta31vh7, delta31vh9, delta31v5;
    double delta44vh3, delta44vh6, delta44vh9, delta44v5;
    double delta43vh3, delta43vh4, delta43vh6;
    double delta42vh3, delta42vh6;
    double delta41vh3, delta41vh4, delta41vh6;
    double delta55vh3, delta55vh6, delta55vh9, delta55v5;
    double delta54vh3, delta54vh4, delta53vh3, delta52vh3, delta52vh4, delta51vh3;
    double delta66vh3, delta65vh3, delta64vh3, delta63vh3, delta62vh3, delta61vh3;
    double delta77vh3, delta75vh3, delta73vh3, delta71vh3;
};

// src/EOB.cu:1946-2462
inline void calc_fac_coeffs(FacCoeffs* c, int use_hm, double m1, double m2, double eta, double a, double chiS, double chiA) {
    std::memset(c, 0, sizeof(FacCoeffs));
    const double PI = 3.141592653589793238462643383279502884;
    const double eta2 = eta * eta, eta3 = eta2 * eta;
    double dM2 = 1.0 - 4.0 * eta;
    const double chiA2 = chiA * chiA, chiS2 = chiS * chiS, chiA3 = chiA2 * chiA, chiS3 = chiS2 * chiS;
    if (dM2 < 0.0 && dM2 > -1e-4) dM2 = 0.0;
    double dM = std::sqrt(dM2);
    if (m1 < m2) dM = -dM;
    const double aDelta = 0.0;
    double a2 = a * a, a3 = a2 * a;
    const double m1p3 = -1.0 + 3.0 * eta, m1p3sq = m1p3 * m1p3, m1p3cu = m1p3 * m1p3sq;
    const bool unequal = (dM2 != 0.0);

    // (2,2): spin enters through the Kerr-like a = tplspin
    c->delta22vh3 = 7.0 / 3.0;
    c->delta22vh6 = -4.0 / 3.0 * (dM * chiA + chiS * (1 - 2 * eta)) + (428.0 * PI) / 105.0;
    c->delta22v8 = (20.0 * aDelta) / 63.0;
    c->delta22vh9 = -2203.0 / 81.0 + (1712.0 * PI * PI) / 315.0;
    c->delta22v5 = -24.0 * eta;
    c->delta22v6 = 0.0;
    c->rho22v2 = -43.0 / 42.0 + (55.0 * eta) / 84.0;
    c->rho22v3 = (-2.0 * (chiS + chiA * dM - chiS * eta)) / 3.0;
    c->rho22v4 = -20555.0 / 10584.0 + 0.5 * (chiS + chiA * dM) * (chiS + chiA * dM) - (33025.0 * eta) / 21168.0 + (19583.0 * eta2) / 42336.0;
    c->rho22v5 = (-34.0 / 21.0 + 49.0 * eta / 18.0 + 209.0 * eta2 / 126.0) * chiS + (-34.0 / 21.0 - 19.0 * eta / 42.0) * dM * chiA;
    c->rho22v6 = 1556919113.0 / 122245200.0 + (89.0 * a2) / 252.0 - (48993925.0 * eta) / 9779616.0 - (6292061.0 * eta2) / 3259872.0
                 + (10620745.0 * eta3) / 39118464.0 + (41.0 * eta * PI * PI) / 192.0;
    c->rho22v6l = -428.0 / 105.0;
    c->rho22v7 = a3 / 3.0 + chiA * dM * (18733.0 / 15876.0 + (50140.0 * eta) / 3969.0 + (97865.0 * eta2) / 63504.0)
                 + chiS * (18733.0 / 15876.0 + (74749.0 * eta) / 5292.0 - (245717.0 * eta2) / 63504.0 + (50803.0 * eta3) / 63504.0);
    c->rho22v8 = -387216563023.0 / 160190110080.0 + (18353.0 * a2) / 21168.0 - a2 * a2 / 8.0;
    c->rho22v8l = 9202.0 / 2205.0;
    c->rho22v10 = -16094530514677.0 / 533967033600.0;
    c->rho22v10l = 439877.0 / 55566.0;

    if (use_hm) { a = 0.; a2 = 0.; a3 = 0.; }

    // (2,1)
    if (unequal) {
        c->rho21v1 = 0.0;
        c->rho21v2 = -59.0 / 56 + (23.0 * eta) / 84.0;
        c->rho21v3 = 0.0;
        c->rho21v4 = -47009.0 / 56448.0 - (865.0 * a2) / 1792.0 - (405.0 * a2 * a2) / 2048.0 - (10993.0 * eta) / 14112.0 + (617.0 * eta2) / 4704.0;
        c->rho21v5 = (-98635.0 * a) / 75264.0 + (2031.0 * a * a2) / 7168.0 - (1701.0 * a2 * a3) / 8192.0;
        c->rho21v6 = 7613184941.0 / 2607897600.0 + (9032393.0 * a2) / 1806336.0 + (3897.0 * a2 * a2) / 16384.0 - (15309.0 * a3 * a3) / 65536.0;
        c->rho21v6l = -107.0 / 105.0;
        c->rho21v7 = (-3859374457.0 * a) / 1159065600.0 - (55169.0 * a3) / 16384.0 + (18603.0 * a2 * a3) / 65536.0 - (72171.0 * a2 * a2 * a3) / 262144.0;
        c->rho21v7l = 107.0 * a / 140.0;
        c->rho21v8 = -1168617463883.0 / 911303737344.0;
        c->rho21v8l = 6313.0 / 5880.0;
        c->rho21v10 = -63735873771463.0 / 16569158860800.0;
        c->rho21v10l = 5029963.0 / 5927040.0;
        c->f21v1 = (-3.0 * (chiS + chiA / dM)) / 2.0;
        c->f21v3 = (chiS * dM * (427.0 + 79.0 * eta) + chiA * (147.0 + 280.0 * dM * dM + 1251.0 * eta)) / 84.0 / dM;
        if (use_hm) {
            c->f21v4 = (-3.0 - 2.0 * eta) * chiA2 + (-3.0 + eta / 2.0) * chiS2 + (-6.0 + 21.0 * eta / 2.0) * chiS * chiA / dM;
            c->f21v5 = (3.0 / 4.0 - 3.0 * eta) * chiA3 / dM
                       + (-81.0 / 16.0 + 1709.0 * eta / 1008.0 + 613.0 * eta2 / 1008.0 + (9.0 / 4.0 - 3 * eta) * chiA2) * chiS + 3.0 / 4.0 * chiS3
                       + (-81.0 / 16.0 - 703.0 * eta2 / 112.0 + 8797.0 * eta / 1008.0 + (9.0 / 4.0 - 6.0 * eta) * chiS2) * chiA / dM;
            c->f21v6 = (4163.0 / 252.0 - 9287.0 * eta / 1008.0 - 85.0 * eta2 / 112.0) * chiA2
                       + (4163.0 / 252.0 - 2633.0 * eta / 1008.0 + 461.0 * eta2 / 1008.0) * chiS2
                       + (4163.0 / 126.0 - 1636.0 * eta / 21.0 + 1088.0 * eta2 / 63.0) * chiS * chiA / dM;
        }
    } else {
        c->f21v1 = -3.0 * chiA / 2.0;
        c->f21v3 = (chiS * dM * (427.0 + 79.0 * eta) + chiA * (147.0 + 280.0 * dM * dM + 1251.0 * eta)) / 84.0;
        if (use_hm) {
            c->f21v4 = (-6 + 21 * eta / 2.0) * chiS * chiA;
            c->f21v5 = (3.0 / 4.0 - 3.0 * eta) * chiA3 + (-81.0 / 16.0 - 703.0 * eta2 / 112.0 + 8797.0 * eta / 1008.0 + (9.0 / 4.0 - 6.0 * eta) * chiS2) * chiA;
            c->f21v6 = (4163.0 / 126.0 - 1636.0 * eta / 21.0 + 1088.0 * eta2 / 63.0) * chiS * chiA;
        }
    }

    // (3,3)
    if (unequal) {
        c->rho33v2 = -7.0 / 6.0 + (2.0 * eta) / 3.0;
        c->rho33v3 = 0.0;
        c->rho33v4 = -6719.0 / 3960.0 + a2 / 2.0 - (1861.0 * eta) / 990.0 + (149.0 * eta2) / 330.0;
        c->rho33v5 = (-4.0 * a) / 3.0;
        c->rho33v6 = 3203101567.0 / 227026800.0 + (5.0 * a2) / 36.0;
        c->rho33v6l = -26.0 / 7.0;
        c->rho33v7 = (5297.0 * a) / 2970.0 + a * a2 / 3.0;
        c->rho33v8 = -57566572157.0 / 8562153600.0;
        c->rho33v8l = 13.0 / 3.0;
        if (use_hm) {
            c->rho33v6 = 3203101567.0 / 227026800.0 + (5.0 * a2) / 36.0 + (-129509.0 / 25740.0 + 41.0 / 192.0 * PI * PI) * eta
                         - 274621.0 / 154440.0 * eta2 + 12011.0 / 46332.0 * eta3;
            c->rho33v10 = -903823148417327.0 / 30566888352000.0;
            c->rho33v10l = 87347.0 / 13860.0;
        }
        c->f33v3 = (chiS * dM * (-4.0 + 5.0 * eta) + chiA * (-4.0 + 19.0 * eta)) / (2.0 * dM);
        if (use_hm) {
            c->f33v4 = (3.0 / 2.0 * chiS2 * dM + (3.0 - 12 * eta) * chiA * chiS + dM * (3.0 / 2.0 - 6.0 * eta) * chiA2) / dM;
            c->f33v5 = (dM * (241.0 / 30.0 * eta2 + 11.0 / 20.0 * eta + 2.0 / 3.0) * chiS + (407.0 / 30.0 * eta2 - 593.0 / 60.0 * eta + 2.0 / 3.0) * chiA) / dM;
            c->f33v6 = (dM * (6.0 * eta2 - 27.0 / 2.0 * eta - 7.0 / 4.0) * chiS2 + (44.0 * eta2 - 1.0 * eta - 7.0 / 2.0) * chiA * chiS
                        + dM * (-12 * eta2 + 11.0 / 2.0 * eta - 7.0 / 4.0) * chiA2) / dM;
            c->f33vh6 = (dM * (593.0 / 108.0 * eta - 81.0 / 20.0) * chiS + (7339.0 / 540.0 * eta - 81.0 / 20.0) * chiA) / dM;
        }
    } else {
        c->f33v3 = chiA * 3.0 / 8.0;
        if (use_hm) {
            c->f33v4 = (3.0 - 12 * eta) * chiA * chiS;
            c->f33v5 = (407.0 / 30.0 * eta2 - 593.0 / 60.0 * eta + 2.0 / 3.0) * chiA;
            c->f33v6 = (44.0 * eta2 - 1.0 * eta - 7.0 / 2.0) * chiA * chiS;
            c->f33vh6 = (7339.0 / 540.0 * eta - 81.0 / 20.0) * chiA;
        }
    }

    // (3,2): the first rho32v3 assignment in the reference is overwritten by 2a/9
    c->rho32v = (4.0 * chiS * eta) / (-3.0 * m1p3);
    c->rho32v2 = (328.0 - 1115.0 * eta + 320.0 * eta2) / (270.0 * m1p3);
    (void)m1p3cu;
    c->rho32v3 = 2.0 / 9.0 * a;
    c->rho32v4 = a2 / 3.0 + (-1444528.0 + 8050045.0 * eta - 4725605.0 * eta2 - 20338960.0 * eta3 + 3085640.0 * eta2 * eta2) / (1603800.0 * m1p3sq);
    c->rho32v5 = (-2788.0 * a) / 1215.0;
    c->rho32v6 = 5849948554.0 / 940355325.0 + (488.0 * a2) / 405.0;
    c->rho32v6l = -104.0 / 63.0;
    c->rho32v8 = -10607269449358.0 / 3072140846775.0;
    c->rho32v8l = 17056.0 / 8505.0;

    // (3,1)
    if (unequal) {
        c->rho31v2 = -13.0 / 18.0 - (2.0 * eta) / 9.0;
        c->rho31v3 = 0.0;
        c->rho31v4 = 101.0 / 7128.0 - (5.0 * a2) / 6.0 - (1685.0 * eta) / 1782.0 - (829.0 * eta2) / 1782.0;
        c->rho31v5 = (4.0 * a) / 9.0;
        c->rho31v6 = 11706720301.0 / 6129723600.0 - (49.0 * a2) / 108.0;
        c->rho31v6l = -26.0 / 63.0;
        c->rho31v7 = (-2579.0 * a) / 5346.0 + a * a2 / 9.0;
        c->rho31v8 = 2606097992581.0 / 4854741091200.0;
        c->rho31v8l = 169.0 / 567.0;
        c->f31v3 = (chiA * (-4.0 + 11.0 * eta) + chiS * dM * (-4.0 + 13.0 * eta)) / (2.0 * dM);
    } else {
        c->f31v3 = -chiA * 5.0 / 8.0;
    }

    // (4,4)
    c->rho44v2 = (1614.0 - 5870.0 * eta + 2625.0 * eta2) / (1320.0 * m1p3);
    c->rho44v3 = (chiA * (10.0 - 39.0 * eta) * dM + chiS * (10.0 - 41.0 * eta + 42.0 * eta2)) / (15.0 * m1p3);
    c->rho44v4 = a2 / 2.0 + (-511573572.0 + 2338945704.0 * eta - 313857376.0 * eta2 - 6733146000.0 * eta3 + 1252563795.0 * eta2 * eta2) / (317116800.0 * m1p3sq);
    c->rho44v5 = (-69.0 * a) / 55.0;
    if (use_hm) {
        c->rho44v4 = (-511573572.0 + 2338945704.0 * eta - 313857376.0 * eta2 - 6733146000.0 * eta3 + 1252563795.0 * eta2 * eta2) / (317116800.0 * m1p3sq)
                     + chiS2 / 2.0 + dM * chiS * chiA + dM2 * chiA2 / 2.0;
        c->rho44v5 = chiA * dM * (-8280.0 + 42716.0 * eta - 57990.0 * eta2 + 8955 * eta3) / (6600.0 * m1p3sq)
                     + chiS * (-8280.0 + 66284.0 * eta - 176418.0 * eta2 + 128085.0 * eta3 + 88650 * eta2 * eta2) / (6600.0 * m1p3sq);
        c->rho44v8 = -172066910136202271.0 / 19426955708160000.0;
        c->rho44v8l = 845198.0 / 190575.0;
        c->rho44v10 = -17154485653213713419357.0 / 568432724020761600000.0;
        c->rho44v10l = 22324502267.0 / 3815311500.0;
    }
    c->rho44v6 = 16600939332793.0 / 1098809712000.0 + (217.0 * a2) / 3960.0;
    c->rho44v6l = -12568.0 / 3465.0;

    // (4,3)
    if (unequal) {
        c->rho43v = 0.0;
        c->rho43v2 = (222.0 - 547.0 * eta + 160.0 * eta2) / (176.0 * (-1.0 + 2.0 * eta));
        c->rho43v4 = -6894273.0 / 7047040.0 + (3.0 * a2) / 8.0;
        c->rho43v5 = (-12113.0 * a) / 6160.0;
        c->rho43v6 = 1664224207351.0 / 195343948800.0;
        c->rho43v6l = -1571.0 / 770.0;
        c->f43v = (5.0 * (chiA - chiS * dM) * eta) / (2.0 * dM * (-1.0 + 2.0 * eta));
    } else {
        c->f43v = -5.0 * chiA / 4.0;
    }

    // (4,2)
    c->rho42v2 = (1146.0 - 3530.0 * eta + 285.0 * eta2) / (1320.0 * m1p3);
    c->rho42v3 = (chiA * (10.0 - 21.0 * eta) * dM + chiS * (10.0 - 59.0 * eta + 78.0 * eta2)) / (15.0 * m1p3);
    c->rho42v4 = a2 / 2.0 + (-114859044.0 + 295834536.0 * eta + 1204388696.0 * eta2 - 3047981160.0 * eta3 - 379526805.0 * eta2 * eta2) / (317116800.0 * m1p3sq);
    c->rho42v5 = (-7.0 * a) / 110.0;
    c->rho42v6 = 848238724511.0 / 219761942400.0 + (2323.0 * a2) / 3960.0;
    c->rho42v6l = -3142.0 / 3465.0;

    // (4,1)
    if (unequal) {
        c->rho41v = 0.0;
        c->rho41v2 = (602.0 - 1385.0 * eta + 288.0 * eta2) / (528.0 * (-1.0 + 2.0 * eta));
        c->rho41v4 = -7775491.0 / 21141120.0 + (3.0 * a2) / 8.0;
        c->rho41v5 = (-20033.0 * a) / 55440.0 - (5 * a * a2) / 6.0;
        c->rho41v6 = 1227423222031.0 / 1758095539200.0;
        c->rho41v6l = -1571.0 / 6930.0;
        c->f41v = (5.0 * (chiA - chiS * dM) * eta) / (2.0 * dM * (-1.0 + 2.0 * eta));
    } else {
        c->f41v = -5.0 * chiA / 4.0;
    }

    // (5,5)
    if (unequal) {
        c->rho55v2 = (487.0 - 1298.0 * eta + 512.0 * eta2) / (390.0 * (-1.0 + 2.0 * eta));
        c->rho55v3 = (-2.0 * a) / 3.0;
        c->rho55v4 = -3353747.0 / 2129400.0 + a2 / 2.0;
        c->rho55v5 = -241.0 * a / 195.0;
        if (use_hm) {
            c->rho55v6 = 190606537999247.0 / 11957879934000.0;
            c->rho55v6l = -1546.0 / 429.0;
            c->rho55v8 = -1213641959949291437.0 / 118143853747920000.0;
            c->rho55v8l = 376451.0 / 83655.0;
            c->rho55v10 = -150082616449726042201261.0 / 4837990810977324000000.0;
            c->rho55v10l = 2592446431.0 / 456756300.0;
            c->f55v3 = chiA / dM * (10.0 / (3.0 * (-1.0 + 2.0 * eta)) - 70.0 * eta / (3.0 * (-1.0 + 2.0 * eta)) + 110.0 * eta2 / (3.0 * (-1.0 + 2.0 * eta)))
                       + chiS * (10.0 / (3.0 * (-1.0 + 2.0 * eta)) - 10.0 * eta / (-1.0 + 2.0 * eta) + 10 * eta2 / (-1.0 + 2.0 * eta));
            c->f55v4 = chiS2 * (-5.0 / (2.0 * (-1.0 + 2.0 * eta)) + 5.0 * eta / (-1.0 + 2.0 * eta))
                       + chiA * chiS / dM * (-5.0 / (-1.0 + 2.0 * eta) + 30.0 * eta / (-1.0 + 2.0 * eta) - 40.0 * eta2 / (-1.0 + 2.0 * eta))
                       + chiA2 * (-5.0 / (2.0 * (-1.0 + 2.0 * eta)) + 15.0 * eta / (-1.0 + 2.0 * eta) - 20.0 * eta2 / (-1.0 + 2.0 * eta));
        }
    } else if (use_hm) {
        c->f55v3 = chiA * (10.0 / (3.0 * (-1.0 + 2.0 * eta)) - 70.0 * eta / (3.0 * (-1.0 + 2.0 * eta)) + 110.0 * eta2 / (3.0 * (-1.0 + 2.0 * eta)));
        c->f55v4 = chiA * chiS * (-5.0 / (-1.0 + 2.0 * eta) + 30.0 * eta / (-1.0 + 2.0 * eta) - 40.0 * eta2 / (-1.0 + 2.0 * eta));
    }

    // (5,4) .. (5,1)
    c->rho54v2 = (-17448.0 + 96019.0 * eta - 127610.0 * eta2 + 33320.0 * eta3) / (13650.0 * (1.0 - 5.0 * eta + 5.0 * eta2));
    c->rho54v3 = (-2.0 * a) / 15.0;
    c->rho54v4 = -16213384.0 / 15526875.0 + (2.0 * a2) / 5.0;
    if (unequal) {
        c->rho53v2 = (375.0 - 850.0 * eta + 176.0 * eta2) / (390.0 * (-1.0 + 2.0 * eta));
        c->rho53v3 = (-2.0 * a) / 3.0;
        c->rho53v4 = -410833.0 / 709800.0 + a2 / 2.0;
        c->rho53v5 = -103.0 * a / 325.0;
    }
    c->rho52v2 = (-15828.0 + 84679.0 * eta - 104930.0 * eta2 + 21980.0 * eta3) / (13650.0 * (1.0 - 5.0 * eta + 5.0 * eta2));
    c->rho52v3 = (-2.0 * a) / 15.0;
    c->rho52v4 = -7187914.0 / 15526875.0 + (2.0 * a2) / 5.0;
    if (unequal) {
        c->rho51v2 = (319.0 - 626.0 * eta + 8.0 * eta2) / (390.0 * (-1.0 + 2.0 * eta));
        c->rho51v3 = (-2.0 * a) / 3.0;
        c->rho51v4 = -31877.0 / 304200.0 + a2 / 2.0;
        c->rho51v5 = 139.0 * a / 975.0;
    }

    // (6,m)
    c->rho66v2 = (-106.0 + 602.0 * eta - 861.0 * eta2 + 273.0 * eta3) / (84.0 * (1.0 - 5.0 * eta + 5.0 * eta2));
    c->rho66v3 = (-2.0 * a) / 3.0;
    c->rho66v4 = -1025435.0 / 659736.0 + a2 / 2.0;
    if (unequal) {
        c->rho65v2 = (-185.0 + 838.0 * eta - 910.0 * eta2 + 220.0 * eta3) / (144.0 * (dM2 + 3.0 * eta2));
        c->rho65v3 = -2.0 * a / 9.0;
    }
    c->rho64v2 = (-86.0 + 462.0 * eta - 581.0 * eta2 + 133.0 * eta3) / (84.0 * (1.0 - 5.0 * eta + 5.0 * eta2));
    c->rho64v3 = (-2.0 * a) / 3.0;
    c->rho64v4 = -476887.0 / 659736.0 + a2 / 2.0;
    if (unequal) {
        c->rho63v2 = (-169.0 + 742.0 * eta - 750.0 * eta2 + 156.0 * eta3) / (144.0 * (dM2 + 3.0 * eta2));
        c->rho63v3 = -2.0 * a / 9.0;
    }
    c->rho62v2 = (-74.0 + 378.0 * eta - 413.0 * eta2 + 49.0 * eta3) / (84.0 * (1.0 - 5.0 * eta + 5.0 * eta2));
    c->rho62v3 = (-2.0 * a) / 3.0;
    c->rho62v4 = -817991.0 / 3298680.0 + a2 / 2.0;
    if (unequal) {
        c->rho61v2 = (-161.0 + 694.0 * eta - 670.0 * eta2 + 124.0 * eta3) / (144.0 * (dM2 + 3.0 * eta2));
        c->rho61v3 = -2.0 * a / 9.0;
    }

    // (7,m)
    if (unequal) {
        c->rho77v2 = (-906.0 + 4246.0 * eta - 4963.0 * eta2 + 1380.0 * eta3) / (714.0 * (dM2 + 3.0 * eta2));
        c->rho77v3 = -2.0 * a / 3.0;
    }
    c->rho76v2 = (2144.0 - 16185.0 * eta + 37828.0 * eta2 - 29351.0 * eta3 + 6104.0 * eta2 * eta2) / (1666.0 * (-1 + 7 * eta - 14 * eta2 + 7 * eta3));
    if (unequal) {
        c->rho75v2 = (-762.0 + 3382.0 * eta - 3523.0 * eta2 + 804.0 * eta3) / (714.0 * (dM2 + 3.0 * eta2));
        c->rho75v3 = -2.0 * a / 3.0;
    }
    c->rho74v2 = (17756.0 - 131805.0 * eta + 298872.0 * eta2 - 217959.0 * eta3 + 41076.0 * eta2 * eta2) / (14994.0 * (-1.0 + 7.0 * eta - 14.0 * eta2 + 7.0 * eta3));
    if (unequal) {
        c->rho73v2 = (-666.0 + 2806.0 * eta - 2563.0 * eta2 + 420.0 * eta3) / (714.0 * (dM2 + 3.0 * eta2));
        c->rho73v3 = -2.0 * a / 3.0;
    }
    c->rho72v2 = (16832.0 - 123489.0 * eta + 273924.0 * eta2 - 190239.0 * eta3 + 32760.0 * eta2 * eta2) / (14994.0 * (-1.0 + 7.0 * eta - 14.0 * eta2 + 7.0 * eta3));
    if (unequal) {
        c->rho71v2 = (-618.0 + 2518.0 * eta - 2083.0 * eta2 + 228.0 * eta3) / (714.0 * (dM2 + 3.0 * eta2));
        c->rho71v3 = -2.0 * a / 3.0;
    }

    // (8,m)
    c->rho88v2 = (3482.0 - 26778.0 * eta + 64659.0 * eta2 - 53445.0 * eta3 + 12243.0 * eta2 * eta2) / (2736.0 * (-1.0 + 7.0 * eta - 14.0 * eta2 + 7.0 * eta3));
    if (unequal) c->rho87v2 = (23478.0 - 154099.0 * eta + 309498.0 * eta2 - 207550.0 * eta3 + 38920 * eta2 * eta2) / (18240.0 * (-1 + 6 * eta - 10 * eta2 + 4 * eta3));
    c->rho86v2 = (1002.0 - 7498.0 * eta + 17269.0 * eta2 - 13055.0 * eta3 + 2653.0 * eta2 * eta2) / (912.0 * (-1.0 + 7.0 * eta - 14.0 * eta2 + 7.0 * eta3));
    if (unequal) c->rho85v2 = (4350.0 - 28055.0 * eta + 54642.0 * eta2 - 34598.0 * eta3 + 6056.0 * eta2 * eta2) / (3648.0 * (-1.0 + 6.0 * eta - 10.0 * eta2 + 4.0 * eta3));
    c->rho84v2 = (2666.0 - 19434.0 * eta + 42627.0 * eta2 - 28965.0 * eta3 + 4899.0 * eta2 * eta2) / (2736.0 * (-1.0 + 7.0 * eta - 14.0 * eta2 + 7.0 * eta3));
    if (unequal) c->rho83v2 = (20598.0 - 131059.0 * eta + 249018.0 * eta2 - 149950.0 * eta3 + 24520.0 * eta2 * eta2) / (18240.0 * (-1.0 + 6.0 * eta - 10.0 * eta2 + 4.0 * eta3));
    c->rho82v2 = (2462.0 - 17598.0 * eta + 37119.0 * eta2 - 22845.0 * eta3 + 3063.0 * eta2 * eta2) / (2736.0 * (-1.0 + 7.0 * eta - 14.0 * eta2 + 7.0 * eta3));
    if (unequal) c->rho81v2 = (20022.0 - 126451.0 * eta + 236922.0 * eta2 - 138430.0 * eta3 + 21640.0 * eta2 * eta2) / (18240.0 * (-1.0 + 6.0 * eta - 10.0 * eta2 + 4.0 * eta3));

    // phases delta_lm of (2,1), (3,3), (4,4), (5,5) (src/EOB.cu:2019-2023, 2080-2083,
    // 2178-2185, 2263-2271); aDelta = 0 as in the reference.  delta21v7, delta44v5
    // and delta55v5 are read by the waveform but never set by the reference (its
    // record is uninitialised memory); they are 0 here.
    c->delta21vh3 = 2.0 / 3.0;
    c->delta21vh6 = (107.0 * PI) / 105.0;
    c->delta21vh7 = 0.0;
    c->delta21vh9 = -272.0 / 81.0 + (214.0 * PI * PI) / 315.0;
    c->delta21v5 = -493.0 * eta / 42.0;
    c->delta33vh3 = 13.0 / 10.0;
    c->delta33vh6 = (39.0 * PI) / 7.0;
    c->delta33vh9 = -227827.0 / 3000.0 + (78.0 * PI * PI) / 7.0;
    c->delta33v5 = -80897.0 * eta / 2430.0;
    c->delta44vh3 = (112.0 + 219.0 * eta) / (-120.0 * m1p3);
    c->delta44vh6 = (25136.0 * PI) / 3465.0;
    c->delta44vh9 = use_hm ? -55144.0 / 375.0 + 201088.0 * PI * PI / 10395.0 : 0.0;
    c->delta55vh3 = (96875.0 + 857528.0 * eta) / (131250.0 * (1 - 2 * eta));
    c->delta55vh6 = use_hm ? 3865.0 / 429.0 * PI : 0.0;
    c->delta55vh9 = use_hm ? (-7686949127.0 + 954500400.0 * PI * PI) / 31783752.0 : 0.0;
    // the other modes' phases (src/EOB.cu:2134-2137, 2150-2154, 2213-2215, 2231-2232,
    // 2243-2245, 2317-2341, 2353-2424), aDelta = 0
    c->delta32vh3 = (10.0 + 33.0 * eta) / (-15.0 * m1p3);
    c->delta32vh4 = 4.0 * aDelta;
    c->delta32vh6 = (52.0 * PI) / 21.0;
    c->delta32vh9 = -9112.0 / 405.0 + (208.0 * PI * PI) / 63.0;
    c->delta31vh3 = 13.0 / 30.0;
    c->delta31vh6 = (13.0 * PI) / 21.0;
    c->delta31vh7 = 0.0;
    c->delta31vh9 = -227827.0 / 81000.0 + (26.0 * PI * PI) / 63.0;
    c->delta31v5 = -17.0 * eta / 10.0;
    c->delta43vh3 = (486.0 + 4961.0 * eta) / (810.0 * (1.0 - 2.0 * eta));
    c->delta43vh4 = 0.0;
    c->delta43vh6 = 1571.0 * PI / 385.0;
    c->delta42vh3 = (7.0 * (1.0 + 6.0 * eta)) / (-15.0 * m1p3);
    c->delta42vh6 = (6284.0 * PI) / 3465.0;
    c->delta41vh3 = (2.0 + 507.0 * eta) / (10.0 * (1.0 - 2.0 * eta));
    c->delta41vh4 = 0.0;
    c->delta41vh6 = 1571.0 * PI / 3465.0;
    c->delta54vh3 = 8.0 / 15.0;
    c->delta54vh4 = 0.0;
    c->delta53vh3 = 31.0 / 70.0;
    c->delta52vh3 = 4.0 / 15.0;
    c->delta52vh4 = 0.0;
    c->delta51vh3 = 31.0 / 210.0;
    c->delta66vh3 = 43.0 / 70.0;
    c->delta65vh3 = 10.0 / 21.0;
    c->delta64vh3 = 43.0 / 105.0;
    c->delta63vh3 = 2.0 / 7.0;
    c->delta62vh3 = 43.0 / 210.0;
    c->delta61vh3 = 2.0 / 21.0;
    c->delta77vh3 = 19.0 / 36.0;
    c->delta75vh3 = 95.0 / 252.0;
    c->delta73vh3 = 19.0 / 84.0;
    c->delta71vh3 = 19.0 / 252.0;
}

// |Y_{l}^{-m}(theta=pi/2, phi=0)| for 0<=m<=l<=8 from the closed form
// sqrt((2l+1)/(4pi) (l-m)!/(l+m)!) |P_l^m(0)|, |P_l^m(0)| = (l+m-1)!!/(l-m)!! for
// l+m even and 0 otherwise.  Replaces the numeric table at src/EOB.cu:1680-1684.
inline double ylm_abs_equator(int l, int m) {
    if ((l + m) % 2) return 0.0;
    double fact_lmm = 1.0, fact_lpm = 1.0;
    for (int i = 2; i <= l - m; ++i) fact_lmm *= i;
    for (int i = 2; i <= l + m; ++i) fact_lpm *= i;
    double dfa = 1.0, dfb = 1.0;
    for (int i = l + m - 1; i > 1; i -= 2) dfa *= i;
    for (int i = l - m; i > 1; i -= 2) dfb *= i;
    const double PI = 3.141592653589793238462643383279502884;
    return std::sqrt((2.0 * l + 1.0) / (4.0 * PI) * fact_lmm / fact_lpm) * dfa / dfb;
}

// rho_lm(v) exactly as nested in src/EOB.cu:1757-1930 (flux variant)
inline double rho_lm_flux(const FacCoeffs& h, int l, int m, double v, double v2, double elog, double h22_calib, double* auxflm) {
    double rho = 0.0;
    *auxflm = 0.0;
    switch (l * 10 + m) {
    case 22: rho = 1.0 + v2 * (h.rho22v2 + v * (h.rho22v3 + v * (h.rho22v4 + v * (h.rho22v5 + v * (h.rho22v6 + h.rho22v6l * elog + v * (h.rho22v7 + v * (h.rho22v8 + h.rho22v8l * elog + (h.rho22v10 + h.rho22v10l * elog) * v2))))))) + h22_calib * std::pow(v, 12); break;
    case 21: rho = 1.0 + v * (h.rho21v1 + v * (h.rho21v2 + v * (h.rho21v3 + v * (h.rho21v4 + v * (h.rho21v5 + v * (h.rho21v6 + h.rho21v6l * elog + v * (h.rho21v7 + h.rho21v7l * elog + v * (h.rho21v8 + h.rho21v8l * elog + (h.rho21v10 + h.rho21v10l * elog) * v2))))))));
             *auxflm = v * h.f21v1 + v2 * v * h.f21v3; break;
    case 33: rho = 1.0 + v2 * (h.rho33v2 + v * (h.rho33v3 + v * (h.rho33v4 + v * (h.rho33v5 + v * (h.rho33v6 + h.rho33v6l * elog + v * (h.rho33v7 + (h.rho33v8 + h.rho33v8l * elog) * v))))));
             *auxflm = v * v2 * h.f33v3; break;
    case 32: rho = 1.0 + v * (h.rho32v + v * (h.rho32v2 + v * (h.rho32v3 + v * (h.rho32v4 + v * (h.rho32v5 + v * (h.rho32v6 + h.rho32v6l * elog + (h.rho32v8 + h.rho32v8l * elog) * v2)))))); break;
    case 31: rho = 1.0 + v2 * (h.rho31v2 + v * (h.rho31v3 + v * (h.rho31v4 + v * (h.rho31v5 + v * (h.rho31v6 + h.rho31v6l * elog + v * (h.rho31v7 + (h.rho31v8 + h.rho31v8l * elog) * v))))));
             *auxflm = v * v2 * h.f31v3; break;
    case 44: rho = 1.0 + v2 * (h.rho44v2 + v * (h.rho44v3 + v * (h.rho44v4 + v * (h.rho44v5 + (h.rho44v6 + h.rho44v6l * elog) * v)))); break;
    case 43: rho = 1.0 + v * (h.rho43v + v * (h.rho43v2 + v2 * (h.rho43v4 + v * (h.rho43v5 + (h.rho43v6 + h.rho43v6l * elog) * v))));
             *auxflm = v * h.f43v; break;
    case 42: rho = 1.0 + v2 * (h.rho42v2 + v * (h.rho42v3 + v * (h.rho42v4 + v * (h.rho42v5 + (h.rho42v6 + h.rho42v6l * elog) * v)))); break;
    case 41: rho = 1.0 + v * (h.rho41v + v * (h.rho41v2 + v2 * (h.rho41v4 + v * (h.rho41v5 + (h.rho41v6 + h.rho41v6l * elog) * v))));
             *auxflm = v * h.f41v; break;
    case 55: rho = 1.0 + v2 * (h.rho55v2 + v * (h.rho55v3 + v * (h.rho55v4 + v * (h.rho55v5 + h.rho55v6 * v)))); break;
    case 54: rho = 1.0 + v2 * (h.rho54v2 + v * (h.rho54v3 + h.rho54v4 * v)); break;
    case 53: rho = 1.0 + v2 * (h.rho53v2 + v * (h.rho53v3 + v * (h.rho53v4 + h.rho53v5 * v))); break;
    case 52: rho = 1.0 + v2 * (h.rho52v2 + v * (h.rho52v3 + h.rho52v4 * v)); break;
    case 51: rho = 1.0 + v2 * (h.rho51v2 + v * (h.rho51v3 + v * (h.rho51v4 + h.rho51v5 * v))); break;
    case 66: rho = 1.0 + v2 * (h.rho66v2 + v * (h.rho66v3 + h.rho66v4 * v)); break;
    case 65: rho = 1.0 + v2 * (h.rho65v2 + h.rho65v3 * v); break;
    case 64: rho = 1.0 + v2 * (h.rho64v2 + v * (h.rho64v3 + h.rho64v4 * v)); break;
    case 63: rho = 1.0 + v2 * (h.rho63v2 + h.rho63v3 * v); break;
    case 62: rho = 1.0 + v2 * (h.rho62v2 + v * (h.rho62v3 + h.rho62v4 * v)); break;
    case 61: rho = 1.0 + v2 * (h.rho61v2 + h.rho61v3 * v); break;
    case 77: rho = 1.0 + v2 * (h.rho77v2 + h.rho77v3 * v); break;
    case 76: rho = 1.0 + h.rho76v2 * v2; break;
    case 75: rho = 1.0 + v2 * (h.rho75v2 + h.rho75v3 * v); break;
    case 74: rho = 1.0 + h.rho74v2 * v2; break;
    case 73: rho = 1.0 + v2 * (h.rho73v2 + h.rho73v3 * v); break;
    case 72: rho = 1.0 + h.rho72v2 * v2; break;
    case 71: rho = 1.0 + v2 * (h.rho71v2 + h.rho71v3 * v); break;
    case 88: rho = 1.0 + h.rho88v2 * v2; break;
    case 87: rho = 1.0 + h.rho87v2 * v2; break;
    case 86: rho = 1.0 + h.rho86v2 * v2; break;
    case 85: rho = 1.0 + h.rho85v2 * v2; break;
    case 84: rho = 1.0 + h.rho84v2 * v2; break;
    case 83: rho = 1.0 + h.rho83v2 * v2; break;
    case 82: rho = 1.0 + h.rho82v2 * v2; break;
    case 81: rho = 1.0 + h.rho81v2 * v2; break;
    }
    (void)v2;
    return rho;
}

// |h_lm|^2 of one flux mode, src/EOB.cu:1706-1941 (only the modulus enters the flux)
inline double flux_mode_abs2(const FacCoeffs& h, int l, int m, double v, double Hreal, double pp, double prefix_abs,
                             double eta, double vPhi2, double Omega, double v2, double h22_calib) {
    const double PI = 3.141592653589793238462643383279502884;
    const double eg = 0.57721566490153286060651209008240243104215933593992;
    if (eta > 0.25 && eta < 0.25 + 1e-4) eta = 0.25;
    const double elog = eg + std::log(2.0 * m * v);
    const int eps = (l + m) % 2;
    const double hN = prefix_abs * std::pow(vPhi2, (l + eps) / 2.0) * ylm_abs_equator(l - eps, m);
    double Slm;
    if (((l + m) % 2) == 0) Slm = (Hreal * Hreal - 1.0) / (2.0 * eta) + 1.0;
    else Slm = v * pp;
    const double k = m * Omega;
    const double hk = Hreal * k;
    const double hksq4 = 4.0 * hk * hk;
    const double hk4pi = 4.0 * PI * hk;
    const double z2 = std::tgamma((double)(l + 1));
    const double pref = std::sqrt(hk4pi / (1.0 - std::exp(-hk4pi))) / z2;
    double prod = 1.0;
    for (int i = 1; i < l + 1; ++i) prod *= hksq4 + std::pow(i, 2);
    const double Tlm = pref * std::sqrt(prod);
    double auxflm;
    const double rho = rho_lm_flux(h, l, m, v, v2, elog, h22_calib, &auxflm);
    double rpl = std::pow(rho, l);
    if (eta == 0.25 && (m % 2)) rpl = auxflm;
    else rpl += auxflm;
    const double habs = Tlm * Slm * rpl * hN;
    return habs * habs;
}

}  // namespace orc
