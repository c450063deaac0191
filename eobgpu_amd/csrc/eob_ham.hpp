// Aligned-spin EOB Hamiltonian in per-binary hoisted form.
//
// The reference evaluates the whole expression, including every polynomial in
// nu, chi1, chi2, K, d5, dSO, dSS, at every call (src/EOB.cu:17-107, and again
// inside each derivative expression, :196-209).  Here all binary-only factors
// are folded once per binary into HamCoef (ham_setup), and the per-call work is
// the r / pr / pphi dependence only:
//
//   u = 1/r, lr = ln r, su = sqrt(u)          (half-integer powers of u)
//   A      = (AK0 + AKa) (1 - e1 u + e2 u^2) (1 + as4 u^4 + as5 u^5 + sc6 u^6) / Dn
//   Dn     = 1 + ap2 (u^2 + 2 u^3)            (shared by A, Bnpa and the Hodd denominator)
//   Bnp    = AK0 Dbpm (1 - e1 u) + bs3 u^3 + bs4 u^4 - 1 + ap2 u^2
//   Bnpa   = -(u^2 + 2 u^3) / Dn
//   Heven  = sqrt(A (Bnp pr^2 + Bnpa pphi^2 apsq u^2 + pphi^2 u^2 + pr^2 + Q + 1))
//   Hodd   = pphi u^3 (G(pr^2, u, lr) + gs3 u^2 + so4 u^4) / Dn
//   H      = M sqrt(1 + 2 nu (Heven + Hodd - 1)) / nu
//
// where G combines the gyro-gravitomagnetic factors gS and gS* with their spin
// weights.  The coefficient tables are the physics of the reference's
// Hamiltonian and are identical to it; only the factoring differs.
//
// ham_eval is templated on the scalar: double (value), Dual (one directional
// derivative), HDual (mixed second derivative).
#pragma once
#include "eob_math.hpp"

namespace eob {

struct HamCoef {
    double M, nu, ap2, apsq;
    // Q: pr^4 {u^2, u^3, u^4, u^4 lr, u^4 su}; pr^6 {u^2, u^3, u^3 su}; pr^8 {u, u^2, u^2 su}
    double q4_2, q4_3, q4_4, q4_4l, q4_92, q6_2, q6_3, q6_72, q8_1, q8_2, q8_52;
    // Dbar: 1 + u^2 (d2 + u (d3 + u (d4 + d4l lr + u (d5 + d5l lr + d112 su))))
    double d2, d3, d4, d4l, d5, d5l, d112;
    // AK0: 1 + u (a1 + u (a2 + u (a3 + u (a4 + u (a5 + a5l lr + u (a6l lr + a132 su))))))
    double a1, a2, a3, a4, a5, a5l, a6l, a132;
    // AKa: u^2 (ka2 + u (ka3 + u (ka4 + u (ka5 + u ka6))))
    double ka2, ka3, ka4, ka5, ka6;
    double e1, e2, as4, as5, sc6, bs3, bs4;
    // G[j][k]: coefficient of pr^(2j) u^k; logs at pr^0 u^4 and pr^2 u^3
    double g00, g01, g02, g03, g0l, g10, g11, g12, g13, g1l, g20, g21, g22, g30, g31, g40;
    double M_nu;   // M / nu: H = (M / nu) sqrt(...) as one product
    double gs3, so4;
};

// binary-only factors; restates the nu/chi polynomials of src/EOB.cu:51-97
EOB_HD void ham_setup(HamCoef& c, double m1, double m2, double chi1, double chi2, double K, double d5, double dSO, double dSS) {
    const double M = m1 + m2;
    const double mu = m1 * m2 / M;
    const double nu = mu / M;
    const double X1 = m1 / M, X2 = m2 / M;
    const double c1 = chi1, c2 = chi2;
    const double nu2 = nu * nu, nu3 = nu2 * nu, nu4 = nu3 * nu, nu5 = nu4 * nu;
    const double PI = kPi, PI2 = kPi * kPi, eg = kEulerGamma, ln2 = kLn2, ln3 = kLn3, ln5 = kLn5, ln7 = kLn7;
    const double s11 = c1 * c1 * X1 * X1, s22 = c2 * c2 * X2 * X2, s12 = c1 * c2 * nu;
    c.M = M;
    c.nu = nu;
    c.M_nu = M / nu;
    c.ap2 = s11 + 2 * s12 + s22;
    const double ap = c1 * X1 + c2 * X2;
    c.apsq = ap * ap;

    // Q = QSalign2 + Qpm
    const double cQS = -5 * s11 * (X2 * (18 * nu - 1) + nu * (11.0 * nu - 21.0)) / 8 - 5 * s12 * (10 * nu2 - 9 * nu - 4) / 8
                       - 5 * s22 * (-18 * X2 * nu + X2 + 11.0 * nu2 - 3.0 * nu - 1) / 8;
    c.q8_1 = nu * (-693359375.0 * ln5 - 643896 + 59326749.0 * ln3 + 1516829440.0 * ln2) / 3150;
    c.q8_52 = 5994461 * PI * nu / 12700800;
    c.q8_2 = -6 * nu4 + 24 * nu3 / 7
             + nu2 * (-11712255859375.0 * ln5 - 3391115364245.0 * ln7 + 1755887616.0 + 8965819883403.0 * ln3 + 22502063927296.0 * ln2) / 1058400
             + nu * (-43480265170944.0 * ln2 - 10632236726043.0 * ln3 + 10005778368.0 + 3391115364245.0 * ln7 + 21877274609375.0 * ln5) / 4233600;
    c.q6_72 = -2723471 * PI * nu / 756000;
    c.q6_2 = 6 * nu3 - 27 * nu2 / 5 + nu * (-2358912 * ln2 / 25 - 827. / 3. + 1399437 * ln3 / 50 + 390625 * ln5 / 18);
    c.q6_3 = -14 * nu4 + 116 * nu3 + nu2 * (-9996617728.0 * ln2 - 408682503 * ln3 + 6681738 + 4580078125.0 * ln5) / 3150
             + nu * (-5084375000.0 * ln5 - 1249161912 * ln3 + 23517747 + 13751491072.0 * ln2) / 9450;
    c.q4_92 = 88703 * PI * nu / 1890;
    c.q4_2 = -6 * nu2 + 8 * nu;
    c.q4_3 = 10 * nu3 - 83 * nu2 + 4 * nu * (-74358 * ln3 - 3981 + 124064 * ln2) / 45 + cQS;
    c.q4_4 = nu3 * (640 - 615 * PI2 / 32)
             + nu2 * (-9765625 * ln5 / 126 - 6396489 * ln3 / 70 - 1184 * eg / 5 + 31633 * PI2 / 512 + 150683. / 105. + 33693536 * ln2 / 105)
             + nu * (-40979464 * ln2 / 315 - 93031 * PI2 / 1536 + 10856 * eg / 105 + 1295219. / 350. + 9765625 * ln5 / 504 + 14203593 * ln3 / 280);
    c.q4_4l = 592 * nu2 / 5 - 5428 * nu / 105;

    // Dbar (Dbpm)
    c.d2 = 6 * nu;
    c.d3 = -6 * nu2 + 52 * nu;
    c.d4 = nu2 * (-260 + 123 * PI2 / 16) + nu * (-6496 * ln2 / 15 - 23761 * PI2 / 1536 - 533. / 45. + 1184 * eg / 15 + 2916 * ln3 / 5);
    c.d4l = -592 * nu / 15;
    c.d5 = d5;
    c.d5l = 3392 * nu2 / 15 + 1420 * nu / 7;
    c.d112 = 264932 * PI * nu / 1575;

    // AK0
    const double K2 = K * K, K3 = K2 * K, K4 = K3 * K, K5 = K4 * K;
    c.a1 = -2. * K * nu;
    c.a2 = 4. * K * nu * (K * nu - 1.);
    c.a3 = 2. * nu * (-4. * K3 * nu2 + 8. * K2 * nu - 4. * K + 1.);
    c.a4 = nu * (16. * K4 * nu3 - 48. * K3 * nu2 + 48. * K2 * nu - 4 * K * (nu + 4) - 41. * PI2 / 32. + 106. / 3.);
    c.a5 = -32. * K5 * nu5 + 128. * K4 * nu4 + K2 * nu3 * (8. - 192. * K) + nu2 * (12288. * K2 - 7552. * K + 246. * PI2 * K - 3536. + 123. * PI2) / 96.
           + nu * (-32. * K + 1. / 20. + 128. * eg / 5. + 963. * PI2 / 512. + 256. * ln2 / 5.);
    c.a5l = -64. * nu / 5.;
    c.a6l = 4. * nu * (nu * (672. * K + 756.) + 1079.) / 105.;
    c.a132 = 13696. * PI * nu / 525.;

    // AKa
    const double kn = K * nu, ap2 = c.ap2;
    c.ka2 = kn * (-kn + 2) * ap2;
    c.ka3 = 2 * kn * ap2 * (2 * K2 * nu2 - 5 * kn + 3);
    c.ka4 = kn * (kn - 2) * (kn - 1) * (kn - 1) * ap2 * ap2 - 4 * kn * (kn - 1) * (kn - 1) * (3 * kn - 4) * ap2;
    c.ka5 = -2 * kn * (kn - 1) * (kn - 1) * (kn - 1) * (3 * kn - 5) * ap2 * ap2 + 2 * nu * (kn - 1) * (kn - 1) * ap2 * (16 * K3 * nu2 - 36 * K2 * nu + 20 * K - 1);
    c.ka6 = -kn * (kn - 2) * ipow(kn - 1, 4) * ap2 * ap2 * ap2
            - nu * (kn - 1) * (kn - 1) * ap2 * (7680 * K4 * nu3 - 24576 * K3 * nu2 + 26112 * K2 * nu - 768 * K * (nu + 12) - 123 * PI2 + 3776) / 96;

    c.e1 = -2 * K * nu + 2;
    c.e2 = (-K * nu + 1) * (-K * nu + 1) * ap2;
    c.as4 = s11 * (5.0 * X2 + nu) / 2 + s12 * (2 - nu) + s22 * (-5.0 * X2 + nu + 5.0) / 2;
    c.as5 = s11 * (-X2 * (294 * nu + 85) + 428.0 * X2 - 21 * nu2 + 987.0 * nu) / 56 + s12 * (6 * nu2 + 145 * nu + 68) / 8
            + s22 * (X2 * (294 * nu + 85) - 428.0 * X2 - 21 * nu2 + 693.0 * nu + 343.0) / 56;
    c.sc6 = dSS * (s11 + c1 * c2 * (1 - 2 * nu) + s22);
    c.bs3 = 3 * s11 * (7 * X2 + 5.0 * nu) / 4 + s12 * (9 * nu / 2 + 6) + 3 * s22 * (-7 * X2 + 5.0 * nu + 7.0) / 4;
    c.bs4 = s11 * (1.0 * X2 * (38 - 48 * nu) - X2 * (80 * nu + 49) - 20 * nu2 + 214.0 * nu) / 4 + s12 * (10 * nu2 + 29 * nu + 8)
            + s22 * (1.0 * X2 * (48 * nu - 38) + X2 * (80 * nu + 49) - 20 * nu2 + 86.0 * nu - 11.0) / 4;

    // gyro-gravitomagnetic factors, weighted: G = gS (chi1 X1^2 + chi2 X2^2) + gS* nu (chi1 + chi2)
    const double wS = c1 * X1 * X1 + c2 * X2 * X2;
    const double wT = nu * (c1 + c2);
    // gS (src/EOB.cu:97)
    const double gS00 = 2, gS01 = -5 * nu / 8, gS02 = (-nu2 - 102 * nu) / 8,
                 gS03 = (-63 * nu3 + 13644 * nu2 - 80399 * nu + 1446 * PI2 * nu) / 1152, gS0l = 584 * nu / 15;
    const double gS10 = -27 * nu / 8, gS11 = (23 * nu2 - 84 * nu) / 8, gS12 = (69 * nu3 + 6228 * nu2 - 5283 * nu) / 64,
                 gS13 = (14400 * nu4 - 2477700 * nu3 - 326100 * PI2 * nu2 + 19813300 * nu2 - 89579520 * nu * ln3 - 1672575 * PI2 * nu
                         - 3993600 * eg * nu + 54396338 * nu + 83824640 * nu * ln2) / 19200,
                 gS1l = 1996800 * nu / 19200;
    const double gS20 = (35 * nu2 + 5 * nu) / 8, gS21 = (-771 * nu3 + 3324 * nu2 + 781 * nu) / 128,
                 gS22 = (-38520 * nu4 - 3309615 * nu3 + 2872170 * nu2 - 1626767360 * nu * ln2 + 19954938 * nu + 1008889344 * nu * ln3) / 11520;
    const double gS30 = -(665 * nu3 + 252 * nu2 - 7 * nu) / 128,
                 gS31 = (567720 * nu4 - 2564325 * nu3 - 1148850 * nu2 - 28998570240.0 * nu * ln3 - 19300000000.0 * nu * ln5 + 175835718 * nu
                         + 90521206784.0 * nu * ln2) / 57600;
    const double gS40 = (7144200 * nu4 + 4465125 * nu3 + 935550 * nu2 - 11506980454400.0 * nu * ln2 - 188268367392.0 * nu * ln3 + 3476909988.0 * nu
                         + 5082137500000.0 * nu * ln5) / 1209600;
    // gS* (src/EOB.cu:95)
    const double gT00 = 3. / 2., gT01 = 3 * (-2 * nu - 3) / 8., gT02 = 3 * (-nu2 - 52 * nu - 9) / 16.,
                 gT03 = (-36 * nu3 + 4266 * nu2 - 15254 * nu + 492 * PI2 * nu - 1215) / 384, gT0l = 24 * nu;
    const double gT10 = 3 * (-6 * nu - 5) / 8., gT11 = 3 * (19 * nu2 - 12 * nu + 23) / 16., gT12 = 3 * (36 * nu3 + 1574 * nu2 - 558 * nu - 15) / 64.,
                 gT13 = (5130 * nu4 - 469440 * nu3 - 59040 * PI2 * nu2 + 2316030 * nu2 - 12317184 * nu * ln3 - 72435 * PI2 * nu - 393216 * eg * nu
                         + 5731444 * nu + 12025856 * nu * ln2 - 3240) / 3840,
                 gT1l = 196608 * nu / 3840;
    const double gT20 = (45 * nu2 + 40 * nu + 35) / 16, gT21 = (-972 * nu3 + 702 * nu2 - 106 * nu - 1105) / 128,
                 gT22 = (-20655 * nu4 - 834300 * nu3 + 182025 * nu2 - 381263872 * nu * ln2 + 3596448 * nu + 237385728 * nu * ln3 + 37875) / 3840;
    const double gT30 = -(420 * nu3 + 462 * nu2 + 350 * nu + 315) / 128,
                 gT31 = (2170395 * nu4 - 1291140 * nu3 - 593325 * nu2 - 62204578560.0 * nu * ln3 - 40300000000.0 * nu * ln5 + 343473264 * nu
                         + 191671894016.0 * nu * ln2 + 2399625) / 172800;
    const double gT40 = (4465125 * nu4 + 5953500 * nu3 + 4975425 * nu2 - 8132450787328.0 * nu * ln2 - 90415828800.0 * nu * ln3 + 2254036356.0 * nu
                         + 3562775000000.0 * nu * ln5 + 3274425) / 1209600;
    c.g00 = gS00 * wS + gT00 * wT; c.g01 = gS01 * wS + gT01 * wT; c.g02 = gS02 * wS + gT02 * wT; c.g03 = gS03 * wS + gT03 * wT;
    c.g0l = gS0l * wS + gT0l * wT;
    c.g10 = gS10 * wS + gT10 * wT; c.g11 = gS11 * wS + gT11 * wT; c.g12 = gS12 * wS + gT12 * wT; c.g13 = gS13 * wS + gT13 * wT;
    c.g1l = gS1l * wS + gT1l * wT;
    c.g20 = gS20 * wS + gT20 * wT; c.g21 = gS21 * wS + gT21 * wT; c.g22 = gS22 * wS + gT22 * wT;
    c.g30 = gS30 * wS + gT30 * wT; c.g31 = gS31 * wS + gT31 * wT;
    c.g40 = gS40 * wS + gT40 * wT;
    // GS31 / GS32 (src/EOB.cu:91-93) and SOcalib (:93)
    const double GS31 = -0.5 * s11 * X2 + s12 * (-2.0 * X2 - 5.0) / 12 + s22 * (4.0 * X2 - 5) / 12;
    const double GS32 = s11 * (-4.0 * X2 - 1.0) / 12 + s12 * (0.166666666666667 * X2 - 0.583333333333333) + s22 * (2.0 * X2 - 2.0) / 4;
    c.gs3 = c1 * X1 * GS31 + c2 * X2 * GS32;
    c.so4 = dSO * ap + (X1 - X2) * (c1 * X1 - c2 * X2)
                           * (7 * nu4 / 512 + 31 * nu3 / 512 + nu2 * (208493 - 5136 * PI2) / 4608
                              - nu * (-13288125 * PI2 + 13762560 * eg + 26542080 * ln2 + 168816592) / 921600 + 1701. / 512.);
}

template <class T>
EOB_HD T ham_eval(const HamCoef& c, T r, T pr, T L) {
    const T u = 1.0 / r;
    const T lr = dlog(r);
    const T su = dsqrt(u);
    const T u2 = u * u;
    const T u3 = u2 * u;
    const T P = pr * pr;
    const T L2 = L * L;

    const T Dn = 1.0 + c.ap2 * (u2 + 2.0 * u3);
    const T iDn = 1.0 / Dn;

    // radial potentials
    const T AK0 = 1.0 + u * (c.a1 + u * (c.a2 + u * (c.a3 + u * (c.a4 + u * (c.a5 + c.a5l * lr + u * (c.a6l * lr + c.a132 * su))))));
    const T AKa = u2 * (c.ka2 + u * (c.ka3 + u * (c.ka4 + u * (c.ka5 + u * c.ka6))));
    const T Dbpm = 1.0 + u2 * (c.d2 + u * (c.d3 + u * (c.d4 + c.d4l * lr + u * (c.d5 + c.d5l * lr + c.d112 * su))));
    const T AS = 1.0 + u2 * u2 * (c.as4 + u * (c.as5 + u * c.sc6));
    const T f1 = 1.0 - c.e1 * u;
    const T A = (AK0 + AKa) * (f1 + c.e2 * u2) * AS * iDn;
    const T Bnp = AK0 * Dbpm * f1 + u3 * (c.bs3 + c.bs4 * u) - 1.0 + c.ap2 * u2;
    const T Bnpa = -(u2 + 2.0 * u3) * iDn;

    // radial-momentum potential
    const T Q = P * P * (u2 * (c.q4_2 + u * (c.q4_3 + u * (c.q4_4 + c.q4_4l * lr + c.q4_92 * su))))
              + P * P * P * (u2 * (c.q6_2 + u * (c.q6_3 + c.q6_72 * su)))
              + P * P * P * P * (u * (c.q8_1 + u * (c.q8_2 + c.q8_52 * su)));

    const T Heven = dsqrt(A * (Bnp * P + Bnpa * L2 * c.apsq * u2 + L2 * u2 + P + Q + 1.0));

    // spin-orbit part
    const T G0 = c.g00 + u * (c.g01 + u * (c.g02 + u * (c.g03 + u * (c.g0l * lr))));
    const T G1 = c.g10 + u * (c.g11 + u * (c.g12 + u * (c.g13 + c.g1l * lr)));
    const T G2 = c.g20 + u * (c.g21 + u * c.g22);
    const T G3 = c.g30 + u * c.g31;
    const T Gt = G0 + P * (G1 + P * (G2 + P * (G3 + P * c.g40))) + u2 * (c.gs3 + c.so4 * u2);
    const T Hodd = L * u3 * Gt * iDn;

    const T Heff = Heven + Hodd;
    return c.M * dsqrt(1.0 + 2.0 * c.nu * (Heff - 1.0)) / c.nu;
}

// ------------------------------------------------------------------ pieces
// The u-dependence of H as 13 independent pieces of one common form,
//   piece(u) = sum_{k=0}^{6} (a_k + b_k ln r + s_k sqrt(u)) u^k,
// so that the value and d/du of every piece can be evaluated together on
// different lanes with the same instructions (piece_eval), after which
// ham_combine forms H and its derivatives in closed form.  The pieces are the
// potentials of ham_eval: AK0, AKa, Dbar, the spin factor of A, Dn, the pr^4,
// pr^6, pr^8 parts of Q, the gyro factors G0..G3 and gs3 u^2 + so4 u^4.
enum HamPiece { kAK0, kAKa, kDb, kAS, kDn, kQ4, kQ6, kQ8, kG0, kG1, kG2, kG3, kGS, kNumPieces };
constexpr int kPieceDeg = 7;
struct PieceCoef {
    double c[kNumPieces][kPieceDeg][3];   // (a_k, b_k, s_k) per power u^k
};

EOB_HD void piece_setup(PieceCoef& p, const HamCoef& c) {
    for (int q = 0; q < kNumPieces; ++q)
        for (int k = 0; k < kPieceDeg; ++k) p.c[q][k][0] = p.c[q][k][1] = p.c[q][k][2] = 0.0;
    auto put = [&](int q, int k, double a, double b = 0.0, double s = 0.0) {
        p.c[q][k][0] = a; p.c[q][k][1] = b; p.c[q][k][2] = s;
    };
    put(kAK0, 0, 1.0); put(kAK0, 1, c.a1); put(kAK0, 2, c.a2); put(kAK0, 3, c.a3); put(kAK0, 4, c.a4);
    put(kAK0, 5, c.a5, c.a5l); put(kAK0, 6, 0.0, c.a6l, c.a132);
    put(kAKa, 2, c.ka2); put(kAKa, 3, c.ka3); put(kAKa, 4, c.ka4); put(kAKa, 5, c.ka5); put(kAKa, 6, c.ka6);
    put(kDb, 0, 1.0); put(kDb, 2, c.d2); put(kDb, 3, c.d3); put(kDb, 4, c.d4, c.d4l); put(kDb, 5, c.d5, c.d5l, c.d112);
    put(kAS, 0, 1.0); put(kAS, 4, c.as4); put(kAS, 5, c.as5); put(kAS, 6, c.sc6);
    put(kDn, 0, 1.0); put(kDn, 2, c.ap2); put(kDn, 3, 2.0 * c.ap2);
    put(kQ4, 2, c.q4_2); put(kQ4, 3, c.q4_3); put(kQ4, 4, c.q4_4, c.q4_4l, c.q4_92);
    put(kQ6, 2, c.q6_2); put(kQ6, 3, c.q6_3, 0.0, c.q6_72);
    put(kQ8, 1, c.q8_1); put(kQ8, 2, c.q8_2, 0.0, c.q8_52);
    put(kG0, 0, c.g00); put(kG0, 1, c.g01); put(kG0, 2, c.g02); put(kG0, 3, c.g03); put(kG0, 4, 0.0, c.g0l);
    put(kG1, 0, c.g10); put(kG1, 1, c.g11); put(kG1, 2, c.g12); put(kG1, 3, c.g13, c.g1l);
    put(kG2, 0, c.g20); put(kG2, 1, c.g21); put(kG2, 2, c.g22);
    put(kG3, 0, c.g30); put(kG3, 1, c.g31);
    put(kGS, 2, c.gs3); put(kGS, 4, c.so4);
}

// value and d/du of one piece; r = 1/u, lr = ln r, su = sqrt(u), hs = d su/du = su r / 2
template <class T>
EOB_HD void piece_eval(const double (*cf)[3], T u, T r, T lr, T su, T hs, T& v, T& dv) {
    T acc = cf[6][0] + cf[6][1] * lr + cf[6][2] * su;
    T dacc = -cf[6][1] * r + cf[6][2] * hs;
#pragma unroll
    for (int k = 5; k >= 0; --k) {
        const T ck = cf[k][0] + cf[k][1] * lr + cf[k][2] * su;
        const T dck = -cf[k][1] * r + cf[k][2] * hs;
        dacc = dck + acc + u * dacc;
        acc = ck + u * acc;
    }
    v = acc;
    dv = dacc;
}

// Everything the flux and the RHS need at one state.
template <class T>
struct HamPackT {
    T H, dHdr, dHdpr, dHdL, omc;   // omc: dH/dpphi at pr = 0 (circular frequency)
};
using HamPack = HamPackT<double>;

// Where dH/dpphi at pr = 0 (the circular frequency) comes from: SeparateP0
// evaluates that branch beside the full one on the same lane; a wavefront can
// instead let one of its lanes evaluate the whole expression at pr = 0 (LaneP0
// in eob_device.hpp), since it is the same expression with P = 0.
struct SeparateP0 {
    static constexpr bool kLaneP0 = false;
};

// H, dH/dr, dH/dpr, dH/dpphi and dH/dpphi at pr = 0 from the piece values and
// u-derivatives (v, d indexed by HamPiece); the closed-form derivatives of
// ham_eval's expression.
template <class T, class Br = SeparateP0>
EOB_HD HamPackT<T> ham_combine(const HamCoef& c, T r, T u, T pr, T L, const T* v, const T* d, const Br& br = Br()) {
    (void)r;
    T P = pr * pr;
    if constexpr (Br::kLaneP0) P = br.P(P);
    const T u2 = u * u, u3 = u2 * u, L2 = L * L;
    const T iDn = qrcp(v[kDn]);
    const T diDn = -d[kDn] * iDn * iDn;
    const T f1 = 1.0 - c.e1 * u, df1 = -c.e1;
    const T fe = f1 + c.e2 * u2, dfe = -c.e1 + 2.0 * c.e2 * u;
    const T Ks = v[kAK0] + v[kAKa], dKs = d[kAK0] + d[kAKa];
    const T A0 = Ks * fe * v[kAS];
    const T A = A0 * iDn;
    const T dA = (dKs * fe * v[kAS] + Ks * dfe * v[kAS] + Ks * fe * d[kAS]) * iDn + A0 * diDn;
    const T Bsp = u3 * (c.bs3 + c.bs4 * u), dBsp = u2 * (3.0 * c.bs3 + 4.0 * c.bs4 * u);
    const T Bnp = v[kAK0] * v[kDb] * f1 + Bsp - 1.0 + c.ap2 * u2;
    const T dBnp = d[kAK0] * v[kDb] * f1 + v[kAK0] * d[kDb] * f1 + v[kAK0] * v[kDb] * df1 + dBsp + 2.0 * c.ap2 * u;
    const T m = u2 + 2.0 * u3, dm = 2.0 * u + 6.0 * u2;
    const T Bnpa = -m * iDn, dBnpa = -dm * iDn - m * diDn;
    const T W = (Bnpa * c.apsq + 1.0) * u2;   // coefficient of pphi^2
    const T dW = dBnpa * c.apsq * u2 + (Bnpa * c.apsq + 1.0) * 2.0 * u;
    const T P2 = P * P;
    const T Q = P2 * (v[kQ4] + P * (v[kQ6] + P * v[kQ8]));
    const T dQ = P2 * (d[kQ4] + P * (d[kQ6] + P * d[kQ8]));
    const T QP = P * (2.0 * v[kQ4] + P * (3.0 * v[kQ6] + 4.0 * P * v[kQ8]));   // dQ/dP
    const T X = (Bnp + 1.0) * P + W * L2 + Q + 1.0;
    T Heven, iHe;
    qsqrt_recip(A * X, Heven, iHe);
    const T ihe = 0.5 * iHe;
    const T Gt = v[kG0] + P * (v[kG1] + P * (v[kG2] + P * (v[kG3] + P * c.g40))) + v[kGS];
    const T dGt = d[kG0] + P * (d[kG1] + P * (d[kG2] + P * d[kG3])) + d[kGS];
    const T GtP = v[kG1] + P * (2.0 * v[kG2] + P * (3.0 * v[kG3] + 4.0 * P * c.g40));
    const T U3i = u3 * iDn, dU3i = 3.0 * u2 * iDn + u3 * diDn;
    const T Heff = Heven + L * U3i * Gt;
    T s, is;
    qsqrt_recip(1.0 + 2.0 * c.nu * (Heff - 1.0), s, is);
    const T kH = c.M * is;   // dH/dHeff
    const T dHe_du = (dA * X + A * (dBnp * P + dW * L2 + dQ)) * ihe + L * (dU3i * Gt + U3i * dGt);
    const T dHe_dP = A * (Bnp + 1.0 + QP) * ihe + L * U3i * GtP;
    const T dHe_dL = A * L * W * iHe + U3i * Gt;
    HamPackT<T> p;
    p.H = s * c.M_nu;
    p.dHdr = -u2 * kH * dHe_du;
    p.dHdpr = kH * dHe_dP * 2.0 * pr;
    p.dHdL = kH * dHe_dL;
    if constexpr (Br::kLaneP0) {
        p.omc = br.omc(p.dHdL);
    } else {
        T He0, iHe0, s0, is0;
        qsqrt_recip(A * (W * L2 + 1.0), He0, iHe0);
        const T Gt0 = v[kG0] + v[kGS];
        qsqrt_recip(1.0 + 2.0 * c.nu * (He0 + L * U3i * Gt0 - 1.0), s0, is0);
        p.omc = (c.M * is0) * (A * L * W * iHe0 + U3i * Gt0);
    }
    return p;
}

}  // namespace eob
