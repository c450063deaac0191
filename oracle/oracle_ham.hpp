// Oracle test infrastructure (CPU only; never shipped, never measured).
//
// Aligned-spin EOB Hamiltonian restated term by term after the reference's
// evaluate_Ham_align_AD (src/EOB.cu:17-107).  The numerical coefficient tables
// are physics (post-Newtonian potentials A, Dbar, Q, gyro-gravitomagnetic
// factors gS/gS*) and must equal the reference's; the arrangement is ours.
// The function is a template so that the same text evaluates H (double), its
// gradient (Dual<double,4>) and its Hessian (Dual<Dual<double,4>,4>).
//
// Decimal literals that the reference spells as truncated decimals
// (0.166666666666667, 0.583333333333333 in GS32, src/EOB.cu:91) are kept as
// written, because they are part of the reference's arithmetic.
#pragma once
#include <cmath>
#include "oracle_dual.hpp"

namespace orc {

static const double kPi = 3.141592653589793238462643383279502884;
static const double kEulerGamma = 0.57721566490153286060651209008240243104215933593992;

struct HamParams {
    double m1, m2, chi1, chi2, K, d5, dSO, dSS;
};

template <class T> inline T sq(const T& x) { return x * x; }
using std::pow;
using std::sqrt;
using std::log;

// H/nu * M (i.e. the reference's returned "H_real/nu"), src/EOB.cu:17-107.
template <class T>
T hamiltonian(const T& r, const T& pr, const T& L, const HamParams& p) {
    const double M = p.m1 + p.m2;
    const double mu = p.m1 * p.m2 / M;
    const double nu = mu / M;
    const double X1 = p.m1 / M, X2 = p.m2 / M;
    const double c1 = p.chi1, c2 = p.chi2;
    const double K = p.K, d5 = p.d5, dSO = p.dSO, dSS = p.dSS;
    const double ln2 = std::log(2.0), ln3 = std::log(3.0), ln5 = std::log(5.0), ln7 = std::log(7.0);
    const double PI = kPi, eg = kEulerGamma;
    const double nu2 = nu * nu, nu3 = nu2 * nu, nu4 = nu3 * nu;
    // Kerr-like combination (chi1 X1 + chi2 X2)^2 written as the reference expands it
    const double ap2 = c1 * c1 * X1 * X1 + 2 * c1 * c2 * nu + c2 * c2 * X2 * X2;
    const double ap = c1 * X1 + c2 * X2;

    const T lnr = log(r);
    const T r2 = r * r, r3 = r2 * r, r4 = r3 * r, r5 = r4 * r, r6 = r5 * r;
    const T pr2 = pr * pr, pr4 = pr2 * pr2, pr6 = pr4 * pr2, pr8 = pr4 * pr4;
    const T L2 = L * L;

    const T psq = L2 / r2 + pr2;

    // QSalign2 (src/EOB.cu:51)
    const double cQS = -5 * c1 * c1 * X1 * X1 * (X2 * (18 * nu - 1) + nu * (11.0 * nu - 21.0)) / 8
                       - 5 * c1 * c2 * nu * (10 * nu2 - 9 * nu - 4) / 8
                       - 5 * c2 * c2 * X2 * X2 * (-18 * X2 * nu + X2 + 11.0 * nu2 - 3.0 * nu - 1) / 8;
    const T QSalign2 = pr4 * cQS / r3;

    // Qpm (src/EOB.cu:53): radial-momentum potential through pr^8
    const double q8a = nu * (-693359375.0 * ln5 - 643896 + 59326749.0 * ln3 + 1516829440.0 * ln2);
    const double q8b = 5994461 * PI * nu;
    const double q8c = -6 * nu4 + 24 * nu3 / 7
                       + nu2 * (-11712255859375.0 * ln5 - 3391115364245.0 * ln7 + 1755887616.0 + 8965819883403.0 * ln3 + 22502063927296.0 * ln2) / 1058400
                       + nu * (-43480265170944.0 * ln2 - 10632236726043.0 * ln3 + 10005778368.0 + 3391115364245.0 * ln7 + 21877274609375.0 * ln5) / 4233600;
    const double q6a = -2723471 * PI * nu;
    const double q6b = 6 * nu3 - 27 * nu2 / 5 + nu * (-2358912 * ln2 / 25 - 827. / 3. + 1399437 * ln3 / 50 + 390625 * ln5 / 18);
    const double q6c = -14 * nu4 + 116 * nu3 + nu2 * (-9996617728.0 * ln2 - 408682503 * ln3 + 6681738 + 4580078125.0 * ln5) / 3150
                       + nu * (-5084375000.0 * ln5 - 1249161912 * ln3 + 23517747 + 13751491072.0 * ln2) / 9450;
    const double q4a = 88703 * PI * nu;
    const double q4b = -6 * nu2 + 8 * nu;
    const double q4c = 10 * nu3 - 83 * nu2 + 4 * nu * (-74358 * ln3 - 3981 + 124064 * ln2) / 45;
    const T q4d = nu3 * (640 - 615 * PI * PI / 32)
                  + nu2 * (592 * lnr / 5 - 9765625 * ln5 / 126 - 6396489 * ln3 / 70 - 1184 * eg / 5 + 31633 * PI * PI / 512 + 150683. / 105. + 33693536 * ln2 / 105)
                  + nu * (-5428 * lnr / 105 - 40979464 * ln2 / 315 - 93031 * PI * PI / 1536 + 10856 * eg / 105 + 1295219. / 350. + 9765625 * ln5 / 504 + 14203593 * ln3 / 280);
    const T Qpm = pr8 * (q8a / (3150 * r) + q8b / (12700800 * pow(r, 5. / 2.)) + q8c / r2)
                  + pr6 * (q6a / (756000 * pow(r, 7. / 2.)) + q6b / r2 + q6c / r3)
                  + pr4 * (q4a / (1890 * pow(r, 9. / 2.)) + q4b / r2 + q4c / r3 + q4d / r4);

    const T Q = QSalign2 + Qpm;

    // Bnpa (src/EOB.cu:59)
    const T Bnpa = -r * (r + 2) / (r4 + r * (r + 2) * ap2);

    // BnpSalign2 (src/EOB.cu:63)
    const double bs3 = 3 * c1 * c1 * X1 * X1 * (7 * X2 + 5.0 * nu) / 4 + c1 * c2 * nu * (9 * nu / 2 + 6)
                       + 3 * c2 * c2 * X2 * X2 * (-7 * X2 + 5.0 * nu + 7.0) / 4;
    const double bs4 = c1 * c1 * X1 * X1 * (1.0 * X2 * (38 - 48 * nu) - X2 * (80 * nu + 49) - 20 * nu2 + 214.0 * nu) / 4
                       + c1 * c2 * nu * (10 * nu2 + 29 * nu + 8)
                       + c2 * c2 * X2 * X2 * (1.0 * X2 * (48 * nu - 38) + X2 * (80 * nu + 49) - 20 * nu2 + 86.0 * nu - 11.0) / 4;
    const T BnpSalign2 = bs3 / r3 + bs4 / r4;

    // Dbpm (src/EOB.cu:65)
    const T Dbpm = 6 * nu / r2 + 264932 * PI * nu / (1575 * pow(r, 11. / 2.)) + 1 + (-6 * nu2 + 52 * nu) / r3
                   + (nu2 * (-260 + 123 * PI * PI / 16)
                      + nu * (-592 * lnr / 15 - 6496 * ln2 / 15 - 23761 * PI * PI / 1536 - 533. / 45. + 1184 * eg / 15 + 2916 * ln3 / 5)) / r4
                   + (d5 + 3392 * nu2 * lnr / 15 + 1420 * nu * lnr / 7) / r5;

    // AK0 (src/EOB.cu:67)
    const T AK0 = -2. * K * nu / r + 4. * K * nu * (K * nu - 1.) / r2
                  + 2. * nu * (-4. * pow(K, 3) * nu2 + 8. * K * K * nu - 4. * K + 1.) / r3
                  + nu * (16. * pow(K, 4) * nu3 - 48. * pow(K, 3) * nu2 + 48. * K * K * nu - 4 * K * (nu + 4) - 41. * PI * PI / 32. + 106. / 3.) / r4
                  + 4. * nu * (nu * (672. * K + 756.) + 1079.) * lnr / (105. * r6)
                  + 13696. * PI * nu / (525. * pow(r, 13. / 2.)) + 1.
                  + (-32. * pow(K, 5) * pow(nu, 5) + 128. * pow(K, 4) * nu4 + K * K * nu3 * (8. - 192. * K)
                     + nu2 * (12288. * K * K - 7552. * K + 246. * PI * PI * K - 3536. + 123. * PI * PI) / 96.
                     + nu * (-32. * K + 1. / 20. + 128. * eg / 5. + 963. * PI * PI / 512. + 256. * ln2 / 5.)
                     - 64. * nu * lnr / 5.) / r5;

    // Bnp (src/EOB.cu:69)
    const T Bnp = AK0 * Dbpm * (1 - (-2 * K * nu + 2) / r) + BnpSalign2 - 1 + ap2 / r2;
    const double Bp = 1.0;  // 1/(BpSprec2 + 1) with BpSprec2 = 0

    // AScalib2 (src/EOB.cu:75) and ASalign2 (:83)
    const double sc6 = dSS * (c1 * c1 * X1 * X1 + c1 * c2 * (1 - 2 * nu) + c2 * c2 * X2 * X2);
    const T AScalib2 = sc6 / r6;
    const double as4 = c1 * c1 * X1 * X1 * (5.0 * X2 + nu) / 2 + c1 * c2 * nu * (2 - nu) + c2 * c2 * X2 * X2 * (-5.0 * X2 + nu + 5.0) / 2;
    const double as5 = c1 * c1 * X1 * X1 * (-X2 * (294 * nu + 85) + 428.0 * X2 - 21 * nu2 + 987.0 * nu) / 56
                       + c1 * c2 * nu * (6 * nu2 + 145 * nu + 68) / 8
                       + c2 * c2 * X2 * X2 * (X2 * (294 * nu + 85) - 428.0 * X2 - 21 * nu2 + 693.0 * nu + 343.0) / 56;
    const T ASalign2 = as4 / r4 + as5 / r5;

    // AKa (src/EOB.cu:85): spin part of the A potential, polynomial in K nu and ap2
    const double kn = K * nu;
    const double ka2 = kn * (-kn + 2) * ap2;
    const double ka3 = 2 * kn * ap2 * (2 * K * K * nu2 - 5 * kn + 3);
    const double ka4 = kn * (kn - 2) * sq(kn - 1) * ap2 * ap2 - 4 * kn * sq(kn - 1) * (3 * kn - 4) * ap2;
    const double ka5 = -2 * kn * pow(kn - 1, 3) * (3 * kn - 5) * ap2 * ap2
                       + 2 * nu * sq(kn - 1) * ap2 * (16 * pow(K, 3) * nu2 - 36 * K * K * nu + 20 * K - 1);
    const double ka6 = -kn * (kn - 2) * pow(kn - 1, 4) * ap2 * ap2 * ap2
                       - nu * sq(kn - 1) * ap2 * (7680 * pow(K, 4) * nu3 - 24576 * pow(K, 3) * nu2 + 26112 * K * K * nu - 768 * K * (nu + 12) - 123 * PI * PI + 3776) / 96;
    const T AKa = ka2 / r2 + ka3 / r3 + ka4 / r4 + ka5 / r5 + ka6 / r6;

    // A (src/EOB.cu:87)
    const T A = (AK0 + AKa) * (1 - (-2 * K * nu + 2) / r + sq(-K * nu + 1) * ap2 / r2)
                * (ASalign2 + AScalib2 + 1) / (1 + (1 + 2. / r) * ap2 / r2);

    // Heven (src/EOB.cu:89)
    const T Heven = sqrt(A * (Bnp * pr2 + Bnpa * L2 * (ap * ap) / r2 + Bp * psq + Q + 1));

    // GS32, GS31 (src/EOB.cu:91-93)
    const T GS32 = (c1 * c1 * X1 * X1 * (-4.0 * X2 - 1.0) / 12 + c1 * c2 * nu * (0.166666666666667 * X2 - 0.583333333333333)
                    + c2 * c2 * X2 * X2 * (2.0 * X2 - 2.0) / 4) / r2;
    const T GS31 = (-0.5 * c1 * c1 * X1 * X1 * X2 + c1 * c2 * nu * (-2.0 * X2 - 5.0) / 12 + c2 * c2 * X2 * X2 * (4.0 * X2 - 5) / 12) / r2;

    // SOcalib (src/EOB.cu:93)
    const double so4 = dSO * ap + (X1 - X2) * (c1 * X1 - c2 * X2)
                                      * (7 * nu4 / 512 + 31 * nu3 / 512 + nu2 * (208493 - 5136 * PI * PI) / 4608
                                         - nu * (-13288125 * PI * PI + 13762560 * eg + 26542080 * ln2 + 168816592) / 921600 + 1701. / 512.);
    const T SOcalib = L * so4 / r4;

    // gSstr (src/EOB.cu:95)
    const T gSstr = 24 * nu * lnr / r4
        + pr8 * (4465125 * nu4 + 5953500 * nu3 + 4975425 * nu2 - 8132450787328.0 * nu * ln2 - 90415828800.0 * nu * ln3 + 2254036356.0 * nu + 3562775000000.0 * nu * ln5 + 3274425) / 1209600
        - pr6 * (420 * nu3 + 462 * nu2 + 350 * nu + 315) / 128
        + pr6 * (2170395 * nu4 - 1291140 * nu3 - 593325 * nu2 - 62204578560.0 * nu * ln3 - 40300000000.0 * nu * ln5 + 343473264 * nu + 191671894016.0 * nu * ln2 + 2399625) / (172800 * r)
        + pr4 * (45 * nu2 + 40 * nu + 35) / 16
        + pr4 * (-972 * nu3 + 702 * nu2 - 106 * nu - 1105) / (128 * r)
        + pr4 * (-20655 * nu4 - 834300 * nu3 + 182025 * nu2 - 381263872 * nu * ln2 + 3596448 * nu + 237385728 * nu * ln3 + 37875) / (3840 * r2)
        + 3 * pr2 * (-6 * nu - 5) / 8
        + 3 * pr2 * (19 * nu2 - 12 * nu + 23) / (16 * r)
        + 3 * pr2 * (36 * nu3 + 1574 * nu2 - 558 * nu - 15) / (64 * r2)
        + pr2 * (5130 * nu4 - 469440 * nu3 - 59040 * PI * PI * nu2 + 2316030 * nu2 + 196608 * nu * lnr - 12317184 * nu * ln3 - 72435 * PI * PI * nu - 393216 * eg * nu + 5731444 * nu + 12025856 * nu * ln2 - 3240) / (3840 * r3)
        + 3. / 2. + 3 * (-2 * nu - 3) / (8 * r) + 3 * (-nu2 - 52 * nu - 9) / (16 * r2)
        + (-36 * nu3 + 4266 * nu2 - 15254 * nu + 492 * PI * PI * nu - 1215) / (384 * r3);

    // gS (src/EOB.cu:97)
    const T gS = -27 * nu * pr2 / 8 - 5 * nu / (8 * r) + 584 * nu * lnr / (15 * r4)
        + pr8 * (7144200 * nu4 + 4465125 * nu3 + 935550 * nu2 - 11506980454400.0 * nu * ln2 - 188268367392.0 * nu * ln3 + 3476909988.0 * nu + 5082137500000.0 * nu * ln5) / 1209600
        - pr6 * (665 * nu3 + 252 * nu2 - 7 * nu) / 128
        + pr6 * (567720 * nu4 - 2564325 * nu3 - 1148850 * nu2 - 28998570240.0 * nu * ln3 - 19300000000.0 * nu * ln5 + 175835718 * nu + 90521206784.0 * nu * ln2) / (57600 * r)
        + pr4 * (35 * nu2 + 5 * nu) / 8
        + pr4 * (-771 * nu3 + 3324 * nu2 + 781 * nu) / (128 * r)
        + pr4 * (-38520 * nu4 - 3309615 * nu3 + 2872170 * nu2 - 1626767360 * nu * ln2 + 19954938 * nu + 1008889344 * nu * ln3) / (11520 * r2)
        + pr2 * (23 * nu2 - 84 * nu) / (8 * r)
        + pr2 * (69 * nu3 + 6228 * nu2 - 5283 * nu) / (64 * r2)
        + pr2 * (14400 * nu4 - 2477700 * nu3 - 326100 * PI * PI * nu2 + 19813300 * nu2 + 1996800 * nu * lnr - 89579520 * nu * ln3 - 1672575 * PI * PI * nu - 3993600 * eg * nu + 54396338 * nu + 83824640 * nu * ln2) / (19200 * r3)
        + 2 + (-nu2 - 102 * nu) / (8 * r2)
        + (-63 * nu3 + 13644 * nu2 - 80399 * nu + 1446 * PI * PI * nu) / (1152 * r3);

    // Hodd (src/EOB.cu:99)
    const T Hodd = (L * gS * (c1 * X1 * X1 + c2 * X2 * X2) + L * gSstr * nu * (c1 + c2) + L * (c1 * GS31 * X1 + c2 * GS32 * X2) + SOcalib)
                   / (2 * ap2 + 2 * r2 + r * (ap2 + r2 - 2 * r));

    const T Heff = Heven + Hodd;
    return M * sqrt(1 + 2 * nu * (Heff - 1)) / nu;
}

}  // namespace orc
