// Merger-ringdown attachment: the NR peak fits that fix the non-quasi-circular
// (NQC) corrections, the ringdown fits, the remnant's mass and spin and its
// quasi-normal mode (QNM) frequencies.  __host__ __device__ and FP64 throughout;
// every function works on one (binary, mode) and is called by one thread.
//
// The fit tables are calibration data of SEOBNRv4 / SEOBNRv4HM and are the
// reference's numbers (they must be, for its waveform): NR peak fits
// eob/seobnrv4phm.py:878-1462, ringdown fits eob/eobstuff.py:44-342, remnant
// mass eob/eobstuff.py:6-21.  The remnant spin and the QNM frequencies come
// from outside the reference (pesummary's HBR2016 fit, lalsimulation's QNM
// tables); here they are the published HBR2016 formula and Leaver's continued-
// fraction solution of the Teukolsky equation.
#pragma once
#include <hip/hip_runtime.h>
#include <cmath>

#define IMR_HD __host__ __device__ inline

namespace imr {

// ------------------------------------------------------------------ complex
struct Cx {
    double re, im;
};
IMR_HD Cx cx(double r, double i = 0.0) { return Cx{r, i}; }
IMR_HD Cx operator+(Cx a, Cx b) { return {a.re + b.re, a.im + b.im}; }
IMR_HD Cx operator-(Cx a, Cx b) { return {a.re - b.re, a.im - b.im}; }
IMR_HD Cx operator-(Cx a) { return {-a.re, -a.im}; }
IMR_HD Cx operator*(Cx a, Cx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
IMR_HD Cx operator*(double s, Cx a) { return {s * a.re, s * a.im}; }
IMR_HD Cx operator*(Cx a, double s) { return {s * a.re, s * a.im}; }
IMR_HD Cx operator+(Cx a, double s) { return {a.re + s, a.im}; }
IMR_HD Cx operator-(Cx a, double s) { return {a.re - s, a.im}; }
IMR_HD Cx operator/(Cx a, Cx b) {
    // scaled (Smith) division: no overflow for the continued fractions' large terms
    if (fabs(b.re) >= fabs(b.im)) {
        const double r = b.im / b.re, d = b.re + b.im * r;
        return {(a.re + a.im * r) / d, (a.im - a.re * r) / d};
    }
    const double r = b.re / b.im, d = b.re * r + b.im;
    return {(a.re * r + a.im) / d, (a.im * r - a.re) / d};
}
IMR_HD double cabs2(Cx a) { return a.re * a.re + a.im * a.im; }
IMR_HD Cx csqrt_(Cx z) {
    const double m = sqrt(sqrt(cabs2(z)));
    const double h = 0.5 * atan2(z.im, z.re);
    return {m * cos(h), m * sin(h)};
}

// ------------------------------------------------------------- NR peak fits
// Quantities of the binary the fits take, masses scaled to M = 1 or not.
struct FitIn {
    double m1, m2, chi1, chi2;
};

// Time of the (2,2) amplitude peak before the orbital-frequency peak,
// eob/seobnrv4phm.py:878-929 with (l, m) = (2, 2) (the only call,
// :2376); the (5,5) mode's 10 M are added by the callers.
IMR_HD double nr_peak_delta_t(const FitIn& b) {
    const double eta = b.m1 * b.m2 / (b.m1 + b.m2) / (b.m1 + b.m2);
    const double chi = 0.5 * (b.chi1 + b.chi2) + 0.5 * (b.chi1 - b.chi2) * (b.m1 - b.m2) / (b.m1 + b.m2) / (1.0 - 2.0 * eta);
    const double e2 = eta * eta, e3 = e2 * eta, c2 = chi * chi, c3 = c2 * chi;
    return 2.50499 + 13.0064 * chi + 11.5435 * c2 + 45.8838 * eta - 40.3183 * eta * chi - 19.0538 * eta * c3 + 13.0879 * e2
           + 0.192775 * e2 * c3 - 716.044 * e3;
}

// A0 + A1 eta + A2 eta^2 with the test-particle fit fTPL and the equal-mass fit
// fEQ recovered exactly (eob/seobnrv4phm.py:932-942)
IMR_HD double tpl_eqm(double eta, double A1, double fEQ, double fTPL) {
    const double A0 = -0.00099601593625498 * A1 - 0.00001600025600409607 * fEQ + 1.000016000256004 * fTPL;
    const double A2 = -3.984063745019967 * A1 + 16.00025600409607 * fEQ - 16.0002560041612 * fTPL;
    return A0 + A1 * eta + A2 * eta * eta;
}

// The spin combinations of the peak fits, in the fits' convention m1 >= m2
// (lighter body second: chiA changes sign when the masses are swapped).
struct PeakSpins {
    double eta, dM, chiS, chiA, chi;
};
IMR_HD PeakSpins peak_spins(const FitIn& b) {
    PeakSpins p;
    p.eta = b.m1 * b.m2 / ((b.m1 + b.m2) * (b.m1 + b.m2));
    p.dM = sqrt(1.0 - 4.0 * p.eta);
    p.chiS = 0.5 * (b.chi1 + b.chi2);
    p.chiA = 0.5 * (b.chi1 - b.chi2);
    double m1 = b.m1, m2 = b.m2;
    if (m1 < m2) {
        const double t = m1;
        m1 = m2;
        m2 = t;
        p.chiA = -p.chiA;
    }
    p.chi = p.chiS + p.chiA * (m1 - m2) / (m1 + m2) / (1.0 - 2.0 * p.eta);
    return p;
}

// |h_lm| at the attachment time, eob/seobnrv4phm.py:945-1075 (before the abs
// the NQC solve takes)
IMR_HD double nr_peak_amp(int l, int m, const FitIn& b) {
    const PeakSpins p = peak_spins(b);
    const double eta = p.eta, dM = p.dM, e2 = eta * eta, chi = p.chi, c2 = chi * chi, c3 = c2 * chi;
    const double chi21 = p.chiS * dM / (1.0 - 1.3 * eta) + p.chiA;
    const double chi33 = p.chiS * dM + p.chiA;
    const double chi44 = p.chiS * (1 - 5 * eta) + p.chiA * dM;
    if (l == 2 && m == 2) {
        const double fTPL = 1.4528573105413543 + 0.16613449160880395 * chi + 0.027355646661735258 * c2 - 0.020072844926136438 * c3;
        const double fEQ = 1.577457498227 - 0.0076949474494639085 * chi + 0.02188705616693344 * c2 + 0.023268366492696667 * c3;
        const double A1 = -0.03442402416125921 - 1.218066264419839 * chi - 0.5683726304811634 * c2 + 0.4011143761465342 * c3;
        return eta * tpl_eqm(eta, A1, fEQ, fTPL);
    }
    if (l == 2 && m == 1) {
        return -((0.29256703361640224 - 0.19710255145276584 * eta) * eta * chi21
                 + dM * eta * (-0.42817941710649793 + 0.11378918021042442 * eta - 0.7736772957051212 * e2
                               + chi21 * chi21 * (0.047004057952214004 - eta * 0.09326128322462478))
                 + dM * eta * chi21 * (-0.010195081244587765 + 0.016876911550777807 * chi21 * chi21));
    }
    if (l == 3 && m == 3) {
        return (0.10109183988848384 * eta - 0.4704095462146807 * e2 + 1.0735457779890183 * e2 * eta) * chi33
               + dM * (0.5636580081367962 * eta - 0.054609013952480856 * e2 + 2.3093699480319234 * e2 * eta
                       + chi33 * chi33 * (0.029812986680919126 * eta - 0.09688097244145283 * e2));
    }
    if (l == 4 && m == 4) {
        return eta * (0.2646580063832686 + 0.067584186955327 * chi44 + 0.02925102905737779 * chi44 * chi44)
               + e2 * (-0.5658246076387973 - 0.8667455348964268 * chi44 + 0.005234192027729502 * chi44 * chi44)
               + eta * e2 * (-2.5008294352355405 + 6.880772754797872 * chi44 - 1.0234651570264885 * chi44 * chi44)
               + e2 * e2 * (7.6974501716202735 - 16.551524307203252 * chi44);
    }
    if (l == 5 && m == 5) {
        return 0.128621 * dM * eta - 0.474201 * dM * eta * eta + 1.0833 * dM * eta * eta * eta + 0.0322784 * eta * chi33
               - 0.134511 * chi33 * eta * eta + 0.0990202 * chi33 * eta * eta * eta;
    }
    return NAN;
}

// d|h_lm|/dt at the attachment time, eob/seobnrv4phm.py:1078-1158
IMR_HD double nr_peak_adot(int l, int m, const FitIn& b) {
    const PeakSpins p = peak_spins(b);
    const double eta = p.eta, dM = p.dM, dM2 = dM * dM, e2 = eta * eta;
    const double chi21 = p.chiS * dM / (1.0 - 2.0 * eta) + p.chiA;
    const double chi33 = p.chiS * dM + p.chiA;
    const double chi44 = p.chiS * (1 - 7 * eta) + p.chiA * dM;
    if (l == 2 && m == 2) return 0.0;
    if (l == 2 && m == 1) {
        return dM * eta * (0.007147528020812309 - eta * 0.035644027582499495)
               + dM * eta * chi21 * (-0.0087785131749995 + eta * 0.03054672006241107)
               + eta * 0.00801714459112299 * fabs(-dM * (0.7875612917853588 + eta * 1.161274164728927 + e2 * 11.306060006923605) + chi21);
    }
    if (l == 3 && m == 3) {
        return dM * eta * (-0.00309943555972098 + eta * 0.010076527264663805) * chi33 * chi33
               + eta * 0.0016309606446766923
                     * sqrt(dM2 * (8.811660714437027 + 104.47752236009688 * eta) + dM * chi33 * (-5.352043503655119 + eta * 49.68621807460999)
                            + chi33 * chi33);
    }
    if (l == 4 && m == 4) {
        return eta * (0.004347588211099233 - 0.0014612210699052148 * chi44 - 0.002428047910361957 * chi44 * chi44)
               + e2 * (0.023320670701084355 - 0.02240684127113227 * chi44 + 0.011427087840231389 * chi44 * chi44)
               + eta * e2 * (-0.46054477257132803 + 0.433526632115367 * chi44) + e2 * e2 * (1.2796262150829425 - 1.2400051122897835 * chi44);
    }
    if (l == 5 && m == 5) {
        return eta
               * (dM * (-0.008389798844109389 + 0.04678354680410954 * eta) + dM * chi33 * (-0.0013605616383929452 + 0.004302712487297126 * eta)
                  + dM * chi33 * chi33 * (-0.0011412109287400596 + 0.0018590391891716925 * eta)
                  + 0.0002944221308683548 * fabs(dM * (37.11125499129578 - 157.79906814398277 * eta) + chi33));
    }
    return NAN;
}

// d^2|h_lm|/dt^2 at the attachment time, eob/seobnrv4phm.py:1162-1258
IMR_HD double nr_peak_addot(int l, int m, const FitIn& b) {
    const PeakSpins p = peak_spins(b);
    const double eta = p.eta, dM = p.dM, e2 = eta * eta, chi = p.chi;
    const double chi21 = p.chiS * dM / (1.0 - 2.0 * eta) + p.chiA;
    const double chi33 = p.chiS * dM + p.chiA;
    const double cm1 = -1.0 + chi;
    if (l == 2 && m == 2) {
        const double fTPL = 0.002395610769995033 * cm1 - 0.00019273850675004356 * cm1 * cm1 - 0.00029666193167435337 * cm1 * cm1 * cm1;
        const double fEQ = -0.004126509071377509 + 0.002223999138735809 * chi;
        const double A1 = -0.005776537350356959 + 0.001030857482885267 * chi;
        return eta * tpl_eqm(eta, A1, fEQ, fTPL);
    }
    if (l == 2 && m == 1) {
        return eta * dM * 0.00037132201959950333
               - fabs(dM * eta * (-0.0003650874948532221 - eta * 0.003054168419880019)
                      + dM * eta * chi21 * chi21 * (-0.0006306232037821514 - eta * 0.000868047918883389 + e2 * 0.022306229435339213)
                      + eta * chi21 * chi21 * chi21 * 0.0003402427901204342 + dM * eta * chi21 * 0.00028398490492743);
    }
    if (l == 3 && m == 3) {
        return dM * eta * (0.0009605689249339088 - 0.00019080678283595965 * eta) * chi33
               - 0.00015623760412359145 * eta
                     * fabs(dM * (4.676662024170895 + 79.20189790272218 * eta - 1097.405480250759 * e2 + 6512.959044311574 * eta * e2
                                  - 13263.36920919937 * e2 * e2)
                            + chi33);
    }
    if (l == 4 && m == 4) {
        return eta * (-0.000301722928925693 + 0.0003215952388023551 * chi) + e2 * (0.006283048344165004 + 0.0011598784110553046 * chi)
               + e2 * eta * (-0.08143521096050622 - 0.013819464720298994 * chi) + e2 * e2 * (0.22684871200570564 + 0.03275749240408555 * chi);
    }
    if (l == 5 && m == 5) {
        return eta
               * (dM * (0.00012727220842255978 + 0.0003211670856771251 * eta) + dM * chi33 * (-0.00006621677859895541 + 0.000328855327605536 * eta)
                  + chi33 * chi33 * (-0.00005824622885648688 + 0.00013944293760663706 * eta));
    }
    return NAN;
}

// the spin the frequency fits take (eob/seobnrv4phm.py:1660-1666 passes
// chiS + chiA (m1 - m2) / (m1 + m2) / (1 - 2 eta), unswapped)
IMR_HD double fit_chi(const FitIn& b) {
    const double eta = b.m1 * b.m2 / ((b.m1 + b.m2) * (b.m1 + b.m2));
    const double chiS = 0.5 * (b.chi1 + b.chi2), chiA = 0.5 * (b.chi1 - b.chi2);
    return chiS + chiA * (b.m1 - b.m2) / (b.m1 + b.m2) / (1.0 - 2.0 * eta);
}

// mode frequency at the attachment time, eob/seobnrv4phm.py:1261-1359
IMR_HD double nr_peak_omega(int l, int m, double eta, double chi) {
    const double e2 = eta * eta;
    if (l == 2 && m == 2) {
        const double c0 = 0.5626787200433265, c1 = -0.08706198756945482, c2 = 25.81979479453255, c3 = 25.85037751197443;
        const double d2 = 7.629921628648589, d3 = 10.26207326082448;
        const double A4 = d2 + 4 * (d2 - c2) * (eta - 0.25);
        const double A3 = d3 + 4 * (d3 - c3) * (eta - 0.25);
        const double c4 = 0.00174345193125868;
        return c0 + (c1 + c4 * chi) * log(A3 - A4 * chi);
    }
    if (l == 2 && m == 1) {
        return 0.1743194440996283 + eta * 0.1938944514123048 + 0.1670063050527942 * e2 + 0.053508705425291826 * chi
               - eta * chi * 0.18460213023455802 + e2 * chi * 0.2187305149636044 + chi * chi * 0.030228846150378793
               - eta * chi * chi * 0.11222178038468673;
    }
    if (l == 3 && m == 3) {
        return 0.3973947703114506 + 0.16419332207671075 * chi + 0.1635531186118689 * chi * chi + 0.06140164491786984 * chi * chi * chi
               + eta * (0.6995063984915486 - 0.3626744855912085 * chi - 0.9775469868881651 * chi * chi)
               + e2 * (-0.3455328417046369 + 0.31952307610699876 * chi + 1.9334166149686984 * chi * chi);
    }
    if (l == 4 && m == 4) {
        return 0.5389359134370971 + 0.16635177426821202 * chi + 0.2075386047689103 * chi * chi + 0.15268115749910835 * chi * chi * chi
               + eta * (0.7617423831337586 + 0.009587856087825369 * chi - 1.302303785053009 * chi * chi - 0.5562751887042064 * chi * chi * chi)
               + e2 * (0.9675153069365782 - 0.22059322127958586 * chi + 2.678097398558074 * chi * chi) - e2 * eta * 4.895381222514275;
    }
    if (l == 5 && m == 5) {
        return 0.6437545281817488 + 0.22315530037902315 * chi + 0.2956893357624277 * chi * chi + 0.17327819169083758 * chi * chi * chi
               + eta * (-0.47017798518175785 - 0.3929010618358481 * chi - 2.2653368626130654 * chi * chi - 0.5512998466154311 * chi * chi * chi)
               + e2 * (2.311483807604238 + 0.8829339243493562 * chi + 5.817595866020152 * chi * chi);
    }
    return NAN;
}

// time derivative of the mode frequency at the attachment time, eob/seobnrv4phm.py:1363-1462
IMR_HD double nr_peak_omegadot(int l, int m, double eta, double chi) {
    const double e2 = eta * eta;
    if (l == 2 && m == 2) {
        const double fTPL = -0.011209791668428353 + (0.0040867958978563915 + 0.0006333925136134493 * chi) * log(68.47466578100956 - 58.301487557007206 * chi);
        const double fEQ = 0.01128156666995859 + 0.0002869276768158971 * chi;
        const double A1 = 0.01574321112717377 + 0.02244178140869133 * chi;
        return tpl_eqm(eta, A1, fEQ, fTPL);
    }
    if (l == 2 && m == 1) {
        return 0.0070987396362959514 + eta * 0.024816844694685373 - e2 * 0.050428973182277494 + eta * e2 * 0.03442040062259341
               - chi * 0.0017751850002442097 + eta * chi * 0.004244058872768811 - e2 * chi * 0.031996494883796855
               - chi * chi * 0.0035627260615894584 + eta * chi * chi * 0.01471807973618255 - chi * chi * chi * 0.0019020967877681962;
    }
    if (l == 3 && m == 3) {
        return 0.010337157192240338 - 0.0053067782526697764 * chi * chi - 0.005087932726777773 * chi * chi * chi
               + eta * (0.027735564986787684 + 0.018864151181629343 * chi + 0.021754491131531044 * chi * chi + 0.01785477515931398 * chi * chi * chi)
               + e2 * (0.018084233854540898 - 0.08204268775495138 * chi);
    }
    if (l == 4 && m == 4) {
        return 0.013997911323773867 - 0.0051178205260273574 * chi - 0.0073874256262988 * chi * chi
               + eta * (0.0528489379269367 + 0.01632304766334543 * chi + 0.02539072293029433 * chi * chi)
               + e2 * (-0.06529992724396189 + 0.05782894076431308 * chi);
    }
    if (l == 5 && m == 5) {
        return 0.01763430670755021 - 0.00024925743340389135 * chi - 0.009240404217656968 * chi * chi - 0.007907831334704586 * chi * chi * chi
               + eta * (-0.1366002854361568 + 0.0561378177186783 * chi + 0.16406275673019852 * chi * chi + 0.07736232247880881 * chi * chi * chi)
               + e2 * (0.9875890632901151 - 0.31392112794887855 * chi - 0.5926145463423832 * chi * chi) - 1.6943356548192614 * e2 * eta;
    }
    return NAN;
}

// ------------------------------------------------------------ ringdown fits
// The four free coefficients of the ringdown ansatz as polynomials in eta and
// chi: sum over the table's (i, j) of k_ij eta^i chi^j (eob/eobstuff.py:44-322).
// Row order of each mode: {00, 01, 02, 10, 11, 12, 20, 21, 31}; absent terms 0.
struct RdTable {
    double c[9];
};
IMR_HD double rd_poly(const RdTable& t, double eta, double chi) {
    return t.c[0] + t.c[1] * chi + t.c[2] * chi * chi + t.c[3] * eta + t.c[4] * eta * chi + t.c[5] * chi * chi * eta
           + t.c[6] * eta * eta + t.c[7] * eta * eta * chi + t.c[8] * eta * eta * eta * chi;
}
// index of (l, m) among (2,2), (2,1), (3,3), (4,4), (5,5); -1 otherwise
IMR_HD int rd_mode_index(int l, int m) {
    if (l == 2 && m == 2) return 0;
    if (l == 2 && m == 1) return 1;
    if (l == 3 && m == 3) return 2;
    if (l == 4 && m == 4) return 3;
    if (l == 5 && m == 5) return 4;
    return -1;
}
// amplitude coefficient c1f (eob/eobstuff.py:44-108)
IMR_HD RdTable rd_table_a1(int k) {
    switch (k) {
        case 0: return {{0.0830664, -0.0196758, -0.0136459, 0.0612892, 0.00146142, 0.0, -0.0893454, 0.0, 0.0}};
        case 1: return {{0.07780330893915006, -0.05070638166864379, 0.0, 0.24091006920322164, 0.38582622780596576, 0.0, -0.7456327888190485,
                         -0.9695534075470388, 0.0}};
        case 2: return {{0.07638733045623343, -0.030993441267953236, 0.0, 0.2543447497371546, 0.2516879591102584, 0.0, -1.0892686061231245,
                         -0.7980907313033606, 0.0}};
        case 3: return {{-0.06392710223439678, -0.03646167590514318, 0.0, 0.345195277237925, 1.2777441574118558, 0.0, -1.764352185878576,
                         -14.825262897834696, 40.67135475479875}};
        default: return {{-0.06704614393611373, 0.021905949257025503, 0.0, -0.24754936787743445, -0.0943771022698497, 0.0, 0.7588042862093705,
                          0.4357768883690394, 0.0}};
    }
}
// amplitude coefficient c2f (eob/eobstuff.py:111-178)
IMR_HD RdTable rd_table_a2(int k) {
    switch (k) {
        case 0: return {{-0.623953, -0.371365, 0.0, 1.39777, 2.40203, 0.0, -1.82173, -5.25339, 0.0}};
        case 1: return {{-1.2451852641667298, -1.195786238319961, 0.0, 6.134202504312409, 15.66696619631313, 0.0, -14.67251127485556,
                         -44.41982201432511, 0.0}};
        case 2: return {{-0.8325292359346013, -0.598880303198448, 0.0, 2.767989795032018, 5.904371617277156, 0.0, -7.028151926115957,
                         -18.232606706124482, 0.0}};
        case 3: return {{0.7813275473485185, 0.8094706044462984, 0.0, -5.18689829943586, -5.38343327318501, 0.0, 14.026415859369477,
                         0.1051625997942299, 46.978434956814006}};
        default: return {{1.6763424265367357, 0.4925695499534606, 0.0, -5.604559311983177, -6.209095657439377, 0.0, 16.751319143123386,
                          16.778452555342554, 0.0}};
    }
}
// phase coefficient d1f (eob/eobstuff.py:181-248)
IMR_HD RdTable rd_table_p1(int k) {
    switch (k) {
        case 0: return {{0.147584, 0.00779176, -0.0244358, 0.263456, -0.120853, 0.0, -0.808987, 0.0, 0.0}};
        case 1: return {{0.15601401627613815, 0.10219957917717622, 0.0, 0.023346852452720928, -0.9435308286367039, 0.0, 0.15326558178697175,
                         1.7979082057513565, 0.0}};
        case 2: return {{0.11085299117493969, 0.018959099261813613, 0.0, 0.9999800463662053, -0.729149797691242, 0.0, -3.3983315694441125,
                         2.5192011762934037, 0.0}};
        case 3: return {{0.11498976343440313, 0.008389519706605305, 0.0, 1.6126522800609633, -0.8069979888526699, 0.0, -6.255895564079467,
                         7.595651881827078, -19.32367406125053}};
        default: return {{0.16465380962882128, -0.026574817803812007, 0.0, -0.19184523794508765, -0.05519618962738479, 0.0, 0.33328424135336965,
                          0.3194274548351241, 0.0}};
    }
}
// phase coefficient d2f (eob/eobstuff.py:251-322)
IMR_HD RdTable rd_table_p2(int k) {
    switch (k) {
        case 0: return {{2.46654, 3.13067, 0.581626, -6.99396, -9.61861, 0.0, 17.5646, 0.0, 0.0}};
        case 1: return {{2.7886287922318105, 4.29290053494256, 2.5582321247274726, -0.8145406685320334, -15.93796979597706, -10.232928498772893,
                         5.549338798935832, 12.649775582333442, 0.0}};
        case 2: return {{2.7825237371542735, 2.8796835808075003, 0.0, -7.844741660437831, -34.7670039322078, 0.0, 27.181024362399302,
                         127.13948436435182, 0.0}};
        case 3: return {{3.111817347262856, 5.399341180960216, 2.3832321567874686, 15.885333959709488, -87.92421137153823, -9.532928476043567,
                         -79.64931908155609, 657.7156442271963, -1555.2968529739226}};
        default: return {{11.102447263357977, 6.015112119742853, 0.0, -58.605776859097084, -81.68032025902797, 0.0, 176.60643662729498,
                          266.47345742836745, 0.0}};
    }
}

// ------------------------------------------------------------------ remnant
// Kerr ISCO radius (Bardeen, Press & Teukolsky 1972) for spin a in [-1, 1];
// a < 0 is the retrograde orbit
IMR_HD double kerr_isco_radius(double a) {
    const double z1 = 1.0 + cbrt(1.0 - a * a) * (cbrt(1.0 + a) + cbrt(1.0 - a));
    const double z2 = sqrt(3.0 * a * a + z1 * z1);
    const double sg = (a > 0.0) ? 1.0 : ((a < 0.0) ? -1.0 : 0.0);
    return 3.0 + z2 - sg * sqrt((3.0 - z1) * (3.0 + z1 + 2.0 * z2));
}
// specific energy of the circular orbit at the ISCO radius r: sqrt(1 - 2/(3 r))
IMR_HD double kerr_isco_energy(double r) { return sqrt(1.0 - 2.0 / (3.0 * r)); }

// Remnant mass in units of m1 + m2, eob/eobstuff.py:6-19 (SEOBNRv4's fit)
IMR_HD double final_mass_seobnrv4(double m1, double m2, double a1, double a2) {
    const double q = m1 / m2;
    const double eta = m1 * m2 / ((m1 + m2) * (m1 + m2));
    const double iq = 1.0 / q, opi = 1.0 + iq, iq2 = iq * iq;
    const double atl = (a1 + a2 * iq2) / (opi * opi);
    const double tmp = (a1 + a2 * iq2) / (1.0 + iq2);
    const double r = kerr_isco_radius(atl);
    const double e = kerr_isco_energy(r);
    (void)tmp;
    return 1.0 - ((1.0 - e) * eta + 16.0 * eta * eta * (0.00258 - 0.0773 / (1.0 / ((1.0 + iq2) / (opi * opi)) * atl - 1.6939) - 0.25 * (1.0 - e)));
}

// Remnant spin for aligned spins: Hofmann, Barausse & Rezzolla 2016 (ApJL 825
// L19), Eq. 7 with n_M = 3, n_J = 4 (their Table 1; pesummary's
// bbh_final_spin_non_precessing_HBR2016(..., version="M3J4"), which the
// reference calls at eob/seobnrv4phm.py:2459-2461):
//   a_f = a_tot + nu (L_isco(a_eff) - 2 a_tot (E_isco(a_eff) - 1))
//         + sum_{i<=3, j<=4} k_ij nu^(2+i) a_eff^j,
//   a_tot = (chi1 m1^2 + chi2 m2^2) / M^2,  a_eff = a_tot + xi nu (chi1 + chi2),
//   L_isco = 2 (1 + 2 sqrt(3 r - 2)) / (3 sqrt 3), E_isco = sqrt(1 - 2 / (3 r)).
IMR_HD double final_spin_hbr2016(double m1, double m2, double chi1, double chi2) {
    constexpr double k[4][5] = {{-5.97723, 3.39221, 4.48865, -5.77101, -13.0459},
                                {35.1278, -72.9336, -86.0036, 93.7371, 200.975},
                                {-146.822, 387.184, 447.009, -467.383, -884.339},
                                {223.911, -648.502, -697.177, 753.738, 1166.89}};
    constexpr double xi = 0.474046;
    const double M = m1 + m2;
    const double nu = m1 * m2 / (M * M);
    const double atot = (chi1 * m1 * m1 + chi2 * m2 * m2) / (M * M);
    const double aeff = atot + xi * nu * (chi1 + chi2);
    const double r = kerr_isco_radius(aeff);
    const double Lisco = 2.0 / (3.0 * sqrt(3.0)) * (1.0 + 2.0 * sqrt(3.0 * r - 2.0));
    const double Eisco = kerr_isco_energy(r);
    double sum = 0.0, nup = nu * nu;
    for (int i = 0; i < 4; ++i) {
        double aj = 1.0, s = 0.0;
        for (int j = 0; j < 5; ++j) {
            s += k[i][j] * aj;
            aj *= aeff;
        }
        sum += s * nup;
        nup *= nu;
    }
    return atot + nu * (Lisco - 2.0 * atot * (Eisco - 1.0)) + sum;
}

// ---------------------------------------------------------------- QNMs (Leaver)
// Kerr QNM frequencies of spin weight s = -2 by Leaver's method (Proc. R. Soc.
// A 402, 285 (1985)), in his units 2 M = 1 (|a| <= 1/2, time dependence
// e^{-i omega t}).  The angular separation constant A solves the angular
// continued fraction for c = a omega; omega solves the radial one.
constexpr int kLeaverAngDepth = 60;
constexpr int kLeaverRadDepth = 300;   // n = 0 converged to 5e-15 up to a = 0.98

// angular continued fraction beta_0 - alpha_0 gamma_1 / (beta_1 - ...)
IMR_HD Cx leaver_angular_cf(Cx A, Cx c, int s, int m) {
    const double k1 = 0.5 * fabs((double)(m - s)), k2 = 0.5 * fabs((double)(m + s));
    const Cx cc = c * c;
    // beta_n = n(n-1) + 2n(k1+k2+1-2c) - [2c(2k1+s+1) - (k1+k2)(k1+k2+1)] - [c^2 + s(s+1) + A]
    auto beta = [&](int n) {
        const double dn = n;
        const Cx t = cx(dn * (dn - 1.0) + 2.0 * dn * (k1 + k2 + 1.0)) - (4.0 * dn) * c;
        const Cx u = (2.0 * (2.0 * k1 + s + 1.0)) * c - (k1 + k2) * (k1 + k2 + 1.0);
        return t - u - (cc + (double)(s * (s + 1)) + A);
    };
    auto alpha = [&](int n) { return cx(-2.0 * (n + 1.0) * (n + 2.0 * k1 + 1.0)); };
    auto gamma = [&](int n) { return (2.0 * (n + k1 + k2 + s)) * c; };
    Cx t = beta(kLeaverAngDepth);
    for (int n = kLeaverAngDepth - 1; n >= 0; --n) t = beta(n) - alpha(n) * gamma(n + 1) / t;
    return t;
}

// the separation constant nearest A0 (Newton with a complex-step-free secant)
IMR_HD Cx leaver_angular_solve(Cx A0, Cx c, int s, int m) {
    Cx a = A0, b = A0 + cx(1e-3, 1e-3);
    Cx fa = leaver_angular_cf(a, c, s, m), fb = leaver_angular_cf(b, c, s, m);
    for (int it = 0; it < 60; ++it) {
        const Cx d = fb - fa;
        if (cabs2(d) == 0.0) break;
        const Cx nb = b - fb * (b - a) / d;
        a = b;
        fa = fb;
        b = nb;
        fb = leaver_angular_cf(b, c, s, m);
        if (cabs2(b - a) < 1e-28 * (1.0 + cabs2(b))) break;
    }
    return b;
}

// radial continued fraction at omega for spin a (2 M = 1), with A solved from
// the guess *A (updated)
IMR_HD Cx leaver_radial_cf(Cx w, double a, int s, int m, Cx* A) {
    *A = leaver_angular_solve(*A, a * w, s, m);
    const double b = sqrt(1.0 - 4.0 * a * a);
    const Cx I = cx(0.0, 1.0);
    const Cx wm = 0.5 * w - a * m;                // omega/2 - a m
    const Cx iwm_b = (2.0 / b) * (I * wm);        // (2i/b)(omega/2 - a m)
    const Cx c0 = cx(1.0 - s) - I * w - iwm_b;
    const Cx c1 = cx(-4.0) + (2.0 + b) * 2.0 * (I * w) + 2.0 * iwm_b;
    const Cx c2 = cx(s + 3.0) - 3.0 * (I * w) + (-1.0) * iwm_b;
    const Cx f4 = (4.0 * w + 2.0 * I) * wm * (1.0 / b);   // ((4 omega + 2i)/b)(omega/2 - a m)
    const Cx c3 = w * w * (4.0 + 2.0 * b - a * a) - (2.0 * a * m) * w - (double)(s + 1) + (2.0 + b) * (I * w) - *A + f4;
    const Cx c4 = cx(s + 1.0) - 2.0 * (w * w) - (2.0 * s + 3.0) * (I * w) - f4;
    auto alpha = [&](int n) { const double d = n; return cx(d * d + d) + (c0 * d) + c0; };
    auto beta = [&](int n) { const double d = n; return cx(-2.0 * d * d + 2.0 * d) + (c1 * d) + c3; };
    auto gamma = [&](int n) { const double d = n; return cx(d * d - 3.0 * d + 2.0) + (c2 * d) + c4 - c2; };
    Cx t = beta(kLeaverRadDepth);
    for (int n = kLeaverRadDepth - 1; n >= 0; --n) t = beta(n) - alpha(n) * gamma(n + 1) / t;
    return t;
}

// The root of the radial continued fraction near w (Leaver units) at spin a,
// with the separation constant followed from A: secant steps from w.
struct QnmRoot {
    Cx w, A;
};
IMR_HD QnmRoot leaver_refine(int m, double a, Cx w, Cx A, double dw) {
    const int s = -2;
    Cx x0 = w, x1 = w + cx(dw, -dw);
    Cx A0 = A, A1 = A;
    Cx f0 = leaver_radial_cf(x0, a, s, m, &A0);
    Cx f1 = leaver_radial_cf(x1, a, s, m, &A1);
    for (int it = 0; it < 50; ++it) {
        const Cx d = f1 - f0;
        if (cabs2(d) == 0.0) break;
        const Cx x2 = x1 - f1 * (x1 - x0) / d;
        x0 = x1;
        f0 = f1;
        A0 = A1;
        x1 = x2;
        f1 = leaver_radial_cf(x1, a, s, m, &A1);
        if (cabs2(x1 - x0) < 1e-30 * (1.0 + cabs2(x1))) break;
    }
    return QnmRoot{x1, A1};
}

// The fundamental (l, m) root at spin af (units of M, |af| <= 0.9999) followed
// from the Schwarzschild root in spin steps of 0.04; Leaver units out.  Only
// the lowest l of each m (l = max(2, |m|)) is followed reliably: for higher l
// the angular solve can settle on a lower l's separation constant.
IMR_HD QnmRoot kerr_qnm_path(int l, int m, double af) {
    const int s = -2;
    // Schwarzschild fundamental modes, l = 2..5, as the start (M omega)
    const Cx w0 = (l == 2) ? cx(0.3736716844180, -0.0889623156889)
                : (l == 3) ? cx(0.5994432884375, -0.0927030479449)
                : (l == 4) ? cx(0.8091783775322, -0.0941639609889)
                           : cx(1.0122953121354, -0.0948705160816);
    QnmRoot r{2.0 * w0, cx((double)((l - s) * (l + s + 1)))};   // Leaver units
    const double a_end = 0.5 * af;
    const int nstep = 1 + (int)(fabs(a_end) / 0.02);
    Cx wprev = r.w;
    for (int k = 1; k <= nstep; ++k) {
        const double a = a_end * (double)k / nstep;
        const Cx guess = (k > 1) ? (2.0 * r.w - wprev) : r.w;   // linear extrapolation along the path
        wprev = r.w;
        r = leaver_refine(m, a, guess, r.A, 1e-4);
    }
    return r;
}

// M omega_lm0 (M = 1, e^{-i omega t}: Im < 0) of a Kerr hole of spin af.
IMR_HD Cx kerr_qnm(int l, int m, double af) {
    if (!(fabs(af) <= 0.9999) || l < 2 || l > 5) return cx(NAN, NAN);
    const QnmRoot r = kerr_qnm_path(l, m, af);
    return cx(0.5 * r.w.re, 0.5 * r.w.im);
}

}  // namespace imr
