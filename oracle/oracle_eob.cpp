// ORACLE — test infrastructure only.  A plain CPU restatement of the
// reference's batched SEOBNR hot path, used by tests/ (and by bench.py's
// cpu_baseline leg and __graft_entry__.smoke()) as the results to compare against.  Nothing in the
// shipped eobgpu_amd library links, loads or calls this file.
//
// Each function cites the reference file:line it restates.  Parity pins (what
// checks this restatement independently of itself) are listed in the tests:
//   - Hamiltonian: Kerr test-mass limit, published 4PN A/Dbar/Q potentials and
//     NLO spin-orbit gyro factors, finite differences of H for the gradient and
//     Hessian; coefficients beyond those orders are "parity unpinned".
//   - DOP853 tableau: scipy.integrate DOP853 coefficients; error norm and
//     controller: the reference's own src/DOPR853.cu compiled unmodified into
//     oracle/_ref.
//   - spline: scipy CubicSpline(bc_type='not-a-knot'); likelihood: numpy rfft.
//   - special functions: scipy.special.loggamma / sph_harm.
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <vector>
#include <algorithm>

#include "oracle_dual.hpp"
#include "oracle_ham.hpp"
#include "oracle_flux.hpp"

using namespace orc;
typedef std::complex<double> cplx;

namespace {

const double PI = 3.141592653589793238462643383279502884;
const double EG = 0.57721566490153286060651209008240243104215933593992;
const int NADD = 80;

// ---------------------------------------------------------------- tableau
// Hairer-Wanner DOP853 (same published table as scipy.integrate DOP853 and
// src/DOPR853.cu:11-185).  Pinned against scipy by tests/test_oracle_dop853.py.
const double kC[12] = {0.0, 0.05260015195876773, 0.0789002279381516, 0.1183503419072274, 0.2816496580927726, 0.3333333333333333, 0.25, 0.3076923076923077, 0.6512820512820513, 0.6, 0.8571428571428571, 1.0};
const double kA[12][12] = {
    {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0},
    {0.05260015195876773, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0},
    {0.0197250569845379, 0.0591751709536137, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0},
    {0.02958758547680685, 0.0, 0.08876275643042054, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0},
    {0.2413651341592667, 0.0, -0.8845494793282861, 0.924834003261792, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0},
    {0.037037037037037035, 0.0, 0.0, 0.17082860872947386, 0.12546768756682242, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0},
    {0.037109375, 0.0, 0.0, 0.17025221101954405, 0.06021653898045596, -0.017578125, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0},
    {0.03709200011850479, 0.0, 0.0, 0.17038392571223998, 0.10726203044637328, -0.015319437748624402, 0.008273789163814023, 0.0, 0.0, 0.0, 0.0, 0.0},
    {0.6241109587160757, 0.0, 0.0, -3.3608926294469414, -0.868219346841726, 27.59209969944671, 20.154067550477894, -43.48988418106996, 0.0, 0.0, 0.0, 0.0},
    {0.47766253643826434, 0.0, 0.0, -2.4881146199716677, -0.590290826836843, 21.230051448181193, 15.279233632882423, -33.28821096898486, -0.020331201708508627, 0.0, 0.0, 0.0},
    {-0.9371424300859873, 0.0, 0.0, 5.186372428844064, 1.0914373489967295, -8.149787010746927, -18.52006565999696, 22.739487099350505, 2.4936055526796523, -3.0467644718982196, 0.0, 0.0},
    {2.273310147516538, 0.0, 0.0, -10.53449546673725, -2.0008720582248625, -17.9589318631188, 27.94888452941996, -2.8589982771350235, -8.87285693353063, 12.360567175794303, 0.6433927460157636, 0.0}};
const double kB[12] = {0.054293734116568765, 0.0, 0.0, 0.0, 0.0, 4.450312892752409, 1.8915178993145003, -5.801203960010585, 0.3111643669578199, -0.1521609496625161, 0.20136540080403034, 0.04471061572777259};
const double kE5[12] = {0.01312004499419488, 0.0, 0.0, 0.0, 0.0, -1.2251564463762044, -0.4957589496572502, 1.6643771824549864, -0.35032884874997366, 0.3341791187130175, 0.08192320648511571, -0.022355307863886294};
const double kBHH1 = 0.244094488188976377952755905512, kBHH2 = 0.733846688281611857341361741547, kBHH3 = 0.220588235294117647058823529412e-01;
// controller constants, src/DOPR853.cu:188-192
const double kBeta = 0.0, kAlpha = 1.0 / 8.0 - kBeta * 0.2, kSafe = 0.9, kMinScale = 1.0 / 3.0, kMaxScale = 6.0;

inline HamParams params_of(const double* add, int n, int i) {
    HamParams p;
    p.m1 = add[0 * n + i]; p.m2 = add[1 * n + i];
    p.chi1 = add[2 * n + i]; p.chi2 = add[3 * n + i];
    p.K = add[5 * n + i]; p.d5 = add[6 * n + i]; p.dSO = add[7 * n + i]; p.dSS = add[8 * n + i];
    return p;
}

typedef Dual<double, 4> D4;

// gradient of H at (r, pr, L); returns H too.  src/EOB.cu:140-222
inline double grad1(double r, double pr, double L, const HamParams& p, double g[4]) {
    D4 R = make_var<double, 4>(r, 0), PR = make_var<double, 4>(pr, 2), LL = make_var<double, 4>(L, 3);
    D4 H = hamiltonian(R, PR, LL, p);
    g[0] = H.d[0]; g[1] = 0.0; g[2] = H.d[2]; g[3] = H.d[3];
    return H.v;
}

// Newtonian prefix modulus per flux mode, stored Re/Im in add rows 10..79
inline double prefix_abs(const double* add, int n, int i, int ind_lm) {
    double re = add[(10 + 2 * ind_lm) * n + i], im = add[(10 + 2 * ind_lm + 1) * n + i];
    return std::abs(cplx(re, im));
}

// Single-lane RR force.  src/EOB.cu:2564-2642 and the flux src/EOB.cu:2506-2560.
// omega is the value the caller left in add row 4 (the ODE overwrites it with
// dH/dpphi first, src/EOB.cu:3112).
void rr_force1(const double y[4], const double* add, int n, int i, double omega, const FacCoeffs& hc, double F[2]) {
    HamParams p = params_of(add, n, i);
    const double h22_calib = add[9 * n + i];
    const double r = y[0], pr = y[2], L = y[3];
    const double M = p.m1 + p.m2, mu = p.m1 * p.m2 / M, nu = mu / M;
    double gc[4];
    grad1(r, 0.0, L, p, gc);
    const double omega_circ = gc[3];
    const double vPhiIn = 1. / (std::pow(omega_circ, 2) * std::pow(r, 3));
    const double H_val = nu * hamiltonian<double>(r, pr, L, p);
    // EOBSpinFactorizedFlux
    const double omegaSq = omega * omega;
    const double v = std::pow(omega, 1.0 / 3.);
    const double v2 = v * v;
    double vPhi = r * std::pow(vPhiIn, 1.0 / 3.);
    vPhi *= omega;
    const double vPhi2 = vPhi * vPhi;
    double flux = 0.0;
    int ind_lm = 0;
    for (int l = 2; l <= 8; ++l) {
        for (int m = 1; m <= l; ++m) {
            double pa = prefix_abs(add, n, i, ind_lm);
            double h2 = flux_mode_abs2(hc, l, m, v, H_val, L, pa, nu, vPhi2, omega, v2, h22_calib);
            flux += (m * m) * omegaSq * h2;
            ++ind_lm;
        }
    }
    flux = flux / PI / 8.0;
    flux /= nu;
    const double f_over_om = flux / omega;
    F[0] = -pr / L * f_over_om;
    F[1] = -f_over_om;
}

// ODE right-hand side for one lane, src/EOB.cu:3047-3144
void rhs1(const double y[4], const double* add, int n, int i, const FacCoeffs& hc, double k[4]) {
    HamParams p = params_of(add, n, i);
    double g[4];
    grad1(y[0], y[2], y[3], p, g);
    double F[2];
    rr_force1(y, add, n, i, g[3], hc, F);
    k[0] = g[2];
    k[1] = g[3];
    k[2] = -g[0] + F[0];
    k[3] = -g[1] + F[1];
}

struct Lane {
    const double* add; int n; int i; const FacCoeffs* hc;
};

// One DOP853 attempted step in the reference's arithmetic order
// (src/DOPR853.cu:357-517 stages, :539-618 error).  k[s][j] holds stage s.
void dopr_attempt(const Lane& ln, const double y[4], double h, double k[12][4], bool have_k1,
                  double ynew[4], double* err_out, double rtol, double atol) {
    if (!have_k1) rhs1(y, ln.add, ln.n, ln.i, *ln.hc, k[0]);
    for (int s = 1; s < 12; ++s) {
        double arg[4];
        for (int j = 0; j < 4; ++j) {
            double ak = 0.0;
            for (int q = 0; q < s; ++q) {
                if (kA[s][q] != 0.0) ak += kA[s][q] * k[q][j];
            }
            arg[j] = y[j] + h * ak;
        }
        rhs1(arg, ln.add, ln.n, ln.i, *ln.hc, k[s]);
    }
    const double nargs = 4.0;
    double err_temp = 0.0, err2 = 0.0;
    for (int j = 0; j < 4; ++j) {
        double temp = kB[0] * k[0][j] + kB[5] * k[5][j] + kB[6] * k[6][j] + kB[7] * k[7][j] + kB[8] * k[8][j] + kB[9] * k[9][j] + kB[10] * k[10][j] + kB[11] * k[11][j];
        double yn = y[j] + h * temp;
        ynew[j] = yn;
        double mx = (y[j] > yn) ? y[j] : yn;   // signed max, src/DOPR853.cu:600-602
        double sk = 1.0 / (atol + rtol * mx);
        err2 += std::pow((temp - kBHH1 * k[0][j] - kBHH2 * k[8][j] - kBHH3 * k[11][j]) * sk, 4);
        err_temp += std::pow((kE5[0] * k[0][j] + kE5[5] * k[5][j] + kE5[6] * k[6][j] + kE5[7] * k[7][j] + kE5[8] * k[8][j] + kE5[9] * k[9][j] + kE5[10] * k[10][j] + kE5[11] * k[11][j]) * sk, 4);
    }
    double denom = err_temp + 0.01 * err2;
    denom = (denom < 0.0) ? 1.0 : denom;
    *err_out = std::fabs(h) * err_temp * std::sqrt(1.0 / (nargs * denom));
}

// src/DOPR853.cu:675-762
bool controller(double err, double* errOld, bool* prevReject, double* h) {
    if (err <= 1.0) {
        double scale;
        if (err == 0.0) scale = kMaxScale;
        else {
            scale = kSafe * std::pow(err, -kAlpha) * std::pow(*errOld, kBeta);
            if (scale > kMaxScale) scale = kMaxScale;
            else if (scale < kMinScale) scale = kMinScale;
        }
        if (*prevReject && scale > 1.0) scale = 1.0;
        *h = *h * scale;
        double e = err;
        if (e < 1e-4) e = 1e-4; else if (e > 1e300) e = 1e300;
        *errOld = e;
        *prevReject = false;
        return true;
    }
    *prevReject = true;
    double t = kSafe * std::pow(err, -kAlpha);
    if (t < kMinScale) t = kMinScale; else if (t > 1e300) t = 1e300;
    *h = *h * t;
    return false;
}

}  // namespace

extern "C" {

int orc_coeff_doubles() { return (int)(sizeof(FacCoeffs) / sizeof(double)); }

// c[12], a[12][12], b[12], e5[12], bhh[3] (the product exports the same layout)
void orc_tableau(double* out) {
    int o = 0;
    for (int i = 0; i < 12; ++i) out[o++] = kC[i];
    for (int i = 0; i < 12; ++i)
        for (int j = 0; j < 12; ++j) out[o++] = kA[i][j];
    for (int i = 0; i < 12; ++i) out[o++] = kB[i];
    for (int i = 0; i < 12; ++i) out[o++] = kE5[i];
    out[o++] = kBHH1; out[o++] = kBHH2; out[o++] = kBHH3;
}

// src/EOB.cu:109-137
void orc_evaluate_H(double* out, const double* r, const double* phi, const double* pr, const double* pphi,
                    const double* m1, const double* m2, const double* chi1, const double* chi2,
                    double K, double d5, double dSO, double dSS, int n) {
    (void)phi;
    for (int i = 0; i < n; ++i) {
        HamParams p = {m1[i], m2[i], chi1[i], chi2[i], K, d5, dSO, dSS};
        out[i] = hamiltonian<double>(r[i], pr[i], pphi[i], p);
    }
}

// src/EOB.cu:224-252
void orc_grad_H(const double* arg, double* grad, const double* add, int n) {
    for (int i = 0; i < n; ++i) {
        HamParams p = params_of(add, n, i);
        double g[4];
        grad1(arg[0 * n + i], arg[2 * n + i], arg[3 * n + i], p, g);
        for (int j = 0; j < 4; ++j) grad[j * n + i] = g[j];
    }
}

// src/EOB.cu:255-401 (row-major 4x4 per lane at (a*4+b)*n+i)
void orc_hessian_H(const double* arg, double* hess, const double* add, int n) {
    typedef Dual<D4, 4> DD;
    for (int i = 0; i < n; ++i) {
        HamParams p = params_of(add, n, i);
        const int vars[3] = {0, 2, 3};
        DD X[4];
        for (int q = 0; q < 4; ++q) {
            D4 inner = make_var<double, 4>(arg[q * n + i], q);
            X[q] = DD(inner, 0);
            X[q].d[q] = D4(1.0);
        }
        DD H = hamiltonian(X[0], X[2], X[3], p);
        for (int a = 0; a < 4; ++a)
            for (int b = 0; b < 4; ++b) hess[(a * 4 + b) * n + i] = 0.0;
        for (int ia = 0; ia < 3; ++ia)
            for (int ib = 0; ib < 3; ++ib) {
                int a = vars[ia], b = vars[ib];
                hess[(a * 4 + b) * n + i] = H.d[a].d[b];
            }
    }
}

// src/EOB.cu:2464-2502; coeffs laid out as n consecutive FacCoeffs records
void orc_fac_coeffs(double* coeffs, int use_hm, const double* m1, const double* m2, const double* eta,
                    const double* a, const double* chiS, const double* chiA, int n) {
    for (int i = 0; i < n; ++i)
        calc_fac_coeffs(reinterpret_cast<FacCoeffs*>(coeffs) + i, use_hm, m1[i], m2[i], eta[i], a[i], chiS[i], chiA[i]);
}

// src/EOB.cu:2644-2662 + SEOBNRv5::RR_force_wrap
void orc_rr_force(double* force, const double* arg, const double* add, int n, const double* coeffs, const int* idx) {
    const FacCoeffs* hc = reinterpret_cast<const FacCoeffs*>(coeffs);
    for (int i = 0; i < n; ++i) {
        double y[4] = {arg[i], arg[n + i], arg[2 * n + i], arg[3 * n + i]};
        double F[2];
        rr_force1(y, add, n, i, add[4 * n + i], hc[idx[i]], F);
        force[i] = F[0]; force[n + i] = F[1];
    }
}

// src/EOB.cu:3047-3144 + :3486-3496
void orc_rhs(const double* arg, double* k, const double* add, int n, const double* coeffs, const int* idx) {
    const FacCoeffs* hc = reinterpret_cast<const FacCoeffs*>(coeffs);
    for (int i = 0; i < n; ++i) {
        double y[4] = {arg[i], arg[n + i], arg[2 * n + i], arg[3 * n + i]};
        double kk[4];
        rhs1(y, add, n, i, hc[idx[i]], kk);
        for (int j = 0; j < 4; ++j) k[j * n + i] = kk[j];
    }
}

// One attempted DOP853 step per lane (src/DOPR853.cu:357-618): the twelve
// stages kout [12][4][n], y_new [4][n] and the error norm err[n].
void orc_dopr_attempt(const double* y, const double* h, const double* add, int n, const double* coeffs, double* kout, double* ynew,
                      double* err, double rtol, double atol) {
    const FacCoeffs* hc = reinterpret_cast<const FacCoeffs*>(coeffs);
    for (int i = 0; i < n; ++i) {
        Lane ln = {add, n, i, &hc[i]};
        double yl[4] = {y[i], y[n + i], y[2 * n + i], y[3 * n + i]};
        double k[12][4], yn[4];
        dopr_attempt(ln, yl, h[i], k, false, yn, &err[i], rtol, atol);
        for (int s = 0; s < 12; ++s)
            for (int j = 0; j < 4; ++j) kout[((size_t)s * 4 + j) * n + i] = k[s][j];
        for (int j = 0; j < 4; ++j) ynew[(size_t)j * n + i] = yn[j];
    }
}

// error norm restated from src/DOPR853.cu:539-618 (same argument order as error_wrap)
void orc_error(double* err, const double* solOld, double* solNew, const double* h, const double* k1, const double* k2,
               const double* k3, const double* k6, const double* k7, const double* k8, const double* k9, const double* k10,
               int numEq, int nargs, double rtol, double atol) {
    for (int e = 0; e < numEq; ++e) {
        double et = 0.0, e2 = 0.0;
        for (int j = 0; j < nargs; ++j) {
            int x = j * numEq + e;
            double temp = kB[0] * k1[x] + kB[5] * k6[x] + kB[6] * k7[x] + kB[7] * k8[x] + kB[8] * k9[x] + kB[9] * k10[x] + kB[10] * k2[x] + kB[11] * k3[x];
            double yn = solOld[x] + h[e] * temp;
            solNew[x] = yn;
            double mx = (solOld[x] > yn) ? solOld[x] : yn;
            double sk = 1.0 / (atol + rtol * mx);
            e2 += std::pow((temp - kBHH1 * k1[x] - kBHH2 * k9[x] - kBHH3 * k3[x]) * sk, nargs);
            et += std::pow((kE5[0] * k1[x] + kE5[5] * k6[x] + kE5[6] * k7[x] + kE5[7] * k8[x] + kE5[8] * k9[x] + kE5[9] * k10[x] + kE5[10] * k2[x] + kE5[11] * k3[x]) * sk, nargs);
        }
        double denom = et + 0.01 * e2;
        denom = (denom < 0.0) ? 1.0 : denom;
        err[e] = std::fabs(h[e]) * et * std::sqrt(1.0 / ((double)nargs * denom));
    }
}

// src/DOPR853.cu:675-762 (flags as 1-byte bools)
void orc_controller(unsigned char* flag, const double* err, double* errOld, unsigned char* prevReject, double* h, int numEq) {
    for (int e = 0; e < numEq; ++e) {
        bool pr = prevReject[e] != 0;
        bool ok = controller(err[e], &errOld[e], &pr, &h[e]);
        prevReject[e] = pr; flag[e] = ok;
    }
}

// The DOPR853.integrate driver with the StoppingCriterion, for a batch of n
// lanes, restated from eob/new_Dormand_horiz.py:424-778 and
// eob/seobnrv4phm.py:61-150.
//
// batch_coupled = 1 reproduces the reference's batch semantics exactly,
// including how StoppingCriterion indexes old_omegas with positions inside
// index_update (eob/seobnrv4phm.py:124-138).  batch_coupled = 0 integrates each
// lane as if it were alone in the batch (the same code with n = 1 per lane).
//
// denseOut: [max_step][4][n], denseLoc: [max_step][n], step_num: [n].
// restart = 0: fresh start from condBound (rows 0 are written);
// restart = 1: continue from row step_num[i] of denseOut/denseLoc.
// Returns the number of driver iterations.
long orc_integrate(const double* condBound, const double* add, const double* coeffs, int n,
                   double hInit, int max_step, double tmax, int fix_step, double rtol, double atol,
                   int* step_num, double* denseOut, double* denseLoc, int restart, int* stop_info,
                   int batch_coupled) {
    const FacCoeffs* hc = reinterpret_cast<const FacCoeffs*>(coeffs);
    if (!batch_coupled && n > 1) {
        long iters = 0;
        std::vector<double> cb(4), ad(NADD), dO((size_t)max_step * 4), dL(max_step);
        for (int i = 0; i < n; ++i) {
            for (int j = 0; j < 4; ++j) cb[j] = condBound[j * n + i];
            for (int j = 0; j < NADD; ++j) ad[j] = add[j * n + i];
            for (int s = 0; s < max_step; ++s) {
                dL[s] = denseLoc[(size_t)s * n + i];
                for (int j = 0; j < 4; ++j) dO[(size_t)s * 4 + j] = denseOut[((size_t)s * 4 + j) * n + i];
            }
            int sn = step_num[i], si = 0;
            long it = orc_integrate(cb.data(), ad.data(), reinterpret_cast<const double*>(hc + i), 1, hInit, max_step, tmax,
                                    fix_step, rtol, atol, &sn, dO.data(), dL.data(), restart, &si, 1);
            iters = std::max(iters, it);
            step_num[i] = sn; stop_info[i] = si;
            for (int s = 0; s < max_step; ++s) {
                denseLoc[(size_t)s * n + i] = dL[s];
                for (int j = 0; j < 4; ++j) denseOut[((size_t)s * 4 + j) * n + i] = dO[(size_t)s * 4 + j];
            }
        }
        return iters;
    }

    std::vector<double> x(n), h(n), y((size_t)4 * n), errOld(n, 1e-4), hInitOrig(n, hInit);
    std::vector<char> prevReject(n, 0), active(n, 1), stop(n, 1), failed(n, 0);
    std::vector<long long> attempts(n, 0);
    const long long max_attempts = 16LL * max_step + 4096;
    std::vector<double> old_omegas((size_t)max_step * n, 0.0);
    for (int i = 0; i < n; ++i) {
        if (!restart) {
            step_num[i] = 0;
            x[i] = 0.0;
            for (int j = 0; j < 4; ++j) { y[j * n + i] = condBound[j * n + i]; denseOut[(size_t)j * n + i] = y[j * n + i]; }
            denseLoc[i] = 0.0;
        } else {
            int s = step_num[i];
            if (s < 0) s += max_step;  // Python negative index
            x[i] = denseLoc[(size_t)s * n + i];
            for (int j = 0; j < 4; ++j) y[j * n + i] = denseOut[((size_t)s * 4 + j) * n + i];
        }
        h[i] = hInit;
        stop_info[i] = 0;
    }
    long iter = 0;
    std::vector<int> index_update;
    std::vector<double> knew((size_t)4 * n);
    while (true) {
        index_update.clear();
        for (int i = 0; i < n; ++i) {
            if (!active[i]) continue;
            Lane ln = {add, n, i, &hc[i]};
            double yl[4], yn[4], k[12][4], err;
            for (int j = 0; j < 4; ++j) yl[j] = y[j * n + i];
            double hOld = h[i];
            dopr_attempt(ln, yl, h[i], k, false, yn, &err, rtol, atol);
            bool pr = prevReject[i] != 0;
            bool ok = controller(err, &errOld[i], &pr, &h[i]);
            prevReject[i] = pr;
            if (fix_step && !ok) hInitOrig[i] /= 2;
            if (fix_step) h[i] = hInitOrig[i];
            // A non-finite error or step (or a step halved to 0) would make the
            // reference's loop run forever; the lane ends with stop code -1
            // (-2 after max_attempts attempts), as the integrator kernel does.
            if (!(err >= 0.0) || !(std::fabs(h[i]) < 1e300) || h[i] == 0.0) failed[i] = 1;
            else if (++attempts[i] >= max_attempts) failed[i] = 2;
            if (ok) {
                for (int j = 0; j < 4; ++j) y[j * n + i] = yn[j];
                x[i] = x[i] + hOld;
                index_update.push_back(i);
            }
        }
        // dense output and stopping criterion for the accepted lanes
        const int nu_ = (int)index_update.size();
        for (int q = 0; q < nu_; ++q) {
            int i = index_update[q];
            step_num[i] += 1;
            int s = step_num[i];
            for (int j = 0; j < 4; ++j) denseOut[((size_t)s * 4 + j) * n + i] = y[j * n + i];
            denseLoc[(size_t)s * n + i] = x[i];
        }
        if (nu_ > 0) {
            // StoppingCriterion.__call__ (eob/seobnrv4phm.py:82-150)
            std::vector<double> rr(nu_), drdt(nu_), om(nu_), dprdt(nu_);
            std::vector<int> code(nu_, 0);
            for (int q = 0; q < nu_; ++q) {
                int i = index_update[q];
                double yl[4] = {y[i], y[n + i], y[2 * n + i], y[3 * n + i]};
                double kk[4];
                rhs1(yl, add, n, i, hc[i], kk);
                rr[q] = yl[0]; drdt[q] = kk[0]; om[q] = kk[1]; dprdt[q] = kk[2];
            }
            std::vector<char> inds_here(nu_, 0);
            int any_here = 0;
            for (int q = 0; q < nu_; ++q) {
                if (rr[q] < 1.0) code[q] = 2;
                else if (rr[q] < 6.0 && drdt[q] > 0.0) code[q] = 3;
                else if (rr[q] < 6.0 && dprdt[q] > 0.0) code[q] = 4;
                inds_here[q] = (code[q] == 0) && (step_num[index_update[q]] > 5);
                any_here |= inds_here[q];
            }
            std::vector<int> still_have;
            if (any_here) {
                for (int q = 0; q < nu_; ++q) if (inds_here[q]) still_have.push_back(q);
                for (int w = 0; w < (int)still_have.size(); ++w) {
                    int q = still_have[w];
                    int i = index_update[q];
                    int s = step_num[i];
                    bool first = rr[q] < 6.0;
                    bool second = om[q] < old_omegas[(size_t)(s - 1) * n + i];
                    // third check reads lane column `q` (a position inside index_update), as the reference does
                    int col = batch_coupled ? q : i;
                    double o[4];
                    for (int t = 0; t < 4; ++t) o[t] = old_omegas[(size_t)(s - 4 + t) * n + col];
                    bool third = (o[1] - o[0] < 0.0) && (o[2] - o[1] < 0.0) && (o[3] - o[2] < 0.0);
                    if (first && second && third) code[q] += 5;
                }
                for (int w = 0; w < (int)still_have.size(); ++w) {
                    int q = still_have[w];
                    int i = index_update[q];
                    old_omegas[(size_t)step_num[i] * n + i] = om[q];
                }
            } else {
                for (int q = 0; q < nu_; ++q) {
                    int i = index_update[q];
                    old_omegas[(size_t)step_num[i] * n + i] = om[q];
                }
            }
            for (int q = 0; q < nu_; ++q) {
                int i = index_update[q];
                stop_info[i] = code[q];
                stop[i] = code[q] != 0;
            }
        }
        bool any = false;
        for (int i = 0; i < n; ++i) {
            if (!active[i]) continue;
            if (failed[i]) stop_info[i] = -failed[i];
            if (x[i] >= tmax || step_num[i] >= max_step - 1 || stop[i] || failed[i]) active[i] = 0;
            any |= active[i] != 0;
        }
        ++iter;
        if (!any) break;
    }
    return iter;
}

// SEOBNRv4PHM.run_trajectory (eob/seobnrv4phm.py:2072-2136): an adaptive pass,
// then a walk back by step_back (in M) and a fixed-step pass of fine_step.
// t_out: [max_step][n] (units of M), traj: [max_step][4][n]; num_steps[n] is
// the number of stored points (accepted steps of pass 2 + 1).  fix_step1 = 1
// runs the first pass at fixed step hInit (the reference's pass 1 is adaptive).
void orc_run_trajectory(const double* condBound, const double* add, const double* coeffs, int n,
                        int max_step, double rtol, double atol, double hInit, double step_back, double fine_step,
                        double* t_out, double* traj, int* num_steps, int* stop_info, int batch_coupled, int fix_step1) {
    std::vector<int> sn(n, 0);
    std::memset(t_out, 0, sizeof(double) * (size_t)max_step * n);
    std::memset(traj, 0, sizeof(double) * (size_t)max_step * 4 * n);
    orc_integrate(condBound, add, coeffs, n, hInit, max_step, 1e13, fix_step1, rtol, atol, sn.data(), traj, t_out, 0, stop_info,
                  batch_coupled);
    std::vector<int> keep;                  // lanes that go on to the fine pass
    for (int i = 0; i < n; ++i) {
        if (stop_info[i] < 0) continue;     // ended by the non-finite guard: no second pass
        keep.push_back(i);
        // t_stop is t[num_steps - 1]: the reference finds the last non-zero row of
        // t[1:] and then indexes the unshifted t with it (eob/seobnrv4phm.py:2095-2097)
        int last = sn[i] - 1;
        if (last < 0) last = 0;
        double t_stop = t_out[(size_t)last * n + i];
        double t_desired = t_stop - step_back;
        int best = 0;
        double bestv = std::fabs(t_out[i] - t_desired);
        for (int s = 1; s < max_step; ++s) {
            double d = std::fabs(t_out[(size_t)s * n + i] - t_desired);
            if (d < bestv) { bestv = d; best = s; }
        }
        for (int s = best; s < max_step; ++s) {
            t_out[(size_t)s * n + i] = 0.0;
            for (int j = 0; j < 4; ++j) traj[((size_t)s * 4 + j) * n + i] = 0.0;
        }
        sn[i] = best - 1;
    }
    if ((int)keep.size() == n) {
        orc_integrate(condBound, add, coeffs, n, fine_step, max_step, 1e13, 1, rtol, atol, sn.data(), traj, t_out, 1, stop_info,
                      batch_coupled);
    } else if (!keep.empty()) {
        // run the fine pass on the kept lanes only (gather, integrate, scatter)
        const int m = (int)keep.size();
        std::vector<double> cb((size_t)4 * m), ad((size_t)NADD * m), to((size_t)max_step * m), tr((size_t)max_step * 4 * m),
            co((size_t)m * sizeof(FacCoeffs) / sizeof(double));
        std::vector<int> sm(m), sim(m);
        const FacCoeffs* hc = reinterpret_cast<const FacCoeffs*>(coeffs);
        for (int q = 0; q < m; ++q) {
            const int i = keep[q];
            for (int j = 0; j < 4; ++j) cb[(size_t)j * m + q] = condBound[(size_t)j * n + i];
            for (int j = 0; j < NADD; ++j) ad[(size_t)j * m + q] = add[(size_t)j * n + i];
            for (int s = 0; s < max_step; ++s) {
                to[(size_t)s * m + q] = t_out[(size_t)s * n + i];
                for (int j = 0; j < 4; ++j) tr[((size_t)s * 4 + j) * m + q] = traj[((size_t)s * 4 + j) * n + i];
            }
            reinterpret_cast<FacCoeffs*>(co.data())[q] = hc[i];
            sm[q] = sn[i];
        }
        orc_integrate(cb.data(), ad.data(), co.data(), m, fine_step, max_step, 1e13, 1, rtol, atol, sm.data(), tr.data(), to.data(), 1,
                      sim.data(), batch_coupled);
        for (int q = 0; q < m; ++q) {
            const int i = keep[q];
            for (int s = 0; s < max_step; ++s) {
                t_out[(size_t)s * n + i] = to[(size_t)s * m + q];
                for (int j = 0; j < 4; ++j) traj[((size_t)s * 4 + j) * n + i] = tr[((size_t)s * 4 + j) * m + q];
            }
            sn[i] = sm[q];
            stop_info[i] = sim[q];
        }
    }
    for (int i = 0; i < n; ++i) num_steps[i] = sn[i] + 1;
}

// ------------------------------------------------------------------ waveform
// principal-branch-free complex Gamma via lnGamma(z+N) (Stirling) and the
// recurrence; only exp(lnGamma) enters Tlm, so any branch gives the same value.
static cplx cgamma(cplx z) {
    cplx prod = 1.0;
    while (z.real() < 10.0) { prod *= z; z += 1.0; }
    cplx iz = 1.0 / z, iz2 = iz * iz;
    // Stirling: (z-1/2)ln z - z + ln(2pi)/2 + sum B_2k/(2k(2k-1) z^(2k-1))
    const double b[8] = {1.0 / 12, -1.0 / 360, 1.0 / 1260, -1.0 / 1680, 1.0 / 1188, -691.0 / 360360, 1.0 / 156, -3617.0 / 122400};
    cplx ser = 0.0, pw = iz;
    for (int k = 0; k < 8; ++k) { ser += b[k] * pw; pw *= iz2; }
    cplx lg = (z - 0.5) * std::log(z) - z + 0.5 * std::log(2 * PI) + ser;
    return std::exp(lg) / prod;
}

extern "C" void orc_cgamma(double re, double im, double* out) {
    cplx g = cgamma(cplx(re, im));
    out[0] = g.real(); out[1] = g.imag();
}

// Y_l^m(theta = pi/2, phi) with the Condon-Shortley phase (scipy sph_harm(m, l, phi, pi/2))
static cplx ylm_equator(int l, int m, double phi) {
    int am = std::abs(m);
    if ((l + am) % 2) return 0.0;
    double fl = 1.0, fh = 1.0;
    for (int i = 2; i <= l - am; ++i) fl *= i;
    for (int i = 2; i <= l + am; ++i) fh *= i;
    double dfa = 1.0, dfb = 1.0;
    for (int i = l + am - 1; i > 1; i -= 2) dfa *= i;
    for (int i = l - am; i > 1; i -= 2) dfb *= i;
    double P = dfa / dfb;                        // |P_l^|m|(0)|
    if (((l + am) / 2) % 2) P = -P;              // (-1)^((l+|m|)/2), Condon-Shortley phase included
    double val = std::sqrt((2.0 * l + 1.0) / (4.0 * PI) * fl / fh) * P;
    if (m < 0 && (am % 2)) val = -val;           // Y_l^{-m} = (-1)^m conj(Y_l^m)
    return val * std::exp(cplx(0.0, m * phi));
}

extern "C" void orc_ylm_equator(int l, int m, double phi, double* out) {
    cplx y = ylm_equator(l, m, phi);
    out[0] = y.real(); out[1] = y.imag();
}

// (2,2) mode on trajectory knots, src/EOB.cu:1236-1613 for l = m = 2.
static cplx hlm22(double r, double phi, double v, double Hreal, const FacCoeffs& hc, cplx prefix, double eta,
                  double vPhiInput, double h22_calib) {
    const int l = 2, m = 2;
    if (eta > 0.25 && eta < 0.25 + 1e-4) eta = 0.25;
    const double v2 = v * v;
    const double Omega = v2 * v;
    const double vh3 = Hreal * Omega;
    const double vh = std::pow(vh3, 1. / 3.);
    const double elog = EG + std::log(2.0 * m * v);
    double vPhi = r * std::pow(vPhiInput, 1.0 / 3.);
    vPhi *= Omega;
    const double vPhi2 = vPhi * vPhi;
    // Newtonian multipole, EOBFluxCalculateNewtonianMultipole (src/EOB.cu:1012-1027)
    cplx hN = prefix * std::pow(vPhi2, (l + 0) / 2.0) * ylm_equator(l, -m, phi);
    const double Slm = (Hreal * Hreal - 1.0) / (2.0 * eta) + 1.0;
    const double k = m * Omega;
    const double hk = Hreal * k;
    const double z2 = std::tgamma((double)(l + 1));
    cplx G = cgamma(cplx(l + 1.0, -2.0 * hk));
    cplx Tlm = std::exp(cplx(PI * hk, 2.0 * hk * std::log(4.0 * k / std::sqrt(std::exp(1.0))))) * G;
    Tlm /= z2;
    const double delta = vh3 * (hc.delta22vh3 + vh3 * (hc.delta22vh6 + vh * vh * (hc.delta22vh9 * vh)))
                         + hc.delta22v5 * v * v2 * v2 + hc.delta22v6 * v2 * v2 * v2 + hc.delta22v8 * v2 * v2 * v2 * v2;
    const double rho = 1. + v2 * (hc.rho22v2 + v * (hc.rho22v3 + v * (hc.rho22v4 + v * (hc.rho22v5 + v * (hc.rho22v6 + hc.rho22v6l * elog + v * (hc.rho22v7 + v * (hc.rho22v8 + hc.rho22v8l * elog + (hc.rho22v10 + hc.rho22v10l * elog) * v2))))))) + h22_calib * std::pow(v, 12);
    cplx rpl = std::pow(rho, l);
    cplx h = Tlm * std::exp(cplx(0.0, delta)) * Slm * rpl;
    return h * hN;
}

// EOBGetSpinFactorizedWaveform (src/EOB.cu:1236-1613) for every (l, m), l = 2..8,
// m = 1..l: h_lm = T_lm e^{i delta_lm} S_lm (rho_lm^l + f_lm) h^N_lm, with rho^l
// replaced by f_lm alone for odd m at eta = 1/4; S = H_eff for even l + m, v pphi
// for odd; h^N = prefix vPhi2^((l+eps)/2) Y_{l-eps}^{-m}(pi/2, phi).  Each mode's
// delta, rho and f are the nested polynomials of the reference, term for term.
static cplx hlm_mode(int l, int m, double r, double phi, double pphi, double v, double Hreal, const FacCoeffs& hc, cplx prefix,
                     double eta, double vPhiInput, double h22_calib) {
    if (eta > 0.25 && eta < 0.25 + 1e-4) eta = 0.25;
    const cplx I(0.0, 1.0);
    const double v2 = v * v, Omega = v2 * v, vh3 = Hreal * Omega, vh = std::pow(vh3, 1. / 3.);
    const double el = EG + std::log(2.0 * m * v);
    double vPhi = r * std::pow(vPhiInput, 1.0 / 3.);
    vPhi *= Omega;
    const double vPhi2 = vPhi * vPhi;
    const int eps = (l + m) % 2;
    const cplx hN = prefix * std::pow(vPhi2, (l + eps) / 2.0) * ylm_equator(l - eps, -m, phi);
    const double Slm = (eps == 0) ? (Hreal * Hreal - 1.0) / (2.0 * eta) + 1.0 : v * pphi;
    const double k = m * Omega, hk = Hreal * k;
    const double z2 = std::tgamma((double)(l + 1));
    cplx Tlm = std::exp(cplx(PI * hk, 2.0 * hk * std::log(4.0 * k / std::sqrt(std::exp(1.0))))) * cgamma(cplx(l + 1.0, -2.0 * hk));
    Tlm /= z2;
    const FacCoeffs& c = hc;
    double d = 0.0, rho = 1.0;
    cplx f = 0.0;
    const int lm = 10 * l + m;
    switch (lm) {
    case 22:
        d = vh3 * (c.delta22vh3 + vh3 * (c.delta22vh6 + vh * vh * (c.delta22vh9 * vh))) + c.delta22v5 * v * v2 * v2 + c.delta22v6 * v2 * v2 * v2
            + c.delta22v8 * v2 * v2 * v2 * v2;
        rho = 1. + v2 * (c.rho22v2 + v * (c.rho22v3 + v * (c.rho22v4 + v * (c.rho22v5 + v * (c.rho22v6 + c.rho22v6l * el + v * (c.rho22v7
              + v * (c.rho22v8 + c.rho22v8l * el + (c.rho22v10 + c.rho22v10l * el) * v2))))))) + h22_calib * std::pow(v, 12);
        break;
    case 21:
        d = vh3 * (c.delta21vh3 + vh3 * (c.delta21vh6 + vh * (c.delta21vh7 + c.delta21vh9 * vh * vh))) + c.delta21v5 * v * v2 * v2
            + c.delta21v7 * v2 * v2 * v2 * v;
        rho = 1. + v * (c.rho21v1 + v * (c.rho21v2 + v * (c.rho21v3 + v * (c.rho21v4 + v * (c.rho21v5 + v * (c.rho21v6 + c.rho21v6l * el
              + v * (c.rho21v7 + c.rho21v7l * el + v * (c.rho21v8 + c.rho21v8l * el + (c.rho21v10 + c.rho21v10l * el) * v2))))))));
        f = v * (c.f21v1 + v2 * (c.f21v3 + v * c.f21v4 + v2 * (c.f21v5 + v * c.f21v6 + v2 * c.f21v7c)));
        break;
    case 33:
        d = vh3 * (c.delta33vh3 + vh3 * (c.delta33vh6 + c.delta33vh9 * vh3)) + c.delta33v5 * v * v2 * v2;
        rho = 1. + v2 * (c.rho33v2 + v * (c.rho33v3 + v * (c.rho33v4 + v * (c.rho33v5 + v * (c.rho33v6 + c.rho33v6l * el
              + v * (c.rho33v7 + v * (c.rho33v8 + c.rho33v8l * el + v2 * (c.rho33v10 + c.rho33v10l * el))))))));
        f = v * (v2 * (c.f33v3 + v * (c.f33v4 + v * (c.f33v5 + v * c.f33v6)))) + I * vh3 * vh3 * c.f33vh6;
        break;
    case 32:
        d = vh3 * (c.delta32vh3 + vh * (c.delta32vh4 + vh * vh * (c.delta32vh6 + c.delta32vh9 * vh3)));
        rho = 1. + v * (c.rho32v + v * (c.rho32v2 + v * (c.rho32v3 + v * (c.rho32v4 + v * (c.rho32v5 + v * (c.rho32v6 + c.rho32v6l * el
              + (c.rho32v8 + c.rho32v8l * el) * v2))))));
        break;
    case 31:
        d = vh3 * (c.delta31vh3 + vh3 * (c.delta31vh6 + vh * (c.delta31vh7 + c.delta31vh9 * vh * vh))) + c.delta31v5 * v * v2 * v2;
        rho = 1. + v2 * (c.rho31v2 + v * (c.rho31v3 + v * (c.rho31v4 + v * (c.rho31v5 + v * (c.rho31v6 + c.rho31v6l * el
              + v * (c.rho31v7 + (c.rho31v8 + c.rho31v8l * el) * v))))));
        f = v * v2 * c.f31v3;
        break;
    case 44:
        d = vh3 * (c.delta44vh3 + vh3 * (c.delta44vh6 + vh3 * c.delta44vh9)) + c.delta44v5 * v2 * v2 * v;
        rho = 1. + v2 * (c.rho44v2 + v * (c.rho44v3 + v * (c.rho44v4 + v * (c.rho44v5 + v * (c.rho44v6 + c.rho44v6l * el
              + v2 * (c.rho44v8 + c.rho44v8l * el + v2 * (c.rho44v10 + c.rho44v10l * el)))))));
        break;
    case 43:
        d = vh3 * (c.delta43vh3 + vh * (c.delta43vh4 + c.delta43vh6 * vh * vh));
        rho = 1. + v * (c.rho43v + v * (c.rho43v2 + v2 * (c.rho43v4 + v * (c.rho43v5 + (c.rho43v6 + c.rho43v6l * el) * v))));
        f = v * c.f43v;
        break;
    case 42:
        d = vh3 * (c.delta42vh3 + c.delta42vh6 * vh3);
        rho = 1. + v2 * (c.rho42v2 + v * (c.rho42v3 + v * (c.rho42v4 + v * (c.rho42v5 + (c.rho42v6 + c.rho42v6l * el) * v))));
        break;
    case 41:
        d = vh3 * (c.delta41vh3 + vh * (c.delta41vh4 + c.delta41vh6 * vh * vh));
        rho = 1. + v * (c.rho41v + v * (c.rho41v2 + v2 * (c.rho41v4 + v * (c.rho41v5 + (c.rho41v6 + c.rho41v6l * el) * v))));
        f = v * c.f41v;
        break;
    case 55:
        d = vh3 * (c.delta55vh3 + vh3 * (c.delta55vh6 + vh3 * c.delta55vh9)) + c.delta55v5 * v2 * v2 * v;
        rho = 1. + v2 * (c.rho55v2 + v * (c.rho55v3 + v * (c.rho55v4 + v * (c.rho55v5 + v * (c.rho55v6 + c.rho55v6l * el
              + v2 * (c.rho55v8 + c.rho55v8l * el + v2 * (c.rho55v10 + c.rho55v10l * el)))))));
        f = v2 * v * (c.f55v3 + v * (c.f55v4 + v * c.f55v5c));
        break;
    case 54: d = vh3 * (c.delta54vh3 + c.delta54vh4 * vh); rho = 1. + v2 * (c.rho54v2 + v * (c.rho54v3 + c.rho54v4 * v)); break;
    case 53: d = c.delta53vh3 * vh3; rho = 1. + v2 * (c.rho53v2 + v * (c.rho53v3 + v * (c.rho53v4 + c.rho53v5 * v))); break;
    case 52: d = vh3 * (c.delta52vh3 + c.delta52vh4 * vh); rho = 1. + v2 * (c.rho52v2 + v * (c.rho52v3 + c.rho52v4 * v)); break;
    case 51: d = c.delta51vh3 * vh3; rho = 1. + v2 * (c.rho51v2 + v * (c.rho51v3 + v * (c.rho51v4 + c.rho51v5 * v))); break;
    case 66: d = c.delta66vh3 * vh3; rho = 1. + v2 * (c.rho66v2 + v * (c.rho66v3 + c.rho66v4 * v)); break;
    case 65: d = c.delta65vh3 * vh3; rho = 1. + v2 * (c.rho65v2 + c.rho65v3 * v); break;
    case 64: d = c.delta64vh3 * vh3; rho = 1. + v2 * (c.rho64v2 + v * (c.rho64v3 + c.rho64v4 * v)); break;
    case 63: d = c.delta63vh3 * vh3; rho = 1. + v2 * (c.rho63v2 + c.rho63v3 * v); break;
    case 62: d = c.delta62vh3 * vh3; rho = 1. + v2 * (c.rho62v2 + v * (c.rho62v3 + c.rho62v4 * v)); break;
    case 61: d = c.delta61vh3 * vh3; rho = 1. + v2 * (c.rho61v2 + c.rho61v3 * v); break;
    case 77: d = c.delta77vh3 * vh3; rho = 1. + v2 * (c.rho77v2 + c.rho77v3 * v); break;
    case 76: rho = 1. + c.rho76v2 * v2; break;
    case 75: d = c.delta75vh3 * vh3; rho = 1. + v2 * (c.rho75v2 + c.rho75v3 * v); break;
    case 74: rho = 1. + c.rho74v2 * v2; break;
    case 73: d = c.delta73vh3 * vh3; rho = 1. + v2 * (c.rho73v2 + c.rho73v3 * v); break;
    case 72: rho = 1. + c.rho72v2 * v2; break;
    case 71: d = c.delta71vh3 * vh3; rho = 1. + v2 * (c.rho71v2 + c.rho71v3 * v); break;
    case 88: rho = 1. + c.rho88v2 * v2; break;
    case 87: rho = 1. + c.rho87v2 * v2; break;
    case 86: rho = 1. + c.rho86v2 * v2; break;
    case 85: rho = 1. + c.rho85v2 * v2; break;
    case 84: rho = 1. + c.rho84v2 * v2; break;
    case 83: rho = 1. + c.rho83v2 * v2; break;
    case 82: rho = 1. + c.rho82v2 * v2; break;
    case 81: rho = 1. + c.rho81v2 * v2; break;
    default: return cplx(NAN, NAN);
    }
    cplx rpl = (eta == 0.25 && (m % 2)) ? f : cplx(std::pow(rho, l), 0.0) + f;
    cplx h = Tlm * std::exp(I * d) * Slm * rpl;
    return h * hN;
}

// |h_lm|^2 as the flux evaluates it (EOBFluxGetSpinFactorizedWaveform,
// src/EOB.cu:1706-1941) at npts points of one binary (coefficient record coeffs)
void orc_flux_mode_abs2(double* out, int l, int m, const double* v, const double* Hreal, const double* pphi, const double* vPhi2,
                        int npts, double prefix_abs, double eta, double h22_calib, const double* coeffs) {
    const FacCoeffs& hc = *reinterpret_cast<const FacCoeffs*>(coeffs);
    for (int i = 0; i < npts; ++i) {
        const double vv = v[i];
        out[i] = flux_mode_abs2(hc, l, m, vv, Hreal[i], pphi[i], prefix_abs, eta, vPhi2[i], vv * vv * vv, vv * vv, h22_calib);
    }
}

// EOBGetSpinFactorizedWaveform_wrap for any (l, m) with 2 <= l <= 8, 1 <= m <= l:
// inputs per knot [numSys][traj_length]; prefix complex per binary.
void orc_hlm(double* out, int l, int m, const double* r, const double* phi, const double* pphi, const double* v, const double* Hreal,
             const double* eta, const double* vPhiIn, const double* prefix, double h22_calib, int numSys, int traj_length,
             const double* coeffs, const int* idx) {
    const FacCoeffs* hc = reinterpret_cast<const FacCoeffs*>(coeffs);
    for (int b = 0; b < numSys; ++b) {
        cplx pf(prefix[2 * b], prefix[2 * b + 1]);
        for (int s = 0; s < traj_length; ++s) {
            size_t q = (size_t)b * traj_length + s;
            cplx h = hlm_mode(l, m, r[q], phi[q], pphi[q], v[q], Hreal[q], hc[idx[b]], pf, eta[b], vPhiIn[q], h22_calib);
            out[2 * q] = h.real(); out[2 * q + 1] = h.imag();
        }
    }
}

// EOBGetSpinFactorizedWaveform_wrap restricted to (l,m) = (2,2),
// src/EOB.cu:1616-1677: inputs per knot [numSys][traj_length].
void orc_hlm22(double* out /* complex interleaved */, const double* r, const double* phi, const double* v, const double* Hreal,
               const double* eta, const double* vPhiIn, const double* prefix /* complex interleaved per binary */, double h22_calib,
               int numSys, int traj_length, const double* coeffs, const int* idx) {
    const FacCoeffs* hc = reinterpret_cast<const FacCoeffs*>(coeffs);
    for (int b = 0; b < numSys; ++b) {
        cplx pf(prefix[2 * b], prefix[2 * b + 1]);
        for (int s = 0; s < traj_length; ++s) {
            size_t q = (size_t)b * traj_length + s;
            cplx h = hlm22(r[q], phi[q], v[q], Hreal[q], hc[idx[b]], pf, eta[b], vPhiIn[q], h22_calib);
            out[2 * q] = h.real(); out[2 * q + 1] = h.imag();
        }
    }
}

// get_hlms for (2,2) (eob/seobnrv4phm.py:2138-2221) on a trajectory laid out
// [nbin][4][len] with lengths[b] valid points; out is complex [nbin][len].
void orc_get_hlm22(double* out, const double* traj, const int* lengths, int nbin, int len, const double* add,
                   const double* coeffs, double h22_calib) {
    const FacCoeffs* hc = reinterpret_cast<const FacCoeffs*>(coeffs);
    for (int b = 0; b < nbin; ++b) {
        HamParams p = params_of(add, nbin, b);
        const double M = p.m1 + p.m2, nu = (p.m1 * p.m2 / M) / M;
        // (2,2) is entry 1 of the (l,m) list l=2..8, m=1..l (after (2,1))
        cplx pf(add[12 * nbin + b], add[13 * nbin + b]);
        for (int s = 0; s < len; ++s) {
            size_t q = (size_t)b * len + s;
            if (s >= lengths[b]) { out[2 * q] = 0.0; out[2 * q + 1] = 0.0; continue; }
            const double r = traj[((size_t)b * 4 + 0) * len + s], phi = traj[((size_t)b * 4 + 1) * len + s];
            const double pr = traj[((size_t)b * 4 + 2) * len + s], L = traj[((size_t)b * 4 + 3) * len + s];
            double g[4], gc[4];
            grad1(r, pr, L, p, g);
            grad1(r, 0.0, L, p, gc);
            const double omega = g[3], omega_c = gc[3];
            const double vPhi = 1 / (omega_c * omega_c * r * r * r);
            const double Hv = nu * hamiltonian<double>(r, pr, L, p);
            const double v = std::pow(omega, 1. / 3);
            cplx h = hlm22(r, phi, v, Hv, hc[b], pf, nu, vPhi, h22_calib);
            out[2 * q] = h.real(); out[2 * q + 1] = h.imag();
        }
    }
}

// get_hlms (eob/seobnrv4phm.py:2138-2221) for any (l, m): the mode on every knot of
// traj [nbin][4][len]; the prefix is additionalArgs' entry of (l, m).
void orc_get_hlm(double* out, int l, int m, const double* traj, const int* lengths, int nbin, int len, const double* add,
                 const double* coeffs, double h22_calib) {
    const FacCoeffs* hc = reinterpret_cast<const FacCoeffs*>(coeffs);
    const int fi = (l - 1) * l / 2 - 1 + m - 1;   // flux order l = 2..8, m = 1..l
    for (int b = 0; b < nbin; ++b) {
        HamParams p = params_of(add, nbin, b);
        const double M = p.m1 + p.m2, nu = (p.m1 * p.m2 / M) / M;
        cplx pf(add[(10 + 2 * fi) * nbin + b], add[(11 + 2 * fi) * nbin + b]);
        for (int s = 0; s < len; ++s) {
            size_t q = (size_t)b * len + s;
            if (s >= lengths[b]) { out[2 * q] = 0.0; out[2 * q + 1] = 0.0; continue; }
            const double r = traj[((size_t)b * 4 + 0) * len + s], phi = traj[((size_t)b * 4 + 1) * len + s];
            const double pr = traj[((size_t)b * 4 + 2) * len + s], L = traj[((size_t)b * 4 + 3) * len + s];
            double g[4], gc[4];
            grad1(r, pr, L, p, g);
            grad1(r, 0.0, L, p, gc);
            const double omega = g[3], omega_c = gc[3];
            const double vPhi = 1 / (omega_c * omega_c * r * r * r);
            const double Hv = nu * hamiltonian<double>(r, pr, L, p);
            const double v = std::pow(omega, 1. / 3);
            cplx h = hlm_mode(l, m, r, phi, L, v, Hv, hc[b], pf, nu, vPhi, h22_calib);
            out[2 * q] = h.real(); out[2 * q + 1] = h.imag();
        }
    }
}

// ------------------------------------------------------------------ splines
// interpolate_TD (src/interpolate.cu:11-308): not-a-knot system, Thomas solve,
// coefficients.  t: [maxlen][nbin]; y, B, c1(ud), c2(diag), c3(ld): [maxlen][nsubs*nbin].
void orc_interpolate_TD(const double* t, const double* y, double* B, double* ud, double* diag, double* ld,
                        const int* lengths, int nbin, int nsubs) {
    const int ni = nbin * nsubs;
    for (int it = 0; it < ni; ++it) {
        const int b = it % nbin;
        const int len = lengths[b];
        for (int i = 0; i < len; ++i) {
            int i0, i1, i2;
            if (i == len - 1) { i0 = len - 3; i1 = len - 2; i2 = len - 1; }
            else if (i == 0) { i0 = 0; i1 = 1; i2 = 2; }
            else { i0 = i - 1; i1 = i; i2 = i + 1; }
            const double x0 = t[(size_t)i0 * nbin + b], x1 = t[(size_t)i1 * nbin + b], x2 = t[(size_t)i2 * nbin + b];
            const double dx1 = x1 - x0, dx2 = x2 - x1, d = x2 - x0;
            const double y0 = y[(size_t)i0 * ni + it], y1 = y[(size_t)i1 * ni + it], y2 = y[(size_t)i2 * ni + it];
            const double s1 = (y1 - y0) / dx1, s2 = (y2 - y1) / dx2;
            const size_t o = (size_t)i * ni + it;
            if (i == len - 1) {
                B[o] = (dx2 * dx2 * s1 + (2 * d + dx2) * dx1 * s2) / d;
                diag[o] = dx1; ld[o] = d; ud[o] = 0.0;
            } else if (i == 0) {
                B[o] = ((dx1 + 2 * d) * dx2 * s1 + dx1 * dx1 * s2) / d;
                ud[o] = d; ld[o] = 0.0; diag[o] = dx2;
            } else {
                B[o] = 3.0 * (dx2 * s1 + dx1 * s2);
                diag[o] = 2 * (dx1 + dx2); ud[o] = dx1; ld[o] = dx2;
            }
        }
        // Thomas (a = ld, b = diag, c = ud, d = B), src/interpolate.cu:161-205
        for (int i = 1; i < len; ++i) {
            size_t ii = (size_t)i * ni + it, im = (size_t)(i - 1) * ni + it;
            double w = ld[ii] / diag[im];
            diag[ii] = diag[ii] - w * ud[im];
            B[ii] = B[ii] - w * B[im];
        }
        size_t last = (size_t)(len - 1) * ni + it;
        B[last] = B[last] / diag[last];
        for (int i = len - 2; i >= 0; --i) {
            size_t ii = (size_t)i * ni + it, ip = (size_t)(i + 1) * ni + it;
            B[ii] = (B[ii] - ud[ii] * B[ip]) / diag[ii];
        }
        // coefficients into ud (c1), diag (c2), ld (c3), src/interpolate.cu:208-267
        for (int i = 0; i < len - 1; ++i) {
            const double dt = t[(size_t)(i + 1) * nbin + b] - t[(size_t)i * nbin + b];
            size_t ii = (size_t)i * ni + it, ip = (size_t)(i + 1) * ni + it;
            const double slope = (y[ip] - y[ii]) / dt;
            const double dy = B[ii];
            const double tt = (dy + B[ip] - 2 * slope) / dt;
            ud[ii] = dy;
            diag[ii] = (slope - dy) / dt - tt;
            ld[ii] = tt / dt;
        }
    }
}

// CubicSplineInterpolantTD.__call__ (eob/seobnrv4phm.py:241-386), deriv_order 0..3,
// for one interpolant: knots t[len] (row stride tstride), coefficients with stride cstride.
void orc_spline_eval(double* out, const double* tnew, int nnew, const double* t, int tstride, const double* y,
                     const double* c1, const double* c2, const double* c3, int cstride, int len, int deriv) {
    for (int k = 0; k < nnew; ++k) {
        // searchsorted(t[:len] padded with 2*max, side='right') - 1, then ind == len -> len-1
        int lo = 0, hi = len;
        while (lo < hi) { int mid = (lo + hi) / 2; if (t[(size_t)mid * tstride] <= tnew[k]) lo = mid + 1; else hi = mid; }
        int ind = lo - 1;
        if (ind == len) ind -= 1;
        if (ind < 0) ind = 0;
        const double x = tnew[k] - t[(size_t)ind * tstride];
        const double x2 = x * x, x3 = x2 * x;
        const size_t o = (size_t)ind * cstride;
        double v;
        if (deriv == 0) v = y[o] + c1[o] * x + c2[o] * x2 + c3[o] * x3;
        else if (deriv == 1) v = c1[o] + 2 * c2[o] * x + 3 * c3[o] * x2;
        else if (deriv == 2) v = 2 * c2[o] + 6 * c3[o] * x;
        else v = 6 * c3[o];
        out[k] = v;
    }
}

// all_in_one_likelihood (src/interpolate.cu:553-628): per-block partial sums of
// |d - (F+ hp + Fx hc) e^{-2 pi i f dt}|^2 / psd, 1024 bins per block.
void orc_likelihood(double* temp_sum /* complex [nblk][nbin][nchan] */, const double* hp, const double* hcr,
                    const double* data, const double* psd, const double* Fp, const double* Fc, const double* ts,
                    double df, int nbin, int nchan, int F) {
    const int nblk = (F + 1023) / 1024;
    for (int blk = 0; blk < nblk; ++blk)
        for (int b = 0; b < nbin; ++b)
            for (int c = 0; c < nchan; ++c) {
                double acc = 0.0;
                for (int i = blk * 1024; i < std::min(F, (blk + 1) * 1024); ++i) {
                    cplx hpv(hp[2 * ((size_t)b * F + i)], hp[2 * ((size_t)b * F + i) + 1]);
                    cplx hcv(hcr[2 * ((size_t)b * F + i)], hcr[2 * ((size_t)b * F + i) + 1]);
                    cplx d(data[2 * ((size_t)c * F + i)], data[2 * ((size_t)c * F + i) + 1]);
                    double f = df * i;
                    cplx hh = (Fp[c * nbin + b] * hpv + Fc[c * nbin + b] * hcv) * std::exp(cplx(0.0, -2. * PI * f * ts[c * nbin + b]));
                    cplx r = d - hh;
                    acc += std::norm(r) / psd[(size_t)c * F + i];
                }
                size_t o = ((size_t)blk * nbin + b) * nchan + c;
                temp_sum[2 * o] = acc; temp_sum[2 * o + 1] = 0.0;
            }
}

// ----------------------------------------------------------- initial state
// Circular orbit at radius r0: dH/dr(r0, pr=0, pphi) = 0 solved by Newton in
// pphi with the exact Hessian element; then pr0 from the reference's IC_diss
// false-position (src/EOB.cu:2666-2834) on [-3e-2, 0].  omega0 is recorded.
static double ic_diss1(double pr, double r, double L, double* add, int n, int i, const FacCoeffs& hc) {
    HamParams p = params_of(add, n, i);
    double g[4];
    grad1(r, pr, L, p, g);
    double arg[4] = {r, 0.0, pr, L};
    double hs[16];
    orc_hessian_H(arg, hs, add, 1);  // n == 1 here
    double y[4] = {r, 0.0, pr, L};
    double F[2];
    rr_force1(y, add, n, i, g[3], hc, F);
    const double dLdr = -hs[0] / hs[12];
    const double rdot = F[1] / dLdr;
    return rdot - g[2];
}

void orc_initial_conditions(const double* r0, double* add, const double* coeffs, int n, double* pphi0, double* pr0,
                            double* omega0, int max_iter, double err) {
    const FacCoeffs* hc = reinterpret_cast<const FacCoeffs*>(coeffs);
    for (int i = 0; i < n; ++i) {
        double a1[NADD];
        for (int j = 0; j < NADD; ++j) a1[j] = add[j * n + i];
        HamParams p = params_of(a1, 1, 0);
        double L = std::sqrt(r0[i]);
        for (int it = 0; it < 100; ++it) {
            double arg[4] = {r0[i], 0.0, 0.0, L};
            double g[4], hs[16];
            orc_grad_H(arg, g, a1, 1);
            orc_hessian_H(arg, hs, a1, 1);
            double dL = g[0] / hs[3];
            L -= dL;
            if (std::fabs(dL) < 1e-15 * std::fabs(L)) break;
        }
        double g[4];
        grad1(r0[i], 0.0, L, p, g);
        omega0[i] = g[3];
        a1[4] = g[3];
        // root_find_scalar (src/EOB.cu:2736-2776)
        double a = -3e-2, b = 0.0;
        double fa = ic_diss1(a, r0[i], L, a1, 1, 0, hc[i]);
        double fb = ic_diss1(b, r0[i], L, a1, 1, 0, hc[i]);
        double c = 0.0;
        for (int it = 0; it < max_iter; ++it) {
            c = (a * fb - b * fa) / (fb - fa);
            double fc = ic_diss1(c, r0[i], L, a1, 1, 0, hc[i]);
            if ((fb < 0.0 && fc < 0.0) || (fb >= 0.0 && fc >= 0.0)) { b = c; fb = fc; }
            else { a = c; fa = fc; }
            if (std::fabs(fc) < err) break;
        }
        pphi0[i] = L;
        pr0[i] = c;
        add[4 * n + i] = omega0[i];
    }
}

// IC_cons (src/EOB.cu:405-430): res = (dH/dpphi - omega0, dH/dr) at (r, 0, 0, pphi),
// omega0 from additional-args row 4; one lane of a batch of n.
static void ic_cons1(const double x[2], const double* add, int n, int i, double res[2]) {
    HamParams p = params_of(add, n, i);
    double g[4];
    grad1(x[0], 0.0, x[1], p, g);
    res[0] = g[3] - add[4 * n + i];
    res[1] = g[0];
}

// root_find_all (src/EOB.cu:2853-3031): per binary, a discrete Newton iteration
// on IC_cons in x = (r, pphi): forward differences of step sqrt(2.22e-16) |x_j|,
// the 2x2 Jacobian inverted in closed form, x -= J^-1 f; it stops once every
// component's change relative to its old value is below err, or after max_iter.
// x0In, xOut: [2][n].
void orc_root_find_all(double* xOut, const double* x0In, const double* add, int max_iter, double err, int n) {
    const double eps = 2.22e-16;
    for (int i = 0; i < n; ++i) {
        double x0[2] = {x0In[i], x0In[n + i]}, x1[2] = {x0[0], x0[1]};
        for (int it = 0; it < max_iter; ++it) {
            x1[0] = x0[0]; x1[1] = x0[1];
            double f[2], J[2][2];
            ic_cons1(x1, add, n, i, f);
            for (int k = 0; k < 2; ++k)
                for (int j = 0; j < 2; ++j) {
                    const double dj = std::sqrt(eps) * std::fabs(x1[j]);
                    double xt[2] = {x1[0], x1[1]};
                    xt[j] += dj;
                    double f2[2];
                    ic_cons1(xt, add, n, i, f2);
                    J[k][j] = (f2[k] - f[k]) / dj;
                }
            const double det = 1. / (J[0][0] * J[1][1] - J[0][1] * J[1][0]);
            const double inv[2][2] = {{det * J[1][1], det * -J[0][1]}, {det * -J[1][0], det * J[0][0]}};
            for (int k = 0; k < 2; ++k) {
                double t = 0.0;
                for (int j = 0; j < 2; ++j) t += inv[k][j] * f[j];
                x1[k] -= t;
            }
            double rel = 0.0;
            for (int j = 0; j < 2; ++j) {
                const double d = std::fabs(x1[j] - x0[j]) / x0[j];
                if (d > rel) rel = d;
            }
            x0[0] = x1[0]; x0[1] = x1[1];
            if (rel < err) break;
        }
        xOut[i] = x1[0];
        xOut[n + i] = x1[1];
    }
}

// IC_diss (src/EOB.cu:2666-2688) through its kernel (IC_diss_kernel): per
// binary, args row 2 := pr, the gradient and Hessian of H there, additional-args
// row 4 := dH/dpphi, the radiation-reaction force with that omega (force_out
// [2][n], grad_temp_force [4][n] = the gradient at pr = 0 RR_force leaves), and
// out = F_phi / (dpphi/dr) - dH/dpr with dpphi/dr = -H_rr / H_rpphi.  args and
// add are updated in place as the reference's kernel does.
void orc_ic_diss(double* out, const double* pr, double* args, double* add, double* grad_out, double* grad_temp_force,
                 double* hess_out, double* force_out, const double* coeffs, int n) {
    const FacCoeffs* hc = reinterpret_cast<const FacCoeffs*>(coeffs);
    for (int i = 0; i < n; ++i) {
        args[2 * n + i] = pr[i];
        HamParams p = params_of(add, n, i);
        double g[4], a1[4] = {args[i], args[n + i], args[2 * n + i], args[3 * n + i]}, hs[16];
        grad1(a1[0], a1[2], a1[3], p, g);
        double add1[NADD];
        for (int j = 0; j < NADD; ++j) add1[j] = add[j * n + i];
        orc_hessian_H(a1, hs, add1, 1);
        for (int j = 0; j < 4; ++j) grad_out[j * n + i] = g[j];
        for (int j = 0; j < 16; ++j) hess_out[j * n + i] = hs[j];
        add[4 * n + i] = g[3];
        double F[2];
        rr_force1(a1, add, n, i, g[3], hc[i], F);
        double gc[4];
        grad1(a1[0], 0.0, a1[3], p, gc);
        for (int j = 0; j < 4; ++j) grad_temp_force[j * n + i] = gc[j];
        force_out[i] = F[0];
        force_out[n + i] = F[1];
        const double dLdr = -hs[0] / hs[12];
        out[i] = F[1] / dLdr - g[2];
    }
}

// root_find_scalar_all (src/EOB.cu:2712-2834): per binary, false position on
// IC_diss(pr) in [start_bounds[0], start_bounds[1]] (rows [2][n]) with the
// binary's args and additional args copied (IC_diss's writes stay in the copy);
// it stops when |IC_diss(c)| < err or after max_iter steps.
void orc_root_find_scalar_all(double* pr_res, const double* bounds, const double* args, const double* add, const double* coeffs,
                              int max_iter, double err, int n) {
    const FacCoeffs* hc = reinterpret_cast<const FacCoeffs*>(coeffs);
    for (int i = 0; i < n; ++i) {
        double add1[NADD];
        for (int j = 0; j < NADD; ++j) add1[j] = add[j * n + i];
        const double r = args[i], L = args[3 * n + i];
        double a = bounds[i], b = bounds[n + i];
        // ic_diss1 evaluates IC_diss with omega = dH/dpphi at that pr, as the reference's copy is updated
        double fa = ic_diss1(a, r, L, add1, 1, 0, hc[i]);
        double fb = ic_diss1(b, r, L, add1, 1, 0, hc[i]);
        double c = 0.0;
        for (int it = 0; it < max_iter; ++it) {
            c = (a * fb - b * fa) / (fb - fa);
            const double fc = ic_diss1(c, r, L, add1, 1, 0, hc[i]);
            if ((fb < 0.0 && fc < 0.0) || (fb >= 0.0 && fc >= 0.0)) { b = c; fb = fc; }
            else { a = c; fa = fc; }
            if (std::fabs(fc) < err) break;
        }
        pr_res[i] = c;
    }
}

// Newtonian multipole prefixes (eob/seobnrv4phm.py:1703-1776, Damour-Iyer-Nagar
// 2008 Eqs. 5-7) for the 35 flux modes l=2..8, m=1..l, complex interleaved [35][n].
void orc_newton_prefixes(double* out, const double* m1, const double* m2, int n) {
    for (int i = 0; i < n; ++i) {
        const double M = m1[i] + m2[i], x1 = m1[i] / M, x2 = m2[i] / M, eta = m1[i] * m2[i] / (M * M);
        int q = 0;
        for (int l = 2; l <= 8; ++l)
            for (int m = 1; m <= l; ++m, ++q) {
                const int eps = (l + m) % 2;
                const int sign = (m % 2 == 0) ? 1 : -1;
                const bool t1 = (m1[i] != m2[i]) || (sign == 1);
                double c;
                if (t1) c = std::pow(x2, l + eps - 1) + sign * std::pow(x1, l + eps - 1);
                else c = (l == 2 || l == 3) ? -1.0 : (l == 4 || l == 5) ? -0.5 : 0.0;
                double df = 1.0;
                for (int k = 2 * l + 1; k > 1; k -= 2) df *= k;
                cplx nn = std::pow(cplx(0.0, (double)m), l);
                if (eps == 0) {
                    nn *= 8.0 * PI / df;
                    nn *= std::sqrt(((l + 1.0) * (l + 2.0)) / (l * (l - 1.0)));
                } else {
                    nn = -nn;
                    double m2v = (2. * l + 1) * (l + 2) * (l * l - m * m);
                    m2v /= (2. * l - 1) * (l + 1) * l * (l - 1);
                    nn *= cplx(0.0, 16.0 * PI / df);
                    nn *= std::sqrt(m2v);
                }
                cplx pf = nn * eta * c;
                out[2 * ((size_t)q * n + i)] = pf.real();
                out[2 * ((size_t)q * n + i) + 1] = pf.imag();
            }
    }
}

}  // extern "C"
