// FP64 operation count of one ODE right-hand side and of one attempted DOP853
// step, taken by evaluating the library's own Hamiltonian and flux code on
// counting scalar types (host build of the same headers the kernels use).
// Conventions: + - * / count 1 each; sqrt, log, exp, cbrt, pow count 1 each and
// are also reported separately as "transcendental".
//
//   g++ -x c++ -D__HIP_PLATFORM_AMD__ -I/opt/rocm/include -Ieobgpu_amd/csrc -O1 -std=c++17 -o flop_count tests/flop_count.cpp
// prints one JSON object; bench.py carries the frozen step count, and
// tests/test_derivations.py builds this file and checks that the two agree.
#include <cmath>
#include <cstdio>
#include "eob_ham.hpp"
#include "eob_flux.hpp"

using namespace eob;

static long g_ops = 0, g_trans = 0;

// counting scalar (for the flux)
struct CS {
    double v;
    CS(double x = 0.0) : v(x) {}
};
static CS operator+(CS a, CS b) { ++g_ops; return CS(a.v + b.v); }
static CS operator-(CS a, CS b) { ++g_ops; return CS(a.v - b.v); }
static CS operator*(CS a, CS b) { ++g_ops; return CS(a.v * b.v); }
static CS operator/(CS a, CS b) { ++g_ops; return CS(a.v / b.v); }
static CS operator+(CS a, double b) { ++g_ops; return CS(a.v + b); }
static CS operator+(double b, CS a) { ++g_ops; return CS(b + a.v); }
static CS operator-(double b, CS a) { ++g_ops; return CS(b - a.v); }
static CS operator*(CS a, double b) { ++g_ops; return CS(a.v * b); }
static CS operator*(double b, CS a) { ++g_ops; return CS(b * a.v); }
static CS operator/(CS a, double b) { ++g_ops; return CS(a.v / b); }
static CS operator/(double a, CS b) { ++g_ops; return CS(a / b.v); }
static CS operator-(CS a, double b) { ++g_ops; return CS(a.v - b); }
static CS operator-(CS a) { return CS(-a.v); }
static CS sqrt(CS a) { ++g_ops; ++g_trans; return CS(std::sqrt(a.v)); }
static CS log(CS a) { ++g_ops; ++g_trans; return CS(std::log(a.v)); }
static CS exp(CS a) { ++g_ops; ++g_trans; return CS(std::exp(a.v)); }

// counting dual (one tangent): each op counts the FP64 operations Dual performs
struct CD {
    double v, d;
    CD(double x = 0.0, double t = 0.0) : v(x), d(t) {}
};
static CD operator+(CD a, CD b) { g_ops += 2; return CD(a.v + b.v, a.d + b.d); }
static CD operator-(CD a, CD b) { g_ops += 2; return CD(a.v - b.v, a.d - b.d); }
static CD operator-(CD a) { return CD(-a.v, -a.d); }
static CD operator*(CD a, CD b) { g_ops += 4; return CD(a.v * b.v, a.d * b.v + a.v * b.d); }
static CD operator+(CD a, double b) { g_ops += 1; return CD(a.v + b, a.d); }
static CD operator+(double b, CD a) { g_ops += 1; return CD(b + a.v, a.d); }
static CD operator-(CD a, double b) { g_ops += 1; return CD(a.v - b, a.d); }
static CD operator-(double b, CD a) { g_ops += 1; return CD(b - a.v, -a.d); }
static CD operator*(CD a, double b) { g_ops += 2; return CD(a.v * b, a.d * b); }
static CD operator*(double b, CD a) { g_ops += 2; return CD(b * a.v, b * a.d); }
static CD operator/(CD a, CD b) { g_ops += 5; double q = a.v / b.v; return CD(q, (a.d - q * b.d) / b.v); }
static CD operator/(double a, CD b) { g_ops += 4; double q = a / b.v; return CD(q, -q * b.d / b.v); }
static CD operator/(CD a, double b) { g_ops += 2; return CD(a.v / b, a.d / b); }
static CD dsqrt(CD a) { g_ops += 2; ++g_trans; double s = sqrt(a.v); return CD(s, 0.5 * a.d / s); }
static CD dlog(CD a) { g_ops += 1; ++g_trans; return CD(log(a.v), a.d / a.v); }

int main() {
    // a representative binary: q = 3, chi1 = 0.4, chi2 = -0.2 (all tables are dense)
    const double m1 = 0.75, m2 = 0.25;
    HamCoef hc;
    ham_setup(hc, m1, m2, 0.4, -0.2, 0.0, 0.0, 0.0, 0.0);
    ModeCoef mc[kNumModes];
    WaveCoef wc;
    double pf[70];
    for (int i = 0; i < 70; ++i) pf[i] = 0.1;
    const double eta = m1 * m2;
    flux_setup(mc, wc, m1, m2, eta, 0.1, 0.1, 0.3, pf, 0.0, 1);

    // Hamiltonian and its four derivatives per RHS: the 13 pieces (value and
    // d/du) once each, then the closed-form combination; u, ln r, sqrt(u) and
    // d sqrt(u)/du once.  (Four single-tangent dual evaluations of ham_eval
    // would take the count reported as ham_dual_4tangents.)
    PieceCoef pc;
    piece_setup(pc, hc);
    g_ops = g_trans = 0;
    {
        CS r(12.0), pr(-0.003), L(3.9);
        CS u = 1.0 / r, lr = log(r), su = sqrt(u), hs = 0.5 * su * r;
        CS v[kNumPieces], d[kNumPieces];
        for (int q = 0; q < kNumPieces; ++q) piece_eval<CS>(pc.c[q], u, r, lr, su, hs, v[q], d[q]);
        (void)ham_combine<CS>(hc, r, u, pr, L, v, d);
    }
    const long ham_ops = g_ops, ham_trans = g_trans;
    g_ops = g_trans = 0;
    for (int td = 0; td < 4; ++td) {
        CD r(12.0, td == 0), pr(td == 3 ? 0.0 : -0.003, td == 1), L(3.9, td >= 2);
        (void)ham_eval<CD>(hc, r, pr, L);
    }
    const long ham_dual_ops = g_ops;

    // flux: preamble (cbrt, log, vPhi, Heff, exp) + 35 modes + the RR-force tail
    g_ops = g_trans = 0;
    CS v(0.2), logv(-1.6), H(0.98), Om(0.008), Heff(0.99), L(3.9), vPhi2(0.04);
    CS sum(0.0);
    for (int i = 0; i < kNumModes; ++i) sum = sum + flux_mode<CS>(mc[i], v, logv, H, Om, Heff, L, vPhi2);
    const long modes_ops = g_ops, modes_trans = g_trans;
    // preamble: v = cbrt, logv = log, vPhiIn = 1/(omc^2 r^3) (4 mul + 1 div), vPhi = r cbrt(.) omega (1 T + 2 mul),
    // vPhi2 (1), Heff (5); Hreal = nu H (1); tail: x 1/(8 pi nu), /omega, -pr/L*f (2), k2 and k3 sums (2);
    // the modes' exp(-4 pi m H Omega) are in modes_ops
    const long pre_ops = 4 + 1 + 2 + 1 + 5 + 1, pre_trans = 3;
    const long tail_ops = 1 + 1 + 2 + 2;
    const long rhs_ops = ham_ops + modes_ops + pre_ops + tail_ops;
    const long rhs_trans = ham_trans + modes_trans + pre_trans;

    // DOP853 step: stage sums (50 nonzero a_sj, 2 ops each) + 11 updates y + h ak (2 each), per component;
    // error norm per component (temp 15, y_new 2, sk 3, bhh combination 7, e5 combination 16, 4th powers and sums 6),
    // then denom / err (5 + 1 sqrt), controller (pow + 4)
    const long stage_ops = 4 * (50 * 2 + 11 * 2);
    const long err_ops = 4 * (15 + 2 + 3 + 7 + 16 + 6) + 5 + 4;
    const long step_ops = 12 * rhs_ops + stage_ops + err_ops;
    const long step_trans = 12 * rhs_trans + 2;
    printf("{\"ham_pieces\": %ld, \"ham_trans\": %ld, \"ham_dual_4tangents\": %ld, \"flux_modes\": %ld, \"rhs_ops\": %ld, "
           "\"rhs_trans\": %ld, \"step_ops\": %ld, \"step_trans\": %ld, \"step_flops\": %ld}\n",
           ham_ops, ham_trans, ham_dual_ops, modes_ops, rhs_ops, rhs_trans, step_ops, step_trans, step_ops + step_trans);
    return 0;
}
