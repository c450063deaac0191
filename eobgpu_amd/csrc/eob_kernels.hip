// HIP kernels and C-ABI entry points for the EOB Hamiltonian, flux, ODE
// right-hand side, DOP853 stepper and the batched integrators.
//
// Entry points mirror the reference's native wraps (see include/eobhip.h for the
// file:line of each); the batched integrators are this project's design.
#include <hip/hip_runtime.h>
#include <cstring>
#include <cstdlib>
#include <string>
#include <vector>

#include "eobhip_common.hpp"
#include "eob_integrator.hpp"
#include "eob_wave.hpp"

using namespace eob;

namespace eobhip_detail {
static thread_local char g_err[512] = "";
void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}
}  // namespace eobhip_detail
using namespace eobhip_detail;

struct eobhip_seobnr {
    ModeCoef* modes = nullptr;   // [numSys][35], prefix-independent (N2 set per call)
    WaveCoef* wave = nullptr;    // [numSys]
    WaveMode* wmodes = nullptr;  // [numSys][35]: the waveform's tables (EOBGetSpinFactorizedWaveform)
    const double* prefixes_hlms = nullptr;
    int numSys = 0;
    bool allocated = false;
};

namespace {

constexpr int kBlock = 256;
static_assert(kBlock / 64 <= kMaxWavesPerBlock, "the piece-gather slots hold kMaxWavesPerBlock wavefronts");

__device__ __forceinline__ HamCoef ham_from_add(const double* add, int n, int i) {
    HamCoef c;
    ham_setup(c, add[0 * n + i], add[1 * n + i], add[2 * n + i], add[3 * n + i], add[5 * n + i], add[6 * n + i], add[7 * n + i],
              add[8 * n + i]);
    return c;
}

__device__ __forceinline__ double nu_of(double m1, double m2) {
    const double M = m1 + m2;
    return (m1 * m2 / M) / M;
}

// ------------------------------------------------------------- Hamiltonian
__global__ __launch_bounds__(kBlock) void k_evaluate_H(double* out, const double* r, const double* pr, const double* pphi,
                                                       const double* m1, const double* m2, const double* c1, const double* c2,
                                                       double K, double d5, double dSO, double dSS, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        HamCoef c;
        ham_setup(c, m1[i], m2[i], c1[i], c2[i], K, d5, dSO, dSS);
        out[i] = ham_eval<double>(c, r[i], pr[i], pphi[i]);
    }
}

__global__ __launch_bounds__(kBlock) void k_grad_H(const double* arg, double* g, const double* add, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const HamCoef c = ham_from_add(add, n, i);
        const double r = arg[i], pr = arg[2 * n + i], L = arg[3 * n + i];
        g[0 * n + i] = ham_tangent(c, r, pr, L, 0).d;
        g[1 * n + i] = 0.0;
        g[2 * n + i] = ham_tangent(c, r, pr, L, 1).d;
        g[3 * n + i] = ham_tangent(c, r, pr, L, 2).d;
    }
}

__device__ __forceinline__ double ham_second(const HamCoef& c, double r, double pr, double L, int a, int b) {
    // a, b in {0: r, 2: pr, 3: pphi}
    const HDual R(r, a == 0 ? 1.0 : 0.0, b == 0 ? 1.0 : 0.0, 0.0);
    const HDual PR(pr, a == 2 ? 1.0 : 0.0, b == 2 ? 1.0 : 0.0, 0.0);
    const HDual LL(L, a == 3 ? 1.0 : 0.0, b == 3 ? 1.0 : 0.0, 0.0);
    return ham_eval<HDual>(c, R, PR, LL).ab;
}

__global__ __launch_bounds__(kBlock) void k_hessian_H(const double* arg, double* hs, const double* add, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const HamCoef c = ham_from_add(add, n, i);
        const double r = arg[i], pr = arg[2 * n + i], L = arg[3 * n + i];
        for (int q = 0; q < 16; ++q) hs[q * n + i] = 0.0;
        const int v[3] = {0, 2, 3};
        for (int ia = 0; ia < 3; ++ia)
            for (int ib = ia; ib < 3; ++ib) {
                const double val = ham_second(c, r, pr, L, v[ia], v[ib]);
                hs[(v[ia] * 4 + v[ib]) * n + i] = val;
                hs[(v[ib] * 4 + v[ia]) * n + i] = val;
            }
    }
}

__global__ __launch_bounds__(kBlock) void k_IC_cons(double* res, const double* x, double* args, const double* add,
                                                    double* grad, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double r = x[i], L = x[n + i];
        args[i] = r; args[n + i] = 0.0; args[2 * n + i] = 0.0; args[3 * n + i] = L;
        const HamCoef c = ham_from_add(add, n, i);
        const double dHdr = ham_tangent(c, r, 0.0, L, 0).d;
        const double dHdL = ham_tangent(c, r, 0.0, L, 2).d;
        grad[i] = dHdr; grad[n + i] = 0.0; grad[2 * n + i] = ham_tangent(c, r, 0.0, L, 1).d; grad[3 * n + i] = dHdL;
        res[i] = dHdL - add[4 * n + i];
        res[n + i] = dHdr;
    }
}

// ------------------------------------------------- reference-layout RHS (G = 1)
// Flux with the handle's rho tables and this call's prefixes / h22_calib
// (src/EOB.cu:2506-2560); hm points at the 35 modes of the binary.
__device__ double flux_dropin(const ModeCoef* hm, const double* add, int n, int i, double eta_f, double r, double L, double Hreal,
                              double omega, double omc) {
    const double v = cbrt(omega), logv = log(v);
    const double vPhiIn = 1.0 / (omc * omc * (r * r * r));
    const double vPhi = r * cbrt(vPhiIn) * omega;
    const double vPhi2 = vPhi * vPhi;
    const double Heff = (Hreal * Hreal - 1.0) / (2.0 * eta_f) + 1.0;
    const double h22c = add[9 * n + i];
    double sum = 0.0;
    for (int mi = 0; mi < kNumModes; ++mi) {
        ModeCoef z = hm[mi];
        const double pre = add[(10 + 2 * mi) * n + i], pim = add[(11 + 2 * mi) * n + i];
        const double y = ylm_abs_eq(z.l - (z.lpe - z.l), z.m);
        z.N2 = (pre * pre + pim * pim) * y * y;
        if (mi == 1) z.c[10] = h22c;
        int fl = ((z.l + z.m) & 1) ? 0 : 1;
        if (eta_f == 0.25 && (z.m & 1)) fl |= 2;
        z.flags = fl;
        sum += flux_mode(z, v, logv, Hreal, omega, Heff, L, vPhi2);
    }
    return sum / kPi / 8.0;
}

__device__ void rr_force_dropin(const ModeCoef* hm, const double* add, int n, int i, const HamCoef& c, double r, double pr, double L,
                                double omega, double F[2]) {
    const double nu = nu_of(add[i], add[n + i]);
    double eta_f = nu;
    if (eta_f > 0.25 && eta_f < 0.25 + 1e-4) eta_f = 0.25;
    const double omc = ham_tangent(c, r, pr, L, 3).d;
    const double Hreal = nu * ham_eval<double>(c, r, pr, L);
    double flux = flux_dropin(hm, add, n, i, eta_f, r, L, Hreal, omega, omc);
    flux /= nu;
    const double f_over_om = flux / omega;
    F[0] = -pr / L * f_over_om;
    F[1] = -f_over_om;
}

__global__ __launch_bounds__(kBlock) void k_ode_dropin(const double* arg, double* k, const double* add, int n, const ModeCoef* modes,
                                                       const int* idx) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const HamCoef c = ham_from_add(add, n, i);
        const double r = arg[i], pr = arg[2 * n + i], L = arg[3 * n + i];
        const Dual d0 = ham_tangent(c, r, pr, L, 0);
        const double dHdpr = ham_tangent(c, r, pr, L, 1).d;
        const double dHdL = ham_tangent(c, r, pr, L, 2).d;
        double F[2];
        rr_force_dropin(modes + (size_t)idx[i] * kNumModes, add, n, i, c, r, pr, L, dHdL, F);
        k[i] = dHdpr;
        k[n + i] = dHdL;
        k[2 * n + i] = -d0.d + F[0];
        k[3 * n + i] = -0.0 + F[1];
    }
}

__global__ __launch_bounds__(kBlock) void k_rr_force_dropin(double* force, double* grad, const double* arg, const double* add, int n,
                                                            const ModeCoef* modes, const int* idx) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const HamCoef c = ham_from_add(add, n, i);
        const double r = arg[i], pr = arg[2 * n + i], L = arg[3 * n + i];
        double F[2];
        rr_force_dropin(modes + (size_t)idx[i] * kNumModes, add, n, i, c, r, pr, L, add[4 * n + i], F);
        force[i] = F[0];
        force[n + i] = F[1];
        // the reference leaves the pr = 0 gradient in grad_out
        grad[i] = ham_tangent(c, r, 0.0, L, 0).d;
        grad[n + i] = 0.0;
        grad[2 * n + i] = ham_tangent(c, r, 0.0, L, 1).d;
        grad[3 * n + i] = ham_tangent(c, r, 0.0, L, 3).d;
    }
}

// ------------------------------------------- initial-condition solvers (drop-ins)
// IC_cons at x = (r, pphi): (dH/dpphi - omega0, dH/dr) at (r, 0, 0, pphi), src/EOB.cu:405-430
__device__ __forceinline__ void ic_cons_dev(const HamCoef& c, double omega0, double r, double L, double res[2]) {
    res[0] = ham_tangent(c, r, 0.0, L, 2).d - omega0;
    res[1] = ham_tangent(c, r, 0.0, L, 0).d;
}

// root_find_all (src/EOB.cu:2853-3031), one binary per thread: discrete Newton on
// IC_cons with forward differences of step sqrt(2.22e-16) |x_j|, the 2x2
// Jacobian inverted in closed form; it stops once every component's change
// relative to its old value is below err, or after max_iter steps.
__global__ __launch_bounds__(kBlock) void k_root_find_all(double* xOut, const double* x0In, const double* add, int max_iter, double err,
                                                          int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const HamCoef c = ham_from_add(add, n, i);
        const double om0 = add[4 * n + i];
        const double eps = 2.22e-16;
        double x0[2] = {x0In[i], x0In[n + i]}, x1[2] = {x0[0], x0[1]};
        for (int it = 0; it < max_iter; ++it) {
            x1[0] = x0[0];
            x1[1] = x0[1];
            double f[2], J[2][2];
            ic_cons_dev(c, om0, x1[0], x1[1], f);
            for (int j = 0; j < 2; ++j) {
                const double dj = sqrt(eps) * fabs(x1[j]);
                double xt[2] = {x1[0], x1[1]};
                xt[j] += dj;
                double f2[2];
                ic_cons_dev(c, om0, xt[0], xt[1], f2);
                J[0][j] = (f2[0] - f[0]) / dj;
                J[1][j] = (f2[1] - f[1]) / dj;
            }
            const double det = 1. / (J[0][0] * J[1][1] - J[0][1] * J[1][0]);
            const double inv[2][2] = {{det * J[1][1], det * -J[0][1]}, {det * -J[1][0], det * J[0][0]}};
            for (int k = 0; k < 2; ++k) x1[k] -= inv[k][0] * f[0] + inv[k][1] * f[1];
            double rel = 0.0;
            for (int j = 0; j < 2; ++j) {
                const double d = fabs(x1[j] - x0[j]) / x0[j];
                rel = (d > rel) ? d : rel;
            }
            x0[0] = x1[0];
            x0[1] = x1[1];
            if (rel < err) break;
        }
        xOut[i] = x1[0];
        xOut[n + i] = x1[1];
    }
}

// IC_diss (src/EOB.cu:2666-2688) for binary i at pr: F_phi / (dpphi/dr) - dH/dpr
// with dpphi/dr = -H_rr / H_rpphi and the force's omega = dH/dpphi at (r, pr, pphi).
__device__ double ic_diss_dev(const ModeCoef* hm, const double* add, int n, int i, const HamCoef& c, double r, double pr, double L,
                              double g[4], double F[2], double* hrr_out = nullptr, double* hrl_out = nullptr) {
    g[0] = ham_tangent(c, r, pr, L, 0).d;
    g[1] = 0.0;
    g[2] = ham_tangent(c, r, pr, L, 1).d;
    g[3] = ham_tangent(c, r, pr, L, 2).d;
    const double hrr = ham_second(c, r, pr, L, 0, 0);
    const double hrl = ham_second(c, r, pr, L, 0, 3);
    if (hrr_out) *hrr_out = hrr;
    if (hrl_out) *hrl_out = hrl;
    rr_force_dropin(hm, add, n, i, c, r, pr, L, g[3], F);
    const double dLdr = -hrr / hrl;
    return F[1] / dLdr - g[2];
}

// IC_diss_kernel: args row 2 := pr, additional-args row 4 := dH/dpphi (both in
// place, as the reference), grad_out [4][n], hess_out [16][n], force_out [2][n],
// grad_temp_force [4][n] the gradient at pr = 0 that RR_force leaves.
__global__ __launch_bounds__(kBlock) void k_ic_diss(double* out, const double* pr, double* args, double* add, double* grad,
                                                    double* gtf, double* hess, double* force, int n, const ModeCoef* modes) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const HamCoef c = ham_from_add(add, n, i);
        const double r = args[i], L = args[3 * n + i], p = pr[i];
        args[2 * n + i] = p;
        double g[4], F[2];
        out[i] = ic_diss_dev(modes + (size_t)i * kNumModes, add, n, i, c, r, p, L, g, F);
        for (int q = 0; q < 4; ++q) grad[q * n + i] = g[q];
        const int v[3] = {0, 2, 3};
        for (int q = 0; q < 16; ++q) hess[q * n + i] = 0.0;
        for (int ia = 0; ia < 3; ++ia)
            for (int ib = ia; ib < 3; ++ib) {
                const double val = ham_second(c, r, p, L, v[ia], v[ib]);
                hess[(v[ia] * 4 + v[ib]) * n + i] = val;
                hess[(v[ib] * 4 + v[ia]) * n + i] = val;
            }
        add[4 * n + i] = g[3];
        force[i] = F[0];
        force[n + i] = F[1];
        gtf[i] = ham_tangent(c, r, 0.0, L, 0).d;
        gtf[n + i] = 0.0;
        gtf[2 * n + i] = ham_tangent(c, r, 0.0, L, 1).d;
        gtf[3 * n + i] = ham_tangent(c, r, 0.0, L, 3).d;
    }
}

// root_find_scalar_all (src/EOB.cu:2712-2834), one binary per thread: false
// position on IC_diss(pr) from start_bounds [2][n]; it stops when |IC_diss(c)| <
// err or after max_iter steps.  The caller's args and additional args are read only.
__global__ __launch_bounds__(kBlock) void k_root_find_scalar_all(double* pr_res, const double* bounds, const double* args,
                                                                 const double* add, int max_iter, double err, int n,
                                                                 const ModeCoef* modes) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const HamCoef c = ham_from_add(add, n, i);
        const ModeCoef* hm = modes + (size_t)i * kNumModes;
        const double r = args[i], L = args[3 * n + i];
        double a = bounds[i], b = bounds[n + i], g[4], F[2];
        double fa = ic_diss_dev(hm, add, n, i, c, r, a, L, g, F);
        double fb = ic_diss_dev(hm, add, n, i, c, r, b, L, g, F);
        double cc = 0.0;
        for (int it = 0; it < max_iter; ++it) {
            cc = (a * fb - b * fa) / (fb - fa);
            const double fc = ic_diss_dev(hm, add, n, i, c, r, cc, L, g, F);
            if ((fb < 0.0 && fc < 0.0) || (fb >= 0.0 && fc >= 0.0)) { b = cc; fb = fc; }
            else { a = cc; fa = fc; }
            if (fabs(fc) < err) break;
        }
        pr_res[i] = cc;
    }
}

__global__ __launch_bounds__(kBlock) void k_update_info(ModeCoef* modes, WaveCoef* wave, const double* m1, const double* m2,
                                                        const double* eta, const double* a, const double* chiS, const double* chiA,
                                                        int use_hm, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        flux_setup(modes + (size_t)i * kNumModes, wave[i], m1[i], m2[i], eta[i], a[i], chiS[i], chiA[i], nullptr, 0.0, use_hm);
    }
}

// The waveform's per-mode tables of every binary (one thread per binary and mode),
// from the flux tables k_update_info built and the same inputs.
__global__ __launch_bounds__(kBlock) void k_wave_tables(WaveMode* wm, const ModeCoef* modes, const WaveCoef* wave, const double* m1,
                                                        const double* m2, const double* eta, const double* chiS, const double* chiA,
                                                        int use_hm, int n) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)n * kNumModes) return;
    const int i = (int)(t / kNumModes), q = (int)(t % kNumModes);
    int l, m;
    mode_lm(q, l, m);
    wave_mode_setup(wm[t], modes[t], l, m, m1[i], m2[i], eta[i], chiS[i], chiA[i], use_hm, wave + i);
}

// EOBGetSpinFactorizedWaveform_kernel for any (l, m): one block per binary,
// threads over knots; h22_calib replaces the (2,2) table's v^12 coefficient
__global__ __launch_bounds__(kBlock) void k_hlm_dropin(double* out, const double* r, const double* phi, const double* pphi,
                                                       const double* v, const double* Hreal, const double* eta, const double* vPhiIn,
                                                       const double* prefix, double h22c, int numSys, int L, const WaveMode* wmodes,
                                                       const int* idx, int fi, int nsys) {
    for (int b = blockIdx.x; b < numSys; b += gridDim.x) {
        const int hi = idx[b];
        if (hi < 0 || hi >= nsys) {   // an index outside the handle's binaries: NaN, not a stray read
            for (int s = threadIdx.x; s < L; s += blockDim.x) {
                out[2 * ((size_t)b * L + s)] = NAN;
                out[2 * ((size_t)b * L + s) + 1] = NAN;
            }
            continue;
        }
        WaveMode w = wmodes[(size_t)hi * kNumModes + fi];
        if (w.l == 2 && w.m == 2) w.c[10] = h22c;
        const double pre_re = prefix[2 * b], pre_im = prefix[2 * b + 1];
        for (int s = threadIdx.x; s < L; s += blockDim.x) {
            const size_t q = (size_t)b * L + s;
            const Cplx h = hlm_knot(w, pre_re, pre_im, eta[b], r[q], phi[q], pphi[q], v[q], Hreal[q], vPhiIn[q]);
            out[2 * q] = h.re;
            out[2 * q + 1] = h.im;
        }
    }
}

// --------------------------------------------------- reference stepper kernels
// ak_term = sum_j a_sj k_j accumulated from 0 in stage order; arg = solOld + h ak;
// xCurrent = x + c_s h.  (src/DOPR853.cu:197-355, 381-516)
struct StagePtrs {
    const double* k[11];
};
__constant__ double dA[12][12] = DOP_A;
__global__ __launch_bounds__(kBlock) void k_stage_arg(double* arg, const double* solOld, const double* x, double* xCur, double* ak,
                                                      const double* h, StagePtrs kp, int stage, int nargs, int n) {
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
        const double hh = h[e];
        xCur[e] = x[e] + dC[stage] * hh;
        for (int j = 0; j < nargs; ++j) {
            double acc = 0.0;
            for (int q = 0; q < stage; ++q) {
                const double a = dA[stage][q];
                if (a != 0.0) acc += a * kp.k[q][j * n + e];
            }
            ak[j * n + e] = acc;
            arg[j * n + e] = solOld[j * n + e] + hh * acc;
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_error(double* err, const double* solOld, double* solNew, const double* h, const double* k1,
                                                  const double* k2, const double* k3, const double* k6, const double* k7,
                                                  const double* k8, const double* k9, const double* k10, int n, int nargs,
                                                  double rtol, double atol) {
    // The error estimate is a sum of stages that cancels by many orders of
    // magnitude, so its value follows the rounding of every product: without
    // contraction into fused multiply-adds the sums round as the reference's
    // own host build of src/DOPR853.cu:586-616 rounds them.
#pragma clang fp contract(off)
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
        double et = 0.0, e2 = 0.0;
        const double hh = h[e];
        for (int j = 0; j < nargs; ++j) {
            const int x = j * n + e;
            const double temp = kB1 * k1[x] + kB6 * k6[x] + kB7 * k7[x] + kB8 * k8[x] + kB9 * k9[x] + kB10 * k10[x] + kB11 * k2[x] + kB12 * k3[x];
            const double yo = solOld[x];
            const double yn = yo + hh * temp;
            solNew[x] = yn;
            const double mx = (yo > yn) ? yo : yn;
            const double sk = 1.0 / (atol + rtol * mx);
            e2 += pow((temp - kBHH1 * k1[x] - kBHH2 * k9[x] - kBHH3 * k3[x]) * sk, (double)nargs);
            et += pow((kE1 * k1[x] + kE6 * k6[x] + kE7 * k7[x] + kE8 * k8[x] + kE9 * k9[x] + kE10 * k10[x] + kE11 * k2[x] + kE12 * k3[x]) * sk,
                      (double)nargs);
        }
        double denom = et + 0.01 * e2;
        denom = (denom < 0.0) ? 1.0 : denom;
        err[e] = fabs(hh) * et * sqrt(1.0 / ((double)nargs * denom));
    }
}

__global__ __launch_bounds__(kBlock) void k_controller(unsigned char* flag, const double* err, double* errOld, unsigned char* prevRej,
                                                       double* h, int n) {
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
        const double er = err[e];
        if (er <= 1.0) {
            double scale;
            if (er == 0.0) scale = kMaxScale;
            else {
                scale = kSafe * pow(er, -kAlpha);   // errOld^beta with beta = 0
                scale = (scale > kMaxScale) ? kMaxScale : ((scale < kMinScale) ? kMinScale : scale);
            }
            if (prevRej[e] && scale > 1.0) scale = 1.0;
            h[e] = h[e] * scale;
            double v = er;
            v = (v < 1e-4) ? 1e-4 : ((v > 1e300) ? 1e300 : v);
            errOld[e] = v;
            prevRej[e] = 0;
            flag[e] = 1;
        } else {
            prevRej[e] = 1;
            flag[e] = 0;
            double t = kSafe * pow(er, -kAlpha);
            t = (t < kMinScale) ? kMinScale : ((t > 1e300) ? 1e300 : t);
            h[e] = h[e] * t;
        }
    }
}

// ------------------------------------------------------------ batch pipeline
// Tables of one binary per wavefront: lane q < 35 sets up flux mode q's
// binary-independent part (flux_mode_init, with its flags), then lane 0 the
// Hamiltonian coefficients, the modes' PN coefficient tables and the 13 pieces,
// straight into the binary's BinData.  With one binary per lane the 35 modes
// ran in turn; this kernel starts every pass, so its latency is on the pass's
// critical path.
__device__ __forceinline__ void bin_setup_one(const double* add, int n, int i, int lane, int use_hm, BinData* bd, ModeCoef* modes) {
    const double m1 = add[i], m2 = add[n + i], c1 = add[2 * n + i], c2 = add[3 * n + i];
    const double nu = nu_of(m1, m2);
    double ef = nu;
    if (ef > 0.25 && ef < 0.25 + 1e-4) ef = 0.25;
    ModeCoef* mc = modes + (size_t)i * kNumModes;
    if (lane < kNumModes) {
        ModeCoef& z = mc[lane];
        flux_mode_init(z, lane, add[(10 + 2 * lane) * n + i], add[(11 + 2 * lane) * n + i], true);
        // the flux's mode flags follow the flux eta
        int fl = ((z.l + z.m) & 1) ? 0 : 1;
        if (ef == 0.25 && (z.m & 1)) fl |= 2;
        z.flags = fl;
    }
    __syncthreads();   // lane 0's tables below overwrite coefficients the modes' lanes zeroed
    if (lane == 0) {
        BinData& d = bd[i];
        ham_setup(d.h, m1, m2, c1, c2, add[5 * n + i], add[6 * n + i], add[7 * n + i], add[8 * n + i]);
        d.nu = nu;
        d.flux_scale = 1.0 / (8.0 * kPi * nu);
        d.eta_f = ef;
        // setup_class inputs (eob/seobnrv4phm.py:1924-1929)
        const double eta = m1 * m2 / ((m1 + m2) * (m1 + m2));
        const double chiS = 0.5 * (c1 + c2), chiA = 0.5 * (c1 - c2);
        const double a = (1 - 2 * eta) * chiS + (m1 - m2) / (m1 + m2) * chiA;
        const double pf22[4] = {add[10 * n + i], add[11 * n + i], add[12 * n + i], add[13 * n + i]};   // (2,1) and (2,2) prefixes
        flux_setup(mc, d.w, m1, m2, eta, a, chiS, chiA, pf22, add[9 * n + i], use_hm, true);
        piece_setup(d.pc, d.h);
    }
}

__global__ __launch_bounds__(64) void k_bin_setup(const double* add, int n, int use_hm, BinData* bd, ModeCoef* modes) {
    const int i = blockIdx.x;
    if (i >= n) return;
    bin_setup_one(add, n, i, threadIdx.x, use_hm, bd, modes);
}

// The parameter-independent part of the 35 Newtonian multipole prefixes
// (Damour, Iyer & Nagar 2008 Eqs. 5-7; EOBComputeNewtonMultipolePrefixes,
// eob/seobnrv4phm.py:1703-1776): n_lm = (i m)^l 8 pi / (2l+1)!! sqrt((l+1)(l+2) /
// (l(l-1))) for even l+m, -(i m)^l i 16 pi / (2l+1)!! sqrt(...) for odd, in
// the order of operations of eobgpu_amd.pipeline.newton_prefixes so that the
// device's prefixes equal the host's bit for bit.
struct PrefixConsts {
    double re[kNumModes], im[kNumModes];
};

PrefixConsts make_prefix_consts() {
    PrefixConsts c{};
    const double pi = 3.141592653589793;
    int q = 0;
    for (int l = 2; l <= 8; ++l) {
        double dfac = 1.0;
        for (int k = 2 * l + 1; k > 1; k -= 2) dfac *= k;
        for (int m = 1; m <= l; ++m, ++q) {
            double ml = 1.0;
            for (int k = 0; k < l; ++k) ml *= m;
            // (i m)^l: m^l times i^l
            double nre = 0.0, nim = 0.0;
            switch (l & 3) {
                case 0: nre = ml; break;
                case 1: nim = ml; break;
                case 2: nre = -ml; break;
                default: nim = -ml; break;
            }
            if (((l + m) & 1) == 0) {
                const double a = (8.0 * pi) / dfac;
                const double b = std::sqrt((double)((l + 1) * (l + 2)) / (double)(l * (l - 1)));
                c.re[q] = (nre * a) * b;
                c.im[q] = (nim * a) * b;
            } else {
                const double m2v = ((2.0 * l + 1) * (l + 2) * (l * l - m * m)) / ((2.0 * l - 1) * (l + 1) * l * (l - 1));
                const double cim = (16.0 * pi) / dfac;
                const double b = std::sqrt(m2v);
                // -(i m)^l times (0 + i cim)
                c.re[q] = (nim * cim) * b;
                c.im[q] = (-(nre * cim)) * b;
            }
        }
    }
    return c;
}

// x^k for k >= 1 by repeated multiplication (the host's order)
__device__ __forceinline__ double ipow_seq(double x, int k) {
    double p = x;
    for (int j = 1; j < k; ++j) p = p * x;
    return p;
}

// additionalArgs [80][n] from the binaries' masses and spins, as
// SEOBNRv4PHM.setup_class builds it on the host (eob/seobnrv4phm.py:1922-1964;
// masses scaled to M = 1, omega0 = 0), then the tables of k_bin_setup, one
// wavefront per binary: lane q < 35 writes the prefix of mode q.
__global__ __launch_bounds__(64) void k_params_setup(const double* m1in, const double* m2in, const double* chi1, const double* chi2,
                                                     int n, int use_hm, double K, double d5, double dSO, double dSS, double h22c,
                                                     PrefixConsts pc, double* add, BinData* bd, ModeCoef* modes) {
#pragma clang fp contract(off)
    const int i = blockIdx.x, lane = threadIdx.x;
    if (i >= n) return;
    const double Mt = m1in[i] + m2in[i];
    const double m1 = m1in[i] / Mt, m2 = m2in[i] / Mt;
    if (lane == 0) {
        add[i] = m1;
        add[n + i] = m2;
        add[2 * n + i] = chi1[i];
        add[3 * n + i] = chi2[i];
        add[4 * n + i] = 0.0;
        add[5 * n + i] = K;
        add[6 * n + i] = d5;
        add[7 * n + i] = dSO;
        add[8 * n + i] = dSS;
        add[9 * n + i] = h22c;
    }
    if (lane < kNumModes) {
        int l, m;
        mode_lm(lane, l, m);
        const int eps = (l + m) & 1;
        const double M = m1 + m2;
        const double x1 = m1 / M, x2 = m2 / M;
        const double eta = (m1 * m2) / (M * M);
        const double sign = (m & 1) ? -1.0 : 1.0;
        double cdeg;
        if (m1 != m2 || sign == 1.0) cdeg = ipow_seq(x2, l + eps - 1) + sign * ipow_seq(x1, l + eps - 1);
        else cdeg = (l == 2 || l == 3) ? -1.0 : ((l == 4 || l == 5) ? -0.5 : 0.0);
        add[(10 + 2 * lane) * n + i] = (pc.re[lane] * eta) * cdeg;
        add[(11 + 2 * lane) * n + i] = (pc.im[lane] * eta) * cdeg;
    }
    __syncthreads();   // the column is in memory for the whole workgroup
    bin_setup_one(add, n, i, lane, use_hm, bd, modes);
}

// The dissipative initial condition (IC_diss, src/EOB.cu:2666): dr/dt from the
// pphi loss along circular orbits, dr/dt = (dpphi/dt) / (dpphi/dr), minus the
// right-hand side's dr/dt, with dpphi/dr = -H_rr / H_rpphi; evaluated with the
// right-hand side of a 64-lane group (one flux mode per lane); every lane
// returns the same value (lane 0's k: lane 63 evaluates H at pr = 0).
__device__ double ic_diss64(const BinData& bd, const ModeCoef* mc, double r, double pr, double L, int gl) {
    double k[4];
    group_rhs<64>(bd, mc, r, pr, L, gl, k);
    const double k0 = read_lane(k[0], 0), k3 = read_lane(k[3], 0);
    const double hrr = ham_second(bd.h, r, pr, L, 0, 0);
    const double hrL = ham_second(bd.h, r, pr, L, 0, 3);
    const double dLdr = -hrr / hrL;
    const double rdot = k3 / dLdr;
    return rdot - k0;
}

// Circular initial conditions (IC_cons / IC_diss), one wavefront per binary:
// Newton for pphi (every lane the same values), then the reference's false
// position for pr on the dissipative condition.  One binary per lane took the
// 35 flux modes of every evaluation in turn; this kernel starts every pass, so
// its latency is on the pass's critical path.
__global__ __launch_bounds__(64) void k_initial_conditions(const double* r0, const BinData* bdp, const ModeCoef* modes, int n,
                                                           double* pphi0, double* pr0, double* omega0, int max_iter, double err,
                                                           double* cond) {
    const int i = blockIdx.x;
    const int gl = threadIdx.x;
    if (i >= n) return;
    const BinData& bd = bdp[i];
    const ModeCoef* mc = modes + (size_t)i * kNumModes;
    const double r = r0[i];
    double L = sqrt(r);
    for (int it = 0; it < 100; ++it) {
        const HDual R(r, 1.0, 0.0, 0.0), PR(0.0), LL(L, 0.0, 1.0, 0.0);
        const HDual H = ham_eval<HDual>(bd.h, R, PR, LL);
        const double dL = H.a / H.ab;
        L -= dL;
        if (fabs(dL) < 1e-15 * fabs(L)) break;
    }
    const double om = ham_tangent(bd.h, r, 0.0, L, 2).d;
    double a = -3e-2, b = 0.0;
    double fa = ic_diss64(bd, mc, r, a, L, gl), fb = ic_diss64(bd, mc, r, b, L, gl);
    double c = 0.0;
    for (int it = 0; it < max_iter; ++it) {
        c = (a * fb - b * fa) / (fb - fa);
        const double fc = ic_diss64(bd, mc, r, c, L, gl);
        if ((fb < 0.0 && fc < 0.0) || (fb >= 0.0 && fc >= 0.0)) { b = c; fb = fc; }
        else { a = c; fa = fc; }
        if (fabs(fc) < err) break;
    }
    if (gl == 0) {
        omega0[i] = om;
        pphi0[i] = L;
        pr0[i] = c;
        if (cond) {   // the integrator's initial state [4][n]: r, phi = 0, pr, pphi (pr and pphi rows are pr0 and pphi0)
            cond[i] = r;
            cond[(size_t)n + i] = 0.0;
        }
    }
}

// Copy one binary's tables (BinData and its 35 ModeCoef) into this wavefront's
// LDS slot; the whole wavefront takes part, and the wavefront fence makes the
// copy visible to its own later reads.
__device__ __forceinline__ void stage_tables(const IntegParams& P, int b, BinData* lbd, ModeCoef* lmc) {
    static_assert(sizeof(BinData) % 8 == 0 && sizeof(ModeCoef) % 8 == 0, "tables are copied as 8-byte words");
    constexpr int nb = sizeof(BinData) / 8, nm = kNumModes * sizeof(ModeCoef) / 8;
    const int lane = threadIdx.x & 63;
    const double* gb = reinterpret_cast<const double*>(P.bd + b);
    const double* gm = reinterpret_cast<const double*>(P.mc + (size_t)b * kNumModes);
    double* db = reinterpret_cast<double*>(lbd);
    double* dm = reinterpret_cast<double*>(lmc);
    for (int i = lane; i < nb; i += 64) db[i] = gb[i];
    for (int i = lane; i < nm; i += 64) dm[i] = gm[i];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// LDS per wavefront of a 64-lane-group launch: its binary's tables and the
// stage weights (stage_rows_copy).  The
// launch sizes the dynamic LDS by the wavefronts of a block, so a one-wavefront
// block holds one copy (about 12.6 KB) and leaves the rest of the CU's LDS to
// other kernels.
constexpr size_t kStageBytesPerWave = sizeof(BinData) + kNumModes * sizeof(ModeCoef) + kStageRowsBytes;
static_assert(sizeof(BinData) % 16 == 0 && kStageBytesPerWave % 16 == 0, "16-byte aligned tables");

template <int G>
__global__ __launch_bounds__(kBlock) void k_integrate_persistent(IntegParams P) {
    // The integrating wavefronts issue first on a SIMD they share with the
    // downstream kernels of binaries that have already finished (the streamed
    // pass): the integrator's makespan is the pass time, and the others fill
    // its issue gaps.  Only the one-binary-per-wavefront build takes it: the
    // lane-per-binary build (G = 1) spills about 2 KB per lane, and with the
    // priority instruction at its entry its register allocation changed and two
    // of sixteen fixed-step binaries no longer matched its plain-kernel twin,
    // which shares the step code (the G = 64, 16 builds stayed bit-identical).
    if constexpr (G == 64) __builtin_amdgcn_s_setprio(3);
    // 64-lane groups own a wavefront: each keeps its binary's tables in LDS
    constexpr bool kStage = (G == 64);
    extern __shared__ __align__(16) unsigned char stage_lds[];
    const int wave = threadIdx.x >> 6;
    const int gl = threadIdx.x % G;
    BinData* lbd = nullptr;
    ModeCoef* lmc = nullptr;
    if constexpr (kStage) {
        unsigned char* base = stage_lds + (size_t)wave * kStageBytesPerWave;
        lbd = reinterpret_cast<BinData*>(base);
        lmc = reinterpret_cast<ModeCoef*>(base + sizeof(BinData));
        stage_rows_copy(reinterpret_cast<double*>(lmc + kNumModes), threadIdx.x & 63);   // read after stage_tables' wave barrier
    }
    GState<G> s;
    s.pass = 0;
    s.b = 0;
    bool exhausted = false;
    while (true) {
        if (!exhausted && s.pass == 0) {
            int b = 0;
            if (gl == 0) b = atomicAdd(P.queue, 1);
            b = gshfl_i<G>(b, 0);
            if (b >= P.nbin) exhausted = true;
            else {
                if constexpr (kStage) stage_tables(P, b, lbd, lmc);
                begin_binary<G>(P, s, b, gl);
            }
        }
        if (__all(exhausted)) break;
        if (!exhausted) {
            bool need_new;
            if constexpr (kStage) iterate<G, true>(P, s, gl, need_new, lbd, lmc);
            else iterate<G>(P, s, gl, need_new);
        }
    }
}

// plain mode: state kept in global memory between launches
struct PlainState {
    int *b, *pass, *step_num, *sn1, *code, *flags;   // flags: attempted | just_acc<<1 | stop_flag<<2 | prev_reject<<3
    double *x, *h, *err_old, *h_init, *om, *y;       // om [4][n], y [4][n]
    unsigned long long* att;
};

template <int G>
__device__ __forceinline__ void load_state(const PlainState& S, GState<G>& s, int b, int gl, int n) {
    s.b = b;
    s.pass = S.pass[b]; s.step_num = S.step_num[b]; s.sn1 = S.sn1[b]; s.code = S.code[b];
    const int f = S.flags[b];
    s.attempted = f & 1; s.just_acc = (f >> 1) & 1; s.stop_flag = (f >> 2) & 1; s.prev_reject = (f >> 3) & 1;
    s.fail = (f >> 4) & 3;
    s.k1_ok = 0;
    s.x = S.x[b]; s.h = S.h[b]; s.err_old = S.err_old[b]; s.h_init = S.h_init[b];
    s.om0 = S.om[b]; s.om1 = S.om[n + b]; s.om2 = S.om[2 * n + b]; s.om3 = S.om[3 * n + b];
    s.att = S.att[b];
#pragma unroll
    for (int q = 0; q < GState<G>::NC; ++q) {
        const int c = comp_of<G>(gl, q);
        s.y[q] = (c < 4) ? S.y[(size_t)c * n + b] : 0.0;
    }
}

template <int G>
__device__ __forceinline__ void store_state(const PlainState& S, const GState<G>& s, int gl, int n) {
    const int b = s.b;
    if (gl == 0) {
        S.pass[b] = s.pass; S.step_num[b] = s.step_num; S.sn1[b] = s.sn1; S.code[b] = s.code;
        S.flags[b] = s.attempted | (s.just_acc << 1) | (s.stop_flag << 2) | (s.prev_reject << 3) | (s.fail << 4);
        S.x[b] = s.x; S.h[b] = s.h; S.err_old[b] = s.err_old; S.h_init[b] = s.h_init;
        S.om[b] = s.om0; S.om[n + b] = s.om1; S.om[2 * n + b] = s.om2; S.om[3 * n + b] = s.om3;
        S.att[b] = s.att;
    }
#pragma unroll
    for (int q = 0; q < GState<G>::NC; ++q) {
        const int c = comp_of<G>(gl, q);
        if (c < 4) S.y[(size_t)c * n + b] = s.y[q];
    }
}

template <int G>
__global__ __launch_bounds__(kBlock) void k_plain_init(IntegParams P, PlainState S, int* active) {
    const int gi = (blockIdx.x * blockDim.x + threadIdx.x) / G;
    const int gl = threadIdx.x % G;
    if (gi >= P.nbin) return;
    GState<G> s;
    begin_binary<G>(P, s, gi, gl);
    store_state<G>(S, s, gl, P.nbin);
    if (gl == 0) active[gi] = gi;
}

template <int G>
__global__ __launch_bounds__(kBlock) void k_plain_step(IntegParams P, PlainState S, const int* active_in, int n_active,
                                                       int* active_out, int* n_out) {
    const int gi = (blockIdx.x * blockDim.x + threadIdx.x) / G;
    const int gl = threadIdx.x % G;
    if (gi >= n_active) return;
    const int b = active_in[gi];
    GState<G> s;
    load_state<G>(S, s, b, gl, P.nbin);
    bool need_new;
    iterate<G>(P, s, gl, need_new);
    store_state<G>(S, s, gl, P.nbin);
    if (gl == 0 && (s.pass == 1 || s.pass == 2)) {
        const int slot = atomicAdd(n_out, 1);
        active_out[slot] = b;
    }
}

// ---------------------------------------------- the reference's batch semantics
// The reference's StoppingCriterion couples the binaries of a batch
// (eob/seobnrv4phm.py:82-150): it runs on the binaries accepted in this driver
// iteration, index_update, in ascending binary order; its third check reads
// the omega history old_omegas[step-4 .. step-1] in column `q`, the binary's
// POSITION inside index_update, not its own column; and while any binary of
// index_update has passed step 5 without another stop code, only those
// binaries store their omega (at their own column).  A binary's stop code and
// step count then depend on its batch mates.  The coupled driver reproduces
// this: one launch attempts a step for every active binary, a one-block launch
// applies the criterion to the accepted ones in order, and both passes of
// run_trajectory are batch-synchronous (pass 2 starts for every binary after
// the last binary's pass 1 ends, eob/seobnrv4phm.py:2086-2123), each with its
// omega history zeroed (StoppingCriterion.setup).  Binaries that the
// non-finite guard ends in pass 1 sit out pass 2, whose history columns then
// number the remaining binaries in order.
struct CoupledState {
    double* k1;       // [4][n] k1 at the binary's current state (valid when k1_ok)
    int* k1_ok;       // [n]
    double* chk;      // [4][n] r, dr/dt, omega, dpr/dt at the state just accepted
    int* acc;         // [n] 1 if the binary's last attempt was accepted
    int* code;        // [n] the criterion's code for the accepted binaries of this iteration
    int* col;         // [n] the binary's column in the pass's omega history
    double* old_om;   // [max_step][n] omega history (column stride n)
};

constexpr int kCoupledBlock = 1024;

template <int G>
__global__ __launch_bounds__(kBlock) void k_coupled_attempt(IntegParams P, PlainState S, CoupledState C, const int* active,
                                                            int n_active) {
    const int gi = (blockIdx.x * blockDim.x + threadIdx.x) / G;
    const int gl = threadIdx.x % G;
    if (gi >= n_active) return;
    const int b = active[gi];
    const int n = P.nbin;
    GState<G> s;
    load_state<G>(S, s, b, gl, n);
    const BinData& bd = P.bd[b];
    const ModeCoef* mc = P.mc + (size_t)b * kNumModes;
    double k1[4];
    if (C.k1_ok[b]) {
#pragma unroll
        for (int c = 0; c < 4; ++c) k1[c] = C.k1[(size_t)c * n + b];
    } else {
        double yf[4];
        gather_state<G>(s, gl, yf);
        group_rhs<G>(bd, mc, yf[0], yf[2], yf[3], gl, k1);
        if constexpr (G == 64) {
#pragma unroll
            for (int c = 0; c < 4; ++c) k1[c] = read_lane(k1[c], 0);   // lane 63 holds the pr = 0 branch
        }
    }
    const bool fixed = (s.pass == 2) || P.fix_step;
    attempt_step<G>(P, s, bd, mc, k1, gl, fixed);
    if (s.just_acc) {
        // the criterion's right-hand side at the accepted state, which is also
        // the next attempt's k1 (eob/seobnrv4phm.py:94)
        double yf[4], kn[4];
        gather_state<G>(s, gl, yf);
        group_rhs<G>(bd, mc, yf[0], yf[2], yf[3], gl, kn);
        if constexpr (G == 64) {
#pragma unroll
            for (int c = 0; c < 4; ++c) kn[c] = read_lane(kn[c], 0);
        }
        if (gl == 0) {
#pragma unroll
            for (int c = 0; c < 4; ++c) C.k1[(size_t)c * n + b] = kn[c];
            C.chk[b] = yf[0];
            C.chk[(size_t)n + b] = kn[0];
            C.chk[2 * (size_t)n + b] = kn[1];
            C.chk[3 * (size_t)n + b] = kn[2];
        }
    } else if (gl == 0) {
#pragma unroll
        for (int c = 0; c < 4; ++c) C.k1[(size_t)c * n + b] = k1[c];
    }
    if (gl == 0) {
        C.k1_ok[b] = 1;
        C.acc[b] = s.just_acc;
    }
    store_state<G>(S, s, gl, n);
}

// block-wide exclusive prefix sum of v (one value per thread); *total = the sum
__device__ __forceinline__ int block_exclusive_scan(int v, int* total) {
    __shared__ int part[kCoupledBlock / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int incl = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int o = __shfl_up(incl, off, 64);
        if (lane >= off) incl += o;
    }
    __syncthreads();
    if (lane == 63) part[wave] = incl;
    __syncthreads();
    int before = 0, all = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
        const int p = part[w];
        if (w < wave) before += p;
        all += p;
    }
    *total = all;
    return before + incl - v;
}

// The criterion and the end-of-iteration bookkeeping for the active binaries,
// in one block.  Thread t takes a contiguous run of the active list (which is in
// ascending binary order), so positions inside index_update follow from one scan.
__global__ __launch_bounds__(kCoupledBlock) void k_coupled_check(IntegParams P, PlainState S, CoupledState C, const int* active_in,
                                                                 int n_active, int* active_out, int* n_out) {
    const int n = P.nbin;
    const int per = (n_active + (int)blockDim.x - 1) / (int)blockDim.x;
    const int lo = min((int)threadIdx.x * per, n_active), hi = min(lo + per, n_active);
    // codes 2-4 and which accepted binaries go on to the omega checks (inds_here)
    int n_acc = 0, any = 0;
    for (int j = lo; j < hi; ++j) {
        const int b = active_in[j];
        if (!C.acc[b]) continue;
        ++n_acc;
        const double r = C.chk[b], drdt = C.chk[(size_t)n + b], dprdt = C.chk[3 * (size_t)n + b];
        int code = 0;
        if (r < 1.0) code = 2;
        else if (r < 6.0 && drdt > 0.0) code = 3;
        else if (r < 6.0 && dprdt > 0.0) code = 4;
        if (!P.use_stop) code = 0;
        const int here = P.use_stop && code == 0 && S.step_num[b] > 5;
        any |= here;
        C.code[b] = code | (here << 8);
    }
    int n_acc_all;
    int q = block_exclusive_scan(n_acc, &n_acc_all);
    const int any_here = __syncthreads_or(any);
    // third check against column q; every read of the history happens before any store
    for (int j = lo; j < hi; ++j) {
        const int b = active_in[j];
        if (!C.acc[b]) continue;
        const int cv = C.code[b];
        if ((cv >> 8) && S.step_num[b] <= P.max_step) {
            const int s = S.step_num[b];
            const double r = C.chk[b], om = C.chk[2 * (size_t)n + b];
            const double* h = C.old_om;
            const bool first = r < 6.0;
            const bool second = om < h[(size_t)(s - 1) * n + C.col[b]];
            const double o0 = h[(size_t)(s - 4) * n + q], o1 = h[(size_t)(s - 3) * n + q], o2 = h[(size_t)(s - 2) * n + q],
                         o3 = h[(size_t)(s - 1) * n + q];
            const bool third = (o1 - o0 < 0.0) && (o2 - o1 < 0.0) && (o3 - o2 < 0.0);
            if (first && second && third) C.code[b] = cv + 5;
        }
        ++q;
    }
    __syncthreads();
    // stores, codes, and the end of pass for every active binary
    int n_keep = 0;
    for (int j = lo; j < hi; ++j) {
        const int b = active_in[j];
        int f = S.flags[b];
        if (C.acc[b]) {
            const int cv = C.code[b];
            const int code = cv & 0xff;
            const int s = S.step_num[b];
            const int srow = (s < 0) ? s + P.max_step : s;   // numpy's negative rows
            if (P.use_stop && (!any_here || (cv >> 8)) && srow >= 0 && srow < P.max_step)
                C.old_om[(size_t)srow * n + C.col[b]] = C.chk[2 * (size_t)n + b];
            S.code[b] = code;
            f = (f & ~4) | ((code != 0) << 2);
            S.flags[b] = f;
        }
        const int fail = (f >> 4) & 3;
        const int step_num = S.step_num[b];
        const bool term = (S.x[b] >= P.tmax) || (step_num >= P.max_step - 1) || ((f >> 2) & 1) || fail;
        if (!term) { ++n_keep; continue; }
        const int pass = S.pass[b];
        if (fail) {
            S.code[b] = -fail;
            if (pass == 1) S.sn1[b] = step_num;
        }
        if (pass == 1 && P.two_pass && !fail) {
            S.sn1[b] = step_num;
            S.pass[b] = 3;   // pass 1 over; pass 2 starts with the batch's
        } else {
            P.num_steps[b] = step_num + 1;
            P.stop_info[b] = S.code[b];
            P.pass1_steps[b] = P.two_pass ? S.sn1[b] : step_num;
            P.attempts[b] = S.att[b];
            S.pass[b] = 0;
        }
    }
    int n_keep_all;
    int w = block_exclusive_scan(n_keep, &n_keep_all);
    for (int j = lo; j < hi; ++j) {
        const int b = active_in[j];
        const int p = S.pass[b];
        if (p == 1 || p == 2) active_out[w++] = b;
    }
    if (threadIdx.x == 0) *n_out = n_keep_all;
}

// Pass 2 of the coupled driver: the binaries whose pass 1 ended normally, in
// ascending order, numbered 0.. as the columns of the pass's omega history.
__global__ __launch_bounds__(kCoupledBlock) void k_coupled_pass2_list(IntegParams P, PlainState S, CoupledState C, int* active_out,
                                                                      int* n_out) {
    const int n = P.nbin;
    const int per = (n + (int)blockDim.x - 1) / (int)blockDim.x;
    const int lo = min((int)threadIdx.x * per, n), hi = min(lo + per, n);
    int cnt = 0;
    for (int b = lo; b < hi; ++b) cnt += (S.pass[b] == 3);
    int total;
    int w = block_exclusive_scan(cnt, &total);
    for (int b = lo; b < hi; ++b) {
        if (S.pass[b] != 3) continue;
        C.col[b] = w;
        active_out[w++] = b;
    }
    if (threadIdx.x == 0) *n_out = total;
}

template <int G>
__global__ __launch_bounds__(kBlock) void k_coupled_restart(IntegParams P, PlainState S, CoupledState C, const int* active,
                                                            int n_active) {
    const int gi = (blockIdx.x * blockDim.x + threadIdx.x) / G;
    const int gl = threadIdx.x % G;
    if (gi >= n_active) return;
    const int b = active[gi];
    GState<G> s;
    load_state<G>(S, s, b, gl, P.nbin);
    restart_pass2<G>(P, s, gl);
    if (gl == 0) C.k1_ok[b] = 0;
    store_state<G>(S, s, gl, P.nbin);
}

__global__ void k_zero_ints(int* a, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) a[i] = 0;
}

__global__ void k_iota(int* a, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) a[i] = i;
}

template <int G>
int run_coupled(const IntegParams& P, PlainState S, CoupledState C, int* act0, int* act1, int* cnt, int* host_cnt, hipStream_t st,
                long long* iters) {
    const int gpb = kBlock / G;
    const size_t hist_bytes = sizeof(double) * (size_t)P.max_step * P.nbin;
    hipLaunchKernelGGL(k_plain_init<G>, dim3(blocks_for(P.nbin, gpb)), dim3(kBlock), 0, st, P, S, act0);
    int rc = check_launch("k_plain_init");
    if (rc) return rc;
    hipLaunchKernelGGL(k_iota, dim3(blocks_for(P.nbin, kBlock)), dim3(kBlock), 0, st, C.col, P.nbin);
    if ((rc = check_launch("k_iota"))) return rc;
    if ((rc = hip_status(hipMemsetAsync(C.k1_ok, 0, sizeof(int) * P.nbin, st), "memset k1_ok"))) return rc;
    long long it = 0;
    int n_active = P.nbin;
    for (int pass = 1; pass <= (P.two_pass ? 2 : 1); ++pass) {
        if (pass == 2) {
            hipLaunchKernelGGL(k_coupled_pass2_list, dim3(1), dim3(kCoupledBlock), 0, st, P, S, C, act0, cnt);
            if ((rc = check_launch("k_coupled_pass2_list"))) return rc;
            if ((rc = hip_status(hipMemcpyAsync(host_cnt, cnt, sizeof(int), hipMemcpyDeviceToHost, st), "count copy"))) return rc;
            if ((rc = hip_status(hipStreamSynchronize(st), "sync"))) return rc;
            n_active = *host_cnt;
            if (n_active == 0) break;
            hipLaunchKernelGGL(k_coupled_restart<G>, dim3(blocks_for(n_active, gpb)), dim3(kBlock), 0, st, P, S, C, act0, n_active);
            if ((rc = check_launch("k_coupled_restart"))) return rc;
        }
        if (P.use_stop && (rc = hip_status(hipMemsetAsync(C.old_om, 0, hist_bytes, st), "memset old_omegas"))) return rc;
        while (n_active > 0) {
            hipLaunchKernelGGL(k_coupled_attempt<G>, dim3(blocks_for(n_active, gpb)), dim3(kBlock), 0, st, P, S, C, act0, n_active);
            if ((rc = check_launch("k_coupled_attempt"))) return rc;
            hipLaunchKernelGGL(k_coupled_check, dim3(1), dim3(kCoupledBlock), 0, st, P, S, C, act0, n_active, act1, cnt);
            if ((rc = check_launch("k_coupled_check"))) return rc;
            if ((rc = hip_status(hipMemcpyAsync(host_cnt, cnt, sizeof(int), hipMemcpyDeviceToHost, st), "count copy"))) return rc;
            if ((rc = hip_status(hipStreamSynchronize(st), "sync"))) return rc;
            n_active = *host_cnt;
            int* t = act0; act0 = act1; act1 = t;
            ++it;
        }
    }
    if (iters) *iters = it;
    return EOBHIP_OK;
}

// (2,2) mode on knots, one thread per knot: get_hlms (eob/seobnrv4phm.py:2138-2221).
// Registers capped at three wavefronts per SIMD (168, a few spills): with the
// 284 it would take, a wavefront could not share a SIMD with an integrating one
// (292 of 512), so knots of finished binaries waited for whole free SIMDs.
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3))) void k_hlm22_knots(const double* dense, const int* num_steps, const BinData* bdp,
                                                        const ModeCoef* modes, int n, int max_step, double* knots) {
    const int b = blockIdx.y;
    if (b >= n) return;
    const int len = num_steps[b];
    const BinData& bd = bdp[b];
    const ModeCoef& mc22 = modes[(size_t)b * kNumModes + 1];
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < len; s += gridDim.x * blockDim.x) {
        const double* row = dense + ((size_t)b * max_step + s) * 5;
        const double r = row[1], phi = row[2], pr = row[3], L = row[4];
        const Dual dfull = ham_tangent(bd.h, r, pr, L, 2);
        const double omega = dfull.d;
        const double omc = ham_tangent(bd.h, r, pr, L, 3).d;
        const double vPhiIn = 1.0 / (omc * omc * (r * r * r));
        const double Hreal = bd.nu * dfull.v;
        const double v = cbrt(omega);
        const Cplx h = hlm22_knot(mc22, bd.w, bd.nu, r, phi, v, Hreal, vPhiIn);
        // h e^{2 i phi}: the slowly varying part the reference splines (eob/seobnrv4phm.py:2338)
        const double cp = cos(2.0 * phi), sp = sin(2.0 * phi);
        double* o = knots + ((size_t)b * max_step + s) * 3;
        o[0] = phi;
        o[1] = h.re * cp - h.im * sp;
        o[2] = h.re * sp + h.im * cp;
    }
}

// ----------------------------------------------------------- higher modes
// The modes a batch evaluates on its knots (get_hlms' ells, mms).
constexpr int kMaxWaveModes = 15;
struct ModeList {
    int nm;
    int l[kMaxWaveModes], m[kMaxWaveModes];
};

// Waveform tables of the listed modes for every binary of a batch (one thread
// per binary and mode), from its additional args (masses and spins, M = 1) and
// the flux tables of its setup: wtab [n][nm].
__global__ __launch_bounds__(kBlock) void k_wave_tables_batch(const double* add, int n, int use_hm, const ModeCoef* modes,
                                                              const BinData* bd, ModeList ml, WaveMode* wtab) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)n * ml.nm) return;
    const int i = (int)(t / ml.nm), j = (int)(t % ml.nm);
    const int l = ml.l[j], m = ml.m[j];
    const double m1 = add[i], m2 = add[n + i], c1 = add[2 * n + i], c2 = add[3 * n + i];
    const double eta = m1 * m2 / ((m1 + m2) * (m1 + m2));
    wave_mode_setup(wtab[t], modes[(size_t)i * kNumModes + flux_index(l, m)], l, m, m1, m2, eta, 0.5 * (c1 + c2), 0.5 * (c1 - c2), use_hm,
                    &bd[i].w);
}

// The listed modes on knots, one thread per knot: get_hlms (eob/seobnrv4phm.py:2138-2221)
// and the series compute_full_waveform splines (:2336-2356).  knots
// [n][max_step][2 nm + 1] = Re(h_j e^{i m_j phi}) for j < nm, their Im, phi.
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(2))) void k_hlm_modes_knots(
    const double* dense, const int* num_steps, const BinData* bdp, const WaveMode* wtab, const double* add, int n, int max_step, int nm,
    double* knots) {
    const int b = blockIdx.y;
    if (b >= n) return;
    const int len = num_steps[b];
    const BinData& bd = bdp[b];
    const int S = 2 * nm + 1;
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < len; s += gridDim.x * blockDim.x) {
        const double* row = dense + ((size_t)b * max_step + s) * 5;
        const double r = row[1], phi = row[2], pr = row[3], L = row[4];
        const Dual dfull = ham_tangent(bd.h, r, pr, L, 2);
        const double omega = dfull.d;
        const double omc = ham_tangent(bd.h, r, pr, L, 3).d;
        const double vPhiIn = 1.0 / (omc * omc * (r * r * r));
        const double Hreal = bd.nu * dfull.v;
        const double v = cbrt(omega);
        double* o = knots + ((size_t)b * max_step + s) * S;
        for (int j = 0; j < nm; ++j) {
            const WaveMode& w = wtab[(size_t)b * nm + j];
            const int fi = flux_index(w.l, w.m);
            const Cplx h = hlm_knot(w, add[(10 + 2 * fi) * n + b], add[(11 + 2 * fi) * n + b], bd.nu, r, phi, L, v, Hreal, vPhiIn, true);
            o[j] = h.re;
            o[nm + j] = h.im;
        }
        o[2 * nm] = phi;
    }
}

template <int G>
int launch_persistent(const IntegParams& P, int block, int grid_groups, hipStream_t st) {
    const int groups_per_block = block / G;
    int nblk = (grid_groups + groups_per_block - 1) / groups_per_block;
    if (nblk < 1) nblk = 1;
    const size_t lds = (G == 64) ? (size_t)(block / 64) * kStageBytesPerWave : 0;
    hipLaunchKernelGGL(k_integrate_persistent<G>, dim3(nblk), dim3(block), lds, st, P);
    return check_launch("k_integrate_persistent");
}

template <int G>
int run_plain(const IntegParams& P, PlainState S, int* act0, int* act1, int* cnt, int* host_cnt, int block, hipStream_t st,
              long long* iters) {
    const int gpb = block / G;
    hipLaunchKernelGGL(k_plain_init<G>, dim3(blocks_for(P.nbin, gpb)), dim3(block), 0, st, P, S, act0);
    int rc = check_launch("k_plain_init");
    if (rc) return rc;
    int n_active = P.nbin;
    long long it = 0;
    while (n_active > 0) {
        rc = hip_status(hipMemsetAsync(cnt, 0, sizeof(int), st), "memset");
        if (rc) return rc;
        hipLaunchKernelGGL(k_plain_step<G>, dim3(blocks_for(n_active, gpb)), dim3(block), 0, st, P, S, act0, n_active, act1, cnt);
        rc = check_launch("k_plain_step");
        if (rc) return rc;
        rc = hip_status(hipMemcpyAsync(host_cnt, cnt, sizeof(int), hipMemcpyDeviceToHost, st), "count copy");
        if (rc) return rc;
        rc = hip_status(hipStreamSynchronize(st), "sync");
        if (rc) return rc;
        n_active = *host_cnt;
        int* t = act0; act0 = act1; act1 = t;
        ++it;
    }
    if (iters) *iters = it;
    return EOBHIP_OK;
}

// One wavefront that waits `cycles` shader clocks (bounded: it always ends).
__global__ void k_spin(long long cycles) {
    const long long t0 = (long long)__builtin_amdgcn_s_memtime();
    while ((long long)__builtin_amdgcn_s_memtime() - t0 < cycles) {
    }
}

}  // namespace

// ====================================================================== C ABI
extern "C" {

int eobhip_spin(long long cycles, void* stream) {
    EOBHIP_REQUIRE(cycles >= 0 && cycles <= 4000000000LL, "cycles must be in [0, 4e9]");
    hipLaunchKernelGGL(k_spin, dim3(1), dim3(64), 0, as_stream(stream), cycles);
    return check_launch("spin");
}

const char* eobhip_last_error(void) { return g_err; }
int eobhip_version(void) { return 1; }

int eobhip_dop853_tableau(double* out) {
    EOBHIP_REQUIRE(out, "null output");
    constexpr double A[12][12] = DOP_A;
    const double c[12] = {0.0, 0.05260015195876773, 0.0789002279381516, 0.1183503419072274, 0.2816496580927726, 0.3333333333333333,
                          0.25, 0.3076923076923077, 0.6512820512820513, 0.6, 0.8571428571428571, 1.0};
    const double b[12] = {kB1, 0, 0, 0, 0, kB6, kB7, kB8, kB9, kB10, kB11, kB12};
    const double e[12] = {kE1, 0, 0, 0, 0, kE6, kE7, kE8, kE9, kE10, kE11, kE12};
    int o = 0;
    for (int i = 0; i < 12; ++i) out[o++] = c[i];
    for (int i = 0; i < 12; ++i)
        for (int j = 0; j < 12; ++j) out[o++] = A[i][j];
    for (int i = 0; i < 12; ++i) out[o++] = b[i];
    for (int i = 0; i < 12; ++i) out[o++] = e[i];
    out[o++] = kBHH1; out[o++] = kBHH2; out[o++] = kBHH3;
    return EOBHIP_OK;
}

int eobhip_evaluate_Ham_align_AD(double* out, const double* r, const double* phi, const double* pr, const double* pphi,
                                 const double* m_1, const double* m_2, const double* chi_1, const double* chi_2, double K, double d5,
                                 double dSO, double dSS, int numSys, void* stream) {
    (void)phi;
    EOBHIP_REQUIRE(numSys >= 0, "numSys < 0");
    if (numSys == 0) return EOBHIP_OK;
    hipLaunchKernelGGL(k_evaluate_H, dim3(blocks_for(numSys, kBlock)), dim3(kBlock), 0, as_stream(stream), out, r, pr, pphi, m_1, m_2,
                       chi_1, chi_2, K, d5, dSO, dSS, numSys);
    return check_launch("evaluate_Ham_align_AD");
}

int eobhip_grad_Ham_align_AD(const double* arg, double* grad_out, const double* add, int numSys, void* stream) {
    EOBHIP_REQUIRE(numSys >= 0, "numSys < 0");
    if (numSys == 0) return EOBHIP_OK;
    hipLaunchKernelGGL(k_grad_H, dim3(blocks_for(numSys, kBlock)), dim3(kBlock), 0, as_stream(stream), arg, grad_out, add, numSys);
    return check_launch("grad_Ham_align_AD");
}

int eobhip_hessian_Ham_align_AD(const double* arg, double* hess_out, const double* add, int numSys, void* stream) {
    EOBHIP_REQUIRE(numSys >= 0, "numSys < 0");
    if (numSys == 0) return EOBHIP_OK;
    hipLaunchKernelGGL(k_hessian_H, dim3(blocks_for(numSys, kBlock)), dim3(kBlock), 0, as_stream(stream), arg, hess_out, add, numSys);
    return check_launch("hessian_Ham_align_AD");
}

int eobhip_IC_cons(double* res, const double* x, double* args, const double* add, double* grad_out, int numSys, void* stream) {
    EOBHIP_REQUIRE(numSys >= 0, "numSys < 0");
    if (numSys == 0) return EOBHIP_OK;
    hipLaunchKernelGGL(k_IC_cons, dim3(blocks_for(numSys, kBlock)), dim3(kBlock), 0, as_stream(stream), res, x, args, add, grad_out,
                       numSys);
    return check_launch("IC_cons");
}

int eobhip_root_find_all(double* xOut, const double* x0In, const double* argsIn, const double* add, int max_iter, double err,
                         int numBinAll, int n, int num_args, int num_add_args, void* stream) {
    (void)argsIn;
    EOBHIP_REQUIRE(n == 2, "root_find_all: n must be 2 (r, pphi)");
    EOBHIP_REQUIRE(num_args == 4 && num_add_args == kNAdd, "root_find_all: 4 args and 80 additional args per binary");
    EOBHIP_REQUIRE(numBinAll >= 0 && max_iter >= 0, "bad sizes");
    if (numBinAll == 0) return EOBHIP_OK;
    hipLaunchKernelGGL(k_root_find_all, dim3(blocks_for(numBinAll, kBlock)), dim3(kBlock), 0, as_stream(stream), xOut, x0In, add,
                       max_iter, err, numBinAll);
    return check_launch("root_find_all");
}

int eobhip_seobnr_create(eobhip_seobnr_t* out) {
    EOBHIP_REQUIRE(out, "null handle pointer");
    *out = new eobhip_seobnr();
    return EOBHIP_OK;
}

int eobhip_seobnr_deallocate_information(eobhip_seobnr_t h) {
    EOBHIP_REQUIRE(h, "null handle");
    if (h->allocated) {
        hipFree(h->modes);
        hipFree(h->wave);
        hipFree(h->wmodes);
        h->modes = nullptr;
        h->wave = nullptr;
        h->wmodes = nullptr;
        h->allocated = false;
        h->numSys = 0;
    }
    return EOBHIP_OK;
}

int eobhip_seobnr_destroy(eobhip_seobnr_t h) {
    if (!h) return EOBHIP_OK;
    eobhip_seobnr_deallocate_information(h);
    delete h;
    return EOBHIP_OK;
}

int eobhip_seobnr_update_information(eobhip_seobnr_t h, const double* m1, const double* m2, const double* eta, const double* tplspin,
                                     const double* chi_S, const double* chi_A, int use_hm, int numSys,
                                     const double* newtonian_prefixes_hlms, void* stream) {
    EOBHIP_REQUIRE(h, "null handle");
    EOBHIP_REQUIRE(!h->allocated, "update_information: information already allocated (call deallocate_information first)");
    EOBHIP_REQUIRE(numSys > 0, "numSys must be positive");
    int rc = hip_status(hipMalloc(&h->modes, sizeof(ModeCoef) * (size_t)numSys * kNumModes), "hipMalloc modes");
    if (rc) return rc;
    rc = hip_status(hipMalloc(&h->wave, sizeof(WaveCoef) * (size_t)numSys), "hipMalloc wave");
    if (rc) { hipFree(h->modes); h->modes = nullptr; return rc; }
    rc = hip_status(hipMalloc(&h->wmodes, sizeof(WaveMode) * (size_t)numSys * kNumModes), "hipMalloc wmodes");
    if (rc) { hipFree(h->modes); hipFree(h->wave); h->modes = nullptr; h->wave = nullptr; return rc; }
    h->numSys = numSys;
    h->prefixes_hlms = newtonian_prefixes_hlms;
    h->allocated = true;
    hipLaunchKernelGGL(k_update_info, dim3(blocks_for(numSys, kBlock)), dim3(kBlock), 0, as_stream(stream), h->modes, h->wave, m1, m2,
                       eta, tplspin, chi_S, chi_A, use_hm, numSys);
    rc = check_launch("update_information");
    if (rc) return rc;
    const long long nw = (long long)numSys * kNumModes;
    hipLaunchKernelGGL(k_wave_tables, dim3((unsigned)((nw + kBlock - 1) / kBlock)), dim3(kBlock), 0, as_stream(stream), h->wmodes, h->modes,
                       h->wave, m1, m2, eta, chi_S, chi_A, use_hm, numSys);
    return check_launch("update_information (waveform tables)");
}

int eobhip_seobnr_RR_force(eobhip_seobnr_t h, double* force_out, double* grad_out, const double* args, const double* add, int numSys,
                           const int* idx, void* stream) {
    EOBHIP_REQUIRE(h && h->allocated, "RR_force: update_information has not been called");
    if (numSys == 0) return EOBHIP_OK;
    hipLaunchKernelGGL(k_rr_force_dropin, dim3(blocks_for(numSys, kBlock)), dim3(kBlock), 0, as_stream(stream), force_out, grad_out, args,
                       add, numSys, h->modes, idx);
    return check_launch("RR_force");
}

int eobhip_seobnr_IC_diss(eobhip_seobnr_t h, double* out, const double* pr, double* args, double* add, double* grad_out,
                          double* grad_temp_force, double* hess_out, double* force_out, int numSys, void* stream) {
    EOBHIP_REQUIRE(h && h->allocated, "IC_diss: update_information has not been called");
    EOBHIP_REQUIRE(numSys >= 0 && numSys <= h->numSys, "IC_diss: numSys exceeds the handle's binaries");
    if (numSys == 0) return EOBHIP_OK;
    hipLaunchKernelGGL(k_ic_diss, dim3(blocks_for(numSys, kBlock)), dim3(kBlock), 0, as_stream(stream), out, pr, args, add, grad_out,
                       grad_temp_force, hess_out, force_out, numSys, h->modes);
    return check_launch("IC_diss");
}

int eobhip_seobnr_root_find_scalar_all(eobhip_seobnr_t h, double* pr_res, const double* start_bounds, const double* argsIn,
                                       const double* add, int max_iter, double err, int numBinAll, int num_args, int num_add_args,
                                       void* stream) {
    EOBHIP_REQUIRE(h && h->allocated, "root_find_scalar_all: update_information has not been called");
    EOBHIP_REQUIRE(num_args == 4 && num_add_args == kNAdd, "root_find_scalar_all: 4 args and 80 additional args per binary");
    EOBHIP_REQUIRE(numBinAll >= 0 && numBinAll <= h->numSys && max_iter >= 0, "bad sizes");
    if (numBinAll == 0) return EOBHIP_OK;
    hipLaunchKernelGGL(k_root_find_scalar_all, dim3(blocks_for(numBinAll, kBlock)), dim3(kBlock), 0, as_stream(stream), pr_res,
                       start_bounds, argsIn, add, max_iter, err, numBinAll, h->modes);
    return check_launch("root_find_scalar_all");
}

int eobhip_seobnr_ODE_Ham_align_AD(eobhip_seobnr_t h, const double* x, const double* arg, double* k, const double* add, int numSys,
                                   const int* idx, void* stream) {
    (void)x;
    EOBHIP_REQUIRE(h && h->allocated, "ODE_Ham_align_AD: update_information has not been called");
    if (numSys == 0) return EOBHIP_OK;
    hipLaunchKernelGGL(k_ode_dropin, dim3(blocks_for(numSys, kBlock)), dim3(kBlock), 0, as_stream(stream), arg, k, add, numSys, h->modes,
                       idx);
    return check_launch("ODE_Ham_align_AD");
}

int eobhip_seobnr_EOBGetSpinFactorizedWaveform(eobhip_seobnr_t h, double* out, const double* r, const double* phi, const double* pr,
                                               const double* pphi, const double* v, const double* Hreal, const double* eta,
                                               const double* vPhiInput, int l, int m, const double* newtonian_prefix, double h22_calib,
                                               int numSys, int traj_length, const int* idx, void* stream) {
    (void)pr;
    EOBHIP_REQUIRE(h && h->allocated, "EOBGetSpinFactorizedWaveform: update_information has not been called");
    if (!(l >= 2 && l <= 8 && m >= 1 && m <= l)) {
        set_error("EOBGetSpinFactorizedWaveform: (l, m) = (%d, %d) outside 2 <= l <= 8, 1 <= m <= l", l, m);
        return EOBHIP_EUNSUPPORTED;
    }
    EOBHIP_REQUIRE(numSys >= 0 && traj_length >= 0, "bad sizes");
    if (numSys == 0 || traj_length == 0) return EOBHIP_OK;
    const int blocks = numSys < 65535 ? numSys : 65535;
    hipLaunchKernelGGL(k_hlm_dropin, dim3(blocks), dim3(kBlock), 0, as_stream(stream), out, r, phi, pphi, v, Hreal, eta, vPhiInput,
                       newtonian_prefix, h22_calib, numSys, traj_length, h->wmodes, idx, flux_index(l, m), h->numSys);
    return check_launch("EOBGetSpinFactorizedWaveform");
}

int eobhip_dormandPrinceSteps(double* x, double* xCurrent, double* solOld, double* h, double* arg, double* add, double* k1, double* k2,
                              double* k3, double* k4, double* k5, double* k6, double* k7, double* k8, double* k9, double* k10,
                              double* ak, int nargs, int numEq, int num_add_args, eobhip_seobnr_t ode, const int* idx, void* stream) {
    (void)num_add_args;
    EOBHIP_REQUIRE(ode && ode->allocated, "dormandPrinceSteps: ode handle without information");
    EOBHIP_REQUIRE(nargs == 4, "dormandPrinceSteps: nargs must be 4");
    if (numEq == 0) return EOBHIP_OK;
    hipStream_t st = as_stream(stream);
    // stage s result buffers in reference order: stages 1..10 -> k1..k10, 11 -> k2, 12 -> k3
    double* outs[12] = {k1, k2, k3, k4, k5, k6, k7, k8, k9, k10, k2, k3};
    StagePtrs kp;
    const double* ins[11] = {k1, k2, k3, k4, k5, k6, k7, k8, k9, k10, k2};
    for (int q = 0; q < 11; ++q) kp.k[q] = ins[q];
    const unsigned nb = blocks_for(numEq, kBlock);
    for (int s = 0; s < 12; ++s) {
        // stage 12 reads stage 11 from k2 (ins[10]) and stage 2 never enters it (a_12,2 = 0)
        hipLaunchKernelGGL(k_stage_arg, dim3(nb), dim3(kBlock), 0, st, arg, solOld, x, xCurrent, ak, h, kp, s, nargs, numEq);
        int rc = check_launch("stage args");
        if (rc) return rc;
        hipLaunchKernelGGL(k_ode_dropin, dim3(nb), dim3(kBlock), 0, st, arg, outs[s], add, numEq, ode->modes, idx);
        rc = check_launch("stage rhs");
        if (rc) return rc;
    }
    return EOBHIP_OK;
}

int eobhip_error_tol(double* err, const double* solOld, double* solNew, const double* h, const double* k1, const double* k2,
                     const double* k3, const double* k6, const double* k7, const double* k8, const double* k9, const double* k10, int numEq,
                     int nargs, double rtol, double atol, void* stream) {
    if (numEq == 0) return EOBHIP_OK;
    hipLaunchKernelGGL(k_error, dim3(blocks_for(numEq, kBlock)), dim3(kBlock), 0, as_stream(stream), err, solOld, solNew, h, k1, k2, k3,
                       k6, k7, k8, k9, k10, numEq, nargs, rtol, atol);
    return check_launch("error");
}

int eobhip_error(double* err, const double* solOld, double* solNew, const double* h, const double* k1, const double* k2, const double* k3,
                 const double* k6, const double* k7, const double* k8, const double* k9, const double* k10, int numEq, int nargs,
                 void* stream) {
    return eobhip_error_tol(err, solOld, solNew, h, k1, k2, k3, k6, k7, k8, k9, k10, numEq, nargs, 1e-12, 1e-12, stream);
}

int eobhip_controllerSuccess(unsigned char* flag, const double* err, double* errOld, unsigned char* prevRej, double* h, const double* x,
                             int numEq, int nargs, void* stream) {
    (void)x; (void)nargs;
    if (numEq == 0) return EOBHIP_OK;
    hipLaunchKernelGGL(k_controller, dim3(blocks_for(numEq, kBlock)), dim3(kBlock), 0, as_stream(stream), flag, err, errOld, prevRej, h,
                       numEq);
    return check_launch("controllerSuccess");
}

size_t eobhip_bindata_bytes(void) { return sizeof(BinData); }
size_t eobhip_modes_bytes(void) { return sizeof(ModeCoef) * kNumModes; }

int eobhip_bin_setup(const double* add, int n, int use_hm, void* bindata, void* modes, void* stream) {
    EOBHIP_REQUIRE(n >= 0, "n < 0");
    if (n == 0) return EOBHIP_OK;
    hipLaunchKernelGGL(k_bin_setup, dim3(n), dim3(64), 0, as_stream(stream), add, n, use_hm,
                       static_cast<BinData*>(bindata), static_cast<ModeCoef*>(modes));
    return check_launch("bin_setup");
}

int eobhip_params_setup(const double* m1, const double* m2, const double* chi1, const double* chi2, int n, int use_hm, double K,
                        double d5, double dSO, double dSS, double h22_calib, double* add, void* bindata, void* modes, void* stream) {
    EOBHIP_REQUIRE(n >= 0, "n < 0");
    if (n == 0) return EOBHIP_OK;
    static const PrefixConsts pc = make_prefix_consts();
    hipLaunchKernelGGL(k_params_setup, dim3(n), dim3(64), 0, as_stream(stream), m1, m2, chi1, chi2, n, use_hm, K, d5, dSO, dSS,
                       h22_calib, pc, add, static_cast<BinData*>(bindata), static_cast<ModeCoef*>(modes));
    return check_launch("params_setup");
}

int eobhip_newton_prefix_constants(double* out) {
    EOBHIP_REQUIRE(out, "null output");
    const PrefixConsts pc = make_prefix_consts();
    for (int q = 0; q < kNumModes; ++q) {
        out[2 * q] = pc.re[q];
        out[2 * q + 1] = pc.im[q];
    }
    return EOBHIP_OK;
}

int eobhip_initial_conditions(const double* r0, const void* bindata, const void* modes, int n, double* pphi0, double* pr0,
                              double* omega0, int max_iter, double err, void* stream) {
    EOBHIP_REQUIRE(n >= 0 && max_iter >= 1, "bad arguments");
    if (n == 0) return EOBHIP_OK;
    hipLaunchKernelGGL(k_initial_conditions, dim3(n), dim3(64), 0, as_stream(stream), r0,
                       static_cast<const BinData*>(bindata), static_cast<const ModeCoef*>(modes), n, pphi0, pr0, omega0, max_iter, err,
                       nullptr);
    return check_launch("initial_conditions");
}

int eobhip_initial_state(const double* r0, const void* bindata, const void* modes, int n, double* cond, double* omega0, int max_iter,
                         double err, void* stream) {
    EOBHIP_REQUIRE(n >= 0 && max_iter >= 1, "bad arguments");
    EOBHIP_REQUIRE(cond != nullptr && omega0 != nullptr, "null output");
    if (n == 0) return EOBHIP_OK;
    const size_t nn = (size_t)n;
    hipLaunchKernelGGL(k_initial_conditions, dim3(n), dim3(64), 0, as_stream(stream), r0,
                       static_cast<const BinData*>(bindata), static_cast<const ModeCoef*>(modes), n, cond + 3 * nn, cond + 2 * nn, omega0,
                       max_iter, err, cond);
    return check_launch("initial_state");
}

size_t eobhip_integrate_scratch_bytes(int n) {
    const size_t nn = (size_t)(n > 0 ? n : 1);
    // queue/count ints + plain-mode state: 6 int arrays, 4 + 4 + 4 double rows, att, 2 active lists;
    // the coupled driver's k1 and criterion rows (4 + 4 doubles) and 4 int arrays;
    // 22 sub-arrays, each padded to a 256-byte boundary
    return 256 + nn * (6 * sizeof(int) + 12 * sizeof(double) + sizeof(unsigned long long) + 2 * sizeof(int)) +
           nn * (8 * sizeof(double) + 4 * sizeof(int)) + 22 * 256;
}

}  // extern "C"

namespace {

int integrate_impl(const double* condBound, const void* bindata, const void* modes, int n, int max_step, double rtol, double atol,
                   double hInit, double step_back, double fine_step, double tmax, int two_pass, int fix_step, const double* x0,
                   const int* step0, int use_stop, int mode, int group, int block, int grid_groups, double* dense, int* num_steps,
                   int* stop_info, int* pass1_steps, unsigned long long* attempts, void* scratch, long long* iterations_out, int* done,
                   double* old_om, void* stream) {
    EOBHIP_REQUIRE((x0 == nullptr) == (step0 == nullptr), "x0 and step0 go together");
    EOBHIP_REQUIRE(n >= 0, "n < 0");
    EOBHIP_REQUIRE(max_step >= 2, "max_step must be >= 2");
    EOBHIP_REQUIRE(mode == 0 || mode == 1 || mode == 2, "mode must be 0 (persistent), 1 (plain) or 2 (coupled)");
    EOBHIP_REQUIRE(mode != 2 || old_om != nullptr || !use_stop, "the coupled driver needs old_omegas");
    EOBHIP_REQUIRE(group == 1 || group == 4 || group == 8 || group == 16 || group == 32 || group == 64,
                   "group must be one of 1, 4, 8, 16, 32, 64");
    if (block <= 0) block = kBlock;
    EOBHIP_REQUIRE(block % 64 == 0 && block <= kBlock && block >= group, "block must be a multiple of 64, <= 256 and >= group");
    EOBHIP_REQUIRE(scratch != nullptr, "scratch is required");
    if (iterations_out) *iterations_out = 0;
    if (n == 0) return EOBHIP_OK;
    hipStream_t st = as_stream(stream);

    IntegParams P;
    P.cond = condBound;
    P.bd = static_cast<const BinData*>(bindata);
    P.mc = static_cast<const ModeCoef*>(modes);
    P.dense = dense;
    P.num_steps = num_steps;
    P.stop_info = stop_info;
    P.pass1_steps = pass1_steps;
    P.attempts = attempts;
    P.nbin = n;
    P.max_step = max_step;
    P.rtol = rtol; P.atol = atol; P.hInit = hInit; P.step_back = step_back; P.fine_step = fine_step; P.tmax = tmax;
    P.two_pass = two_pass ? 1 : 0;
    P.fix_step = fix_step ? 1 : 0;
    P.max_attempts = 16ull * (unsigned long long)max_step + 4096ull;
    P.x0 = x0;
    P.step0 = step0;
    P.use_stop = use_stop ? 1 : 0;
    P.done = done;
    EOBHIP_REQUIRE(done == nullptr || mode == 0, "completion flags need the persistent mode");

    // carve scratch
    char* base = static_cast<char*>(scratch);
    size_t off = 0;
    auto take = [&](size_t bytes) { void* p = base + off; off += (bytes + 255) & ~size_t(255); return p; };
    int* queue = static_cast<int*>(take(sizeof(int) * 2));
    P.queue = queue;
    // cleared by a kernel, not a memset node, so the persistent path can be captured in a HIP graph
    hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(64), 0, st, queue, 2);
    int rc = check_launch("k_zero_ints");
    if (rc) return rc;

    if (mode == 0) {
        int gg = grid_groups > 0 ? grid_groups : n;
        switch (group) {
            case 1: return launch_persistent<1>(P, block, gg, st);
            case 4: return launch_persistent<4>(P, block, gg, st);
            case 8: return launch_persistent<8>(P, block, gg, st);
            case 16: return launch_persistent<16>(P, block, gg, st);
            case 32: return launch_persistent<32>(P, block, gg, st);
            default: return launch_persistent<64>(P, block, gg, st);
        }
    }
    const size_t nn = (size_t)n;
    PlainState S;
    S.b = static_cast<int*>(take(nn * sizeof(int)));
    S.pass = static_cast<int*>(take(nn * sizeof(int)));
    S.step_num = static_cast<int*>(take(nn * sizeof(int)));
    S.sn1 = static_cast<int*>(take(nn * sizeof(int)));
    S.code = static_cast<int*>(take(nn * sizeof(int)));
    S.flags = static_cast<int*>(take(nn * sizeof(int)));
    S.x = static_cast<double*>(take(nn * sizeof(double)));
    S.h = static_cast<double*>(take(nn * sizeof(double)));
    S.err_old = static_cast<double*>(take(nn * sizeof(double)));
    S.h_init = static_cast<double*>(take(nn * sizeof(double)));
    S.om = static_cast<double*>(take(4 * nn * sizeof(double)));
    S.y = static_cast<double*>(take(4 * nn * sizeof(double)));
    S.att = static_cast<unsigned long long*>(take(nn * sizeof(unsigned long long)));
    int* act0 = static_cast<int*>(take(nn * sizeof(int)));
    int* act1 = static_cast<int*>(take(nn * sizeof(int)));
    CoupledState C{};
    if (mode == 2) {
        C.k1 = static_cast<double*>(take(4 * nn * sizeof(double)));
        C.chk = static_cast<double*>(take(4 * nn * sizeof(double)));
        C.k1_ok = static_cast<int*>(take(nn * sizeof(int)));
        C.acc = static_cast<int*>(take(nn * sizeof(int)));
        C.code = static_cast<int*>(take(nn * sizeof(int)));
        C.col = static_cast<int*>(take(nn * sizeof(int)));
        C.old_om = old_om;
    }
    if (off > eobhip_integrate_scratch_bytes(n)) {
        set_error("internal: scratch layout exceeds eobhip_integrate_scratch_bytes");
        return EOBHIP_EINVAL;
    }
    int* host_cnt = nullptr;
    rc = hip_status(hipHostMalloc(&host_cnt, sizeof(int)), "hipHostMalloc");
    if (rc) return rc;
    if (mode == 2) {
        switch (group) {
            case 1: rc = run_coupled<1>(P, S, C, act0, act1, queue + 1, host_cnt, st, iterations_out); break;
            case 4: rc = run_coupled<4>(P, S, C, act0, act1, queue + 1, host_cnt, st, iterations_out); break;
            case 8: rc = run_coupled<8>(P, S, C, act0, act1, queue + 1, host_cnt, st, iterations_out); break;
            case 16: rc = run_coupled<16>(P, S, C, act0, act1, queue + 1, host_cnt, st, iterations_out); break;
            case 32: rc = run_coupled<32>(P, S, C, act0, act1, queue + 1, host_cnt, st, iterations_out); break;
            default: rc = run_coupled<64>(P, S, C, act0, act1, queue + 1, host_cnt, st, iterations_out); break;
        }
        hipHostFree(host_cnt);
        return rc;
    }
    switch (group) {
        case 1: rc = run_plain<1>(P, S, act0, act1, queue + 1, host_cnt, block, st, iterations_out); break;
        case 4: rc = run_plain<4>(P, S, act0, act1, queue + 1, host_cnt, block, st, iterations_out); break;
        case 8: rc = run_plain<8>(P, S, act0, act1, queue + 1, host_cnt, block, st, iterations_out); break;
        case 16: rc = run_plain<16>(P, S, act0, act1, queue + 1, host_cnt, block, st, iterations_out); break;
        case 32: rc = run_plain<32>(P, S, act0, act1, queue + 1, host_cnt, block, st, iterations_out); break;
        default: rc = run_plain<64>(P, S, act0, act1, queue + 1, host_cnt, block, st, iterations_out); break;
    }
    hipHostFree(host_cnt);
    return rc;
}

}  // namespace

extern "C" {

int eobhip_integrate(const double* condBound, const void* bindata, const void* modes, int n, int max_step, double rtol, double atol,
                     double hInit, double step_back, double fine_step, double tmax, int two_pass, int fix_step, const double* x0, const int* step0,
                     int use_stop, int mode, int group, int block, int grid_groups, double* dense, int* num_steps, int* stop_info,
                     int* pass1_steps, unsigned long long* attempts, void* scratch, long long* iterations_out, void* stream) {
    return integrate_impl(condBound, bindata, modes, n, max_step, rtol, atol, hInit, step_back, fine_step, tmax, two_pass, fix_step, x0,
                          step0, use_stop, mode, group, block, grid_groups, dense, num_steps, stop_info, pass1_steps, attempts, scratch,
                          iterations_out, nullptr, nullptr, stream);
}

int eobhip_integrate_signal(const double* condBound, const void* bindata, const void* modes, int n, int max_step, double rtol,
                            double atol, double hInit, double step_back, double fine_step, double tmax, int two_pass, int fix_step,
                            const double* x0, const int* step0, int use_stop, int group, int block, int grid_groups, double* dense,
                            int* num_steps, int* stop_info, int* pass1_steps, unsigned long long* attempts, void* scratch, int* done,
                            void* stream) {
    EOBHIP_REQUIRE(done != nullptr, "done is required");
    // the device's address of the flags; fails (instead of a faulting store) for
    // memory that is not pinned host memory
    void* done_dev = nullptr;
    const hipError_t e = hipHostGetDevicePointer(&done_dev, done, 0);
    if (e != hipSuccess || done_dev == nullptr) {
        set_error("done must be pinned host memory (hipHostGetDevicePointer: %s)", hipGetErrorString(e));
        return EOBHIP_EINVAL;
    }
    return integrate_impl(condBound, bindata, modes, n, max_step, rtol, atol, hInit, step_back, fine_step, tmax, two_pass, fix_step, x0,
                          step0, use_stop, 0, group, block, grid_groups, dense, num_steps, stop_info, pass1_steps, attempts, scratch,
                          nullptr, static_cast<int*>(done_dev), nullptr, stream);
}

int eobhip_integrate_batch_coupled(const double* condBound, const void* bindata, const void* modes, int n, int max_step, double rtol,
                                   double atol, double hInit, double step_back, double fine_step, double tmax, int two_pass,
                                   int fix_step, const double* x0, const int* step0, int use_stop, int group, double* dense,
                                   int* num_steps, int* stop_info, int* pass1_steps, unsigned long long* attempts, double* old_omegas,
                                   void* scratch, long long* iterations_out, void* stream) {
    return integrate_impl(condBound, bindata, modes, n, max_step, rtol, atol, hInit, step_back, fine_step, tmax, two_pass, fix_step, x0,
                          step0, use_stop, 2, group, kBlock, 0, dense, num_steps, stop_info, pass1_steps, attempts, scratch,
                          iterations_out, nullptr, old_omegas, stream);
}

size_t eobhip_wave_mode_bytes(void) { return sizeof(WaveMode); }

static int mode_list(int nm, const int* ls, const int* ms, ModeList& ml) {
    EOBHIP_REQUIRE(nm >= 1 && nm <= kMaxWaveModes, "nm must be 1 .. 15");
    ml.nm = nm;
    for (int j = 0; j < nm; ++j) {
        if (!(ls[j] >= 2 && ls[j] <= 8 && ms[j] >= 1 && ms[j] <= ls[j])) {
            set_error("mode %d: (l, m) = (%d, %d) outside 2 <= l <= 8, 1 <= m <= l", j, ls[j], ms[j]);
            return EOBHIP_EUNSUPPORTED;
        }
        ml.l[j] = ls[j];
        ml.m[j] = ms[j];
    }
    return EOBHIP_OK;
}

int eobhip_wave_tables(const double* add, int n, int use_hm, const void* bindata, const void* modes, int nm, const int* ls,
                       const int* ms, void* wtab, void* stream) {
    EOBHIP_REQUIRE(n >= 0, "bad sizes");
    ModeList ml;
    const int rc = mode_list(nm, ls, ms, ml);
    if (rc) return rc;
    if (n == 0) return EOBHIP_OK;
    const long long nt = (long long)n * nm;
    hipLaunchKernelGGL(k_wave_tables_batch, dim3((unsigned)((nt + kBlock - 1) / kBlock)), dim3(kBlock), 0, as_stream(stream), add, n,
                       use_hm, static_cast<const ModeCoef*>(modes), static_cast<const BinData*>(bindata), ml, static_cast<WaveMode*>(wtab));
    return check_launch("wave_tables");
}

int eobhip_hlm_modes_knots(const double* dense, const int* num_steps, const void* bindata, const void* wtab, const double* add, int n,
                           int max_step, int nm, double* knots, void* stream) {
    EOBHIP_REQUIRE(n >= 0 && max_step >= 1 && nm >= 1 && nm <= kMaxWaveModes, "bad arguments");
    if (n == 0) return EOBHIP_OK;
    EOBHIP_REQUIRE(n <= 65535, "n must be <= 65535 (grid y dimension)");
    const int bx = (max_step + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_hlm_modes_knots, dim3(bx < 64 ? bx : 64, n), dim3(kBlock), 0, as_stream(stream), dense, num_steps,
                       static_cast<const BinData*>(bindata), static_cast<const WaveMode*>(wtab), add, n, max_step, nm, knots);
    return check_launch("hlm_modes_knots");
}

int eobhip_hlm22_knots(const double* dense, const int* num_steps, const void* bindata, const void* modes, int n, int max_step,
                       double* knots, void* stream) {
    EOBHIP_REQUIRE(n >= 0 && max_step >= 1, "bad arguments");
    if (n == 0) return EOBHIP_OK;
    EOBHIP_REQUIRE(n <= 65535, "n must be <= 65535 (grid y dimension)");
    const int bx = (max_step + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_hlm22_knots, dim3(bx < 64 ? bx : 64, n), dim3(kBlock), 0, as_stream(stream), dense, num_steps,
                       static_cast<const BinData*>(bindata), static_cast<const ModeCoef*>(modes), n, max_step, knots);
    return check_launch("hlm22_knots");
}

}  // extern "C"
