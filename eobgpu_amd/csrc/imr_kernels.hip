// Merger-ringdown attachment of the modes on the device: the NQC solve, the
// remnant and its QNMs, the ringdown coefficients and the template on the
// uniform grid with the NQC factor applied up to the attachment sample and the
// ringdown after it.
//
// Reference: SEOBNRv4PHM.compute_full_waveform (eob/seobnrv4phm.py:2308-2556),
// EOBCalculateNQCCoefficientsV4_freeattach (:1465-1683), EOBNonQCCorrection
// (:1686-1700), compute_MR_mode_free (:1779-1844), the fits of
// eob/eobstuff.py and the template sum of BBHWaveformTD (:720).
//
// Layout.  The splines of the modes are the batch's mode splines (coef
// [n][max_step][4 (2 nm + 1)]: Re and Im of h_j e^{i m_j phi}, then phi) and a
// second set of r and pr (rp [n][max_step][8]).  The attachment needs the
// modes only around two times per (binary, mode): the NQC time and the
// ringdown match time.  Each is a not-a-knot spline over the uniform-grid
// samples from a fixed start to the series' end; here the end is cut
// kImrAfter rows past the evaluation interval (the end condition's influence
// on that interval falls as 0.268 per row, 5e-18 at 31 rows), so one
// wavefront samples a whole window, one sample per lane, in one pass.
#include <hip/hip_runtime.h>

#include <mutex>

#include "eobhip_common.hpp"
#include "imr.hpp"
#include "knot_window.hpp"
#include "spline_nak.hpp"

using namespace eobhip_detail;
using imr::Cx;

namespace {

constexpr int kImrMaxModes = 8;
struct ImrModes {
    int nm;
    int l[kImrMaxModes];
    int m[kImrMaxModes];
};
constexpr int kImrMaxM = 5;     // the ringdown modes' largest m
constexpr int kImrAfter = 31;   // window rows kept past the evaluation interval
constexpr int kImrNqcBefore = 7; // rows before the NQC time (the reference's N = 7)
constexpr int kImrSeries = 7;   // q1 A, q2 A, q3 A, A, p1, p2, phase
constexpr double kTwoPi = 6.283185307179586;
constexpr double kPiD = 3.141592653589793;

// per (binary, mode) ringdown record
enum RdField { kRdAmp, kRdDamp, kRdPhi, kRdOmega, kRdC1c, kRdC2c, kRdC1f, kRdC2f, kRdD1c, kRdD1f, kRdD2f, kRdWr, kRdWi, kRdEta, kRdPad0, kRdPad1,
               kRdCount };
// per binary doubles and ints
enum BinField { kBinTPeak, kBinTMatch, kBinDt, kBinMf, kBinAf, kBinDT22, kBinCount };
enum BinInt { kBinNpts, kBinIdx, kBinNadd, kBinICount };

// t_k = k dT + t0 with the two roundings numpy makes (arange * delta_T + t_0)
__device__ __forceinline__ double grid_t(long long k, double dT, double t0) { return __dadd_rn(__dmul_rn((double)k, dT), t0); }

// argmin_k |t_k - T| over k < npts, first index on ties
__device__ int nearest_index(double T, double dT, double t0, int npts) {
    double kf = rint((T - t0) / dT);
    long long k = (kf == kf) ? (long long)fmax(fmin(kf, (double)(npts - 1)), 0.0) : 0;
    long long best = k;
    double bd = fabs(grid_t(k, dT, t0) - T);
    for (int d = -1; d <= 1; d += 2) {
        const long long kk = k + d;
        if (kk < 0 || kk >= npts) continue;
        const double dd = fabs(grid_t(kk, dT, t0) - T);
        if (dd < bd || (dd == bd && kk < best)) { best = kk; bd = dd; }
    }
    return (int)best;
}

// searchsorted(t_ws..t_{we-1}, T, 'right') - 1 as a grid index, clamped to [ws, we - 1]
__device__ int interval_index(double T, double dT, double t0, int ws, int we) {
    double kf = floor((T - t0) / dT);
    long long k = (kf == kf) ? (long long)fmax(fmin(kf, (double)(we - 1)), (double)ws) : ws;
    while (k > ws && grid_t(k, dT, t0) > T) --k;
    while (k + 1 < we && grid_t(k + 1, dT, t0) <= T) ++k;
    return (int)k;
}

// searchsorted(knots, t, 'right') - 1 over the binary's knots (row stride 5), >= 0
__device__ __forceinline__ int knot_index(const double* __restrict__ T, int len, double t) {
    int lo = 0, hi = len;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (T[5 * mid] <= t) lo = mid + 1;
        else hi = mid;
    }
    return lo > 0 ? lo - 1 : 0;
}

// A spline piece at offset x; the last knot's coefficients are zero (the
// reference fills knots 0 .. len - 2 only), so there the value is y and the
// derivative 0.
struct Piece {
    double x, x2, x3;
    bool last;
    template <class Ptr>
    __device__ double v(Ptr c) const { return last ? c[0] : c[0] + c[1] * x + c[2] * x2 + c[3] * x3; }
    template <class Ptr>
    __device__ double d(Ptr c) const { return last ? 0.0 : c[1] + 2.0 * c[2] * x + 3.0 * c[3] * x2; }
};

// numpy.mod for a positive divisor
__device__ __forceinline__ double pymod(double x, double y) {
    double r = fmod(x, y);
    if (r != 0.0 && (r < 0.0) != (y < 0.0)) r += y;
    return r;
}

// numpy.unwrap over the wavefront's lanes 0 .. W - 1 (lane i holds sample i)
__device__ double unwrap_lanes(double ang, int lane, int W) {
    const double prev = __shfl_up(ang, 1, 64);
    double corr = 0.0;
    if (lane > 0 && lane < W) {
        const double dd = ang - prev;
        double ddmod = pymod(dd + kPiD, kTwoPi) - kPiD;
        if (ddmod == -kPiD && dd > 0.0) ddmod = kPiD;
        corr = (fabs(dd) < kPiD) ? 0.0 : ddmod - dd;
    }
    // inclusive prefix sum over the lanes
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const double o = __shfl_up(corr, off, 64);
        if (lane >= off) corr += o;
    }
    return ang + corr;
}

// Lane s < nser: the not-a-knot spline of series s over the W window samples
// (sh_t, sh_y[s]) evaluated at T in interval kI: value and first two
// derivatives (the reference's Thomas solve, src/interpolate.cu:11-267).
struct WinLds {
    double t[64];
    double y[kImrSeries][64];
    double ud[kImrSeries][64], dg[kImrSeries][64], b[kImrSeries][64];
    double val[kImrSeries][3];
};
__device__ void window_spline_eval(WinLds& w, int s, int W, int kI, double T) {
    const double* t = w.t;
    const double* y = w.y[s];
    double pdg = 0.0, pud = 0.0, pb = 0.0;
    for (int i = 0; i < W; ++i) {
        int i0, i1, i2;
        if (i == W - 1) { i0 = W - 3; i1 = W - 2; i2 = W - 1; }
        else if (i == 0) { i0 = 0; i1 = 1; i2 = 2; }
        else { i0 = i - 1; i1 = i; i2 = i + 1; }
        Row r = nak_row(i, W, t[i0], t[i1], t[i2], y[i0], y[i1], y[i2]);
        if (i > 0) {
            const double f = r.ld / pdg;
            r.dg = r.dg - f * pud;
            r.b = r.b - f * pb;
        }
        w.ud[s][i] = r.ud;
        w.dg[s][i] = r.dg;
        w.b[s][i] = r.b;
        pdg = r.dg; pud = r.ud; pb = r.b;
    }
    double sn = w.b[s][W - 1] / w.dg[s][W - 1];
    double s_hi = sn, s_lo = sn;
    for (int i = W - 2; i >= kI; --i) {
        s_hi = sn;
        sn = (w.b[s][i] - w.ud[s][i] * sn) / w.dg[s][i];
        s_lo = sn;
    }
    double c1 = 0.0, c2 = 0.0, c3 = 0.0;
    if (kI < W - 1) {
        const double dt = t[kI + 1] - t[kI];
        const double slope = (y[kI + 1] - y[kI]) / dt;
        const double tt = (s_lo + s_hi - 2.0 * slope) / dt;
        c1 = s_lo;
        c2 = (slope - s_lo) / dt - tt;
        c3 = tt / dt;
    }
    const double x = T - t[kI];
    w.val[s][0] = y[kI] + c1 * x + c2 * x * x + c3 * x * x * x;
    w.val[s][1] = c1 + 2.0 * c2 * x + 3.0 * c3 * x * x;
    w.val[s][2] = 2.0 * c2 + 6.0 * c3 * x;
}

// dense solves with partial pivoting (numpy.linalg.solve's LU), n <= 3
template <int N>
__device__ void solve_pivot(double A[N][N], double b[N], double x[N]) {
    for (int c = 0; c < N; ++c) {
        int p = c;
        for (int r = c + 1; r < N; ++r)
            if (fabs(A[r][c]) > fabs(A[p][c])) p = r;
        if (p != c) {
            for (int k = 0; k < N; ++k) { const double t = A[c][k]; A[c][k] = A[p][k]; A[p][k] = t; }
            const double t = b[c]; b[c] = b[p]; b[p] = t;
        }
        for (int r = c + 1; r < N; ++r) {
            const double f = A[r][c] / A[c][c];
            for (int k = c + 1; k < N; ++k) A[r][k] -= f * A[c][k];
            b[r] -= f * b[c];
        }
    }
    for (int r = N - 1; r >= 0; --r) {
        double s = b[r];
        for (int k = r + 1; k < N; ++k) s -= A[r][k] * x[k];
        x[r] = s / A[r][r];
    }
}

struct BinInputs {
    const double *m1, *m2, *chi1, *chi2;
};

// ------------------------------------------------------------- QNM table
// The five ringdown modes' roots (omega and the separation constant, Leaver
// units) on the spins -0.99 .. 0.99 in steps of 0.01, filled once per device;
// a binary's QNM is then a few secant steps from the interpolated entry
// instead of a walk from a = 0.  The mode index is imr::rd_mode_index.
constexpr int kQnmGrid = 199;
constexpr double kQnmA0 = -0.99, kQnmDa = 0.01;
__device__ double g_qnm_tab[5][kQnmGrid][4];

__device__ __forceinline__ void rd_mode_lm(int k, int& l, int& m) {
    l = (k <= 1) ? 2 : k + 1;
    m = (k == 1) ? 1 : l;
}

__global__ void k_qnm_table() {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 5 * kQnmGrid) return;
    const int k = t / kQnmGrid, i = t % kQnmGrid;
    int l, m;
    rd_mode_lm(k, l, m);
    const imr::QnmRoot r = imr::kerr_qnm_path(l, m, kQnmA0 + kQnmDa * i);
    double* o = g_qnm_tab[k][i];
    o[0] = r.w.re;
    o[1] = r.w.im;
    o[2] = r.A.re;
    o[3] = r.A.im;
}

// M omega_lm0 at spin af: from the table for the ringdown modes, |af| <= 0.99;
// otherwise along the spin path from a = 0
__device__ Cx kerr_qnm_fast(int l, int m, double af) {
    const int k = imr::rd_mode_index(l, m);
    if (k < 0 || !(fabs(af) <= 0.99)) return imr::kerr_qnm(l, m, af);
    const double x = (af - kQnmA0) / kQnmDa;
    int i = (int)floor(x);
    i = i < 0 ? 0 : (i > kQnmGrid - 2 ? kQnmGrid - 2 : i);
    const double f = x - i;
    const double* p0 = g_qnm_tab[k][i];
    const double* p1 = g_qnm_tab[k][i + 1];
    const Cx w = imr::cx((1.0 - f) * p0[0] + f * p1[0], (1.0 - f) * p0[1] + f * p1[1]);
    const Cx A = imr::cx((1.0 - f) * p0[2] + f * p1[2], (1.0 - f) * p0[3] + f * p1[3]);
    const imr::QnmRoot r = imr::leaver_refine(m, 0.5 * af, w, A, 1e-6);
    return imr::cx(0.5 * r.w.re, 0.5 * r.w.im);
}

// fill the table on the current device once (stream-ordered, then waited for,
// so that launches on any stream see it)
std::mutex g_qnm_mu;
bool g_qnm_ready[64];
int ensure_qnm_table(void* stream) {
    int dev = 0;
    int rc = hip_status(hipGetDevice(&dev), "hipGetDevice");
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(g_qnm_mu);
    if (dev >= 0 && dev < 64 && g_qnm_ready[dev]) return EOBHIP_OK;
    hipLaunchKernelGGL(k_qnm_table, dim3(blocks_for(5 * kQnmGrid, 64)), dim3(64), 0, as_stream(stream));
    rc = check_launch("qnm_table");
    if (rc) return rc;
    rc = hip_status(hipStreamSynchronize(as_stream(stream)), "qnm_table sync");
    if (rc) return rc;
    if (dev >= 0 && dev < 64) g_qnm_ready[dev] = true;
    return EOBHIP_OK;
}

// ------------------------------------------------------------ QNMs per binary
// Thread (b, j): j < nm the QNM of mode j, j == nm the (2,2) QNM that sets the
// ringdown length (the reference's 30 damping times, :2463-2465 and :2514).
// wl [n][nm] complex in lalsimulation's convention (omega_r + i / tau) in
// units of the total mass; binf/bini as the enums.
__global__ void k_imr_qnm(BinInputs in, const double* __restrict__ dense, const double* __restrict__ dT_M, int n, int max_step,
                          ImrModes mm, double* __restrict__ wl, double* __restrict__ binf, int* __restrict__ bini) {
    const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int per = mm.nm + 1;
    if (gid >= (long long)n * per) return;
    const int b = (int)(gid / per), j = (int)(gid % per);
    const double m1 = in.m1[b], m2 = in.m2[b], c1 = in.chi1[b], c2 = in.chi2[b];
    const double M = m1 + m2;
    const double mf = imr::final_mass_seobnrv4(m1 / M, m2 / M, c1, c2);
    const double af = imr::final_spin_hbr2016(m1 / M, m2 / M, c1, c2);
    const int l = (j < mm.nm) ? mm.l[j] : 2, m = (j < mm.nm) ? mm.m[j] : 2;
    const Cx w = kerr_qnm_fast(l, m, af);
    const double wr = w.re / mf, wi = -w.im / mf;   // conj(omega) / M_f
    if (j < mm.nm) {
        wl[2 * ((size_t)b * mm.nm + j)] = wr;
        wl[2 * ((size_t)b * mm.nm + j) + 1] = wi;
        return;
    }
    binf[(size_t)b * kBinCount + kBinMf] = mf;
    binf[(size_t)b * kBinCount + kBinAf] = af;
    const double t0 = dense[(size_t)b * max_step * 5];
    const double dT = dT_M[b];
    const double dt = grid_t(1, dT, t0) - grid_t(0, dT, t0);
    const double ringdown_time = trunc(30.0 * (1.0 / wi));
    const double nadd = floor(ringdown_time / dt) + 1.0;
    bini[(size_t)b * kBinICount + kBinNadd] = (dt > 0.0 && nadd >= 1.0 && nadd < 1073741824.0) ? (int)nadd : 0;
}

// ------------------------------------------------ NQC and ringdown parameters
// One wavefront per (binary, mode): block b * nm + q.
__global__ __launch_bounds__(64) void k_imr_params(BinInputs in, const double* __restrict__ dense, const int* __restrict__ num_steps,
                                                   const double* __restrict__ coef, const double* __restrict__ rp,
                                                   const double* __restrict__ dT_M, int max_step, ImrModes mm,
                                                   const double* __restrict__ wl, double* __restrict__ nqc, double* __restrict__ rd,
                                                   double* __restrict__ binf, int* __restrict__ bini) {
    __shared__ WinLds w;
    const int b = blockIdx.x / mm.nm, q = blockIdx.x % mm.nm;
    const int lane = threadIdx.x;
    const int nm = mm.nm, cs = 4 * (2 * nm + 1);
    const int l = mm.l[q], m = mm.m[q];
    const int len = num_steps[b];
    double* out_nqc = nqc + ((size_t)b * nm + q) * 5;
    double* out_rd = rd + ((size_t)b * nm + q) * kRdCount;
    const double* T = dense + (size_t)b * max_step * 5;
    const double* C = coef + (size_t)b * max_step * cs;
    const double* RP = rp + (size_t)b * max_step * 8;
    const double t0 = T[0];
    const double dT = dT_M[b];
    // grid samples up to the last knot (a spacing that is not positive, or a
    // grid beyond 2^30 samples, marks the binary as failed)
    const double span = (len >= 1 && dT > 0.0) ? (T[5 * (len - 1)] - t0) / dT : -1.0;
    const int npts = (span >= 0.0 && span < 1073741824.0) ? (int)span + 1 : 0;
    if (len < 4 || npts < 2 * kImrAfter + 16) {
        if (lane < 5) out_nqc[lane] = NAN;
        if (lane < kRdCount) out_rd[lane] = NAN;
        if (q == 0 && lane == 0) {
            bini[(size_t)b * kBinICount + kBinNpts] = npts;
            bini[(size_t)b * kBinICount + kBinIdx] = -1;
        }
        return;
    }
    // time of the orbital-frequency peak over the knots (phi's slope at the
    // knots, 0 at the last one; first maximum)
    double best = -INFINITY;
    int bi = 0;
    for (int i = lane; i < len; i += 64) {
        const double v = (i < len - 1) ? C[(size_t)cs * i + 4 * (2 * nm) + 1] : 0.0;
        if (v > best) { best = v; bi = i; }
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const double ob = __shfl_xor(best, off, 64);
        const int oi = __shfl_xor(bi, off, 64);
        if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
    }
    const double t_peak = T[5 * bi];
    const imr::FitIn fin{in.m1[b], in.m2[b], in.chi1[b], in.chi2[b]};
    const double dT22 = imr::nr_peak_delta_t(fin);
    const bool is55 = (l == 5 && m == 5);

    // sample window [ws, we) of the grid: the mode h, |h|, unwrapped phase, r, pr, omega (one sample per lane)
    auto sample = [&](int k, double& hr, double& hi, double& r, double& pr, double& om) {
        const double tk = grid_t(k, dT, t0);
        const int ind = knot_index(T, len, tk);
        Piece p;
        p.x = tk - T[5 * ind];
        p.x2 = p.x * p.x;
        p.x3 = p.x2 * p.x;
        p.last = (ind == len - 1);
        const double* row = C + (size_t)cs * ind;
        const double phi = p.v(row + 4 * (2 * nm));
        om = p.d(row + 4 * (2 * nm));
        const double re = p.v(row + 4 * q), im = p.v(row + 4 * (nm + q));
        r = p.v(RP + (size_t)8 * ind);
        pr = p.v(RP + (size_t)8 * ind + 4);
        double s, c;
        sincos((double)m * phi, &s, &c);
        hr = re * c + im * s;
        hi = im * c - re * s;
    };

    // ---- NQC (EOBCalculateNQCCoefficientsV4_freeattach)
    const double Tn = t_peak - dT22 - (is55 ? 10.0 : 0.0);
    const int idx = nearest_index(Tn, dT, t0, npts);
    const int ws = idx - kImrNqcBefore > 0 ? idx - kImrNqcBefore : 0;
    const int kI = interval_index(Tn, dT, t0, ws, npts);
    const int we = (kI + kImrAfter + 1 < npts) ? kI + kImrAfter + 1 : npts;
    const int W = we - ws;
    {
        double hr = 0.0, hi = 0.0, r = 1.0, pr = 0.0, om = 1.0;
        if (lane < W) sample(ws + lane, hr, hi, r, pr, om);
        const double amp = hypot(hr, hi);
        const double ph = unwrap_lanes(atan2(hi, hr), lane, W);
        if (lane < W) {
            const double rO = r * om;
            const double q1 = pr * pr / (rO * rO);
            const double q2 = q1 / r;
            const double q3 = q2 / sqrt(r);
            const double p1 = pr / rO;
            w.t[lane] = grid_t(ws + lane, dT, t0);
            w.y[0][lane] = q1 * amp;
            w.y[1][lane] = q2 * amp;
            w.y[2][lane] = q3 * amp;
            w.y[3][lane] = amp;
            w.y[4][lane] = p1;
            w.y[5][lane] = p1 * pr * pr;
            w.y[6][lane] = ph;
        }
    }
    __syncthreads();
    if (lane < kImrSeries) window_spline_eval(w, lane, W, kI - ws, Tn);
    __syncthreads();
    double nq[5];
    {
        double Q[3][3], rhs[3], a[3];
        for (int i = 0; i < 3; ++i)
            for (int k = 0; k < 3; ++k) Q[i][k] = w.val[k][i];
        rhs[0] = fabs(imr::nr_peak_amp(l, m, fin)) - w.val[3][0];
        rhs[1] = imr::nr_peak_adot(l, m, fin) - w.val[3][1];
        rhs[2] = imr::nr_peak_addot(l, m, fin) - w.val[3][2];
        solve_pivot<3>(Q, rhs, a);
        double P[2][2], r2[2], bb[2];
        for (int i = 1; i <= 2; ++i) {
            P[i - 1][0] = -w.val[4][i];
            P[i - 1][1] = -w.val[5][i];
        }
        const double wt = w.val[6][1], wdt = w.val[6][2];
        const double omega = fabs(wt);
        const double omegaDot = (wt * wdt > 0.0) ? fabs(wdt) : -fabs(wdt);
        const double eta = fin.m1 * fin.m2 / ((fin.m1 + fin.m2) * (fin.m1 + fin.m2));
        const double chi = imr::fit_chi(fin);
        r2[0] = imr::nr_peak_omega(l, m, eta, chi) - omega;
        r2[1] = imr::nr_peak_omegadot(l, m, eta, chi) - omegaDot;
        solve_pivot<2>(P, r2, bb);
        nq[0] = a[0]; nq[1] = a[1]; nq[2] = a[2]; nq[3] = bb[0]; nq[4] = bb[1];
    }
    if (lane < 5) out_nqc[lane] = nq[lane];
    __syncthreads();   // the window is rewritten below

    // ---- ringdown attachment (compute_full_waveform :2438-2519, compute_MR_mode_free)
    const double t_attach = t_peak - dT22;
    const int idr = nearest_index(t_attach, dT, t0, npts);
    const double dt = grid_t(1, dT, t0) - grid_t(0, dT, t0);
    const long long Nw = (long long)(20.0 / dt) + 1;
    const double t_match = grid_t(idr, dT, t0);
    const double tm = t_match - (is55 ? 10.0 : 0.0);
    const int ws2 = (idr - Nw > 0) ? (int)(idr - Nw) : 0;
    const int kR = interval_index(tm, dT, t0, ws2, npts);
    const int wlo = (kR - kImrAfter > ws2) ? kR - kImrAfter : ws2;
    const int whi = (kR + kImrAfter + 1 < npts) ? kR + kImrAfter + 1 : npts;
    const int W2 = whi - wlo;
    {
        double hr = 0.0, hi = 0.0, r = 1.0, pr = 0.0, om = 1.0;
        if (lane < W2) sample(wlo + lane, hr, hi, r, pr, om);
        // the NQC factor (EOBNonQCCorrection with a3S = a4 = a5 = b3 = b4 = 0)
        const double sr = sqrt(r), rO = r * om;
        const double mag = 1.0 + (pr * pr / (rO * rO)) * (nq[0] + nq[1] / r + nq[2] / (r * sr));
        const double nph = nq[3] * pr / rO + pr * pr * pr / rO * nq[4];
        double sn, cn;
        sincos(nph, &sn, &cn);
        const double fr = mag * cn, fi = mag * sn;
        const double cr = hr * fr - hi * fi, ci = hr * fi + hi * fr;
        const double amp = hypot(cr, ci);
        const double ph = unwrap_lanes(atan2(ci, cr), lane, W2);
        if (lane < W2) {
            w.t[lane] = grid_t(wlo + lane, dT, t0);
            w.y[0][lane] = amp;
            w.y[1][lane] = ph;
        }
    }
    __syncthreads();
    if (lane < 2) window_spline_eval(w, lane, W2, kR - wlo, tm);
    __syncthreads();
    if (lane == 0) {
        const double A0 = w.val[0][0], dA0 = w.val[0][1], phm = w.val[1][0], wm = w.val[1][1];
        const double wr = wl[2 * ((size_t)b * nm + q)], wi = wl[2 * ((size_t)b * nm + q) + 1];
        const double sR = -wi, sI = -wr;
        const double eta = fin.m1 * fin.m2 / ((fin.m1 + fin.m2) * (fin.m1 + fin.m2));
        const double chi = imr::fit_chi(fin);
        const int ri = imr::rd_mode_index(l, m);
        double c1f = NAN, c2f = NAN, d1f = NAN, d2f = NAN;
        if (ri >= 0) {
            c1f = imr::rd_poly(imr::rd_table_a1(ri), eta, chi);
            c2f = imr::rd_poly(imr::rd_table_a2(ri), eta, chi);
            d1f = imr::rd_poly(imr::rd_table_p1(ri), eta, chi);
            d2f = imr::rd_poly(imr::rd_table_p2(ri), eta, chi);
        }
        if (ri == 0 && sR > 2.0 * c1f * tanh(c2f)) c1f = sR / (2.0 * tanh(c2f));
        const double ch = cosh(c2f);
        const double c1c = 1.0 / (c1f * eta) * (dA0 - sR * A0) * ch * ch;
        const double c2c = A0 / eta - 1.0 / (c1f * eta) * (dA0 - sR * A0) * ch * sinh(c2f);
        const double d1c = (wm - sI) * (1.0 + d2f) / (d1f * d2f);
        double* o = out_rd;
        o[kRdAmp] = A0; o[kRdDamp] = dA0; o[kRdPhi] = phm; o[kRdOmega] = wm;
        o[kRdC1c] = c1c; o[kRdC2c] = c2c; o[kRdC1f] = c1f; o[kRdC2f] = c2f;
        o[kRdD1c] = d1c; o[kRdD1f] = d1f; o[kRdD2f] = d2f;
        o[kRdWr] = wr; o[kRdWi] = wi; o[kRdEta] = eta; o[kRdPad0] = 0.0; o[kRdPad1] = 0.0;
        if (q == 0) {
            double* bf = binf + (size_t)b * kBinCount;
            bf[kBinTPeak] = t_peak;
            bf[kBinTMatch] = t_match;
            bf[kBinDt] = dt;
            bf[kBinDT22] = dT22;
            bini[(size_t)b * kBinICount + kBinNpts] = npts;
            bini[(size_t)b * kBinICount + kBinIdx] = idr;
        }
    }
}

// sin and cos of x: Taylor series through x^9 / x^10 when every active lane of
// the wavefront has |x| < 1/16 (first omitted terms below 2.3e-20 relative),
// through x^17 / x^18 for |x| < 1/2 (below 2e-23), the library's sincos
// otherwise.  The NQC phase of the inspiral samples is small except near the
// attachment.
__device__ __forceinline__ void sincos_nqc(double x, double* s, double* c) {
    const double ax = fabs(x);
    if (__all(ax < 0.0625)) {
        const double x2 = x * x;
        *s = x * (1.0 + x2 * (-1.0 / 6 + x2 * (1.0 / 120 + x2 * (-1.0 / 5040 + x2 * (1.0 / 362880)))));
        *c = 1.0 + x2 * (-0.5 + x2 * (1.0 / 24 + x2 * (-1.0 / 720 + x2 * (1.0 / 40320 + x2 * (-1.0 / 3628800)))));
    } else if (ax < 0.5) {
        const double x2 = x * x;
        *s = x * (1.0 + x2 * (-1.0 / 6 + x2 * (1.0 / 120 + x2 * (-1.0 / 5040 + x2 * (1.0 / 362880 + x2 * (-1.0 / 39916800
                 + x2 * (1.0 / 6227020800.0 + x2 * (-1.0 / 1307674368000.0 + x2 * (1.0 / 355687428096000.0)))))))));
        *c = 1.0 + x2 * (-0.5 + x2 * (1.0 / 24 + x2 * (-1.0 / 720 + x2 * (1.0 / 40320 + x2 * (-1.0 / 3628800 + x2 * (1.0 / 479001600.0
                 + x2 * (-1.0 / 87178291200.0 + x2 * (1.0 / 20922789888000.0 + x2 * (-1.0 / 6402373705728000.0)))))))));
    } else {
        sincos(x, s, c);
    }
}

// ------------------------------------------------------- the IMR template
// h_j = amp (sum over modes of h_lm NQC_lm (t_j) for j <= idx, of the ringdown
// for idx < j < idx + nadd, 0 after), t_j = t_0 + j dT; hplus = Re, hcross = Im.
constexpr int kImrBlock = 256;
constexpr int kImrSpan = 2048;

// One inspiral sample's sum over the NM modes: row holds the interval's
// spline coefficients (Re of each mode, Im of each mode, phi), rpr those of r
// and pr; nq the binary's NQC coefficients (5 per mode).  e^{-i m phi} is
// formed from e^{-i phi} by m - 1 products (the same products as a table of
// its powers, without selecting from the table per mode).
template <int NM, class Ptr>
__device__ __forceinline__ void imr_inspiral_sample(Ptr row, Ptr rpr, const Piece& p, const ImrModes& mm, const double* __restrict__ nq,
                                                    double& hr, double& hi) {
    const double phi = p.v(row + 4 * (2 * NM));
    const double om = p.d(row + 4 * (2 * NM));
    const double r = p.v(rpr), pr = p.v(rpr + 4);
    const double sr = sqrt(r), rO = r * om;
    const double P = pr * pr / (rO * rO);
    const double ir = 1.0 / r, irs = 1.0 / (r * sr);
    const double p1 = pr / rO, p3 = pr * pr * pr / rO;
    double sp, cph;
    sincos(phi, &sp, &cph);
#pragma unroll
    for (int q = 0; q < NM; ++q) {
        const double re = p.v(row + 4 * q), im = p.v(row + 4 * (NM + q));
        double pr_ = cph, pi_ = -sp;   // e^{-i phi}
        for (int k = 1; k < mm.m[q]; ++k) {
            const double t = pr_ * cph + pi_ * sp;
            pi_ = pi_ * cph - pr_ * sp;
            pr_ = t;
        }
        const double h_r = re * pr_ - im * pi_, h_i = re * pi_ + im * pr_;
        const double* c = nq + 5 * q;
        const double mag = 1.0 + P * (c[0] + c[1] * ir + c[2] * irs);
        double sn, cn;
        sincos_nqc(c[3] * p1 + p3 * c[4], &sn, &cn);
        const double fr = mag * cn, fi = mag * sn;
        hr += h_r * fr - h_i * fi;
        hi += h_r * fi + h_i * fr;
    }
}

// grid (blocks, binaries); a block takes spans of kImrSpan samples, stages the
// knot times of their inspiral part in LDS (knot_window) and searches there.
// When all lanes of a wavefront sample one knot interval, its coefficient
// rows are read once through the scalar cache (constant address space: the
// tables are read-only while the kernel runs).
template <int NM>
__global__ __launch_bounds__(kImrBlock) void k_imr_resample(const double* __restrict__ dense, const int* __restrict__ num_steps,
                                                            const double* __restrict__ coef, const double* __restrict__ rp,
                                                            const double* __restrict__ amp, const double* __restrict__ dT_M, int max_step,
                                                            int N, ImrModes mm, const double* __restrict__ nqc, const double* __restrict__ rd,
                                                            const double* __restrict__ binf, const int* __restrict__ bini,
                                                            double* __restrict__ hplus, double* __restrict__ hcross, int* __restrict__ nsamp) {
    typedef const __attribute__((address_space(4))) double* ConstPtr;
    constexpr int cs = 4 * (2 * NM + 1);
    const int b = blockIdx.y;
    const int len = num_steps[b];
    const int idx = bini[(size_t)b * kBinICount + kBinIdx];
    const int nadd = bini[(size_t)b * kBinICount + kBinNadd];
    double* outp = hplus + (size_t)b * N;
    double* outc = hcross + (size_t)b * N;
    const long long nvalid = (idx < 0 || len < 4) ? 0 : (long long)idx + (nadd > 0 ? nadd : 1);
    if (blockIdx.x == 0 && threadIdx.x == 0) nsamp[b] = (int)(nvalid < N ? nvalid : N);
    const double* T = dense + (size_t)b * max_step * 5;
    const double* C = coef + (size_t)b * max_step * cs;
    const double* RP = rp + (size_t)b * max_step * 8;
    const double* nq = nqc + (size_t)b * NM * 5;
    const double t0 = T[0];
    const double dT = dT_M[b];
    const double a = amp[b];
    const double t_match = binf[(size_t)b * kBinCount + kBinTMatch];
    const double dt = binf[(size_t)b * kBinCount + kBinDt];
    __shared__ KnotWindow<kImrBlock> win;
    const long long nspan = ((long long)N + kImrSpan - 1) / kImrSpan;
    for (long long sp0 = blockIdx.x; sp0 < nspan; sp0 += gridDim.x) {
        const long long j0 = sp0 * kImrSpan;
        const long long jinsp = (long long)idx + 1 < N ? (long long)idx + 1 : N;   // inspiral samples j < jinsp
        if (j0 < jinsp && nvalid > 0) {
            const long long jmax = (j0 + kImrSpan < jinsp ? j0 + kImrSpan : jinsp) - 1;
            knot_window<kImrBlock>(win, T, 5, len, grid_t(j0, dT, t0), grid_t(jmax, dT, t0));
        }
        for (int u = 0; u < kImrSpan / kImrBlock; ++u) {
            const long long j = j0 + (long long)u * kImrBlock + threadIdx.x;
            if (j >= N) break;
            double hr = 0.0, hi = 0.0;
            if (nvalid > 0 && j <= idx) {
                const double tj = grid_t(j, dT, t0);
                const int ind = knot_interval<kImrBlock>(win, T, 5, len, tj);
                const int i0 = __builtin_amdgcn_readfirstlane(ind);
                Piece p;
                p.last = (ind == len - 1);
                if (__all(ind == i0)) {
                    p.x = tj - ((ConstPtr)T)[5 * (size_t)i0];
                    p.x2 = p.x * p.x;
                    p.x3 = p.x2 * p.x;
                    imr_inspiral_sample<NM>((ConstPtr)(C + (size_t)cs * i0), (ConstPtr)(RP + (size_t)8 * i0), p, mm, nq, hr, hi);
                } else {
                    p.x = tj - T[5 * (size_t)ind];
                    p.x2 = p.x * p.x;
                    p.x3 = p.x2 * p.x;
                    imr_inspiral_sample<NM>(C + (size_t)cs * ind, RP + (size_t)8 * ind, p, mm, nq, hr, hi);
                }
            } else if (nvalid > 0 && j < nvalid) {
                const long long k = j - idx;
                const double x = (t_match + (double)k * dt) - t_match;
                for (int q = 0; q < NM; ++q) {
                    const double* o = rd + ((size_t)b * NM + q) * kRdCount;
                    const double Alm = o[kRdC1c] * tanh(o[kRdC1f] * x + o[kRdC2f]) + o[kRdC2c];
                    const double philm = o[kRdPhi] - o[kRdD1c] * log((1.0 + o[kRdD2f] * exp(-o[kRdD1f] * x)) / (1.0 + o[kRdD2f]));
                    const double mod = o[kRdEta] * Alm * exp(-o[kRdWi] * x);
                    double s, c;
                    sincos(philm - o[kRdWr] * x, &s, &c);
                    hr += mod * c;
                    hi += mod * s;
                }
            }
            outp[j] = a * hr;
            outc[j] = a * hi;
        }
        __syncthreads();   // the window is rewritten by the next span
    }
}

// ---------------------------------------------------------- drop-in fits
__global__ void k_imr_fit(int which, int l, int m, const double* __restrict__ x1, const double* __restrict__ x2,
                          const double* __restrict__ x3, const double* __restrict__ x4, int n, double* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double a = x1[i], b = x2 ? x2[i] : 0.0, c = x3 ? x3[i] : 0.0, d = x4 ? x4[i] : 0.0;
    double v = NAN;
    switch (which) {
        case EOBHIP_FIT_DT22: v = imr::nr_peak_delta_t({a, b, c, d}); break;
        // amplitude fits take (m1, m2, chiS, chiA) as the reference's functions do
        case EOBHIP_FIT_AMP: v = imr::nr_peak_amp(l, m, {a, b, c + d, c - d}); break;
        case EOBHIP_FIT_ADOT: v = imr::nr_peak_adot(l, m, {a, b, c + d, c - d}); break;
        case EOBHIP_FIT_ADDOT: v = imr::nr_peak_addot(l, m, {a, b, c + d, c - d}); break;
        case EOBHIP_FIT_OMEGA: v = imr::nr_peak_omega(l, m, a, b); break;
        case EOBHIP_FIT_OMEGADOT: v = imr::nr_peak_omegadot(l, m, a, b); break;
        case EOBHIP_FIT_C1F: { const int k = imr::rd_mode_index(l, m); if (k >= 0) v = imr::rd_poly(imr::rd_table_a1(k), a, b); } break;
        case EOBHIP_FIT_C2F: { const int k = imr::rd_mode_index(l, m); if (k >= 0) v = imr::rd_poly(imr::rd_table_a2(k), a, b); } break;
        case EOBHIP_FIT_D1F: { const int k = imr::rd_mode_index(l, m); if (k >= 0) v = imr::rd_poly(imr::rd_table_p1(k), a, b); } break;
        case EOBHIP_FIT_D2F: { const int k = imr::rd_mode_index(l, m); if (k >= 0) v = imr::rd_poly(imr::rd_table_p2(k), a, b); } break;
        case EOBHIP_FIT_FINAL_MASS: v = imr::final_mass_seobnrv4(a, b, c, d); break;
        case EOBHIP_FIT_FINAL_SPIN: v = imr::final_spin_hbr2016(a, b, c, d); break;
        case EOBHIP_FIT_ISCO_RADIUS: v = imr::kerr_isco_radius(a); break;
        default: break;
    }
    out[i] = v;
}

__global__ void k_kerr_qnm(int l, int m, const double* __restrict__ af, int n, double* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Cx w = kerr_qnm_fast(l, m, af[i]);
    out[2 * i] = w.re;
    out[2 * i + 1] = w.im;
}

int make_modes(ImrModes& mm, int nm, const int* ls, const int* ms) {
    EOBHIP_REQUIRE(nm >= 1 && nm <= kImrMaxModes, "1 <= nm <= %d", kImrMaxModes);
    mm.nm = nm;
    for (int i = 0; i < nm; ++i) {
        EOBHIP_REQUIRE(imr::rd_mode_index(ls[i], ms[i]) >= 0, "mode (%d, %d) has no ringdown fit", ls[i], ms[i]);
        mm.l[i] = ls[i];
        mm.m[i] = ms[i];
    }
    return EOBHIP_OK;
}

}  // namespace

extern "C" {

int eobhip_imr_fit(int which, int l, int m, const double* x1, const double* x2, const double* x3, const double* x4, int n, double* out,
                   void* stream) {
    EOBHIP_REQUIRE(n >= 0 && which >= 0 && which < EOBHIP_FIT_COUNT, "bad arguments");
    if (n == 0) return EOBHIP_OK;
    hipLaunchKernelGGL(k_imr_fit, dim3(blocks_for(n, 128)), dim3(128), 0, as_stream(stream), which, l, m, x1, x2, x3, x4, n, out);
    return check_launch("imr_fit");
}

int eobhip_kerr_qnm(int l, int m, const double* af, int n, double* omega, void* stream) {
    EOBHIP_REQUIRE(n >= 0 && l >= 2 && l <= 5 && m >= -l && m <= l && l == (abs(m) > 2 ? abs(m) : 2),
                   "bad arguments (the fundamental mode of l = max(2, |m|), l <= 5)");
    if (n == 0) return EOBHIP_OK;
    const int rc = ensure_qnm_table(stream);
    if (rc) return rc;
    hipLaunchKernelGGL(k_kerr_qnm, dim3(blocks_for(n, 64)), dim3(64), 0, as_stream(stream), l, m, af, n, omega);
    return check_launch("kerr_qnm");
}

int eobhip_imr_params(const double* m1, const double* m2, const double* chi1, const double* chi2, const double* dense,
                      const int* num_steps, const double* coef, const double* rp_coef, const double* dT_M, int n, int max_step, int nm,
                      const int* ls, const int* ms, double* qnm, double* nqc, double* rd, double* binf, int* bini, void* stream) {
    ImrModes mm;
    int rc = make_modes(mm, nm, ls, ms);
    if (rc) return rc;
    EOBHIP_REQUIRE(n >= 0 && max_step >= 4, "bad sizes");
    if (n == 0) return EOBHIP_OK;
    rc = ensure_qnm_table(stream);
    if (rc) return rc;
    const BinInputs in{m1, m2, chi1, chi2};
    const long long nq = (long long)n * (nm + 1);
    hipLaunchKernelGGL(k_imr_qnm, dim3(blocks_for(nq, 64)), dim3(64), 0, as_stream(stream), in, dense, dT_M, n, max_step, mm, qnm, binf,
                       bini);
    rc = check_launch("imr_qnm");
    if (rc) return rc;
    EOBHIP_REQUIRE((long long)n * nm < (1LL << 31), "too many (binary, mode) pairs");
    hipLaunchKernelGGL(k_imr_params, dim3(n * nm), dim3(64), 0, as_stream(stream), in, dense, num_steps, coef, rp_coef, dT_M, max_step, mm,
                       qnm, nqc, rd, binf, bini);
    return check_launch("imr_params");
}

int eobhip_imr_resample(const double* dense, const int* num_steps, const double* coef, const double* rp_coef, const double* amp,
                        const double* dT_M, int n, int max_step, int N, int nm, const int* ls, const int* ms, const double* nqc,
                        const double* rd, const double* binf, const int* bini, double* hplus, double* hcross, int* nsamp, void* stream) {
    ImrModes mm;
    int rc = make_modes(mm, nm, ls, ms);
    if (rc) return rc;
    EOBHIP_REQUIRE(n >= 0 && n <= 65535 && N >= 1, "bad sizes");
    if (n == 0) return EOBHIP_OK;
    int bx = (int)((N + (long long)kImrSpan - 1) / kImrSpan);
    if (bx > 4096) bx = 4096;
    const dim3 grid(bx, n), block(kImrBlock);
    hipStream_t st = as_stream(stream);
#define EOBHIP_IMR_RESAMPLE(NM)                                                                                                             \
    case NM:                                                                                                                              \
        hipLaunchKernelGGL(k_imr_resample<NM>, grid, block, 0, st, dense, num_steps, coef, rp_coef, amp, dT_M, max_step, N, mm, nqc, rd,  \
                           binf, bini, hplus, hcross, nsamp);                                                                             \
        break;
    switch (nm) {
        EOBHIP_IMR_RESAMPLE(1)
        EOBHIP_IMR_RESAMPLE(2)
        EOBHIP_IMR_RESAMPLE(3)
        EOBHIP_IMR_RESAMPLE(4)
        EOBHIP_IMR_RESAMPLE(5)
        EOBHIP_IMR_RESAMPLE(6)
        EOBHIP_IMR_RESAMPLE(7)
        EOBHIP_IMR_RESAMPLE(8)
        default: break;   // make_modes has checked 1 <= nm <= kImrMaxModes
    }
    static_assert(kImrMaxModes == 8, "one instantiation per mode count");
#undef EOBHIP_IMR_RESAMPLE
    return check_launch("imr_resample");
}

size_t eobhip_imr_rd_doubles(void) { return kRdCount; }
size_t eobhip_imr_bin_doubles(void) { return kBinCount; }
size_t eobhip_imr_bin_ints(void) { return kBinICount; }

}  // extern "C"
