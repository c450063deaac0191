// Batched not-a-knot cubic splines, uniform resampling of the (2,2) mode, the
// real-to-complex FFT of the templates and the frequency-domain likelihood.
//
// Reference: interpolate_TD (src/interpolate.cu:11-308), CubicSplineInterpolantTD
// (eob/seobnrv4phm.py:161-386), compute_full_waveform's resampling of
// h_lm e^{i m phi} and phi (eob/seobnrv4phm.py:2319-2372), the likelihood
// all_in_one_likelihood (src/interpolate.cu:553-628) with the host finish in
// get_ll (eob/seobnrv4phm.py:772-823).
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include <map>
#include <mutex>
#include <tuple>
#include <utility>

#include "eobhip_common.hpp"
#include "knot_window.hpp"
#include "spline_nak.hpp"

using namespace eobhip_detail;

namespace {

constexpr int kBlock = 256;
constexpr double kPi = 3.141592653589793;


// Drop-in layout: t [len][nbin], y/B/diagonals [len][ninterp], interpolant
// j = sub * nbin + bin.  One thread per interpolant; neighbouring threads walk
// neighbouring interpolants, so every knot step is one coalesced row access.
__global__ __launch_bounds__(kBlock) void k_interpolate_TD(const double* __restrict__ t, const double* __restrict__ y, double* B,
                                                           double* ud, double* dg, double* ld, const int* __restrict__ lengths,
                                                           int nbin, int ninterp) {
    const int it = blockIdx.x * blockDim.x + threadIdx.x;
    if (it >= ninterp) return;
    const int b = it % nbin;
    const int len = lengths[b];
    if (len < 3) return;
    // forward: assemble and eliminate in one sweep (Thomas, src/interpolate.cu:161-205)
    double prev_dg = 0.0, prev_ud = 0.0, prev_b = 0.0;
    for (int i = 0; i < len; ++i) {
        int i0, i1, i2;
        if (i == len - 1) { i0 = len - 3; i1 = len - 2; i2 = len - 1; }
        else if (i == 0) { i0 = 0; i1 = 1; i2 = 2; }
        else { i0 = i - 1; i1 = i; i2 = i + 1; }
        Row r = nak_row(i, len, t[(size_t)i0 * nbin + b], t[(size_t)i1 * nbin + b], t[(size_t)i2 * nbin + b], y[(size_t)i0 * ninterp + it],
                        y[(size_t)i1 * ninterp + it], y[(size_t)i2 * ninterp + it]);
        const size_t o = (size_t)i * ninterp + it;
        ld[o] = r.ld;
        if (i > 0) {
            const double w = r.ld / prev_dg;
            r.dg = r.dg - w * prev_ud;
            r.b = r.b - w * prev_b;
        }
        ud[o] = r.ud; dg[o] = r.dg; B[o] = r.b;
        prev_dg = r.dg; prev_ud = r.ud; prev_b = r.b;
    }
    // back substitution
    size_t o = (size_t)(len - 1) * ninterp + it;
    double snext = B[o] / dg[o];
    B[o] = snext;
    for (int i = len - 2; i >= 0; --i) {
        o = (size_t)i * ninterp + it;
        snext = (B[o] - ud[o] * snext) / dg[o];
        B[o] = snext;
    }
    // polynomial coefficients into ud (c1), dg (c2), ld (c3) for knots 0..len-2
    double s_i = B[it];
    for (int i = 0; i < len - 1; ++i) {
        const size_t oi = (size_t)i * ninterp + it, op = (size_t)(i + 1) * ninterp + it;
        const double dt = t[(size_t)(i + 1) * nbin + b] - t[(size_t)i * nbin + b];
        const double s_p = B[op];
        const double slope = (y[op] - y[oi]) / dt;
        const double tt = (s_i + s_p - 2 * slope) / dt;
        ud[oi] = s_i;
        dg[oi] = (slope - s_i) / dt - tt;
        ld[oi] = tt / dt;
        s_i = s_p;
    }
    // the last row keeps the eliminated system values (ud 0, reduced diag, ld), as the reference leaves them
}

// ---------------------------------------------------- batched EOB splines
// nser interpolants per binary over its trajectory knots t_k (units of M):
// knots [n][max_step][nser] (for the (2,2) pipeline nser = 3: phi, Re and Im of
// h22 e^{2 i phi}), coef [n][max_step][4 nser] = (y, c1, c2, c3) per interpolant.
//
// One wavefront per binary builds all of them.  Each interpolant's slope
// system is split into P <= 64 / nser chunks of >= kChunkMin rows, one lane
// per (interpolant, chunk); the lane solves its chunk together with kHalo rows on either side as
// a system of its own (cut off where the halo ends) and keeps the slopes of its
// own rows.  Inside the system every row is diagonally dominant (2(dx1 + dx2)
// against dx1 + dx2), so the effect of the cut on a row k rows away is damped
// by at least 2^-k, and by 0.268^k on evenly spaced knots: with 32 halo rows
// it is below 5e-19 of the neighbouring slopes on even knots and 2e-10 in the
// worst case the dominance allows (every spacing shrinking by a large factor
// over 32 rows, which an adaptive step sequence cannot do).  Emulated on the
// CPU against the uncut solve, 32 halo rows give the same coefficients as 64
// to rounding on trajectory knots (spacing ratio 33 where the fixed-step pass
// begins) and on 5k-60k-knot chirps; 24 rows already move them by 1e-14.  The
// not-a-knot end rows are never cut.
//
// Nothing of the system goes to memory: the rows are formed from t and y as the
// sweeps reach them, the forward elimination keeps only its carry (c', d') at
// the start of every kSeg rows (a checkpoint, 16 bytes per kSeg rows), and the
// back substitution walks the segments top down, recomputing each segment's
// (c', d') in registers from its checkpoint and writing the polynomial
// coefficients of each interval as soon as both end slopes are known.  Memory
// traffic is t and y read once and the coefficients written once (the
// recomputation and the halos re-read rows from the caches).
constexpr int kHalo = 32;
constexpr int kChunkMin = 96;
constexpr int kSplineLanes = 64;
constexpr int kSeg = 16;
constexpr int kMaxSeries = 32;

__host__ __device__ inline int spline_chunks(int len, int nser) {
    const int pmax = kSplineLanes / nser;
    int P = len / kChunkMin;
    return P < 1 ? 1 : (P > pmax ? pmax : P);
}
// checkpoints per lane for a binary of len knots (segments of its chunk's sweep)
__host__ __device__ inline int spline_segs(int len, int nser) {
    const int P = spline_chunks(len, nser);
    const int m = (len + P - 1) / P;
    return (m + 2 * kHalo + kSeg - 1) / kSeg + 1;
}
// doubles per binary: nser P lanes x 2 doubles x spline_segs for every len <=
// max_step (nser P 2 segs <= 2 nser len / kSeg + 2 P nser (1 + 2 kHalo) / kSeg +
// 4 P nser with P nser <= 64; checked for nser <= 32 and every len up to 70000)
__host__ __device__ inline long long spline_scratch_per_bin(int max_step, int nser) {
    return (2LL * nser * max_step + 2LL * kSplineLanes * (2 * kHalo + 2 * kSeg)) / kSeg + 4 * kSplineLanes + 64;
}

// t and y of rows start-2 .. start+kSeg+1, clamped into [0, len-1]; issued together
__device__ __forceinline__ void seg_load(const double* __restrict__ T, const double* __restrict__ Y, int ys, int start, int len,
                                         double* wt, double* wy) {
#pragma unroll
    for (int k = 0; k < kSeg + 4; ++k) {
        int i = start - 2 + k;
        i = i < 0 ? 0 : (i > len - 1 ? len - 1 : i);
        wt[k] = T[5 * i];
        wy[k] = Y[(size_t)ys * i];
    }
}

// Forward elimination over rows start .. end-1 (end - start <= kSeg) of the
// system cut to rows lo .. hi-1, from the carry (pc, pd) of the row below;
// kKeep stores every row's (c', d').  Interior row i = start + j reads the
// window at j+1, j+2, j+3 (rows i-1, i, i+1); the two end rows (rfirst, rlast)
// are formed once by the caller and picked by value, so that every window
// index stays a compile-time constant (a select between window entries would
// turn into a computed address and push the window out of registers).
template <bool kKeep>
__device__ __forceinline__ void seg_forward(const double* wt, const double* wy, int start, int end, int lo, int hi, int len,
                                            const Row& rfirst, const Row& rlast, double& pc, double& pd, double* pcs, double* pds) {
#pragma unroll
    for (int j = 0; j < kSeg; ++j) {
        const int i = start + j;
        if (i < end) {
            Row r = nak_row(1, 3, wt[j + 1], wt[j + 2], wt[j + 3], wy[j + 1], wy[j + 2], wy[j + 3]);   // the interior form
            if (i == 0) r = rfirst;
            if (i == len - 1) r = rlast;
            const double a = (i == lo) ? 0.0 : r.ld;        // cut below the halo
            const double c = (i == hi - 1) ? 0.0 : r.ud;    // cut above the halo
            const double den = r.dg - a * pc;
            pc = c / den;
            pd = (r.b - a * pd) / den;
            if constexpr (kKeep) {
                pcs[j] = pc;
                pds[j] = pd;
            }
        }
    }
}

__global__ __launch_bounds__(kSplineLanes) void k_spline_build_seg(const double* __restrict__ dense, const double* __restrict__ knots,
                                                                   const int* __restrict__ num_steps, int n, int max_step, int nser,
                                                                   double* __restrict__ coef, double* __restrict__ scratch) {
    const int b = blockIdx.x;
    if (b >= n) return;
    const int len = num_steps[b];
    const int lane = threadIdx.x;
    const int cs = 4 * nser;   // coefficient row stride
    const double* T = dense + (size_t)b * max_step * 5;
    if (len < 3) {
        if (lane < nser) {
            const double* Y = knots + (size_t)b * max_step * nser + lane;
            double* C = coef + (size_t)b * max_step * cs + 4 * lane;
            for (int i = 0; i < len; ++i) {
                C[(size_t)cs * i] = Y[(size_t)nser * i];
                C[(size_t)cs * i + 1] = C[(size_t)cs * i + 2] = C[(size_t)cs * i + 3] = 0.0;
            }
        }
        return;
    }
    const int P = spline_chunks(len, nser);
    if (lane >= nser * P) return;
    const int q = lane / P, cidx = lane % P;
    const double* Y = knots + (size_t)b * max_step * nser + q;
    double* C = coef + (size_t)b * max_step * cs + 4 * q;
    const int m = (len + P - 1) / P;
    const int s0 = cidx * m, e0 = (s0 + m < len) ? s0 + m : len;
    if (s0 >= e0) return;
    const int lo = (s0 - kHalo > 0) ? s0 - kHalo : 0;
    const int hi = (e0 + kHalo < len) ? e0 + kHalo : len;
    const int nseg = (hi - lo + kSeg - 1) / kSeg;
    double* ck = scratch + (size_t)b * spline_scratch_per_bin(max_step, nser) + (size_t)lane * 2 * spline_segs(len, nser);
    const Row rfirst = nak_row(0, len, T[0], T[5], T[10], Y[0], Y[nser], Y[2 * nser]);
    const Row rlast = nak_row(len - 1, len, T[5 * (len - 3)], T[5 * (len - 2)], T[5 * (len - 1)], Y[(size_t)nser * (len - 3)],
                              Y[(size_t)nser * (len - 2)], Y[(size_t)nser * (len - 1)]);
    // the next segment's rows (and checkpoint) are loaded while this one is
    // eliminated: the sweeps are chains of dependent divisions, and without the
    // prefetch each segment would wait a memory round trip for its rows
    double wt[kSeg + 4], wy[kSeg + 4], nt[kSeg + 4], ny[kSeg + 4];
    double pc = 0.0, pd = 0.0;
    seg_load(T, Y, nser, lo, len, wt, wy);
    for (int sg = 0; sg < nseg; ++sg) {
        const int start = lo + sg * kSeg, end = (start + kSeg < hi) ? start + kSeg : hi;
        ck[2 * sg] = pc;
        ck[2 * sg + 1] = pd;
        if (sg + 1 < nseg) seg_load(T, Y, nser, start + kSeg, len, nt, ny);
        seg_forward<false>(wt, wy, start, end, lo, hi, len, rfirst, rlast, pc, pd, nullptr, nullptr);
        if (sg + 1 < nseg) {
#pragma unroll
            for (int k = 0; k < kSeg + 4; ++k) { wt[k] = nt[k]; wy[k] = ny[k]; }
        }
    }
    // back substitution, segment by segment from the top; xn = slope of the row above
    double xn = 0.0;
    int sg = nseg - 1;
    seg_load(T, Y, nser, lo + sg * kSeg, len, wt, wy);
    double cpc = ck[2 * sg], cpd = ck[2 * sg + 1];
    for (; sg >= 0; --sg) {
        const int start = lo + sg * kSeg, end = (start + kSeg < hi) ? start + kSeg : hi;
        if (end <= s0) break;   // the rows below this lane's chunk need no slopes
        double pcs[kSeg], pds[kSeg];
        pc = cpc;
        pd = cpd;
        const bool more = sg > 0 && start > s0;
        if (more) {
            seg_load(T, Y, nser, start - kSeg, len, nt, ny);
            cpc = ck[2 * (sg - 1)];
            cpd = ck[2 * (sg - 1) + 1];
        }
        seg_forward<true>(wt, wy, start, end, lo, hi, len, rfirst, rlast, pc, pd, pcs, pds);
#pragma unroll
        for (int j = kSeg - 1; j >= 0; --j) {
            const int i = start + j;
            if (i < end) {
                const double xi = pds[j] - pcs[j] * xn;   // the cut row hi-1 has c' = 0
                if (i >= s0 && i < e0) {
                    double* c = C + (size_t)cs * i;
                    if (i < len - 1) {
                        const double yi = wy[j + 2], yp = wy[j + 3];
                        const double dt = wt[j + 3] - wt[j + 2];
                        const double slope = (yp - yi) / dt;
                        const double tt = (xi + xn - 2 * slope) / dt;
                        c[0] = yi;
                        c[1] = xi;
                        c[2] = (slope - xi) / dt - tt;
                        c[3] = tt / dt;
                    } else {
                        c[0] = wy[j + 2];
                        c[1] = c[2] = c[3] = 0.0;
                    }
                }
                xn = xi;
            }
        }
        if (!more) break;
#pragma unroll
        for (int k = 0; k < kSeg + 4; ++k) { wt[k] = nt[k]; wy[k] = ny[k]; }
    }
}

// Uniform resample of h+ = amp * Re[(Re S + i Im S)(t) e^{-2 i phi(t)}] at
// t_j = t_0 + j dT_M (dT_M = dt / (M MTSUN)), j < N_b = floor((t_last - t_0)/dT_M) + 1,
// zero for N_b <= j < N.  Interval: searchsorted(t, t_j, 'right') - 1 over the knots
// (eob/seobnrv4phm.py:241-272).  grid (blocks, binaries); a block takes spans of
// kResSpan consecutive samples, stages the knot times of their intervals in LDS
// (knot_window) and each thread searches there.
constexpr int kResPerThread = 16;
constexpr int kResSpan = kBlock * kResPerThread;
constexpr int kResSearch = 4;   // samples per thread whose knot searches run together
static_assert(kResPerThread % kResSearch == 0, "kResSearch divides kResPerThread");
__global__ __launch_bounds__(kBlock) void k_resample_hplus(const double* __restrict__ dense, const double* __restrict__ coef,
                                                           const int* __restrict__ num_steps, const double* __restrict__ amp,
                                                           const double* __restrict__ dT_M, int n, int max_step, int N,
                                                           double* __restrict__ hplus, int* __restrict__ nsamp,
                                                           double* __restrict__ hcross) {
    const int b = blockIdx.y;
    const int len = num_steps[b];
    double* out = hplus + (size_t)b * N;
    double* outc = hcross ? hcross + (size_t)b * N : nullptr;
    if (len < 2) {
        for (long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x; j < N; j += (long long)gridDim.x * blockDim.x) {
            out[j] = 0.0;
            if (outc) outc[j] = 0.0;
        }
        if (blockIdx.x == 0 && threadIdx.x == 0) nsamp[b] = 0;
        return;
    }
    const double* T = dense + (size_t)b * max_step * 5;
    const double* C = coef + (size_t)b * max_step * 12;
    const double t0 = T[0], tl = T[5 * (len - 1)];
    const double dT = dT_M[b];
    long long nb = (long long)((tl - t0) / dT) + 1;
    if (nb > N) nb = N;
    if (blockIdx.x == 0 && threadIdx.x == 0) nsamp[b] = (int)nb;
    const double a = amp[b];
    __shared__ KnotWindow<kBlock> win;
    const long long nspan = ((long long)N + kResSpan - 1) / kResSpan;
    for (long long sp0 = blockIdx.x; sp0 < nspan; sp0 += gridDim.x) {   // the same trip count on every thread
        const long long j0 = sp0 * kResSpan;
        if (j0 < nb) {
            const long long jmax = (j0 + kResSpan < nb ? j0 + kResSpan : nb) - 1;
            knot_window<kBlock>(win, T, 5, len, t0 + (double)j0 * dT, t0 + (double)jmax * dT);
        }
        for (int u0 = 0; u0 < kResPerThread; u0 += kResSearch) {
            double tj[kResSearch];
            int ind[kResSearch];
#pragma unroll
            for (int k = 0; k < kResSearch; ++k) tj[k] = t0 + (double)(j0 + (long long)(u0 + k) * kBlock + threadIdx.x) * dT;
            if (j0 < nb) knot_intervals<kBlock, kResSearch>(win, T, 5, len, tj, ind);
#pragma unroll
            for (int k = 0; k < kResSearch; ++k) {
                const long long j = j0 + (long long)(u0 + k) * kBlock + threadIdx.x;
                if (j < N) {
                    double val = 0.0, valc = 0.0;
                    if (j < nb) {
                        double tk, c[12];
                        interval_row<12>(T, 5, C, ind[k], tk, c);
                        const double x = tj[k] - tk;
                        const double x2 = x * x, x3 = x2 * x;
                        const double ph = c[0] + c[1] * x + c[2] * x2 + c[3] * x3;
                        const double sr = c[4] + c[5] * x + c[6] * x2 + c[7] * x3;
                        const double si = c[8] + c[9] * x + c[10] * x2 + c[11] * x3;
                        double sp, cp;
                        sincos(2.0 * ph, &sp, &cp);
                        // the template (sr + i si)(cos 2ph - i sin 2ph); h+ and hx are its real and
                        // imaginary parts, the series the reference transforms (eob/seobnrv4phm.py:720, 742-744)
                        val = a * (sr * cp + si * sp);
                        valc = a * (si * cp - sr * sp);
                    }
                    out[j] = val;
                    if (outc) outc[j] = valc;
                }
            }
        }
        __syncthreads();   // the window is rewritten by the next span
    }
}

// Uniform resample of the mode sum h = amp sum_j S_j(t) e^{-i m_j phi(t)} (the
// reference's modes *= exp(-1j m phi_interp) and template = modes.sum,
// eob/seobnrv4phm.py:2366-2372, 720) at t_j = t_0 + j dT_M, j < N_b, zero past;
// hplus, hcross = Re h, Im h (the two series it transforms, :742-744).  coef
// [n][max_step][4 (2 nm + 1)]: series j < nm Re S_j, nm + j Im S_j, 2 nm phi.
// e^{-i m phi} is formed as the m-th power of e^{-i phi}.
constexpr int kMaxModeM = 8;
struct ModeMs {
    int nm;
    int m[15];
};
// One sample's mode sum: row holds the interval's coefficients (Re of each
// mode, Im of each mode, phi; 4 per series).  e^{-i m phi} is m - 1 products
// with e^{-i phi} (the products a table of the powers would hold, without a
// per-mode select from the table).
template <int NMT, class Ptr>
__device__ __forceinline__ void modes_sample(Ptr row, double x, double x2, double x3, int nm, const ModeMs& mm, double& hr,
                                             double& hi) {
    Ptr cp = row + 4 * (2 * nm);
    const double ph = cp[0] + cp[1] * x + cp[2] * x2 + cp[3] * x3;
    double sp, cph;
    sincos(ph, &sp, &cph);
#pragma unroll
    for (int q = 0; q < (NMT > 0 ? NMT : nm); ++q) {
        Ptr cr = row + 4 * q;
        Ptr ci = row + 4 * (nm + q);
        const double sr = cr[0] + cr[1] * x + cr[2] * x2 + cr[3] * x3;
        const double si = ci[0] + ci[1] * x + ci[2] * x2 + ci[3] * x3;
        double pr_ = cph, pi_ = -sp;   // e^{-i phi}
        for (int k = 1; k < mm.m[q]; ++k) {
            const double t = pr_ * cph + pi_ * sp;
            pi_ = pi_ * cph - pr_ * sp;
            pr_ = t;
        }
        hr += sr * pr_ - si * pi_;
        hi += sr * pi_ + si * pr_;
    }
}

// NMT > 0: built for that many modes (the loop over modes unrolled); 0: any.
// When all lanes of a wavefront sample one knot interval, its coefficient row
// is read once through the scalar cache (constant address space: the table
// is read-only while the kernel runs).
template <int NMT>
__global__ __launch_bounds__(kBlock) void k_resample_modes(const double* __restrict__ dense, const double* __restrict__ coef,
                                                           const int* __restrict__ num_steps, const double* __restrict__ amp,
                                                           const double* __restrict__ dT_M, int max_step, int N, ModeMs mm,
                                                           double* __restrict__ hplus, double* __restrict__ hcross,
                                                           int* __restrict__ nsamp) {
    typedef const __attribute__((address_space(4))) double* ConstPtr;
    const int b = blockIdx.y;
    const int len = num_steps[b];
    const int nm = NMT > 0 ? NMT : mm.nm, S = 2 * nm + 1, cs = 4 * S;
    double* out = hplus + (size_t)b * N;
    double* outc = hcross + (size_t)b * N;
    if (len < 2) {
        for (long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x; j < N; j += (long long)gridDim.x * blockDim.x) {
            out[j] = 0.0;
            outc[j] = 0.0;
        }
        if (blockIdx.x == 0 && threadIdx.x == 0) nsamp[b] = 0;
        return;
    }
    const double* T = dense + (size_t)b * max_step * 5;
    const double* C = coef + (size_t)b * max_step * cs;
    const double t0 = T[0], tl = T[5 * (len - 1)];
    const double dT = dT_M[b];
    long long nb = (long long)((tl - t0) / dT) + 1;
    if (nb > N) nb = N;
    if (blockIdx.x == 0 && threadIdx.x == 0) nsamp[b] = (int)nb;
    const double a = amp[b];
    __shared__ KnotWindow<kBlock> win;
    const long long nspan = ((long long)N + kResSpan - 1) / kResSpan;
    for (long long sp0 = blockIdx.x; sp0 < nspan; sp0 += gridDim.x) {
        const long long j0 = sp0 * kResSpan;
        if (j0 < nb) {
            const long long jmax = (j0 + kResSpan < nb ? j0 + kResSpan : nb) - 1;
            knot_window<kBlock>(win, T, 5, len, t0 + (double)j0 * dT, t0 + (double)jmax * dT);
        }
        for (int u = 0; u < kResPerThread; ++u) {
            const long long j = j0 + (long long)u * kBlock + threadIdx.x;
            if (j >= N) break;
            double hr = 0.0, hi = 0.0;
            if (j < nb) {
                const double tj = t0 + (double)j * dT;
                const int ind = knot_interval<kBlock>(win, T, 5, len, tj);
                const int i0 = __builtin_amdgcn_readfirstlane(ind);
                if (__all(ind == i0)) {
                    const double x = tj - ((ConstPtr)T)[5 * (size_t)i0];
                    modes_sample<NMT>((ConstPtr)(C + (size_t)cs * i0), x, x * x, x * x * x, nm, mm, hr, hi);
                } else {
                    const double x = tj - T[5 * (size_t)ind];
                    modes_sample<NMT>(C + (size_t)cs * ind, x, x * x, x * x * x, nm, mm, hr, hi);
                }
                hr *= a;
                hi *= a;
            }
            out[j] = hr;
            outc[j] = hi;
        }
        __syncthreads();   // the window is rewritten by the next span
    }
}

// TD2 (src/interpolate.cu:441-549), TDInterp_wrap2's kernel: per binary b and
// sample i < lengths[b], sum over modes of (re + i im) e^{-i m phi} with every
// series evaluated at the caller's interval inds[b][i] (t - ts[b][ind], data
// time dataTime[i] shared by the binaries).  Series s of binary b starts at
// b (2M + 1) old_length + s * stride: stride old_lengths[b] as the reference
// indexes (correct only when every binary fills old_length), or old_length for
// the padded layout (legacy_stride 0).  Intervals are clamped into
// [0, old_length - 1].  grid (blocks over samples, binaries).
__global__ __launch_bounds__(kBlock) void k_td_interp2(double2* __restrict__ out, const double* __restrict__ dataTime,
                                                       const double* __restrict__ ts, const double* __restrict__ y,
                                                       const double* __restrict__ c1, const double* __restrict__ c2,
                                                       const double* __restrict__ c3, int old_length, const int* __restrict__ old_lengths,
                                                       int data_length, int numModes, const int* __restrict__ ms,
                                                       const int* __restrict__ inds, const int* __restrict__ lengths, int max_length,
                                                       int legacy_stride) {
    const int b = blockIdx.y;
    // sample counts and strides are clamped to the arrays' extents (the reference
    // trusts them; for valid inputs the clamps change nothing)
    int n = lengths[b];
    n = n < max_length ? n : max_length;
    n = n < data_length ? n : data_length;
    int olh = old_lengths[b];
    olh = olh < 0 ? 0 : (olh > old_length ? old_length : olh);
    const long long stride = legacy_stride ? olh : old_length;
    const size_t base = (size_t)b * (2 * numModes + 1) * old_length;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        int ind = inds[(size_t)max_length * b + i];
        ind = ind < 0 ? 0 : (ind > old_length - 1 ? old_length - 1 : ind);
        const double x = dataTime[i] - ts[(size_t)b * old_length + ind];
        const double x2 = x * x, x3 = x * x2;
        size_t k = base + (size_t)(2 * numModes) * stride + ind;
        const double phi = y[k] + c1[k] * x + c2[k] * x2 + c3[k] * x3;
        double hr = 0.0, hi = 0.0;
        for (int q = 0; q < numModes; ++q) {
            k = base + (size_t)q * stride + ind;
            const double re = y[k] + c1[k] * x + c2[k] * x2 + c3[k] * x3;
            k = base + (size_t)(numModes + q) * stride + ind;
            const double im = y[k] + c1[k] * x + c2[k] * x2 + c3[k] * x3;
            double s, c;
            sincos(-((double)ms[q] * phi), &s, &c);
            hr += re * c - im * s;
            hi += re * s + im * c;
        }
        out[(size_t)b * data_length + i] = make_double2(hr, hi);
    }
}

// log L_b = -1/2 * 4 df * sum_k |d_k - h_k|^2 / S(f_k), S(f) = (f/f0)^-4 + 1 + (f/f0)^2,
// 1/S(0) = 0.  Stage 1: per (chunk, binary) partial sums with wave shuffles + LDS;
// chunks of one binary are adjacent in blockIdx.x so the data spectrum is read
// once from HBM and reused from L2 by every binary.
constexpr int kLikeBlock = 256;
constexpr int kLikePerThread = 8;
__global__ __launch_bounds__(kLikeBlock) void k_fd_likelihood_partial(const double2* __restrict__ ht, const double2* __restrict__ dt,
                                                                      int n, int F, double df, double f0, double* __restrict__ partial,
                                                                      int nchunks) {
    const int b = blockIdx.x % n;
    const int chunk = blockIdx.x / n;
    const int base = chunk * kLikeBlock * kLikePerThread;
    const double2* h = ht + (size_t)b * F;
    double acc = 0.0;
#pragma unroll
    for (int u = 0; u < kLikePerThread; ++u) {
        const int k = base + u * kLikeBlock + threadIdx.x;
        if (k < F) {
            const double2 hv = h[k];
            const double2 dv = dt[k];
            const double re = dv.x - hv.x, im = dv.y - hv.y;
            const double x = df * (double)k / f0;
            const double x2 = x * x, x4 = x2 * x2;
            const double w = (k == 0) ? 0.0 : x4 / (1.0 + x4 + x4 * x2);   // 1 / S(f)
            acc += (re * re + im * im) * w;
        }
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) acc += __shfl_xor(acc, off, 64);
    __shared__ double red[kLikeBlock / 64];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < kLikeBlock / 64; ++w) s += red[w];
        partial[(size_t)b * nchunks + chunk] = s;
    }
}

__global__ __launch_bounds__(64) void k_fd_likelihood_final(const double* __restrict__ partial, int n, int nchunks, double df,
                                                            double* __restrict__ logl) {
    const int b = blockIdx.x;
    if (b >= n) return;
    double acc = 0.0;
    for (int c = threadIdx.x; c < nchunks; c += 64) acc += partial[(size_t)b * nchunks + c];
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if (threadIdx.x == 0) logl[b] = -0.5 * 4.0 * df * acc;
}

// Drop-in all_in_one_likelihood: per 1024-bin block partial sums of
// |d - (F+ hp + Fx hc) e^{-2 pi i f dt}|^2 / psd.  One 256-thread block covers
// 1024 bins (4 per thread); grid (bins / 1024, nbin, nchan).
__global__ __launch_bounds__(kBlock) void k_all_in_one_likelihood(double2* temp_sum, const double2* __restrict__ hp,
                                                                  const double2* __restrict__ hc, const double2* __restrict__ data,
                                                                  const double* __restrict__ psd, const double* __restrict__ Fp,
                                                                  const double* __restrict__ Fc, const double* __restrict__ ts,
                                                                  double df, int nbin, int nchan, int F) {
    const int blk = blockIdx.x, b = blockIdx.y, c = blockIdx.z;
    const double fp = Fp[c * nbin + b], fc = Fc[c * nbin + b], tsh = ts[c * nbin + b];
    double acc = 0.0;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int k = blk * 1024 + u * kBlock + threadIdx.x;
        if (k < F) {
            const double2 a = hp[(size_t)b * F + k], e = hc[(size_t)b * F + k], d = data[(size_t)c * F + k];
            const double hr = fp * a.x + fc * e.x, hi = fp * a.y + fc * e.y;
            double s, co;
            sincos(-2.0 * kPi * (df * k) * tsh, &s, &co);
            const double rr = d.x - (hr * co - hi * s), ri = d.y - (hr * s + hi * co);
            acc += (rr * rr + ri * ri) / psd[(size_t)c * F + k];
        }
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) acc += __shfl_xor(acc, off, 64);
    __shared__ double red[kBlock / 64];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < kBlock / 64; ++w) s += red[w];
        temp_sum[((size_t)blk * nbin + b) * nchan + c] = make_double2(s, 0.0);
    }
}

}  // namespace

namespace eobhip_detail {

// hipFFT plans, cached per (N, rows per call, stream, direction).  A plan owns a
// work area, so two streams never share one: work queued on different streams
// may overlap.  Callers pass rows in chunks of kFftChunk and the rest in
// power-of-two chunks, which bounds the plans per stream and size to seven.
namespace {
std::mutex g_plan_mu;
std::map<std::tuple<long long, int, void*, int>, hipfftHandle> g_plans;
}  // namespace

int fft_exec(int inverse, const void* in, void* out, int rows, int N, void* stream) {
    std::lock_guard<std::mutex> g(g_plan_mu);
    hipfftHandle plan;
    const auto key = std::make_tuple((long long)N, rows, stream, inverse);
    auto itp = g_plans.find(key);
    if (itp == g_plans.end()) {
        int dims[1] = {N};
        hipfftResult r = inverse ? hipfftPlanMany(&plan, 1, dims, nullptr, 1, N / 2 + 1, nullptr, 1, N, HIPFFT_Z2D, rows)
                                 : hipfftPlanMany(&plan, 1, dims, nullptr, 1, N, nullptr, 1, N / 2 + 1, HIPFFT_D2Z, rows);
        if (r != HIPFFT_SUCCESS) {
            set_error("hipfftPlanMany failed (%d)", (int)r);
            return EOBHIP_EHIP;
        }
        r = hipfftSetStream(plan, as_stream(stream));
        if (r != HIPFFT_SUCCESS) {
            set_error("hipfftSetStream failed (%d)", (int)r);
            return EOBHIP_EHIP;
        }
        g_plans[key] = plan;
    } else {
        plan = itp->second;
    }
    hipfftResult r = inverse ? hipfftExecZ2D(plan, reinterpret_cast<hipfftDoubleComplex*>(const_cast<void*>(in)),
                                             static_cast<double*>(out))
                             : hipfftExecD2Z(plan, static_cast<double*>(const_cast<void*>(in)),
                                             reinterpret_cast<hipfftDoubleComplex*>(out));
    if (r != HIPFFT_SUCCESS) {
        set_error("hipfftExec failed (%d)", (int)r);
        return EOBHIP_EHIP;
    }
    return EOBHIP_OK;
}

int fft_rows(int inverse, const void* in, void* out, int n, int N, void* stream) {
    const size_t F = (size_t)N / 2 + 1;
    const size_t in_stride = inverse ? 2 * F : (size_t)N;     // doubles per input row
    const size_t out_stride = inverse ? (size_t)N : 2 * F;    // doubles per output row
    const double* src = static_cast<const double*>(in);
    double* dst = static_cast<double*>(out);
    // full chunks of kFftChunk rows, then the rest by its binary digits (32, 16, ..., 1)
    int row = 0;
    for (int chunk = kFftChunk; chunk >= 1; chunk >>= 1) {
        while (row + chunk <= n) {
            const int rc = fft_exec(inverse, src + row * in_stride, dst + row * out_stride, chunk, N, stream);
            if (rc) return rc;
            row += chunk;
            if (chunk < kFftChunk) break;
        }
    }
    return EOBHIP_OK;
}

}  // namespace eobhip_detail

extern "C" {

int eobhip_interpolate_TD(const double* t_arr, const double* propArrays, double* B, double* upper_diag, double* diag, double* lower_diag,
                          const int* lengths, int numBinAll, int nsubs, void* stream) {
    EOBHIP_REQUIRE(numBinAll >= 0 && nsubs >= 0, "bad sizes");
    const int ni = numBinAll * nsubs;
    if (ni == 0) return EOBHIP_OK;
    hipLaunchKernelGGL(k_interpolate_TD, dim3(blocks_for(ni, kBlock)), dim3(kBlock), 0, as_stream(stream), t_arr, propArrays, B,
                       upper_diag, diag, lower_diag, lengths, numBinAll, ni);
    return check_launch("interpolate_TD");
}

int eobhip_all_in_one_likelihood(double* temp_sum, const double* hp_fd, const double* hc_fd, const double* data, const double* psd,
                                 const double* Fplus, const double* Fcross, const double* time_shift, double df, int num_bin_all,
                                 int nchannels, int data_length, void* stream) {
    EOBHIP_REQUIRE(num_bin_all >= 0 && nchannels >= 0 && data_length >= 0, "bad sizes");
    EOBHIP_REQUIRE(num_bin_all <= 65535 && nchannels <= 65535, "num_bin_all and nchannels must be <= 65535");
    if (num_bin_all == 0 || nchannels == 0 || data_length == 0) return EOBHIP_OK;
    const int nblk = (data_length + 1023) / 1024;
    hipLaunchKernelGGL(k_all_in_one_likelihood, dim3(nblk, num_bin_all, nchannels), dim3(kBlock), 0, as_stream(stream),
                       reinterpret_cast<double2*>(temp_sum), reinterpret_cast<const double2*>(hp_fd),
                       reinterpret_cast<const double2*>(hc_fd), reinterpret_cast<const double2*>(data), psd, Fplus, Fcross, time_shift,
                       df, num_bin_all, nchannels, data_length);
    return check_launch("all_in_one_likelihood");
}

size_t eobhip_spline_scratch_doubles(int n, int max_step) {
    if (n < 0 || max_step < 1) return 0;
    return (size_t)n * (size_t)spline_scratch_per_bin(max_step, 3);
}

int eobhip_spline_build_batched(const double* dense, const double* knots, const int* num_steps, int n, int max_step, double* coef,
                                double* scratch, void* stream) {
    return eobhip_spline_build_series(dense, knots, num_steps, n, max_step, 3, coef, scratch, stream);
}

size_t eobhip_spline_series_scratch_doubles(int n, int max_step, int nser) {
    if (n < 0 || max_step < 1 || nser < 1 || nser > kMaxSeries) return 0;
    return (size_t)n * (size_t)spline_scratch_per_bin(max_step, nser);
}

int eobhip_spline_build_series(const double* dense, const double* knots, const int* num_steps, int n, int max_step, int nser,
                               double* coef, double* scratch, void* stream) {
    EOBHIP_REQUIRE(n >= 0 && max_step >= 1, "bad sizes");
    EOBHIP_REQUIRE(nser >= 1 && nser <= kMaxSeries, "nser must be 1 .. 32");
    if (n == 0) return EOBHIP_OK;
    hipLaunchKernelGGL(k_spline_build_seg, dim3(n), dim3(kSplineLanes), 0, as_stream(stream), dense, knots, num_steps, n, max_step, nser,
                       coef, scratch);
    return check_launch("spline_build");
}

int eobhip_resample_hplus(const double* dense, const double* coef, const int* num_steps, const double* amp, const double* dT_M, int n,
                          int max_step, int N, double* hplus, int* nsamp, void* stream) {
    EOBHIP_REQUIRE(n >= 0 && N >= 1 && n <= 65535, "bad sizes");
    if (n == 0) return EOBHIP_OK;
    int bx = (int)((N + (long long)kResSpan - 1) / kResSpan);
    if (bx > 4096) bx = 4096;
    hipLaunchKernelGGL(k_resample_hplus, dim3(bx, n), dim3(kBlock), 0, as_stream(stream), dense, coef, num_steps, amp, dT_M, n, max_step, N,
                       hplus, nsamp, nullptr);
    return check_launch("resample_hplus");
}

int eobhip_resample_modes(const double* dense, const double* coef, const int* num_steps, const double* amp, const double* dT_M, int n,
                          int max_step, int N, int nm, const int* ms, double* hplus, double* hcross, int* nsamp, void* stream) {
    EOBHIP_REQUIRE(n >= 0 && N >= 1 && n <= 65535, "bad sizes");
    EOBHIP_REQUIRE(nm >= 1 && nm <= 15, "nm must be 1 .. 15");
    EOBHIP_REQUIRE(hplus && hcross, "hplus and hcross are required");
    ModeMs mm;
    mm.nm = nm;
    for (int q = 0; q < nm; ++q) {
        EOBHIP_REQUIRE(ms[q] >= 1 && ms[q] <= kMaxModeM, "m must be 1 .. 8");
        mm.m[q] = ms[q];
    }
    if (n == 0) return EOBHIP_OK;
    int bx = (int)((N + (long long)kResSpan - 1) / kResSpan);
    if (bx > 4096) bx = 4096;
    // built for the reference's five modes and for (2,2) alone; any other count runs the general build
    if (nm == 5)
        hipLaunchKernelGGL(k_resample_modes<5>, dim3(bx, n), dim3(kBlock), 0, as_stream(stream), dense, coef, num_steps, amp, dT_M,
                           max_step, N, mm, hplus, hcross, nsamp);
    else if (nm == 1)
        hipLaunchKernelGGL(k_resample_modes<1>, dim3(bx, n), dim3(kBlock), 0, as_stream(stream), dense, coef, num_steps, amp, dT_M,
                           max_step, N, mm, hplus, hcross, nsamp);
    else
        hipLaunchKernelGGL(k_resample_modes<0>, dim3(bx, n), dim3(kBlock), 0, as_stream(stream), dense, coef, num_steps, amp, dT_M,
                           max_step, N, mm, hplus, hcross, nsamp);
    return check_launch("resample_modes");
}

int eobhip_TDInterp2(void* templateChannels, const double* dataTime, const double* ts, const double* propArrays, const double* c1,
                     const double* c2, const double* c3, int old_length, const int* old_lengths, int data_length, int numBinAll,
                     int numModes, const int* ls, const int* ms, const int* inds, const int* lengths, int max_length, int legacy_stride,
                     void* stream) {
    (void)ls;   // the reference reads ls into shared memory but combines with m only
    EOBHIP_REQUIRE(numBinAll >= 0 && numModes >= 1 && old_length >= 1 && max_length >= 0 && data_length >= 0, "bad sizes");
    EOBHIP_REQUIRE(numBinAll <= 65535, "numBinAll must be <= 65535 (grid y dimension)");
    if (numBinAll == 0 || max_length == 0) return EOBHIP_OK;
    int bx = (max_length + kBlock - 1) / kBlock;
    if (bx > 4096) bx = 4096;
    hipLaunchKernelGGL(k_td_interp2, dim3(bx, numBinAll), dim3(kBlock), 0, as_stream(stream), static_cast<double2*>(templateChannels),
                       dataTime, ts, propArrays, c1, c2, c3, old_length, old_lengths, data_length, numModes, ms, inds, lengths, max_length,
                       legacy_stride);
    return check_launch("TDInterp2");
}

int eobhip_resample_hpc(const double* dense, const double* coef, const int* num_steps, const double* amp, const double* dT_M, int n,
                        int max_step, int N, double* hplus, double* hcross, int* nsamp, void* stream) {
    EOBHIP_REQUIRE(n >= 0 && N >= 1 && n <= 65535, "bad sizes");
    EOBHIP_REQUIRE(hcross != nullptr, "hcross is required");
    if (n == 0) return EOBHIP_OK;
    int bx = (int)((N + (long long)kResSpan - 1) / kResSpan);
    if (bx > 4096) bx = 4096;
    hipLaunchKernelGGL(k_resample_hplus, dim3(bx, n), dim3(kBlock), 0, as_stream(stream), dense, coef, num_steps, amp, dT_M, n, max_step, N,
                       hplus, nsamp, hcross);
    return check_launch("resample_hpc");
}

// Batched real-to-complex FFT of n rows of N samples into n rows of N/2+1 bins (no dt factor,
// as the reference's rfft, eob/seobnrv4phm.py:742-744).
int eobhip_rfft_batched(const double* in, double* out, int n, int N, void* stream) {
    EOBHIP_REQUIRE(n >= 1 && N >= 2 && (N % 2) == 0, "bad sizes (n >= 1, N even)");
    return fft_rows(0, in, out, n, N, stream);
}

int eobhip_fd_likelihood(const double* htilde, const double* dtilde, double* partial, int n, int F, double df, double f0,
                         double* logl, void* stream) {
    EOBHIP_REQUIRE(n >= 1 && F >= 1, "bad sizes");
    const int per = kLikeBlock * kLikePerThread;
    const int nchunks = (F + per - 1) / per;
    const long long grid = (long long)nchunks * n;
    EOBHIP_REQUIRE(grid < (1LL << 31), "grid too large");
    hipLaunchKernelGGL(k_fd_likelihood_partial, dim3((unsigned)grid), dim3(kLikeBlock), 0, as_stream(stream),
                       reinterpret_cast<const double2*>(htilde), reinterpret_cast<const double2*>(dtilde), n, F, df, f0, partial, nchunks);
    int rc = check_launch("fd_likelihood_partial");
    if (rc) return rc;
    hipLaunchKernelGGL(k_fd_likelihood_final, dim3(n), dim3(64), 0, as_stream(stream), partial, n, nchunks, df, logl);
    return check_launch("fd_likelihood_final");
}

size_t eobhip_fd_likelihood_partial_doubles(int n, int F) {
    const int per = kLikeBlock * kLikePerThread;
    return (size_t)n * (size_t)((F + per - 1) / per);
}

}  // extern "C"

// ---------------------------------------------- likelihood without allocation
// The partial sums of eobhip_fd_likelihood_static live in file-scope device
// memory: one copy per GPU, held until the process ends.  Host code reaches a
// GPU's copy only through the address hipGetSymbolAddress returns with that GPU
// current (the array's name is a host-side stand-in, not a device address).
namespace fdlik_static {
constexpr size_t kScratchDoubles = size_t(1) << 19;   // 4 MiB
static __device__ __align__(256) double scratch[kScratchDoubles];
constexpr int kMaxDevices = 64;
static std::mutex g_lock;
static double* g_ptr[kMaxDevices] = {};

// this GPU's copy of `scratch`, looked up on the first call on the device
static int scratch_ptr(double** out) {
    int dev = -1;
    int rc = hip_status(hipGetDevice(&dev), "hipGetDevice");
    if (rc) return rc;
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("fd_likelihood_static: device %d outside the %d devices the library tracks", dev, kMaxDevices);
        return EOBHIP_EINVAL;
    }
    std::lock_guard<std::mutex> g(g_lock);
    if (!g_ptr[dev]) {
        void* p = nullptr;
        rc = hip_status(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        if (rc) return rc;
        g_ptr[dev] = static_cast<double*>(p);
    }
    *out = g_ptr[dev];
    return EOBHIP_OK;
}
}  // namespace fdlik_static

extern "C" {

size_t eobhip_fd_likelihood_static_max_partials(void) { return fdlik_static::kScratchDoubles; }

int eobhip_fd_likelihood_static(const double* htilde, const double* dtilde, int n, int F, double df, double f0, double* logl,
                                void* stream) {
    EOBHIP_REQUIRE(n >= 1 && F >= 1, "bad sizes");
    const size_t need = eobhip_fd_likelihood_partial_doubles(n, F);
    EOBHIP_REQUIRE(need <= fdlik_static::kScratchDoubles,
                   "fd_likelihood_static: %zu partial sums (n = %d, F = %d) exceed the library's static buffer of %zu; "
                   "use eobhip_fd_likelihood with a caller-provided buffer",
                   need, n, F, fdlik_static::kScratchDoubles);
    double* partial = nullptr;
    const int rc = fdlik_static::scratch_ptr(&partial);
    if (rc) return rc;
    // every (binary, chunk) partial is written before it is read: nothing to clear
    return eobhip_fd_likelihood(htilde, dtilde, partial, n, F, df, f0, logl, stream);
}

}  // extern "C"
