// Likelihood of many templates against one data stream without N-point
// transforms of the templates.
//
// The reference (all_in_one_likelihood, src/interpolate.cu:553-628; get_ll,
// eob/seobnrv4phm.py:772-823) takes h~ = rfft(h) of every N-sample template
// and sums |d~ - h~|^2 w over the N/2+1 bins, w = 1/S.  With
//   DD    = sum_k |d~_k|^2 w_k                                (once per data set)
//   cross = Re sum_k conj(d~_k) w_k h~_k = sum_j h_j g_j,
//           g_j = Re sum_k d~_k w_k e^{2 pi i jk/N}           (once per data set)
//   hh    = sum_k w_k |h~_k|^2 = (1/M) sum_q |X_q|^2 C^(M)_q,
//           X = DFT_M(h on its support), M >= 2 * support     (C^(M) once per M)
// the same number is -2 df (DD - 2 cross + hh).  A template of L non-zero
// samples then costs one real FFT of M = max(2^m_min, 2^ceil(log2 2L)) <= N
// points, and neither h nor h~ is written at full length.  hh is exact (no
// truncation): the autocorrelation of h has no lags beyond L, so its circular
// form on M >= 2L points equals the linear one, and C^(M) is the N-grid
// weight's lag kernel c_m = sum_k w~_k e^{2 pi i mk/N} on lags |m| < M/2.
// For M = N the formula reduces to the direct sum.
#include <hip/hip_runtime.h>

#include "eobhip_common.hpp"
#include "knot_window.hpp"

using namespace eobhip_detail;

namespace {

constexpr int kBlock = 256;
constexpr int kPerThread = 16;                 // resample: samples per thread per block
constexpr int kSpan = kBlock * kPerThread;     // samples per block
constexpr int kSearch = 4;                     // resample: samples per thread whose knot searches run together
static_assert(kPerThread % kSearch == 0, "kSearch divides kPerThread");
constexpr int kBinsPerThread = 8;              // hh: bins per thread per block
constexpr int kBinSpan = kBlock * kBinsPerThread;

// 1/S(f_k) for S(f) = (f/f0)^-4 + 1 + (f/f0)^2, zero at k = 0 (the same
// expression as the direct likelihood kernel).
__device__ __forceinline__ double psd_weight(long long k, double df, double f0) {
    if (k == 0) return 0.0;
    const double x = df * (double)k / f0;
    const double x2 = x * x, x4 = x2 * x2;
    return x4 / (1.0 + x4 + x4 * x2);
}

__device__ __forceinline__ double block_sum(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    __shared__ double red[kBlock / 64];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < kBlock / 64; ++w) s += red[w];
    return s;   // valid on thread 0
}

// Offset of the weight vector for M = 2^k in the bank: sum_{j=m_min}^{k-1} (2^(j-1) + 1).
__host__ __device__ __forceinline__ long long bank_offset(int k, int m_min) {
    return ((1LL << (k - 1)) - (1LL << (m_min - 1))) + (k - m_min);
}

// ---------------------------------------------------------------- data set
// Z_k = d~_k w_k (DC and Nyquist: 2 Re z, so that Z2D gives 2 g); V_k = the
// two-sided weight folded onto k <= N/2 (w/2 inside, w at the ends) for the lag
// kernel c = Z2D(V); per-block partial sums of |d~|^2 w.
__global__ __launch_bounds__(kBlock) void k_lik_prep(const double2* __restrict__ d, int F, double df, double f0, double2* Z, double2* V,
                                                     double* dd_partial) {
    const long long k = (long long)blockIdx.x * kBlock + threadIdx.x;
    double acc = 0.0;
    if (k < F) {
        const double w = psd_weight(k, df, f0);
        const double2 dv = d[k];
        double2 z = make_double2(dv.x * w, dv.y * w);
        const bool end = (k == 0) || (k == F - 1);
        if (end) z = make_double2(2.0 * z.x, 0.0);
        Z[k] = z;
        V[k] = make_double2(end ? w : 0.5 * w, 0.0);
        acc = (dv.x * dv.x + dv.y * dv.y) * w;
    }
    const double s = block_sum(acc);
    if (threadIdx.x == 0) dd_partial[blockIdx.x] = s;
}

__global__ __launch_bounds__(kBlock) void k_sum_into(const double* __restrict__ part, int count, double* out) {
    double acc = 0.0;
    for (int i = threadIdx.x; i < count; i += kBlock) acc += part[i];
    const double s = block_sum(acc);
    if (threadIdx.x == 0) *out = s;
}

__global__ __launch_bounds__(kBlock) void k_scale(double* x, long long n, double a) {
    const long long i = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) x[i] *= a;
}

// lags m in [-M/2, M/2) of c onto an M-point circle
__global__ __launch_bounds__(kBlock) void k_wrap(const double* __restrict__ c, long long N, long long M, double* cM) {
    const long long m = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (m >= M) return;
    cM[m] = (m < M / 2) ? c[m] : c[N - M + m];
}

// one-sided weights of the M-grid sum: 2 Re C_q / M inside, Re C_q / M at q = 0, M/2
__global__ __launch_bounds__(kBlock) void k_bank(const double2* __restrict__ C, long long M, double* W) {
    const long long q = (long long)blockIdx.x * kBlock + threadIdx.x;
    const long long H = M / 2;
    if (q > H) return;
    const double s = (q == 0 || q == H) ? 1.0 / (double)M : 2.0 / (double)M;
    W[q] = C[q].x * s;
}

// ---------------------------------------------------------------- templates
// support length of each template on the uniform grid (the same count as the
// full-length resample) and the FFT size exponent for its hh term
__global__ __launch_bounds__(kBlock) void k_lik_sizes(const double* __restrict__ dense, const int* __restrict__ num_steps,
                                                      const double* __restrict__ dT_M, int n, int max_step, int N, int log2N, int m_min,
                                                      int* nsamp, int* mlog2) {
    const int b = blockIdx.x * kBlock + threadIdx.x;
    if (b >= n) return;
    const int len = num_steps[b];
    long long nb = 0;
    if (len >= 2) {
        const double* T = dense + (size_t)b * max_step * 5;
        const double t0 = T[0], tl = T[5 * (len - 1)];
        nb = (long long)((tl - t0) / dT_M[b]) + 1;
        if (nb > N) nb = N;
    }
    int k = m_min;
    while (k < log2N && (1LL << k) < 2 * nb) ++k;
    nsamp[b] = (int)nb;
    mlog2[b] = k;
}

// h+ = amp Re[S(t) e^{-2 i phi(t)}] from the spline coefficients (the uniform
// resample of the full-length path) on j < M_b, zero for nsamp <= j < M_b;
// per-block partial sums of h_j g_j.  grid (gx, n); a block covers kSpan
// consecutive samples, stages the knot times of their intervals in LDS
// (knot_window) and each thread searches there.
__global__ __launch_bounds__(kBlock) void k_lik_resample(const double* __restrict__ dense, const double* __restrict__ coef,
                                                         const int* __restrict__ num_steps, const double* __restrict__ amp,
                                                         const double* __restrict__ dT_M, const int* __restrict__ nsamp,
                                                         const int* __restrict__ mlog2, const long long* __restrict__ rowoff, int max_step,
                                                         const double* __restrict__ g, double* __restrict__ hbuf,
                                                         double* __restrict__ cross_partial) {
    const int b = blockIdx.y;
    const long long M = 1LL << mlog2[b];
    const long long j0 = (long long)blockIdx.x * kSpan;
    double* out = hbuf + rowoff[b];
    const long long nb = nsamp[b];
    double acc = 0.0;
    __shared__ KnotWindow<kBlock> win;
    if (j0 < M) {
        const int len = num_steps[b];
        const double* T = dense + (size_t)b * max_step * 5;
        const double* C = coef + (size_t)b * max_step * 12;
        const double dT = dT_M[b];
        const double t0 = T[0];
        if (j0 < nb) {
            const long long jmax = (j0 + kSpan < nb ? j0 + kSpan : nb) - 1;   // last sample with a value in this block
            knot_window<kBlock>(win, T, 5, len, t0 + (double)j0 * dT, t0 + (double)jmax * dT);
        }
        const double a = amp[b];
        for (int u0 = 0; u0 < kPerThread; u0 += kSearch) {
            double tj[kSearch];
            int ind[kSearch];
#pragma unroll
            for (int k = 0; k < kSearch; ++k) tj[k] = t0 + (double)(j0 + (long long)(u0 + k) * kBlock + threadIdx.x) * dT;
            if (j0 < nb) knot_intervals<kBlock, kSearch>(win, T, 5, len, tj, ind);
#pragma unroll
            for (int k = 0; k < kSearch; ++k) {
                const long long j = j0 + (long long)(u0 + k) * kBlock + threadIdx.x;
                if (j < M) {
                    double val = 0.0;
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
                        val = a * (sr * cp + si * sp);
                        acc += val * g[j];
                    }
                    out[j] = val;
                }
            }
        }
    }
    // every thread of the block reaches the reduction's barrier (the window's
    // barriers are taken by all threads or by none: j0 and nb are per block)
    const double s = block_sum(acc);
    if (threadIdx.x == 0) cross_partial[(size_t)b * gridDim.x + blockIdx.x] = s;
}

// per-block partial sums of |X_q|^2 W^(M)_q, q <= M/2; grid (gx2, n)
__global__ __launch_bounds__(kBlock) void k_lik_hh(const double2* __restrict__ X, const long long* __restrict__ xoff,
                                                   const int* __restrict__ mlog2, int m_min, const double* __restrict__ bank,
                                                   double* __restrict__ hh_partial) {
    const int b = blockIdx.y;
    const int k = mlog2[b];
    const long long H = (1LL << k) / 2;
    const double2* x = X + xoff[b];
    const double* W = bank + bank_offset(k, m_min);
    const long long q0 = (long long)blockIdx.x * kBinSpan;
    double acc = 0.0;
    if (q0 <= H) {
#pragma unroll
        for (int u = 0; u < kBinsPerThread; ++u) {
            const long long q = q0 + (long long)u * kBlock + threadIdx.x;
            if (q <= H) {
                const double2 v = x[q];
                acc += (v.x * v.x + v.y * v.y) * W[q];
            }
        }
    }
    const double s = block_sum(acc);
    if (threadIdx.x == 0) hh_partial[(size_t)b * gridDim.x + blockIdx.x] = s;
}

__global__ __launch_bounds__(64) void k_lik_final(const double* __restrict__ cross_partial, int gx, const double* __restrict__ hh_partial,
                                                  int gx2, const double* __restrict__ dd, double df, int n, double* __restrict__ logl,
                                                  double* __restrict__ terms) {
    const int b = blockIdx.x;
    if (b >= n) return;
    double c = 0.0, h = 0.0;
    for (int i = threadIdx.x; i < gx; i += 64) c += cross_partial[(size_t)b * gx + i];
    for (int i = threadIdx.x; i < gx2; i += 64) h += hh_partial[(size_t)b * gx2 + i];
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        c += __shfl_xor(c, off, 64);
        h += __shfl_xor(h, off, 64);
    }
    if (threadIdx.x == 0) {
        logl[b] = -2.0 * df * (dd[0] - 2.0 * c + h);
        if (terms) {
            terms[2 * b] = c;
            terms[2 * b + 1] = h;
        }
    }
}

int log2_exact(long long N) {
    int k = 0;
    while ((1LL << k) < N) ++k;
    return ((1LL << k) == N) ? k : -1;
}

}  // namespace

namespace {

// Downstream gather: binary idx[j] of a batch (rows below its num_steps,
// tables, amplitude, caller position) into slot j of a downstream buffer, with
// dT_M[j] = dt / (M[idx[j]] MTSUN), rounded as torch evaluates dt / (M * MTSUN)
// on the device (the product, its reciprocal, times dt).  One block per binary.
__global__ __launch_bounds__(kBlock) void k_gather_binaries(const long long* __restrict__ idx, int k, int n_src, int max_step,
                                                            const double* __restrict__ dense, const int* __restrict__ num_steps,
                                                            const double* __restrict__ bd, int bd_words, const double* __restrict__ md,
                                                            int md_words, const double* __restrict__ M, const double* __restrict__ amp,
                                                            const long long* __restrict__ pos, double dt, double mtsun,
                                                            double* __restrict__ dense_o, int* __restrict__ ns_o, double* __restrict__ bd_o,
                                                            double* __restrict__ md_o, double* __restrict__ dTM_o, double* __restrict__ amp_o,
                                                            long long* __restrict__ pos_o) {
    const int j = blockIdx.x;
    if (j >= k) return;
    const long long b = idx[j];
    if (b < 0 || b >= n_src) return;
    const int len = num_steps[b];
    const int rows = len < 0 ? 0 : (len > max_step ? max_step : len);
    const double* src = dense + (size_t)b * max_step * 5;
    double* dst = dense_o + (size_t)j * max_step * 5;
    for (int i = threadIdx.x; i < rows * 5; i += blockDim.x) dst[i] = src[i];
    for (int i = threadIdx.x; i < bd_words; i += blockDim.x) bd_o[(size_t)j * bd_words + i] = bd[(size_t)b * bd_words + i];
    for (int i = threadIdx.x; i < md_words; i += blockDim.x) md_o[(size_t)j * md_words + i] = md[(size_t)b * md_words + i];
    if (threadIdx.x == 0) {
        ns_o[j] = len;
        const double mm = M[b] * mtsun;
        dTM_o[j] = (1.0 / mm) * dt;
        amp_o[j] = amp[b];
        pos_o[j] = pos[b];
    }
}

}  // namespace

extern "C" {

int eobhip_gather_binaries(const long long* idx, int k, int n_src, int max_step, const double* dense, const int* num_steps,
                           const void* bindata, size_t bindata_bytes, const void* modes, size_t modes_bytes, const double* total_mass,
                           const double* amp, const long long* pos, double dt, double mtsun, double* dense_out, int* num_steps_out,
                           void* bindata_out, void* modes_out, double* dT_M_out, double* amp_out, long long* pos_out, void* stream) {
    EOBHIP_REQUIRE(k >= 0 && n_src >= 0 && max_step >= 1, "bad sizes");
    EOBHIP_REQUIRE(bindata_bytes % 8 == 0 && modes_bytes % 8 == 0, "table sizes must be multiples of 8 bytes");
    EOBHIP_REQUIRE(bindata_bytes / 8 <= 0x7fffffff && modes_bytes / 8 <= 0x7fffffff, "tables too large");
    if (k == 0) return EOBHIP_OK;
    hipLaunchKernelGGL(k_gather_binaries, dim3(k), dim3(kBlock), 0, as_stream(stream), idx, k, n_src, max_step, dense, num_steps,
                       static_cast<const double*>(bindata), (int)(bindata_bytes / 8), static_cast<const double*>(modes),
                       (int)(modes_bytes / 8), total_mass, amp, pos, dt, mtsun, dense_out, num_steps_out, static_cast<double*>(bindata_out),
                       static_cast<double*>(modes_out), dT_M_out, amp_out, pos_out);
    return check_launch("gather_binaries");
}

size_t eobhip_lik_bank_doubles(int N, int m_min_log2) {
    const int L = log2_exact(N);
    if (L < 0 || m_min_log2 < 1 || m_min_log2 > L) return 0;
    return (size_t)bank_offset(L + 1, m_min_log2);
}

size_t eobhip_lik_setup_scratch_bytes(int N) {
    const size_t F = (size_t)N / 2 + 1;
    const size_t nblk = (F + kBlock - 1) / kBlock;
    // Z, V (complex F each), c (N), cM (N), C (complex F), dd partials
    return 16 * F * 3 + 8 * (size_t)N * 2 + 8 * nblk + 4 * 256;
}

int eobhip_lik_setup(const double* dtilde, int N, double df, double f0, int m_min_log2, double* g, double* dd, double* bank, void* scratch,
                     void* stream) {
    const int L = log2_exact(N);
    EOBHIP_REQUIRE(L >= 2, "N must be a power of two >= 4");
    EOBHIP_REQUIRE(m_min_log2 >= 1 && m_min_log2 <= L, "m_min_log2 must be in [1, log2 N]");
    EOBHIP_REQUIRE(dtilde && g && dd && bank && scratch, "null pointer");
    hipStream_t st = as_stream(stream);
    const long long F = N / 2 + 1;
    const unsigned nblk = blocks_for(F, kBlock);
    char* p = static_cast<char*>(scratch);
    double2* Z = reinterpret_cast<double2*>(p);
    double2* V = Z + F;
    double2* Cq = V + F;
    double* c = reinterpret_cast<double*>(Cq + F);
    double* cM = c + N;
    double* part = cM + N;
    hipLaunchKernelGGL(k_lik_prep, dim3(nblk), dim3(kBlock), 0, st, reinterpret_cast<const double2*>(dtilde), (int)F, df, f0, Z, V, part);
    int rc = check_launch("lik_prep");
    if (rc) return rc;
    hipLaunchKernelGGL(k_sum_into, dim3(1), dim3(kBlock), 0, st, part, (int)nblk, dd);
    if ((rc = check_launch("lik_dd"))) return rc;
    // g = Z2D(Z) / 2, c = Z2D(V)
    if ((rc = fft_exec(1, Z, g, 1, N, stream))) return rc;
    hipLaunchKernelGGL(k_scale, dim3(blocks_for(N, kBlock)), dim3(kBlock), 0, st, g, (long long)N, 0.5);
    if ((rc = check_launch("lik_scale"))) return rc;
    if ((rc = fft_exec(1, V, c, 1, N, stream))) return rc;
    for (int k = m_min_log2; k <= L; ++k) {
        const long long M = 1LL << k;
        hipLaunchKernelGGL(k_wrap, dim3(blocks_for(M, kBlock)), dim3(kBlock), 0, st, c, (long long)N, M, cM);
        if ((rc = check_launch("lik_wrap"))) return rc;
        if ((rc = fft_exec(0, cM, Cq, 1, (int)M, stream))) return rc;
        hipLaunchKernelGGL(k_bank, dim3(blocks_for(M / 2 + 1, kBlock)), dim3(kBlock), 0, st, Cq, M, bank + bank_offset(k, m_min_log2));
        if ((rc = check_launch("lik_bank"))) return rc;
    }
    return EOBHIP_OK;
}

int eobhip_lik_sizes(const double* dense, const int* num_steps, const double* dT_M, int n, int max_step, int N, int m_min_log2,
                     int* nsamp, int* mlog2, void* stream) {
    const int L = log2_exact(N);
    EOBHIP_REQUIRE(L >= 2 && m_min_log2 >= 1 && m_min_log2 <= L, "bad N or m_min_log2");
    EOBHIP_REQUIRE(n >= 0 && max_step >= 1, "bad sizes");
    if (n == 0) return EOBHIP_OK;
    hipLaunchKernelGGL(k_lik_sizes, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, as_stream(stream), dense, num_steps, dT_M, n, max_step, N,
                       L, m_min_log2, nsamp, mlog2);
    return check_launch("lik_sizes");
}

int eobhip_lik_partial_blocks(int max_mlog2, int* gx_resample, int* gx_hh) {
    EOBHIP_REQUIRE(max_mlog2 >= 1 && max_mlog2 < 31, "bad max_mlog2");
    const long long M = 1LL << max_mlog2;
    *gx_resample = (int)((M + kSpan - 1) / kSpan);
    *gx_hh = (int)((M / 2 + 1 + kBinSpan - 1) / kBinSpan);
    return EOBHIP_OK;
}

int eobhip_lik_resample(const double* dense, const double* coef, const int* num_steps, const double* amp, const double* dT_M,
                        const int* nsamp, const int* mlog2, const long long* rowoff, int n, int max_step, int max_mlog2, const double* g,
                        double* hbuf, double* cross_partial, void* stream) {
    EOBHIP_REQUIRE(n >= 0 && n <= 65535 && max_step >= 1, "bad sizes");
    if (n == 0) return EOBHIP_OK;
    int gx, gx2;
    int rc = eobhip_lik_partial_blocks(max_mlog2, &gx, &gx2);
    if (rc) return rc;
    hipLaunchKernelGGL(k_lik_resample, dim3(gx, n), dim3(kBlock), 0, as_stream(stream), dense, coef, num_steps, amp, dT_M, nsamp, mlog2,
                       rowoff, max_step, g, hbuf, cross_partial);
    return check_launch("lik_resample");
}

int eobhip_lik_finish(const double* X, const long long* xoff, const int* mlog2, int n, int max_mlog2, int m_min_log2, const double* bank,
                      const double* cross_partial, double* hh_partial, const double* dd, double df, double* logl, double* terms,
                      void* stream) {
    EOBHIP_REQUIRE(n >= 0 && n <= 65535, "bad sizes");
    if (n == 0) return EOBHIP_OK;
    int gx, gx2;
    int rc = eobhip_lik_partial_blocks(max_mlog2, &gx, &gx2);
    if (rc) return rc;
    hipStream_t st = as_stream(stream);
    hipLaunchKernelGGL(k_lik_hh, dim3(gx2, n), dim3(kBlock), 0, st, reinterpret_cast<const double2*>(X), xoff, mlog2, m_min_log2, bank,
                       hh_partial);
    if ((rc = check_launch("lik_hh"))) return rc;
    hipLaunchKernelGGL(k_lik_final, dim3(n), dim3(64), 0, st, cross_partial, gx, hh_partial, gx2, dd, df, n, logl, terms);
    return check_launch("lik_final");
}

}  // extern "C"
