// Knot-interval lookup for a block of uniformly spaced samples.
//
// A block resamples consecutive samples t_j of one binary; their knot
// intervals (searchsorted(t, t_j, 'right') - 1, eob/seobnrv4phm.py:241-272)
// span a short range of the binary's knots.  knot_window finds that range with
// one load per thread (a coarse probe of every S-th knot, S = ceil(len /
// BLOCK)) and stages the range's knot times in LDS, so that every sample
// searches LDS instead of making a chain of dependent global loads.
#pragma once
#include <hip/hip_runtime.h>

namespace eobhip_detail {

constexpr int kWinMax = 1024;   // knots a window holds

template <int BLOCK>
struct KnotWindow {
    double t[kWinMax];
    double probe[BLOCK];
    int base, cnt;   // the window holds knots base .. base + cnt - 1; cnt == 0: too wide
    int pa, pb;      // the last probes <= ta and <= tb (-1: none)
};

// T: knot times at stride `stride`, sorted, len >= 1; ta <= tb: the first and
// last sample time of the block.  Every thread of the block calls it (it
// synchronises the block).  All knots before `base` are <= ta and all knots
// after the window are > tb, so for ta <= t <= tb the interval is
// base + (entries of the window <= t) - 1.
template <int BLOCK>
__device__ void knot_window(KnotWindow<BLOCK>& w, const double* __restrict__ T, int stride, int len, double ta, double tb) {
    const int S = (len + BLOCK - 1) / BLOCK;
    const int tid = threadIdx.x;
    double pt;
    {
        long long i = (long long)tid * S;
        if (i > len - 1) i = len - 1;
        pt = T[(size_t)stride * (size_t)i];
        w.probe[tid] = pt;
    }
    if (tid == 0) { w.pa = -1; w.pb = -1; }
    __syncthreads();
    // The probes do not decrease, so for each of ta, tb at most one thread
    // holds the last probe <= it: the one whose successor (+inf past the end)
    // is greater.  No thread matches when the first probe is greater (or the
    // bound is NaN), which leaves -1.
    {
        const double pn = tid + 1 < BLOCK ? w.probe[tid + 1] : HUGE_VAL;
        if (pt <= ta && !(pn <= ta)) w.pa = tid;
        if (pt <= tb && !(pn <= tb)) w.pb = tid;
    }
    __syncthreads();
    const int pa = w.pa, pb = w.pb;
    long long base = (long long)(pa < 0 ? 0 : pa) * S;
    if (base > len - 1) base = len - 1;
    long long last = (long long)(pb + 1) * S;
    if (last > len - 1) last = len - 1;
    const int cnt = (last - base + 1 <= kWinMax) ? (int)(last - base + 1) : 0;
    if (tid == 0) { w.base = (int)base; w.cnt = cnt; }
    for (int i = tid; i < cnt; i += BLOCK) w.t[i] = T[(size_t)stride * (size_t)(base + i)];
    __syncthreads();
}

// Number of entries <= x of the sorted array a[0 .. n-1], n >= 1, for K
// values at once.  The bisection keeps pos + m <= n and reads a[pos + half]
// with half < m, so every read is in bounds whatever x is (NaN included).
// The trip count, ceil(log2 n), depends on n only: with n the same on every
// lane the loop is uniform (no divergence) and the K reads of an iteration
// are independent, so their latencies overlap.
template <int K, class Get>
__device__ __forceinline__ void count_le(const Get& a, int n, const double (&x)[K], int (&cnt)[K]) {
    int pos[K];
#pragma unroll
    for (int k = 0; k < K; ++k) pos[k] = 0;
    for (int m = n; m > 1;) {
        const int half = m >> 1;
#pragma unroll
        for (int k = 0; k < K; ++k) pos[k] = (a(pos[k] + half) <= x[k]) ? pos[k] + half : pos[k];
        m -= half;
    }
#pragma unroll
    for (int k = 0; k < K; ++k) cnt[k] = pos[k] + (a(pos[k]) <= x[k] ? 1 : 0);
}

// searchsorted(t, tj, 'right') - 1, clamped at 0, for K samples of the
// block: from the window when it holds the block's range (LDS reads only),
// else by bisection over the whole knot array.  w.cnt is the same on every
// thread of the block, so the branch is uniform.
template <int BLOCK, int K>
__device__ __forceinline__ void knot_intervals(const KnotWindow<BLOCK>& w, const double* __restrict__ T, int stride, int len,
                                               const double (&tj)[K], int (&ind)[K]) {
    int c[K];
    int base;
    if (w.cnt > 0) {
        count_le<K>([&](int i) { return w.t[i]; }, w.cnt, tj, c);
        base = w.base;
    } else {
        count_le<K>([&](int i) { return T[(size_t)stride * (size_t)i]; }, len, tj, c);
        base = 0;
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int i = base + c[k] - 1;
        ind[k] = i < 0 ? 0 : i;
    }
}

// one sample's interval (knot_intervals for K = 1)
template <int BLOCK>
__device__ __forceinline__ int knot_interval(const KnotWindow<BLOCK>& w, const double* __restrict__ T, int stride, int len, double tj) {
    const double x[1] = {tj};
    int ind[1];
    knot_intervals<BLOCK, 1>(w, T, stride, len, x, ind);
    return ind[0];
}

// Knot time and the NC spline coefficients of interval ind (rows of T at
// stride `stride`, of C at stride NC).  When every active lane of the
// wavefront has the same interval (consecutive samples inside one knot
// interval, the common case) the row is read once for the wavefront through
// the scalar cache (constant address space: both tables are read-only while
// the kernel runs), instead of once per lane through the vector memory path.
template <int NC>
__device__ __forceinline__ void interval_row(const double* __restrict__ T, int stride, const double* __restrict__ C, int ind, double& tk,
                                             double (&c)[NC]) {
    typedef const __attribute__((address_space(4))) double* ConstPtr;
    const int i0 = __builtin_amdgcn_readfirstlane(ind);
    if (__all(ind == i0)) {
        const ConstPtr r = (ConstPtr)(C + (size_t)NC * i0);
#pragma unroll
        for (int q = 0; q < NC; ++q) c[q] = r[q];
        tk = ((ConstPtr)T)[(size_t)stride * i0];
    } else {
        const double* r = C + (size_t)NC * ind;
#pragma unroll
        for (int q = 0; q < NC; ++q) c[q] = r[q];
        tk = T[(size_t)stride * ind];
    }
}

}  // namespace eobhip_detail
