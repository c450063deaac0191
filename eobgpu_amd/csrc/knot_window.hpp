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
    {
        long long i = (long long)tid * S;
        if (i > len - 1) i = len - 1;
        w.probe[tid] = T[(size_t)stride * (size_t)i];
    }
    __syncthreads();
    if (tid == 0) {
        int p[2];
        for (int e = 0; e < 2; ++e) {
            const double tt = e == 0 ? ta : tb;
            int lo = 0, hi = BLOCK;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (w.probe[mid] <= tt) lo = mid + 1;
                else hi = mid;
            }
            p[e] = lo - 1;   // last probe <= tt, -1 if none
        }
        long long base = (long long)(p[0] < 0 ? 0 : p[0]) * S;
        if (base > len - 1) base = len - 1;
        long long last = (long long)(p[1] + 1) * S;
        if (last > len - 1) last = len - 1;
        const long long cnt = last - base + 1;
        w.base = (int)base;
        w.cnt = (cnt <= kWinMax) ? (int)cnt : 0;
    }
    __syncthreads();
    const int base = w.base, cnt = w.cnt;
    for (int i = tid; i < cnt; i += BLOCK) w.t[i] = T[(size_t)stride * (size_t)(base + i)];
    __syncthreads();
}

// searchsorted(t, tj, 'right') - 1, clamped at 0: from the window when it
// holds the block's range, else by bisection over the whole knot array
template <int BLOCK>
__device__ __forceinline__ int knot_interval(const KnotWindow<BLOCK>& w, const double* __restrict__ T, int stride, int len, double tj) {
    int lo, hi;
    const bool local = w.cnt > 0;
    if (local) { lo = 0; hi = w.cnt; }
    else { lo = 0; hi = len; }
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        const double tm = local ? w.t[mid] : T[(size_t)stride * mid];
        if (tm <= tj) lo = mid + 1;
        else hi = mid;
    }
    int ind = (local ? w.base : 0) + lo - 1;
    return ind < 0 ? 0 : ind;
}

}  // namespace eobhip_detail
