// One row of the not-a-knot cubic-spline system in the knot slopes, shared by
// the batched spline builders and the attachment's short-window splines.
#pragma once
#include <hip/hip_runtime.h>

namespace eobhip_detail {

// ----------------------------------------------------------------- tridiagonal
// Row i of the not-a-knot system in the slopes s (src/interpolate.cu:11-120):
//   interior: dx2 s_{i-1} + 2 (dx1 + dx2) s_i + dx1 s_{i+1} = 3 (dx2 m1 + dx1 m2)
//   first/last rows: the not-a-knot end conditions.
struct Row {
    double b, ud, dg, ld;
};

__host__ __device__ __forceinline__ Row nak_row(int i, int len, double x0, double x1, double x2, double y0, double y1, double y2) {
    const double dx1 = x1 - x0, dx2 = x2 - x1, d = x2 - x0;
    const double s1 = (y1 - y0) / dx1, s2 = (y2 - y1) / dx2;
    Row r;
    if (i == len - 1) {
        r.b = (dx2 * dx2 * s1 + (2 * d + dx2) * dx1 * s2) / d;
        r.dg = dx1; r.ld = d; r.ud = 0.0;
    } else if (i == 0) {
        r.b = ((dx1 + 2 * d) * dx2 * s1 + dx1 * dx1 * s2) / d;
        r.ud = d; r.ld = 0.0; r.dg = dx2;
    } else {
        r.b = 3.0 * (dx2 * s1 + dx1 * s2);
        r.dg = 2 * (dx1 + dx2); r.ud = dx1; r.ld = dx2;
    }
    return r;
}

}  // namespace eobhip_detail
