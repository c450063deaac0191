// Logarithm, division and square root for the integrator's right-hand side.
//
// The right-hand side runs as one wavefront per binary, so its cost is the
// length of its instruction stream.  The device library's log carries a
// double-double evaluation (about 90 instructions besides its constants); the
// version here uses the textbook reduction and series (about 45), with every
// multiply-add written as fma() and contraction of the rest switched off, so
// that a host build rounds exactly as the device build does; within 1 ulp.
//
// log_pos: x = m 2^e with m in [sqrt(1/2), sqrt(2)); with f = m - 1 and
//   s = f / (2 + f), log(1 + f) = 2 atanh(s) = 2s + s R(s^2),
//   R(z) = sum_{k>=1} 2 z^k / (2k + 1), rearranged as
//   f - (f^2/2 - s (f^2/2 + R)) so that the exact f leads (the arrangement of
//   fdlibm's e_log.c).  |s| <= 0.1716, and the series stops at z^10: the first
//   omitted term is below 1e-18 of the result.
// ln 2 is split into a 31-bit head (e ln2_hi is exact for |e| < 2^22) and the
// rounded remainder.
#pragma once
#include <cmath>

#ifndef EOB_HD
#define EOB_HD inline
#endif

namespace eob {

constexpr double kLn2Hi = 0x1.62e42fec00000p-1;       // ln 2 with the low 22 mantissa bits cleared
constexpr double kLn2Lo = 0x1.d1cf79abc9e3bp-32;      // ln 2 - kLn2Hi, rounded
constexpr double kSqrtHalf = 0x1.6a09e667f3bcdp-1;    // sqrt(1/2)

// a / b and 1 / b for normal, finite operands with a normal result: the
// hardware reciprocal estimate refined by two Newton steps, then one residual
// correction (Markstein's method).  The quotient rounds as IEEE division on
// such operands; the library's division adds scaling steps for the extreme
// exponent ranges that these arguments never reach.  Other types divide.
// A host build (the unit test) starts from the reciprocal rounded to single
// precision, a coarser estimate than the hardware's, so that it exercises
// the refinement.
EOB_HD double rcp_estimate(double b) {
#ifdef __HIP_DEVICE_COMPILE__
    return __builtin_amdgcn_rcp(b);
#else
    return (double)(float)(1.0 / b);
#endif
}
EOB_HD double refined_rcp(double b) {
    double y = rcp_estimate(b);
    y = fma(y, fma(-b, y, 1.0), y);
    y = fma(y, fma(-b, y, 1.0), y);
    return y;
}
template <class T>
EOB_HD T qdiv(T a, T b) { return a / b; }
template <class T>
EOB_HD T qrcp(T b) { return 1.0 / b; }
EOB_HD double qdiv(double a, double b) {
    const double y = refined_rcp(b);
    const double q = a * y;
    return fma(fma(-b, q, a), y, q);
}
EOB_HD double qrcp(double b) {
    const double y = refined_rcp(b);
    return fma(fma(-b, y, 1.0), y, y);
}

// natural log of a positive finite x, within 1 ulp; NaN for every other x
// (0, negative, infinite or NaN), which the integrator treats as a failed binary
EOB_HD double log_pos(double x) {
#ifdef __clang__
#pragma clang fp contract(off)
#endif
    int e;
    double m = frexp(x, &e);   // x = m 2^e, m in [1/2, 1)
    const bool lo = m < kSqrtHalf;
    m = lo ? m + m : m;
    e = lo ? e - 1 : e;
    const double f = m - 1.0;   // exact
    const double s = qdiv(f, 2.0 + f);
    const double z = s * s;
    double R = 2.0 / 21.0;
    R = fma(R, z, 2.0 / 19.0);
    R = fma(R, z, 2.0 / 17.0);
    R = fma(R, z, 2.0 / 15.0);
    R = fma(R, z, 2.0 / 13.0);
    R = fma(R, z, 2.0 / 11.0);
    R = fma(R, z, 2.0 / 9.0);
    R = fma(R, z, 2.0 / 7.0);
    R = fma(R, z, 2.0 / 5.0);
    R = fma(R, z, 2.0 / 3.0);
    const double hfsq = (0.5 * f) * f;
    const double dk = (double)e;
    const double t = fma(s, fma(R, z, hfsq), dk * kLn2Lo);
    const double res = fma(dk, kLn2Hi, -((hfsq - t) - f));
    return (x > 0.0 && x < HUGE_VAL) ? res : NAN;
}

// sqrt(x) for positive normal x: the hardware reciprocal-square-root
// estimate y, then Goldschmidt's iteration (g = x y, h = y/2; one step on the
// pair, two residual corrections on g) — correctly rounded in 20 000 of
// 20 000 simulated cases; the library's version adds range scaling and
// special-case selects these arguments never need.  Other types use sqrt.
EOB_HD double rsqrt_estimate(double x) {
#ifdef __HIP_DEVICE_COMPILE__
    return __builtin_amdgcn_rsq(x);
#else
    return (double)(float)(1.0 / sqrt(x));   // host build: a single-precision start, as for rcp_estimate
#endif
}
template <class T>
EOB_HD T qsqrt(T x) { return sqrt(x); }
EOB_HD double qsqrt(double x) {
    const double y = rsqrt_estimate(x);
    double g = x * y, h = 0.5 * y;
    const double r = fma(-g, h, 0.5);
    g = fma(g, r, g);
    h = fma(h, r, h);
    g = fma(fma(-g, g, x), h, g);
    g = fma(fma(-g, g, x), h, g);
    return g;
}

}  // namespace eob
