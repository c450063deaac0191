// Logarithm, division, square and cube roots for the integrator's right-hand side.
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
//   fdlibm's e_log.c).  |s| <= 0.1716, z <= 0.0295.  R is a polynomial of
//   degree 7 obtained by Chebyshev economization (Lanczos) of the series of
//   R(z)/z on [0, 0.0295]: the series to z^40 in exact rationals, rewritten
//   in Chebyshev polynomials of t = 2z/0.0295 - 1, cut after T_6 and rewritten
//   in powers of z; with the coefficients rounded to double, |R - poly| is
//   below 8.4e-18 on the interval (the Taylor series needs degree 10 for
//   that), which moves log by under 0.02 ulp (log_series_coefficients.py
//   derives and checks them).
// ln 2 is split into a 31-bit head (e ln2_hi is exact for |e| < 2^22) and the
// rounded remainder.
#pragma once
#include <cmath>

#ifndef EOB_HD
#ifdef __HIPCC__
#define EOB_HD __host__ __device__ __forceinline__
#else
#define EOB_HD inline   // host-only build (the unit test's probe)
#endif
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
    double R = 0x1.2b6ada4e6cac0p-3;   // R(z)/z to degree 6, economized (see above)
    R = fma(R, z, 0x1.39fcd91739d16p-3);
    R = fma(R, z, 0x1.7462c0aff40b1p-3);
    R = fma(R, z, 0x1.c71c62ba3869bp-3);
    R = fma(R, z, 0x1.2492492e143e6p-2);
    R = fma(R, z, 0x1.999999999519ap-2);
    R = fma(R, z, 0x1.5555555555558p-1);
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

// sqrt(x) and 1/sqrt(x) together: the square root as qsqrt, and 2h (which
// Goldschmidt's iteration carries as 1/sqrt(x) to about 2^-44) refined by one
// Newton step for 1/g, within 1 ulp -- two fused multiply-adds instead of a
// separate reciprocal.
template <class T>
EOB_HD void qsqrt_recip(T x, T& root, T& inv) {
    root = sqrt(x);
    inv = 1.0 / root;
}
EOB_HD void qsqrt_recip(double x, double& root, double& inv) {
    const double y = rsqrt_estimate(x);
    double g = x * y, h = 0.5 * y;
    const double r = fma(-g, h, 0.5);
    g = fma(g, r, g);
    h = fma(h, r, h);
    g = fma(fma(-g, g, x), h, g);
    g = fma(fma(-g, g, x), h, g);
    const double y2 = h + h;
    root = g;
    inv = fma(y2, fma(-g, y2, 1.0), y2);
}

// cube root of a positive normal x: a single-precision estimate of x^(-1/3)
// (hardware log2 and exp2), two Newton steps for z = x^(-1/3) (z += z (1 - x
// z^3) / 3, no division), y = x z^2 and one residual correction
// y += (x - y^3) z^2 / 3.  Within 1 ulp (0.71 in 20 000 simulated cases),
// about 23 instructions against 34 for the library's; NaN for 0 and NaN (and,
// on the device, for x < 0, whose log2 is NaN).
EOB_HD float rcbrt_estimate(double x) {
#ifdef __HIP_DEVICE_COMPILE__
    return __builtin_amdgcn_exp2f(__builtin_amdgcn_logf((float)x) * (-1.0f / 3.0f));
#else
    return (float)(1.0 / cbrt(x));   // host build: a single-precision start, as for rcp_estimate
#endif
}
EOB_HD double cbrt_pos(double x) {
    constexpr double kThird = 1.0 / 3.0;
    double z = (double)rcbrt_estimate(x);
    z = fma(z * kThird, fma(-x, (z * z) * z, 1.0), z);
    z = fma(z * kThird, fma(-x, (z * z) * z, 1.0), z);
    const double z2 = z * z;
    const double y = x * z2;
    return fma(fma(-(y * y), y, x) * kThird, z2, y);
}

}  // namespace eob
