// Scalar types and small math helpers shared by the HIP kernels.
//
// Dual   : value + one tangent (forward-mode derivative along one seed).
// HDual  : hyper-dual number (value, two first-order tangents, mixed second
//          order), giving an exact second derivative d2H/(dx dy) in one pass.
//
// Everything is FP64.  Functions are __host__ __device__ so host-side unit
// probes can reuse them, but the library only evaluates them on the GPU.
#pragma once
#include <hip/hip_runtime.h>
#include <cmath>

#define EOB_HD __host__ __device__ __forceinline__
#include "eob_fastmath.hpp"

namespace eob {

// correctly rounded doubles of the transcendental constants the tables use
constexpr double kPi = 3.141592653589793;
constexpr double kPi2 = 9.869604401089358;        // pi^2
constexpr double kEulerGamma = 0.5772156649015329;
constexpr double kLn2 = 0.6931471805599453;
constexpr double kLn3 = 1.0986122886681098;
constexpr double kLn5 = 1.6094379124341003;
constexpr double kLn7 = 1.9459101490553132;

struct Dual {
    double v, d;
    EOB_HD Dual() : v(0.0), d(0.0) {}
    EOB_HD Dual(double c) : v(c), d(0.0) {}
    EOB_HD Dual(double c, double t) : v(c), d(t) {}
};
EOB_HD Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
EOB_HD Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.d - b.d); }
EOB_HD Dual operator-(Dual a) { return Dual(-a.v, -a.d); }
EOB_HD Dual operator*(Dual a, Dual b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
EOB_HD Dual operator+(Dual a, double b) { return Dual(a.v + b, a.d); }
EOB_HD Dual operator+(double b, Dual a) { return Dual(b + a.v, a.d); }
EOB_HD Dual operator-(Dual a, double b) { return Dual(a.v - b, a.d); }
EOB_HD Dual operator-(double b, Dual a) { return Dual(b - a.v, -a.d); }
EOB_HD Dual operator*(Dual a, double b) { return Dual(a.v * b, a.d * b); }
EOB_HD Dual operator*(double b, Dual a) { return Dual(b * a.v, b * a.d); }
EOB_HD Dual operator/(Dual a, Dual b) {
    double inv = 1.0 / b.v;
    double q = a.v * inv;
    return Dual(q, (a.d - q * b.d) * inv);
}
EOB_HD Dual operator/(double a, Dual b) {
    double inv = 1.0 / b.v;
    double q = a * inv;
    return Dual(q, -q * b.d * inv);
}
EOB_HD Dual operator/(Dual a, double b) { return Dual(a.v / b, a.d / b); }
EOB_HD Dual dsqrt(Dual a) {
    double s = sqrt(a.v);
    return Dual(s, 0.5 * a.d / s);
}
EOB_HD Dual dlog(Dual a) { return Dual(log(a.v), a.d / a.v); }
EOB_HD double dsqrt(double a) { return sqrt(a); }
EOB_HD double dlog(double a) { return log(a); }

struct HDual {
    double v, a, b, ab;
    EOB_HD HDual() : v(0.0), a(0.0), b(0.0), ab(0.0) {}
    EOB_HD HDual(double c) : v(c), a(0.0), b(0.0), ab(0.0) {}
    EOB_HD HDual(double c, double ta, double tb, double tab) : v(c), a(ta), b(tb), ab(tab) {}
};
EOB_HD HDual operator+(HDual x, HDual y) { return HDual(x.v + y.v, x.a + y.a, x.b + y.b, x.ab + y.ab); }
EOB_HD HDual operator-(HDual x, HDual y) { return HDual(x.v - y.v, x.a - y.a, x.b - y.b, x.ab - y.ab); }
EOB_HD HDual operator-(HDual x) { return HDual(-x.v, -x.a, -x.b, -x.ab); }
EOB_HD HDual operator*(HDual x, HDual y) {
    return HDual(x.v * y.v, x.a * y.v + x.v * y.a, x.b * y.v + x.v * y.b, x.ab * y.v + x.a * y.b + x.b * y.a + x.v * y.ab);
}
EOB_HD HDual operator+(HDual x, double c) { return HDual(x.v + c, x.a, x.b, x.ab); }
EOB_HD HDual operator+(double c, HDual x) { return HDual(c + x.v, x.a, x.b, x.ab); }
EOB_HD HDual operator-(HDual x, double c) { return HDual(x.v - c, x.a, x.b, x.ab); }
EOB_HD HDual operator-(double c, HDual x) { return HDual(c - x.v, -x.a, -x.b, -x.ab); }
EOB_HD HDual operator*(HDual x, double c) { return HDual(x.v * c, x.a * c, x.b * c, x.ab * c); }
EOB_HD HDual operator*(double c, HDual x) { return HDual(c * x.v, c * x.a, c * x.b, c * x.ab); }
// f(x) for a hyper-dual argument from f, f', f'' at the value
EOB_HD HDual hd_apply(HDual x, double f, double f1, double f2) {
    return HDual(f, f1 * x.a, f1 * x.b, f1 * x.ab + f2 * x.a * x.b);
}
EOB_HD HDual operator/(double c, HDual y) {
    double inv = 1.0 / y.v;
    return c * hd_apply(y, inv, -inv * inv, 2.0 * inv * inv * inv);
}
EOB_HD HDual operator/(HDual x, HDual y) { return x * (1.0 / y); }
EOB_HD HDual operator/(HDual x, double c) { return x * (1.0 / c); }
EOB_HD HDual dsqrt(HDual x) {
    double s = sqrt(x.v);
    return hd_apply(x, s, 0.5 / s, -0.25 / (s * x.v));
}
EOB_HD HDual dlog(HDual x) {
    double inv = 1.0 / x.v;
    return hd_apply(x, log(x.v), inv, -inv * inv);
}

// x^n for small non-negative integer n by repeated squaring
template <class T>
EOB_HD T ipow(T x, int n) {
    T r = 1.0;
    while (n > 0) {
        if (n & 1) r = r * x;
        x = x * x;
        n >>= 1;
    }
    return r;
}

// x^n for 0 <= n < 16: the same products as ipow, as straight-line code with
// selects (lanes with different n then share one instruction stream)
template <class T>
EOB_HD T ipow15(T x, int n) {
    T r = (n & 1) ? x : T(1.0);
    x = x * x;
    r = (n & 2) ? r * x : r;
    x = x * x;
    r = (n & 4) ? r * x : r;
    x = x * x;
    r = (n & 8) ? r * x : r;
    return r;
}

template <class T> EOB_HD T sq(T x) { return x * x; }

}  // namespace eob
