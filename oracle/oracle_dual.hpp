// Oracle test infrastructure (CPU only). Not part of the shipped library:
// only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it,
// and only as the results to compare against.
//
// Forward-mode automatic differentiation with a fixed number of tangents.
// Nesting Dual<Dual<double,N>,N> gives exact second derivatives (Hessian).
#pragma once
#include <cmath>

namespace orc {

template <class T, int N>
struct Dual {
    T v;
    T d[N];
    Dual() : v(0.0) { for (int i = 0; i < N; ++i) d[i] = T(0.0); }
    Dual(double c) : v(c) { for (int i = 0; i < N; ++i) d[i] = T(0.0); }
    Dual(const T& c, int) : v(c) { for (int i = 0; i < N; ++i) d[i] = T(0.0); }
};

template <class T> struct is_dual { static const bool value = false; };
template <class T, int N> struct is_dual<Dual<T, N>> { static const bool value = true; };

template <class T, int N>
Dual<T, N> make_var(const T& v, int k) {
    Dual<T, N> x(v, 0);
    x.d[k] = T(1.0);
    return x;
}

template <class T, int N>
Dual<T, N> operator+(const Dual<T, N>& a, const Dual<T, N>& b) {
    Dual<T, N> r(a.v + b.v, 0);
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] + b.d[i];
    return r;
}
template <class T, int N>
Dual<T, N> operator-(const Dual<T, N>& a, const Dual<T, N>& b) {
    Dual<T, N> r(a.v - b.v, 0);
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] - b.d[i];
    return r;
}
template <class T, int N>
Dual<T, N> operator-(const Dual<T, N>& a) {
    Dual<T, N> r(-a.v, 0);
    for (int i = 0; i < N; ++i) r.d[i] = -a.d[i];
    return r;
}
template <class T, int N>
Dual<T, N> operator*(const Dual<T, N>& a, const Dual<T, N>& b) {
    Dual<T, N> r(a.v * b.v, 0);
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
    return r;
}
template <class T, int N>
Dual<T, N> operator/(const Dual<T, N>& a, const Dual<T, N>& b) {
    T inv = T(1.0) / b.v;
    T q = a.v * inv;
    Dual<T, N> r(q, 0);
    for (int i = 0; i < N; ++i) r.d[i] = (a.d[i] - q * b.d[i]) * inv;
    return r;
}
// mixed with plain double
template <class T, int N> Dual<T, N> operator+(const Dual<T, N>& a, double b) { Dual<T, N> r = a; r.v = a.v + b; return r; }
template <class T, int N> Dual<T, N> operator+(double b, const Dual<T, N>& a) { Dual<T, N> r = a; r.v = b + a.v; return r; }
template <class T, int N> Dual<T, N> operator-(const Dual<T, N>& a, double b) { Dual<T, N> r = a; r.v = a.v - b; return r; }
template <class T, int N> Dual<T, N> operator-(double b, const Dual<T, N>& a) { Dual<T, N> r = -a; r.v = b - a.v; return r; }
template <class T, int N> Dual<T, N> operator*(const Dual<T, N>& a, double b) {
    Dual<T, N> r(a.v * b, 0);
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b;
    return r;
}
template <class T, int N> Dual<T, N> operator*(double b, const Dual<T, N>& a) {
    Dual<T, N> r(b * a.v, 0);
    for (int i = 0; i < N; ++i) r.d[i] = b * a.d[i];
    return r;
}
template <class T, int N> Dual<T, N> operator/(const Dual<T, N>& a, double b) {
    Dual<T, N> r(a.v / b, 0);
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] / b;
    return r;
}
template <class T, int N> Dual<T, N> operator/(double b, const Dual<T, N>& a) {
    T q = b / a.v;
    Dual<T, N> r(q, 0);
    for (int i = 0; i < N; ++i) r.d[i] = -(q / a.v) * a.d[i];
    return r;
}

// elementary functions (chain rule on the value part)
inline double dval(double x) { return x; }
template <class T, int N> double dval(const Dual<T, N>& x) { return dval(x.v); }

template <class T, int N>
Dual<T, N> sqrt(const Dual<T, N>& a) {
    using std::sqrt;
    T s = sqrt(a.v);
    Dual<T, N> r(s, 0);
    T f = T(0.5) / s;
    for (int i = 0; i < N; ++i) r.d[i] = f * a.d[i];
    return r;
}
template <class T, int N>
Dual<T, N> log(const Dual<T, N>& a) {
    using std::log;
    Dual<T, N> r(log(a.v), 0);
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] / a.v;
    return r;
}
template <class T, int N>
Dual<T, N> exp(const Dual<T, N>& a) {
    using std::exp;
    T e = exp(a.v);
    Dual<T, N> r(e, 0);
    for (int i = 0; i < N; ++i) r.d[i] = e * a.d[i];
    return r;
}
// x^p for a real exponent p (the reference calls pow(x, p) with p in {2,3,...,5/2,...})
template <class T, int N>
Dual<T, N> pow(const Dual<T, N>& a, double p) {
    using std::pow;
    T val = pow(a.v, p);
    Dual<T, N> r(val, 0);
    T f = p * pow(a.v, p - 1.0);
    for (int i = 0; i < N; ++i) r.d[i] = f * a.d[i];
    return r;
}

}  // namespace orc
