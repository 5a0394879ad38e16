// Forward-mode dual numbers for the local Jacobians of the adjoint: a value
// and N directional derivatives, with the elementary functions the step
// equations use (+ - * /, exp, log, sin, cos, sqrt).
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

namespace qdfn {
// a namespace of its own: the elementary functions below are found for Dual
// arguments by argument-dependent lookup and never hide ::exp etc. for doubles
namespace ad {

template <int N>
struct Dual {
    double v;
    double d[N];
    __device__ __host__ Dual() : v(0.0) {
        for (int k = 0; k < N; ++k) d[k] = 0.0;
    }
    __device__ __host__ Dual(double x) : v(x) {
        for (int k = 0; k < N; ++k) d[k] = 0.0;
    }
    // variable number `dir`
    __device__ __host__ static Dual var(double x, int dir) {
        Dual r(x);
        r.d[dir] = 1.0;
        return r;
    }
};

#define QDFN_DUAL_UNARY(expr_v, expr_dv)                        \
    Dual<N> r;                                                  \
    r.v = (expr_v);                                             \
    const double g = (expr_dv);                                 \
    for (int k = 0; k < N; ++k) r.d[k] = g * a.d[k];            \
    return r;

template <int N>
__device__ __host__ inline Dual<N> operator-(const Dual<N>& a) {
    QDFN_DUAL_UNARY(-a.v, -1.0)
}
template <int N>
__device__ __host__ inline Dual<N> operator+(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r;
    r.v = a.v + b.v;
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] + b.d[k];
    return r;
}
template <int N>
__device__ __host__ inline Dual<N> operator-(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r;
    r.v = a.v - b.v;
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] - b.d[k];
    return r;
}
template <int N>
__device__ __host__ inline Dual<N> operator*(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r;
    r.v = a.v * b.v;
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
    return r;
}
template <int N>
__device__ __host__ inline Dual<N> operator/(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r;
    r.v = a.v / b.v;
    const double ib = 1.0 / b.v;
    for (int k = 0; k < N; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) * ib;
    return r;
}
template <int N>
__device__ __host__ inline Dual<N> operator+(const Dual<N>& a, double b) {
    Dual<N> r = a;
    r.v = a.v + b;
    return r;
}
template <int N>
__device__ __host__ inline Dual<N> operator+(double b, const Dual<N>& a) {
    return a + b;
}
template <int N>
__device__ __host__ inline Dual<N> operator-(const Dual<N>& a, double b) {
    Dual<N> r = a;
    r.v = a.v - b;
    return r;
}
template <int N>
__device__ __host__ inline Dual<N> operator-(double b, const Dual<N>& a) {
    QDFN_DUAL_UNARY(b - a.v, -1.0)
}
template <int N>
__device__ __host__ inline Dual<N> operator*(const Dual<N>& a, double b) {
    QDFN_DUAL_UNARY(a.v * b, b)
}
template <int N>
__device__ __host__ inline Dual<N> operator*(double b, const Dual<N>& a) {
    return a * b;
}
template <int N>
__device__ __host__ inline Dual<N> operator/(const Dual<N>& a, double b) {
    QDFN_DUAL_UNARY(a.v / b, 1.0 / b)
}
template <int N>
__device__ __host__ inline Dual<N> operator/(double b, const Dual<N>& a) {
    QDFN_DUAL_UNARY(b / a.v, -(b / a.v) / a.v)
}
template <int N>
__device__ __host__ inline Dual<N> exp(const Dual<N>& a) {
    const double e = ::exp(a.v);
    QDFN_DUAL_UNARY(e, e)
}
template <int N>
__device__ __host__ inline Dual<N> log(const Dual<N>& a) {
    QDFN_DUAL_UNARY(::log(a.v), 1.0 / a.v)
}
template <int N>
__device__ __host__ inline Dual<N> sin(const Dual<N>& a) {
    QDFN_DUAL_UNARY(::sin(a.v), ::cos(a.v))
}
template <int N>
__device__ __host__ inline Dual<N> cos(const Dual<N>& a) {
    QDFN_DUAL_UNARY(::cos(a.v), -::sin(a.v))
}
template <int N>
__device__ __host__ inline Dual<N> sqrt(const Dual<N>& a) {
    const double s = ::sqrt(a.v);
    QDFN_DUAL_UNARY(s, 0.5 / s)
}

#undef QDFN_DUAL_UNARY

}  // namespace ad
}  // namespace qdfn
