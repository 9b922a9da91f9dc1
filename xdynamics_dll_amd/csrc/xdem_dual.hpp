// xdem_dual.hpp -- forward-mode dual numbers for the calibration's local Jacobians.
//
// Dual<N> carries a value and its derivatives along N input directions. Every operation
// computes the value with exactly the operations of the plain double code (so a function
// evaluated on Dual<N> returns the same value bits as on double) and the derivatives by the
// chain rule. Comparisons and branches look at the value only: derivatives are those of the
// branch taken.
#pragma once

#include <cmath>

namespace xdem {

// the double versions stay visible next to the dual overloads below
using ::exp;
using ::fabs;
using ::fmin;
using ::log;
using ::sqrt;

template <int N> struct Dual {
    double v;
    double d[N];
    __host__ __device__ Dual() : v(0.0)
    {
        for (int k = 0; k < N; k++) d[k] = 0.0;
    }
    __host__ __device__ Dual(double x) : v(x)   // a constant
    {
        for (int k = 0; k < N; k++) d[k] = 0.0;
    }
    // input number k (0 <= k < N) of value x
    __host__ __device__ static Dual input(double x, int k)
    {
        Dual r(x);
        r.d[k] = 1.0;
        return r;
    }
};

template <int N> __host__ __device__ inline Dual<N> operator+(const Dual<N>& a, const Dual<N>& b)
{
    Dual<N> r;
    r.v = a.v + b.v;
    for (int k = 0; k < N; k++) r.d[k] = a.d[k] + b.d[k];
    return r;
}
template <int N> __host__ __device__ inline Dual<N> operator-(const Dual<N>& a, const Dual<N>& b)
{
    Dual<N> r;
    r.v = a.v - b.v;
    for (int k = 0; k < N; k++) r.d[k] = a.d[k] - b.d[k];
    return r;
}
template <int N> __host__ __device__ inline Dual<N> operator-(const Dual<N>& a)
{
    Dual<N> r;
    r.v = -a.v;
    for (int k = 0; k < N; k++) r.d[k] = -a.d[k];
    return r;
}
template <int N> __host__ __device__ inline Dual<N> operator*(const Dual<N>& a, const Dual<N>& b)
{
    Dual<N> r;
    r.v = a.v * b.v;
    for (int k = 0; k < N; k++) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
    return r;
}
template <int N> __host__ __device__ inline Dual<N> operator/(const Dual<N>& a, const Dual<N>& b)
{
    Dual<N> r;
    r.v = a.v / b.v;
    const double ib = 1.0 / b.v;
    for (int k = 0; k < N; k++) r.d[k] = (a.d[k] - r.v * b.d[k]) * ib;
    return r;
}
// mixed operands: a double is a constant
template <int N> __host__ __device__ inline Dual<N> operator+(const Dual<N>& a, double b) { return a + Dual<N>(b); }
template <int N> __host__ __device__ inline Dual<N> operator+(double a, const Dual<N>& b) { return Dual<N>(a) + b; }
template <int N> __host__ __device__ inline Dual<N> operator-(const Dual<N>& a, double b) { return a - Dual<N>(b); }
template <int N> __host__ __device__ inline Dual<N> operator-(double a, const Dual<N>& b) { return Dual<N>(a) - b; }
template <int N> __host__ __device__ inline Dual<N> operator*(const Dual<N>& a, double b)
{
    Dual<N> r;
    r.v = a.v * b;
    for (int k = 0; k < N; k++) r.d[k] = a.d[k] * b;
    return r;
}
template <int N> __host__ __device__ inline Dual<N> operator*(double a, const Dual<N>& b) { return b * a; }
template <int N> __host__ __device__ inline Dual<N> operator/(const Dual<N>& a, double b) { return a / Dual<N>(b); }
template <int N> __host__ __device__ inline Dual<N> operator/(double a, const Dual<N>& b) { return Dual<N>(a) / b; }

template <int N> __host__ __device__ inline bool operator<(const Dual<N>& a, const Dual<N>& b) { return a.v < b.v; }
template <int N> __host__ __device__ inline bool operator>(const Dual<N>& a, const Dual<N>& b) { return a.v > b.v; }
template <int N> __host__ __device__ inline bool operator<=(const Dual<N>& a, const Dual<N>& b) { return a.v <= b.v; }
template <int N> __host__ __device__ inline bool operator>=(const Dual<N>& a, const Dual<N>& b) { return a.v >= b.v; }

template <int N> __host__ __device__ inline Dual<N> sqrt(const Dual<N>& a)
{
    Dual<N> r;
    r.v = ::sqrt(a.v);
    const double g = r.v > 0.0 ? 0.5 / r.v : 0.0;   // the derivative at 0 is taken as 0
    for (int k = 0; k < N; k++) r.d[k] = a.d[k] * g;
    return r;
}
template <int N> __host__ __device__ inline Dual<N> log(const Dual<N>& a)
{
    Dual<N> r;
    r.v = ::log(a.v);
    for (int k = 0; k < N; k++) r.d[k] = a.d[k] / a.v;
    return r;
}
template <int N> __host__ __device__ inline Dual<N> exp(const Dual<N>& a)
{
    Dual<N> r;
    r.v = ::exp(a.v);
    for (int k = 0; k < N; k++) r.d[k] = a.d[k] * r.v;
    return r;
}
template <int N> __host__ __device__ inline Dual<N> fabs(const Dual<N>& a) { return a.v < 0.0 ? -a : a; }
template <int N> __host__ __device__ inline Dual<N> fmin(const Dual<N>& a, const Dual<N>& b)
{
    // fmin(x, y) returns x when x < y, else y (NaN-free operands)
    return a.v < b.v ? a : b;
}

// value of a double or a dual
__host__ __device__ inline double val(double x) { return x; }
template <int N> __host__ __device__ inline double val(const Dual<N>& x) { return x.v; }

}  // namespace xdem
