// shcfd_ad.h -- forward-mode dual numbers and the scalar-generic forms of the
// per-face and per-cell flow functions.
//
// The discrete-adjoint kernels need the exact local Jacobians of the SLAU
// face flux, the MUSCL slope pair and the ideal-gas primitive/thermo map.
// These functions are written once, generic in the scalar type: the primal
// kernels instantiate them with double (same operations in the same order,
// so the primal results do not change) and the adjoint kernels with
// Dual<N>, which carries N directional derivatives through every operation.
// Branches test values only, so a dual evaluation follows the branch the
// primal takes; at a branch point the derivative is the one-sided derivative
// of that branch.
#pragma once
#include <hip/hip_runtime.h>

namespace shcfd {

// keep the double forms visible next to the dual overloads below
using ::fabs;
using ::sqrt;

template <int N>
struct Dual {
    double v;
    double d[N];
    __device__ Dual() {}
    __device__ Dual(double x) : v(x) {
#pragma unroll
        for (int k = 0; k < N; k++) d[k] = 0.0;
    }
    // independent variable number `k`
    __device__ static Dual var(double x, int k) {
        Dual r(x);
        r.d[k] = 1.0;
        return r;
    }
};

template <int N> __device__ inline Dual<N> operator+(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r; r.v = a.v + b.v;
#pragma unroll
    for (int k = 0; k < N; k++) r.d[k] = a.d[k] + b.d[k];
    return r;
}
template <int N> __device__ inline Dual<N> operator-(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r; r.v = a.v - b.v;
#pragma unroll
    for (int k = 0; k < N; k++) r.d[k] = a.d[k] - b.d[k];
    return r;
}
template <int N> __device__ inline Dual<N> operator-(const Dual<N>& a) {
    Dual<N> r; r.v = -a.v;
#pragma unroll
    for (int k = 0; k < N; k++) r.d[k] = -a.d[k];
    return r;
}
template <int N> __device__ inline Dual<N> operator*(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r; r.v = a.v * b.v;
#pragma unroll
    for (int k = 0; k < N; k++) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
    return r;
}
template <int N> __device__ inline Dual<N> operator/(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r; r.v = a.v / b.v;
    const double ib = 1.0 / b.v;
#pragma unroll
    for (int k = 0; k < N; k++) r.d[k] = (a.d[k] - r.v * b.d[k]) * ib;
    return r;
}
template <int N> __device__ inline Dual<N> operator+(const Dual<N>& a, double b) { return a + Dual<N>(b); }
template <int N> __device__ inline Dual<N> operator+(double a, const Dual<N>& b) { return Dual<N>(a) + b; }
template <int N> __device__ inline Dual<N> operator-(const Dual<N>& a, double b) { return a - Dual<N>(b); }
template <int N> __device__ inline Dual<N> operator-(double a, const Dual<N>& b) { return Dual<N>(a) - b; }
template <int N> __device__ inline Dual<N> operator*(const Dual<N>& a, double b) {
    Dual<N> r; r.v = a.v * b;
#pragma unroll
    for (int k = 0; k < N; k++) r.d[k] = a.d[k] * b;
    return r;
}
template <int N> __device__ inline Dual<N> operator*(double a, const Dual<N>& b) { return b * a; }
template <int N> __device__ inline Dual<N> operator/(const Dual<N>& a, double b) { return a / Dual<N>(b); }
template <int N> __device__ inline Dual<N> operator/(double a, const Dual<N>& b) { return Dual<N>(a) / b; }

template <int N> __device__ inline bool operator<(const Dual<N>& a, const Dual<N>& b) { return a.v < b.v; }
template <int N> __device__ inline bool operator<(const Dual<N>& a, double b) { return a.v < b; }
template <int N> __device__ inline bool operator<(double a, const Dual<N>& b) { return a < b.v; }
template <int N> __device__ inline bool operator>(const Dual<N>& a, double b) { return a.v > b; }
template <int N> __device__ inline bool operator>(double a, const Dual<N>& b) { return a > b.v; }
template <int N> __device__ inline bool operator<=(const Dual<N>& a, double b) { return a.v <= b; }
template <int N> __device__ inline bool operator>=(const Dual<N>& a, double b) { return a.v >= b; }

template <int N> __device__ inline Dual<N> sqrt(const Dual<N>& a) {
    Dual<N> r; r.v = ::sqrt(a.v);
    const double h = 0.5 / r.v;
#pragma unroll
    for (int k = 0; k < N; k++) r.d[k] = a.d[k] * h;
    return r;
}
template <int N> __device__ inline Dual<N> fabs(const Dual<N>& a) { return (a.v < 0.0) ? -a : a; }

// value of a scalar (double or dual)
__device__ inline double val(double x) { return x; }
template <int N> __device__ inline double val(const Dual<N>& x) { return x.v; }

// sign with sign(0) = +1 (Fortran sign(1d0, x) for x = +0)
template <typename T> __device__ inline double sgn1(const T& x) { return (val(x) >= 0.0) ? 1.0 : -1.0; }
// selects by value; the derivative follows the selected operand
template <typename T> __device__ inline T tmin(const T& a, const T& b) { return (val(a) < val(b)) ? a : b; }
template <typename T> __device__ inline T tmax(const T& a, const T& b) { return (val(a) > val(b)) ? a : b; }
template <int N> __device__ inline Dual<N> tfabs(const Dual<N>& a) { return (a.v < 0.0) ? -a : a; }
__device__ inline double tfabs(double a) { return ::fabs(a); }

}  // namespace shcfd
