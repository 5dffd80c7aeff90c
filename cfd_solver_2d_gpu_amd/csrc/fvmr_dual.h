// Forward-mode dual numbers for the adjoint kernels: a value and N directional derivatives.
// The solver's face-physics, gradient and limiter templates are instantiated with Dual<N> in
// place of double, so the local Jacobians the adjoint needs come from the same source as the
// residual.  Branches (min/max, limiter cases, the out-of-range fallback, the farfield regimes)
// are taken on the values, so derivatives follow the branch the value takes.
#pragma once
#include <hip/hip_runtime.h>

namespace fvmr {

template <int N>
struct Dual {
    double v;
    double d[N];

    __device__ __forceinline__ Dual() : v(0.0) {
#pragma unroll
        for (int i = 0; i < N; i++) d[i] = 0.0;
    }
    __device__ __forceinline__ Dual(double x) : v(x) {
#pragma unroll
        for (int i = 0; i < N; i++) d[i] = 0.0;
    }
    // value x with derivative 1 in direction k
    __device__ __forceinline__ static Dual seed(double x, int k) {
        Dual r(x);
#pragma unroll
        for (int i = 0; i < N; i++) r.d[i] = (i == k) ? 1.0 : 0.0;
        return r;
    }

    friend __device__ __forceinline__ Dual operator+(const Dual &a, const Dual &b) {
        Dual r;
        r.v = a.v + b.v;
#pragma unroll
        for (int i = 0; i < N; i++) r.d[i] = a.d[i] + b.d[i];
        return r;
    }
    friend __device__ __forceinline__ Dual operator+(const Dual &a, double b) { Dual r = a; r.v = a.v + b; return r; }
    friend __device__ __forceinline__ Dual operator+(double a, const Dual &b) { Dual r = b; r.v = a + b.v; return r; }
    friend __device__ __forceinline__ Dual operator-(const Dual &a, const Dual &b) {
        Dual r;
        r.v = a.v - b.v;
#pragma unroll
        for (int i = 0; i < N; i++) r.d[i] = a.d[i] - b.d[i];
        return r;
    }
    friend __device__ __forceinline__ Dual operator-(const Dual &a, double b) { Dual r = a; r.v = a.v - b; return r; }
    friend __device__ __forceinline__ Dual operator-(double a, const Dual &b) {
        Dual r;
        r.v = a - b.v;
#pragma unroll
        for (int i = 0; i < N; i++) r.d[i] = -b.d[i];
        return r;
    }
    friend __device__ __forceinline__ Dual operator-(const Dual &a) {
        Dual r;
        r.v = -a.v;
#pragma unroll
        for (int i = 0; i < N; i++) r.d[i] = -a.d[i];
        return r;
    }
    friend __device__ __forceinline__ Dual operator*(const Dual &a, const Dual &b) {
        Dual r;
        r.v = a.v * b.v;
#pragma unroll
        for (int i = 0; i < N; i++) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
        return r;
    }
    friend __device__ __forceinline__ Dual operator*(const Dual &a, double b) {
        Dual r;
        r.v = a.v * b;
#pragma unroll
        for (int i = 0; i < N; i++) r.d[i] = a.d[i] * b;
        return r;
    }
    friend __device__ __forceinline__ Dual operator*(double a, const Dual &b) { return b * a; }
    friend __device__ __forceinline__ Dual operator/(const Dual &a, const Dual &b) {
        Dual r;
        const double ib = 1.0 / b.v;
        r.v = a.v * ib;
#pragma unroll
        for (int i = 0; i < N; i++) r.d[i] = (a.d[i] - r.v * b.d[i]) * ib;
        return r;
    }
    friend __device__ __forceinline__ Dual operator/(const Dual &a, double b) {
        Dual r;
        const double ib = 1.0 / b;
        r.v = a.v * ib;
#pragma unroll
        for (int i = 0; i < N; i++) r.d[i] = a.d[i] * ib;
        return r;
    }
    friend __device__ __forceinline__ Dual operator/(double a, const Dual &b) {
        Dual r;
        const double ib = 1.0 / b.v;
        r.v = a * ib;
#pragma unroll
        for (int i = 0; i < N; i++) r.d[i] = -r.v * b.d[i] * ib;
        return r;
    }
    __device__ __forceinline__ Dual &operator+=(const Dual &b) { return *this = *this + b; }
    __device__ __forceinline__ Dual &operator-=(const Dual &b) { return *this = *this - b; }
    __device__ __forceinline__ Dual &operator*=(const Dual &b) { return *this = *this * b; }
    __device__ __forceinline__ Dual &operator/=(const Dual &b) { return *this = *this / b; }

    friend __device__ __forceinline__ bool operator<(const Dual &a, const Dual &b) { return a.v < b.v; }
    friend __device__ __forceinline__ bool operator>(const Dual &a, const Dual &b) { return a.v > b.v; }
    friend __device__ __forceinline__ bool operator<=(const Dual &a, const Dual &b) { return a.v <= b.v; }
    friend __device__ __forceinline__ bool operator>=(const Dual &a, const Dual &b) { return a.v >= b.v; }
    friend __device__ __forceinline__ bool operator==(const Dual &a, const Dual &b) { return a.v == b.v; }
    friend __device__ __forceinline__ bool operator!=(const Dual &a, const Dual &b) { return a.v != b.v; }
    friend __device__ __forceinline__ bool operator<(const Dual &a, double b) { return a.v < b; }
    friend __device__ __forceinline__ bool operator>(const Dual &a, double b) { return a.v > b; }
    friend __device__ __forceinline__ bool operator<=(const Dual &a, double b) { return a.v <= b; }
    friend __device__ __forceinline__ bool operator>=(const Dual &a, double b) { return a.v >= b; }
    friend __device__ __forceinline__ bool operator==(const Dual &a, double b) { return a.v == b; }
    friend __device__ __forceinline__ bool operator!=(const Dual &a, double b) { return a.v != b; }
    friend __device__ __forceinline__ bool operator<(double a, const Dual &b) { return a < b.v; }
    friend __device__ __forceinline__ bool operator>(double a, const Dual &b) { return a > b.v; }

    friend __device__ __forceinline__ Dual sqrt(const Dual &a) {
        Dual r;
        r.v = ::sqrt(a.v);
        const double h = 0.5 / r.v;
#pragma unroll
        for (int i = 0; i < N; i++) r.d[i] = a.d[i] * h;
        return r;
    }
    friend __device__ __forceinline__ Dual fabs(const Dual &a) { return (a.v < 0.0) ? -a : a; }
    friend __device__ __forceinline__ bool isnan(const Dual &a) { return ::isnan(a.v); }
    // fmin/fmax: the operand whose value is returned (the non-NaN one when one is NaN; the
    // first on a tie) carries its derivatives
    friend __device__ __forceinline__ Dual fmin(const Dual &a, const Dual &b) {
        if (::isnan(a.v)) return b;
        if (::isnan(b.v)) return a;
        return (b.v < a.v) ? b : a;
    }
    friend __device__ __forceinline__ Dual fmax(const Dual &a, const Dual &b) {
        if (::isnan(a.v)) return b;
        if (::isnan(b.v)) return a;
        return (b.v > a.v) ? b : a;
    }
    // a^b = exp(b ln a): d = b a^(b-1) da + a^b ln(a) db (the ln term only where b varies)
    friend __device__ __forceinline__ Dual pow(const Dual &a, const Dual &b) {
        Dual r;
        r.v = ::pow(a.v, b.v);
        const double da = b.v * ::pow(a.v, b.v - 1.0);
        bool vb = false;
#pragma unroll
        for (int i = 0; i < N; i++) vb = vb || (b.d[i] != 0.0);
        const double db = vb ? r.v * ::log(a.v) : 0.0;
#pragma unroll
        for (int i = 0; i < N; i++) r.d[i] = da * a.d[i] + db * b.d[i];
        return r;
    }
};

}  // namespace fvmr
