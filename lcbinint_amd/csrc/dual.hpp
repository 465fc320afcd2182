// Forward-mode tangents for the gradient kernels: a value with kTangents
// partial derivatives, propagated through +, -, *, /, sqrt, fabs and clamps.
// Tangent slots: source x, source y, s, q (the lens-frame ratio q_eff),
// rho, linear limb-darkening c.
#pragma once

#include <hip/hip_runtime.h>

namespace lcbg {

constexpr int kTangents = 6;
enum : int { kTx = 0, kTy = 1, kTs = 2, kTq = 3, kTrho = 4, kTc = 5 };

struct Dual {
    double v;
    double d[kTangents];
    Dual() = default;
    // a constant (no tangent)
    __device__ explicit Dual(double value) : v(value)
    {
#pragma unroll
        for (int k = 0; k < kTangents; ++k) d[k] = 0.0;
    }
};

__device__ __forceinline__ Dual dconst(double v)
{
    Dual r;
    r.v = v;
#pragma unroll
    for (int k = 0; k < kTangents; ++k) r.d[k] = 0.0;
    return r;
}

__device__ __forceinline__ Dual dvar(double v, int slot)
{
    Dual r = dconst(v);
    r.d[slot] = 1.0;
    return r;
}

__device__ __forceinline__ Dual operator+(const Dual& a, const Dual& b)
{
    Dual r;
    r.v = a.v + b.v;
#pragma unroll
    for (int k = 0; k < kTangents; ++k) r.d[k] = a.d[k] + b.d[k];
    return r;
}

__device__ __forceinline__ Dual operator-(const Dual& a, const Dual& b)
{
    Dual r;
    r.v = a.v - b.v;
#pragma unroll
    for (int k = 0; k < kTangents; ++k) r.d[k] = a.d[k] - b.d[k];
    return r;
}

__device__ __forceinline__ Dual operator-(const Dual& a)
{
    Dual r;
    r.v = -a.v;
#pragma unroll
    for (int k = 0; k < kTangents; ++k) r.d[k] = -a.d[k];
    return r;
}

__device__ __forceinline__ Dual operator*(const Dual& a, const Dual& b)
{
    Dual r;
    r.v = a.v * b.v;
#pragma unroll
    for (int k = 0; k < kTangents; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
    return r;
}

__device__ __forceinline__ Dual operator*(double a, const Dual& b)
{
    Dual r;
    r.v = a * b.v;
#pragma unroll
    for (int k = 0; k < kTangents; ++k) r.d[k] = a * b.d[k];
    return r;
}

__device__ __forceinline__ Dual operator*(const Dual& a, double b) { return b * a; }

__device__ __forceinline__ Dual operator+(const Dual& a, double b)
{
    Dual r = a;
    r.v = a.v + b;
    return r;
}

__device__ __forceinline__ Dual operator+(double a, const Dual& b) { return b + a; }

__device__ __forceinline__ Dual operator-(const Dual& a, double b)
{
    Dual r = a;
    r.v = a.v - b;
    return r;
}

__device__ __forceinline__ Dual operator-(double a, const Dual& b)
{
    Dual r;
    r.v = a - b.v;
#pragma unroll
    for (int k = 0; k < kTangents; ++k) r.d[k] = -b.d[k];
    return r;
}

__device__ __forceinline__ Dual operator/(const Dual& a, const Dual& b)
{
    Dual r;
    const double inv = 1.0 / b.v;
    r.v = a.v * inv;
#pragma unroll
    for (int k = 0; k < kTangents; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) * inv;
    return r;
}

__device__ __forceinline__ Dual operator/(double a, const Dual& b)
{
    Dual r;
    const double inv = 1.0 / b.v;
    r.v = a * inv;
#pragma unroll
    for (int k = 0; k < kTangents; ++k) r.d[k] = -r.v * b.d[k] * inv;
    return r;
}

__device__ __forceinline__ Dual operator/(const Dual& a, double b)
{
    Dual r;
    const double inv = 1.0 / b;
    r.v = a.v * inv;
#pragma unroll
    for (int k = 0; k < kTangents; ++k) r.d[k] = a.d[k] * inv;
    return r;
}

__device__ __forceinline__ Dual dsqrt(const Dual& a)
{
    Dual r;
    r.v = __builtin_sqrt(a.v);
    const double h = r.v > 0.0 ? 0.5 / r.v : 0.0;
#pragma unroll
    for (int k = 0; k < kTangents; ++k) r.d[k] = a.d[k] * h;
    return r;
}

__device__ __forceinline__ Dual dabs(const Dual& a) { return a.v < 0.0 ? -a : a; }

// clamp(a, lo, hi) with the derivative of the branch taken
__device__ __forceinline__ Dual dclampd(const Dual& a, double lo, double hi)
{
    if (a.v <= lo) return dconst(lo);
    if (a.v >= hi) return dconst(hi);
    return a;
}

__device__ __forceinline__ Dual& operator+=(Dual& a, const Dual& b)
{
    a = a + b;
    return a;
}

// The value-only counterparts, so that one template serves both.
__device__ __forceinline__ double dsqrt(double a) { return __builtin_sqrt(a); }
__device__ __forceinline__ double dabs(double a) { return a < 0.0 ? -a : a; }
__device__ __forceinline__ double dclampd(double a, double lo, double hi)
{
    return a <= lo ? lo : (a >= hi ? hi : a);
}
__device__ __forceinline__ double val(double a) { return a; }
__device__ __forceinline__ double val(const Dual& a) { return a.v; }

}  // namespace lcbg
