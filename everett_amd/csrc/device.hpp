// Device-side helpers shared by the HIP kernels: amplitude types, the
// reference's complex product, and bit-insertion index maps.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace ev {

// Amplitude storage: c128 = double2 {re, im} (src/complex.rs:22-29 layout),
// c64 = float2.
template <class R> struct AmpOf;
template <> struct AmpOf<double> { using type = double2; };
template <> struct AmpOf<float> { using type = float2; };
template <class R> using Amp = typename AmpOf<R>::type;

__device__ __forceinline__ double dfma(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float dfma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

template <class V> __host__ __device__ __forceinline__ V mk(decltype(V::x) re, decltype(V::x) im)
{
    V v;
    v.x = re;
    v.y = im;
    return v;
}

// src/complex.rs:115-125: (a.re*b.re - a.im*b.im, a.re*b.im + a.im*b.re) with
// one mul_add per component, in exactly this operand order.
template <class V> __device__ __forceinline__ V cmul(V a, V b)
{
    return mk<V>(dfma(a.x, b.x, -(a.y * b.y)), dfma(a.x, b.y, a.y * b.x));
}
template <class V> __device__ __forceinline__ V cadd(V a, V b) { return mk<V>(a.x + b.x, a.y + b.y); }

// Gate matrices as kernel arguments, in the state's precision.
template <class R> struct M2 { Amp<R> m[4]; };
template <class R> struct M4 { Amp<R> m[16]; };

// src/kernel.rs:37-44: insert a 0 bit at position k of i.
__device__ __forceinline__ uint64_t insert0(uint64_t i, int k)
{
    const uint64_t lo = i & ((uint64_t(1) << k) - 1);
    return ((i >> k) << (k + 1)) | lo;
}

}  // namespace ev
