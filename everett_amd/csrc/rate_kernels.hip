// Arithmetic-rate probes for gfx950: how fast this GPU runs fused
// multiply-adds on the vector ALU, on the matrix cores, and on both at once,
// in a state precision's arithmetic (f64 for complex128, f32 for complex64).
// The bench divides the fused pass's algorithmic flops by the vector rate
// measured here, as it divides the pass's bytes by the measured copy
// bandwidth, and the mixed rate shows whether the two units add up.
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "device.hpp"
#include "kernels.hpp"

namespace ev::kern {

namespace {

typedef float F4 __attribute__((ext_vector_type(4)));

// One matrix-core step in the shape a 4x4 gate maps onto: f64 4x4x4 (4
// blocks of 4x4 outputs, K = 4) and f32 4x4x1 (16 blocks, K = 1); either is
// 256 multiply-adds per wave.
__device__ __forceinline__ double mfma_step(double a, double b, double c)
{
    return __builtin_amdgcn_mfma_f64_4x4x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ F4 mfma_step(float a, float b, F4 c) { return __builtin_amdgcn_mfma_f32_4x4x1f32(a, b, c, 0, 0, 0); }

template <class T> struct RateAcc;
template <> struct RateAcc<double> {
    using type = double;
    static __device__ double sum(double v) { return v; }
};
template <> struct RateAcc<float> {
    using type = F4;
    static __device__ float sum(F4 v) { return v[0] + v[1] + v[2] + v[3]; }
};

constexpr int kRateBlock = 256;
constexpr unsigned kRateBlocks = 3072;  // 12,288 waves: several rounds of every SIMD's slots

// NV independent vector multiply-add chains and NM independent matrix-core
// accumulations per thread, `iters` rounds.  v <- v x + y with x < 1 stays
// bounded; the accumulations grow by about 4 a b per round.  The sums are
// stored so no chain is dead code.
template <class T, int NV, int NM>
__global__ __launch_bounds__(kRateBlock) void k_rate(T *__restrict__ out, uint32_t iters, T x, T y)
{
    using A = typename RateAcc<T>::type;
    constexpr int nv = NV > 0 ? NV : 1, nm = NM > 0 ? NM : 1;
    T v[nv];
    A m[nm];
    const T t = T(threadIdx.x) * T(1e-6);
#pragma unroll
    for (int i = 0; i < nv; ++i) v[i] = t + T(i);
#pragma unroll
    for (int i = 0; i < nm; ++i) m[i] = A{} + T(i);
    const T a = x + t, b = y - t;
    for (uint32_t it = 0; it < iters; ++it) {
#pragma unroll
        for (int i = 0; i < NM; ++i) m[i] = mfma_step(a, b, m[i]);
#pragma unroll
        for (int i = 0; i < NV; ++i) v[i] = dfma(v[i], x, y);
    }
    T s = 0;
#pragma unroll
    for (int i = 0; i < NV; ++i) s += v[i];
#pragma unroll
    for (int i = 0; i < NM; ++i) s += RateAcc<T>::sum(m[i]);
    out[size_t(blockIdx.x) * kRateBlock + threadIdx.x] = s;
}

// TFLOP/s of one k_rate configuration: a short warm-up launch, then one
// timed launch between HIP events on the stream.  The timed launch runs for
// about 10 ms, as long as a fused pass at 30 qubits: under sustained
// multiply-add load the clock settles below its starting value (a 1.5 ms
// launch measured the f64 vector rate 15% higher), and the sustained rate is
// the one a pass sees.
template <class T, int NV, int NM>
double time_rate(T *out, uint32_t iters, hipStream_t st, hipEvent_t e0, hipEvent_t e1)
{
    hipLaunchKernelGGL((k_rate<T, NV, NM>), dim3(kRateBlocks), dim3(kRateBlock), 0, st, out, 16u, T(0.999), T(1e-3));
    EV_HIP(hipGetLastError());
    EV_HIP(hipEventRecord(e0, st));
    hipLaunchKernelGGL((k_rate<T, NV, NM>), dim3(kRateBlocks), dim3(kRateBlock), 0, st, out, iters, T(0.999), T(1e-3));
    EV_HIP(hipGetLastError());
    EV_HIP(hipEventRecord(e1, st));
    EV_HIP(hipEventSynchronize(e1));
    float ms = 0;
    EV_HIP(hipEventElapsedTime(&ms, e0, e1));
    const double waves = double(kRateBlocks) * (kRateBlock / 64);
    const double fma = waves * iters * (64.0 * NV + 256.0 * NM);  // multiply-adds per wave per round
    return 2.0 * fma / (double(ms) * 1e-3) / 1e12;
}

template <class T> ComputeRates measure(void *scratch, hipStream_t st, hipEvent_t e0, hipEvent_t e1)
{
    T *out = static_cast<T *>(scratch);
    ComputeRates r;
    // 32 vector chains keep every SIMD's issue busy; 16 matrix-core
    // accumulations cover the matrix pipeline's latency; the mixed run gives
    // both units equal flops (32 x 64 = 8 x 256 multiply-adds per wave per round).
    constexpr bool f32 = sizeof(T) == 4;
    r.valu_tflops = time_rate<T, 32, 0>(out, 12000, st, e0, e1);
    r.mfma_tflops = time_rate<T, 0, 16>(out, f32 ? 12000 : 6000, st, e0, e1);
    r.mixed_tflops = time_rate<T, 32, 8>(out, 6000, st, e0, e1);
    return r;
}

}  // namespace

size_t compute_rates_scratch_bytes() { return size_t(kRateBlocks) * kRateBlock * sizeof(double); }

ComputeRates compute_rates(bool f32, void *scratch, hipStream_t st, hipEvent_t e0, hipEvent_t e1)
{
    return f32 ? measure<float>(scratch, st, e0, e1) : measure<double>(scratch, st, e0, e1);
}

}  // namespace ev::kern
