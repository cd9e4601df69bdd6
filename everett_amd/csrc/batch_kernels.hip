// Kernels over a batch arena: many statevectors packed back to back (each a
// power-of-two run of amplitudes starting at a multiple of 256 amplitudes).
// Reductions are chunked by 256 amplitudes (a chunk never straddles two
// states) and combined per state in a fixed order, so results are
// deterministic.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "batch_kernels.hpp"
#include "device.hpp"

namespace ev::batchk {

namespace {

constexpr int kBlock = 256;  // == kChunk

#define GRID_STRIDE(i, count)                                                               \
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < (count);         \
         i += uint64_t(gridDim.x) * blockDim.x)

inline dim3 grid_for(uint64_t work)
{
    uint64_t b = (work + kBlock - 1) / kBlock;
    return dim3(static_cast<unsigned>(std::max<uint64_t>(1, std::min<uint64_t>(b, 1u << 16))));
}

__device__ __forceinline__ double2 widen(double2 a) { return a; }
__device__ __forceinline__ double2 widen(float2 a) { return make_double2(a.x, a.y); }

template <class R> __global__ void k_zero(Amp<R> *a, uint64_t len)
{
    GRID_STRIDE(i, len) a[i] = mk<Amp<R>>(R(0), R(0));
}

template <class R> __global__ void k_set_first(Amp<R> *a, const uint64_t *aoff, int count)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) a[aoff[i]] = mk<Amp<R>>(R(1), R(0));
}

// Block-wide sum of NV doubles; thread 0 gets the totals.
template <int NV> __device__ void block_reduce(double (&v)[NV], double *out)
{
    __shared__ double sh[NV][kBlock];
    for (int c = 0; c < NV; ++c) sh[c][threadIdx.x] = v[c];
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if (threadIdx.x < unsigned(s))
            for (int c = 0; c < NV; ++c) sh[c][threadIdx.x] += sh[c][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int c = 0; c < NV; ++c) out[c] = sh[c][0];
    __syncthreads();
}

// One wave per chunk: each lane sums 4 of the chunk's 256 |a|^2 (lane,
// lane + 64, ...), then a fixed butterfly of shuffles combines the lanes, so
// the partial needs no LDS or barrier and its rounding is deterministic.
template <class R> __global__ void k_chunk_norm(const Amp<R> *a, uint64_t nchunks, double *part)
{
    const uint64_t wave = (uint64_t(blockIdx.x) * blockDim.x + threadIdx.x) / 64;
    const uint64_t nwaves = uint64_t(gridDim.x) * blockDim.x / 64;
    const unsigned lane = threadIdx.x & 63;
    for (uint64_t c = wave; c < nchunks; c += nwaves) {
        const Amp<R> *p = a + c * kChunk + lane;
        double v = 0.0;
#pragma unroll
        for (int k = 0; k < kChunk / 64; ++k) {
            const double2 x = widen(p[64 * k]);
            v += __builtin_fma(x.x, x.x, x.y * x.y);
        }
#pragma unroll
        for (int m = 32; m > 0; m >>= 1) v += __shfl_xor(v, m, 64);
        if (lane == 0) part[c] = v;
    }
}

// Per chunk: <ref|x> (re, im), <x|x>, <ref|ref> and max |x - ref| (the max is
// exact in any order).
template <class R> __global__ void k_chunk_compare(const Amp<R> *x, const double2 *ref, uint64_t nchunks, double *part)
{
    __shared__ double mx[kBlock];
    for (uint64_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
        const double2 a = widen(x[c * kChunk + threadIdx.x]);
        const double2 r = ref[c * kChunk + threadIdx.x];
        double v[4];
        v[0] = __builtin_fma(r.x, a.x, r.y * a.y);   // re conj(r) * a
        v[1] = __builtin_fma(r.x, a.y, -(r.y * a.x));  // im conj(r) * a
        v[2] = __builtin_fma(a.x, a.x, a.y * a.y);
        v[3] = __builtin_fma(r.x, r.x, r.y * r.y);
        const double dx = a.x - r.x, dy = a.y - r.y;
        mx[threadIdx.x] = __builtin_sqrt(dx * dx + dy * dy);
        double out[4];
        block_reduce<4>(v, out);
        for (int s = kBlock / 2; s > 0; s >>= 1) {
            if (threadIdx.x < unsigned(s)) mx[threadIdx.x] = fmax(mx[threadIdx.x], mx[threadIdx.x + s]);
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            for (int k = 0; k < 4; ++k) part[5 * c + k] = out[k];
            part[5 * c + 4] = mx[0];
        }
        __syncthreads();
    }
}

// One thread per state: sequential sum (or max for slot `max_slot`) over the
// state's chunks.
__global__ void k_combine(const double *part, int nv, int max_slot, const uint64_t *chunk_begin, int count, double *out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    for (int k = 0; k < nv; ++k) {
        double s = 0.0;
        for (uint64_t c = chunk_begin[i]; c < chunk_begin[i + 1]; ++c)
            s = (k == max_slot) ? fmax(s, part[nv * c + k]) : s + part[nv * c + k];
        out[nv * i + k] = s;
    }
}

// Each workgroup scales a contiguous range of chunks; the state of the
// first chunk is found by binary search and later ones by stepping forward
// (chunks are in state order), four chunks per iteration so four loads per
// thread are in flight.
template <class R>
__global__ void k_scale_states(Amp<R> *a, uint64_t nchunks, const uint64_t *chunk_begin, int count, const double *scale)
{
    const uint64_t per = (nchunks + gridDim.x - 1) / gridDim.x;
    const uint64_t c0 = uint64_t(blockIdx.x) * per, c1 = c0 + per < nchunks ? c0 + per : nchunks;
    if (c0 >= c1) return;
    int lo = 0, hi = count - 1;  // last state with chunk_begin <= c0
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (chunk_begin[mid] <= c0) lo = mid; else hi = mid - 1;
    }
    int st = lo;
    uint64_t next = chunk_begin[st + 1];
    double s = scale[st];
    constexpr int U = 4;
    for (uint64_t c = c0; c < c1; c += U) {
        double sc[U];
        Amp<R> x[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            while (c + u >= next && st + 1 < count) {
                ++st;
                next = chunk_begin[st + 1];
                s = scale[st];
            }
            sc[u] = s;
            if (c + u < c1) x[u] = a[(c + u) * kChunk + threadIdx.x];
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (c + u < c1) a[(c + u) * kChunk + threadIdx.x] = mk<Amp<R>>(R(x[u].x * sc[u]), R(x[u].y * sc[u]));
    }
}

// One thread per state: the renormalisation factor of its f64 norm (IEEE
// square root and division, as on the host).
__global__ void k_inv_sqrt(const double *nrm, int count, double *scale)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) scale[i] = nrm[i] > 0 ? 1.0 / sqrt(nrm[i]) : 0.0;
}

}  // namespace

void inv_sqrt(const double *d_nrm, int count, double *d_scale, hipStream_t st)
{
    hipLaunchKernelGGL(k_inv_sqrt, dim3((count + 255) / 256), dim3(256), 0, st, d_nrm, count, d_scale);
}

void init(void *arena, bool c64, uint64_t total_amps, const uint64_t *d_aoff, int count, hipStream_t st)
{
    if (!c64) {
        hipLaunchKernelGGL(k_zero<double>, grid_for(total_amps), dim3(kBlock), 0, st, static_cast<double2 *>(arena), total_amps);
        hipLaunchKernelGGL(k_set_first<double>, dim3((count + 255) / 256), dim3(256), 0, st, static_cast<double2 *>(arena),
                           d_aoff, count);
    } else {
        hipLaunchKernelGGL(k_zero<float>, grid_for(total_amps), dim3(kBlock), 0, st, static_cast<float2 *>(arena), total_amps);
        hipLaunchKernelGGL(k_set_first<float>, dim3((count + 255) / 256), dim3(256), 0, st, static_cast<float2 *>(arena),
                           d_aoff, count);
    }
}

void norms(const void *arena, bool c64, uint64_t total_amps, const uint64_t *d_chunk_begin, int count, double *part,
           double *out, hipStream_t st)
{
    const uint64_t nchunks = total_amps / kChunk;
    const dim3 g(static_cast<unsigned>(std::min<uint64_t>((nchunks + 3) / 4, 16384)));  // 4 waves (chunks) per group
    if (!c64) hipLaunchKernelGGL(k_chunk_norm<double>, g, dim3(kBlock), 0, st, static_cast<const double2 *>(arena), nchunks, part);
    else hipLaunchKernelGGL(k_chunk_norm<float>, g, dim3(kBlock), 0, st, static_cast<const float2 *>(arena), nchunks, part);
    hipLaunchKernelGGL(k_combine, dim3((count + 63) / 64), dim3(64), 0, st, part, 1, -1, d_chunk_begin, count, out);
}

void compare(const void *arena, bool c64, const void *ref_arena, uint64_t total_amps, const uint64_t *d_chunk_begin,
             int count, double *part, double *out, hipStream_t st)
{
    const uint64_t nchunks = total_amps / kChunk;
    const dim3 g(static_cast<unsigned>(std::min<uint64_t>(nchunks, 8192)));
    const double2 *ref = static_cast<const double2 *>(ref_arena);
    if (!c64)
        hipLaunchKernelGGL(k_chunk_compare<double>, g, dim3(kBlock), 0, st, static_cast<const double2 *>(arena), ref, nchunks,
                           part);
    else
        hipLaunchKernelGGL(k_chunk_compare<float>, g, dim3(kBlock), 0, st, static_cast<const float2 *>(arena), ref, nchunks,
                           part);
    hipLaunchKernelGGL(k_combine, dim3((count + 63) / 64), dim3(64), 0, st, part, 5, 4, d_chunk_begin, count, out);
}

void scale(void *arena, bool c64, uint64_t total_amps, const uint64_t *d_chunk_begin, int count, const double *d_scale,
           hipStream_t st)
{
    const uint64_t nchunks = total_amps / kChunk;
    const dim3 g(static_cast<unsigned>(std::min<uint64_t>(nchunks, 8192)));
    if (!c64)
        hipLaunchKernelGGL(k_scale_states<double>, g, dim3(kBlock), 0, st, static_cast<double2 *>(arena), nchunks,
                           d_chunk_begin, count, d_scale);
    else
        hipLaunchKernelGGL(k_scale_states<float>, g, dim3(kBlock), 0, st, static_cast<float2 *>(arena), nchunks,
                           d_chunk_begin, count, d_scale);
}

}  // namespace ev::batchk
