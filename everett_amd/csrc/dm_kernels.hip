// Density-matrix kernels for gfx950: diagonal reductions, collapse and the
// one-qubit partial trace.  rho is row-major dim x dim c128 (dim = 2^n), i.e.
// a 2n-qubit vector whose element (i, j) sits at (i << n) | j.  Gate and
// channel application reuse the statevector kernels on that vector (a gate
// with its channel is one 4x4 superoperator on bits (n+t, t)).
#include <hip/hip_runtime.h>

#include "device.hpp"
#include "dm_kernels.hpp"

namespace ev::kern {

namespace {

constexpr int kBlock = 256;

#define GRID_STRIDE(i, count)                                                               \
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < (count);         \
         i += uint64_t(gridDim.x) * blockDim.x)

inline dim3 rgrid(uint64_t work)
{
    uint64_t b = (work + kBlock - 1) / kBlock;
    if (b < 1) b = 1;
    if (b > uint64_t(kReduceBlocks)) b = kReduceBlocks;
    return dim3(static_cast<unsigned>(b));
}

inline dim3 grid_for(uint64_t work)
{
    uint64_t b = (work + kBlock - 1) / kBlock;
    if (b < 1) b = 1;
    if (b > (1u << 16)) b = 1u << 16;
    return dim3(static_cast<unsigned>(b));
}

template <int NV> __device__ void block_sum(double (&v)[NV], double *partials)
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
        for (int c = 0; c < NV; ++c) partials[NV * blockIdx.x + c] = sh[c][0];
}

template <int NV> __global__ void k_finalize_dm(const double *partials, int nblocks, double *out)
{
    __shared__ double sh[NV][1024];
    for (int c = 0; c < NV; ++c) sh[c][threadIdx.x] = threadIdx.x < unsigned(nblocks) ? partials[NV * threadIdx.x + c] : 0.0;
    __syncthreads();
    for (int s = 512; s > 0; s >>= 1) {
        if (threadIdx.x < unsigned(s))
            for (int c = 0; c < NV; ++c) sh[c][threadIdx.x] += sh[c][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int c = 0; c < NV; ++c) out[c] = sh[c][0];
}

// density.rs:121-124 (mask = 0) and density_matrix.rs:175-186 (mask = bit q):
// sum of Re rho[j][j] over j with (j & mask) == mask.
__global__ void k_diag_partial(const double2 *rho, int n, uint64_t mask, uint64_t want, double *partials)
{
    double v[1] = {0.0};
    const uint64_t dim = uint64_t(1) << n;
    GRID_STRIDE(j, dim)
    {
        if ((j & mask) == want) v[0] += rho[(j << n) | j].x;
    }
    block_sum<1>(v, partials);
}

// density_matrix.rs:189-207: zero every entry whose row or column bit q differs
// from the outcome.
__global__ void k_dm_collapse(double2 *rho, int n, int q, int outcome)
{
    const uint64_t len = uint64_t(1) << (2 * n);
    GRID_STRIDE(x, len)
    {
        const bool ok = (((x >> (n + q)) & 1) == uint64_t(outcome)) && (((x >> q) & 1) == uint64_t(outcome));
        if (!ok) rho[x] = make_double2(0.0, 0.0);
    }
}

// k_dm_collapse, then every entry times s (density.rs:219-227 with s = 1 /
// trace): a zeroed entry stays +0, so the result equals the two passes'.
__global__ void k_dm_collapse_scale(double2 *rho, int n, int q, int outcome, double s)
{
    const uint64_t len = uint64_t(1) << (2 * n);
    GRID_STRIDE(x, len)
    {
        const bool ok = (((x >> (n + q)) & 1) == uint64_t(outcome)) && (((x >> q) & 1) == uint64_t(outcome));
        if (ok) {
            const double2 e = rho[x];
            rho[x] = make_double2(e.x * s, e.y * s);
        } else {
            rho[x] = make_double2(0.0, 0.0);
        }
    }
}

// A block of consecutive measurements works on a copy of the diagonal:
// diag[j] = Re rho[j][j].
__global__ void k_dm_extract_diag(const double2 *rho, int n, double *diag)
{
    const uint64_t dim = uint64_t(1) << n;
    GRID_STRIDE(j, dim) diag[j] = rho[(j << n) | j].x;
}

// k_diag_partial over the diagonal copy: the same threads add the same
// entries in the same order, so the sums equal the matrix's bit for bit.
__global__ void k_vec_partial(const double *diag, int n, uint64_t mask, uint64_t want, double *partials)
{
    double v[1] = {0.0};
    const uint64_t dim = uint64_t(1) << n;
    GRID_STRIDE(j, dim)
    {
        if ((j & mask) == want) v[0] += diag[j];
    }
    block_sum<1>(v, partials);
}

// One measurement on the diagonal copy: entries whose bit q differs from the
// outcome become 0, the others are multiplied by s (when scaled).
__global__ void k_diag_collapse(double *diag, int n, int q, int outcome, double s, int scaled)
{
    const uint64_t dim = uint64_t(1) << n;
    GRID_STRIDE(j, dim)
    {
        if (((j >> q) & 1) != uint64_t(outcome)) diag[j] = 0.0;
        else if (scaled) diag[j] = diag[j] * s;
    }
}

// Every measurement of a block in one pass over rho: an entry survives when
// its row and column bits match every outcome ((x & keep_mask) == keep_val),
// and is then multiplied by the block's scales in order (the value each
// measurement's own collapse-and-scale pass would leave); the rest become 0.
__global__ void k_dm_collapse_block(double2 *rho, int n, uint64_t keep_mask, uint64_t keep_val, int none,
                                    DmScales sc)
{
    const uint64_t len = uint64_t(1) << (2 * n);
    GRID_STRIDE(x, len)
    {
        if (!none && (x & keep_mask) == keep_val) {
            double2 e = rho[x];
            for (int k = 0; k < sc.count; ++k) e = make_double2(e.x * sc.s[k], e.y * sc.s[k]);
            rho[x] = e;
        } else {
            rho[x] = make_double2(0.0, 0.0);
        }
    }
}

// density.rs:270-288: reduced[iq*2+jq] = sum over the other bits of rho[i][j].
__global__ void k_ptrace_partial(const double2 *rho, int n, int q, double *partials)
{
    double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const uint64_t half = uint64_t(1) << (n - 1), bit = uint64_t(1) << q;
    GRID_STRIDE(x, half)
    {
        const uint64_t b = insert0(x, q);
#pragma unroll
        for (int iq = 0; iq < 2; ++iq)
#pragma unroll
            for (int jq = 0; jq < 2; ++jq) {
                const uint64_t i = b | (iq ? bit : 0), j = b | (jq ? bit : 0);
                const double2 e = rho[(i << n) | j];
                v[2 * (iq * 2 + jq)] += e.x;
                v[2 * (iq * 2 + jq) + 1] += e.y;
            }
    }
    block_sum<8>(v, partials);
}

}  // namespace

void dm_diag_sum(const void *rho, int n, uint64_t mask, uint64_t want, double *partials, double *out, hipStream_t st)
{
    const dim3 g = rgrid(uint64_t(1) << n);
    hipLaunchKernelGGL(k_diag_partial, g, dim3(kBlock), 0, st, static_cast<const double2 *>(rho), n, mask, want,
                       partials);
    hipLaunchKernelGGL(k_finalize_dm<1>, dim3(1), dim3(1024), 0, st, partials, int(g.x), out);
}

void dm_extract_diag(const void *rho, int n, double *diag, hipStream_t st)
{
    hipLaunchKernelGGL(k_dm_extract_diag, grid_for(uint64_t(1) << n), dim3(kBlock), 0, st,
                       static_cast<const double2 *>(rho), n, diag);
}

void diag_vec_sum(const double *diag, int n, uint64_t mask, uint64_t want, double *partials, double *out, hipStream_t st)
{
    const dim3 g = rgrid(uint64_t(1) << n);
    hipLaunchKernelGGL(k_vec_partial, g, dim3(kBlock), 0, st, diag, n, mask, want, partials);
    hipLaunchKernelGGL(k_finalize_dm<1>, dim3(1), dim3(1024), 0, st, partials, int(g.x), out);
}

void diag_collapse(double *diag, int n, int q, int outcome, double s, bool scaled, hipStream_t st)
{
    hipLaunchKernelGGL(k_diag_collapse, grid_for(uint64_t(1) << n), dim3(kBlock), 0, st, diag, n, q, outcome, s,
                       scaled ? 1 : 0);
}

void dm_collapse_block(void *rho, int n, uint64_t keep_mask, uint64_t keep_val, bool none, const DmScales &sc,
                       hipStream_t st)
{
    hipLaunchKernelGGL(k_dm_collapse_block, grid_for(uint64_t(1) << (2 * n)), dim3(kBlock), 0, st,
                       static_cast<double2 *>(rho), n, keep_mask, keep_val, none ? 1 : 0, sc);
}

void dm_collapse_scale(void *rho, int n, int q, int outcome, double s, hipStream_t st)
{
    hipLaunchKernelGGL(k_dm_collapse_scale, grid_for(uint64_t(1) << (2 * n)), dim3(kBlock), 0, st,
                       static_cast<double2 *>(rho), n, q, outcome, s);
}

void dm_collapse(void *rho, int n, int q, int outcome, hipStream_t st)
{
    hipLaunchKernelGGL(k_dm_collapse, grid_for(uint64_t(1) << (2 * n)), dim3(kBlock), 0, st, static_cast<double2 *>(rho), n,
                       q, outcome);
}

void dm_partial_trace(const void *rho, int n, int q, double *partials, double *out, hipStream_t st)
{
    const dim3 g = rgrid(uint64_t(1) << (n - 1));
    hipLaunchKernelGGL(k_ptrace_partial, g, dim3(kBlock), 0, st, static_cast<const double2 *>(rho), n, q, partials);
    hipLaunchKernelGGL(k_finalize_dm<8>, dim3(1), dim3(1024), 0, st, partials, int(g.x), out);
}

}  // namespace ev::kern
