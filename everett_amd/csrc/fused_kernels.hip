// The fused statevector pass kernel for gfx950 (descriptors in fused.hpp).
//
// One workgroup per tile of 2^t amplitudes (W = 2^(t-4) threads, 16
// amplitudes per thread).  Rows of the tile are contiguous in HBM because the
// lowest qubits are always tile qubits; the load puts consecutive lanes on
// consecutive amplitudes.  The tile lives in LDS (64 KiB at most, XOR-swizzled
// so every stage mapping reads conflict-free); each stage pulls 16 amplitudes
// per thread into registers, applies its ops with compile-time register
// indices (no scratch), and writes them back.  Matrix-vector products are
// chained fused multiply-adds (one rounding per step); results agree with the
// reference's per-product rounding to ~1e-15 relative per gate.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "common.hpp"
#include "device.hpp"
#include "fused.hpp"
#include "fused_kernels.hpp"

namespace ev::fused {

namespace {

template <class V> using Regs = V[kRegs];

// m0 * a0 + m1 * a1 as one fused chain per component (1 mul + 3 fma), which
// rounds once per step instead of rounding each product and then the sum.
template <class V> __device__ __forceinline__ V cmac2(V m0, V a0, V m1, V a1)
{
    V r;
    r.x = dfma(m0.x, a0.x, dfma(-m0.y, a0.y, dfma(m1.x, a1.x, -(m1.y * a1.y))));
    r.y = dfma(m0.x, a0.y, dfma(m0.y, a0.x, dfma(m1.x, a1.y, m1.y * a1.x)));
    return r;
}

// sum_c m[c] * a[c] over four terms, same chained form.
template <class V> __device__ __forceinline__ V cmac4(const V *m, V a0, V a1, V a2, V a3)
{
    V r;
    r.x = dfma(m[0].x, a0.x, dfma(-m[0].y, a0.y, dfma(m[1].x, a1.x, dfma(-m[1].y, a1.y,
          dfma(m[2].x, a2.x, dfma(-m[2].y, a2.y, dfma(m[3].x, a3.x, -(m[3].y * a3.y))))))));
    r.y = dfma(m[0].x, a0.y, dfma(m[0].y, a0.x, dfma(m[1].x, a1.y, dfma(m[1].y, a1.x,
          dfma(m[2].x, a2.y, dfma(m[2].y, a2.x, dfma(m[3].x, a3.y, m[3].y * a3.x)))))));
    return r;
}

template <class V, int B> __device__ __forceinline__ void op_u1(Regs<V> &v, V m0, V m1, V m2, V m3)
{
#pragma unroll
    for (int j = 0; j < kRegs; ++j) {
        if (j & (1 << B)) continue;
        const int j1 = j | (1 << B);
        const V a0 = v[j], a1 = v[j1];
        v[j] = cmac2(m0, a0, m1, a1);
        v[j1] = cmac2(m2, a0, m3, a1);
    }
}

template <class V, int B>
__device__ __forceinline__ void op_cu1(Regs<V> &v, V m0, V m1, V m2, V m3, uint32_t creg, bool tok)
{
#pragma unroll
    for (int j = 0; j < kRegs; ++j) {
        if (j & (1 << B)) continue;
        if ((uint32_t(j) & creg) != creg || !tok) continue;
        const int j1 = j | (1 << B);
        const V a0 = v[j], a1 = v[j1];
        v[j] = cmac2(m0, a0, m1, a1);
        v[j1] = cmac2(m2, a0, m3, a1);
    }
}

template <class V, int B> __device__ __forceinline__ void op_cx(Regs<V> &v, uint32_t creg, bool tok)
{
#pragma unroll
    for (int j = 0; j < kRegs; ++j) {
        if (j & (1 << B)) continue;
        if ((uint32_t(j) & creg) != creg || !tok) continue;
        const int j1 = j | (1 << B);
        const V t = v[j];
        v[j] = v[j1];
        v[j1] = t;
    }
}

template <class V, int B>
__device__ __forceinline__ void op_diagr(Regs<V> &v, V d0, V d1, bool d0_one, uint32_t creg, bool tok)
{
#pragma unroll
    for (int j = 0; j < kRegs; ++j) {
        if ((uint32_t(j) & creg) != creg || !tok) continue;
        if (j & (1 << B)) {
            v[j] = cmul(d1, v[j]);
        } else if (!d0_one) {
            v[j] = cmul(d0, v[j]);
        }
    }
}

template <class V> __device__ __forceinline__ void op_diagt(Regs<V> &v, V d, uint32_t creg, bool tok)
{
#pragma unroll
    for (int j = 0; j < kRegs; ++j) {
        if ((uint32_t(j) & creg) != creg || !tok) continue;
        v[j] = cmul(d, v[j]);
    }
}

// Dense 4x4 in the basis |b_hi b_lo> (b_hi = MSB).
template <class V, int BH, int BL> __device__ __forceinline__ void op_u2(Regs<V> &v, const V *__restrict__ m)
{
#pragma unroll
    for (int j = 0; j < kRegs; ++j) {
        if (j & ((1 << BH) | (1 << BL))) continue;
        const int i0 = j, i1 = j | (1 << BL), i2 = j | (1 << BH), i3 = j | (1 << BH) | (1 << BL);
        const V a0 = v[i0], a1 = v[i1], a2 = v[i2], a3 = v[i3];
        const int idx[4] = {i0, i1, i2, i3};
#pragma unroll
        for (int r = 0; r < 4; ++r) v[idx[r]] = cmac4(m + 4 * r, a0, a1, a2, a3);
    }
}

template <class V, int BH, int BL> __device__ __forceinline__ void op_swap(Regs<V> &v)
{
#pragma unroll
    for (int j = 0; j < kRegs; ++j) {
        if (j & ((1 << BH) | (1 << BL))) continue;
        const int a = j | (1 << BL), b = j | (1 << BH);
        const V t = v[a];
        v[a] = v[b];
        v[b] = t;
    }
}

#define EV_CASE_B(fn, ...)                  \
    switch (b0) {                           \
    case 0: fn<V, 0>(__VA_ARGS__); break;   \
    case 1: fn<V, 1>(__VA_ARGS__); break;   \
    case 2: fn<V, 2>(__VA_ARGS__); break;   \
    default: fn<V, 3>(__VA_ARGS__); break;  \
    }

#define EV_CASE_PAIR(fn, ...)                        \
    switch (b0 * 4 + b1) {                           \
    case 1 * 4 + 0: fn<V, 1, 0>(__VA_ARGS__); break; \
    case 2 * 4 + 0: fn<V, 2, 0>(__VA_ARGS__); break; \
    case 2 * 4 + 1: fn<V, 2, 1>(__VA_ARGS__); break; \
    case 3 * 4 + 0: fn<V, 3, 0>(__VA_ARGS__); break; \
    case 3 * 4 + 1: fn<V, 3, 1>(__VA_ARGS__); break; \
    default: fn<V, 3, 2>(__VA_ARGS__); break;        \
    }

template <class V>
__device__ __forceinline__ void run_op(const DevOp &op, Regs<V> &v, uint64_t gfix, const V *__restrict__ coef)
{
    const uint32_t kind = op.code & 0xff, b0 = (op.code >> 8) & 0xff, b1 = (op.code >> 16) & 0xff;
    const uint32_t flags = op.code >> 24;
    const V *m = (kind == kU2) ? coef + op.coef : reinterpret_cast<const V *>(op.m);
    const bool tok = (gfix & op.cglob) == op.cglob;
    switch (kind) {
    case kU1: {
        const V m0 = m[0], m1 = m[1], m2 = m[2], m3 = m[3];
        EV_CASE_B(op_u1, v, m0, m1, m2, m3);
        break;
    }
    case kCU1: {
        const V m0 = m[0], m1 = m[1], m2 = m[2], m3 = m[3];
        EV_CASE_B(op_cu1, v, m0, m1, m2, m3, op.creg, tok);
        break;
    }
    case kCX: EV_CASE_B(op_cx, v, op.creg, tok); break;
    case kDiagR: {
        const V d0 = m[0], d1 = m[1];
        const bool one = (flags & kD0IsOne) != 0;
        EV_CASE_B(op_diagr, v, d0, d1, one, op.creg, tok);
        break;
    }
    case kDiagT: {
        const V d = ((gfix >> op.tq) & 1) ? m[1] : m[0];
        op_diagt<V>(v, d, op.creg, tok);
        break;
    }
    case kU2: EV_CASE_PAIR(op_u2, v, m); break;
    default: EV_CASE_PAIR(op_swap, v); break;
    }
}

// One tile of one pass.  Threads beyond the tile's 2^(t-4) (a batch launch
// sizes its workgroups for the largest tile) skip the data movement but keep
// every barrier.
template <class R>
__device__ __forceinline__ void run_tile(Amp<R> *__restrict__ psi, const Pass *__restrict__ pass,
                                         const Stage *__restrict__ stages, const DevOp *__restrict__ ops,
                                         const Amp<R> *__restrict__ coef, uint64_t T, Amp<R> *lds)
{
    using V = Amp<R>;
    const int t = pass->t;
    const uint32_t tid = threadIdx.x;
    const uint32_t W = 1u << (t - kRegBits);
    const bool active = tid < W;

    // Tile base: deposit the tile index into the non-tile qubits.
    uint64_t base = 0;
    for (int i = 0; i < pass->ncomp; ++i) base |= ((T >> i) & 1ull) << pass->comp[i];

    // Load: local index l = j*W + tid (lanes on consecutive amplitudes).
    uint64_t gthr = 0;
    for (int i = 0; i < t - kRegBits; ++i)
        if ((tid >> i) & 1u) gthr |= 1ull << pass->q[i];
    if (active) {
#pragma unroll
        for (int j = 0; j < kRegs; ++j) {
            uint64_t gj = 0;
#pragma unroll
            for (int k = 0; k < kRegBits; ++k)
                if ((j >> k) & 1) gj |= 1ull << pass->q[t - kRegBits + k];
            lds[swz(uint32_t(j) * W + tid)] = psi[base | gthr | gj];
        }
    }
    __syncthreads();

    const int nst = pass->nstages;
    for (int s = 0; s < nst; ++s) {
        const Stage &st = stages[pass->stage_begin + s];
        if (active) {
            uint32_t lb = 0;
            uint64_t gfix = base;
            for (int i = 0; i < st.np; ++i)
                if ((tid >> i) & 1u) {
                    lb |= 1u << st.p[i];
                    gfix |= 1ull << st.gpos[i];
                }
            const uint32_t sb = swz(lb);
            Regs<V> v;
#pragma unroll
            for (int j = 0; j < kRegs; ++j) v[j] = lds[sb ^ st.roff[j]];
            for (int o = st.op_begin; o < st.op_end; ++o) run_op<V>(ops[o], v, gfix, coef);
#pragma unroll
            for (int j = 0; j < kRegs; ++j) lds[sb ^ st.roff[j]] = v[j];
        }
        __syncthreads();
    }

    if (active) {
#pragma unroll
        for (int j = 0; j < kRegs; ++j) {
            uint64_t gj = 0;
#pragma unroll
            for (int k = 0; k < kRegBits; ++k)
                if ((j >> k) & 1) gj |= 1ull << pass->q[t - kRegBits + k];
            psi[base | gthr | gj] = lds[swz(uint32_t(j) * W + tid)];
        }
    }
}

template <class R>
__global__ __launch_bounds__(kMaxThreads) void k_fused_pass(Amp<R> *__restrict__ psi, const Pass *__restrict__ pass,
                                                            const Stage *__restrict__ stages,
                                                            const DevOp *__restrict__ ops,
                                                            const Amp<R> *__restrict__ coef)
{
    extern __shared__ unsigned char lds_raw[];
    run_tile<R>(psi, pass, stages, ops, coef, blockIdx.x, reinterpret_cast<Amp<R> *>(lds_raw));
}

// Persistent batch kernel: linear tile ids over every entry of the round
// (entries sorted by tile_begin); workgroups stride over them.
template <class R>
__global__ __launch_bounds__(kMaxThreads) void k_fused_batch(const BatchEntry *__restrict__ entries, int nentries,
                                                             uint64_t total_tiles)
{
    extern __shared__ unsigned char lds_raw[];
    Amp<R> *lds = reinterpret_cast<Amp<R> *>(lds_raw);
    for (uint64_t lin = blockIdx.x; lin < total_tiles; lin += gridDim.x) {
        int lo = 0, hi = nentries - 1;  // last entry with tile_begin <= lin
        while (lo < hi) {
            const int mid = (lo + hi + 1) / 2;
            if (entries[mid].tile_begin <= lin) lo = mid; else hi = mid - 1;
        }
        const BatchEntry &e = entries[lo];
        run_tile<R>(static_cast<Amp<R> *>(e.psi), e.pass, e.stages, e.ops, static_cast<const Amp<R> *>(e.coef),
                    lin - e.tile_begin, lds);
        __syncthreads();  // the next tile's load overwrites this tile's LDS
    }
}

}  // namespace

void launch_pass(void *psi, bool c64, int n, int t, const Pass *pass, const Stage *stages, const DevOp *ops,
                 const void *coef, hipStream_t st)
{
    EV_REQUIRE(t >= kMinTile && t <= kMaxTile && t <= n, EV_E_ARG, "fused pass tile width out of range");
    const dim3 grid(static_cast<unsigned>(uint64_t(1) << (n - t)));
    const dim3 block(1u << (t - kRegBits));  // <= kMaxThreads = the kernel's launch bound
    if (!c64) {
        const size_t lds = (size_t(1) << t) * sizeof(double2);
        hipLaunchKernelGGL(k_fused_pass<double>, grid, block, lds, st, static_cast<double2 *>(psi), pass, stages, ops,
                           static_cast<const double2 *>(coef));
    } else {
        const size_t lds = (size_t(1) << t) * sizeof(float2);
        hipLaunchKernelGGL(k_fused_pass<float>, grid, block, lds, st, static_cast<float2 *>(psi), pass, stages, ops,
                           static_cast<const float2 *>(coef));
    }
}

void launch_batch(bool c64, const BatchEntry *entries, int nentries, uint64_t total_tiles, int max_t, unsigned grid,
                  hipStream_t st)
{
    EV_REQUIRE(max_t >= kMinTile && max_t <= kMaxTile, EV_E_ARG, "batch tile width out of range");
    if (nentries == 0 || total_tiles == 0) return;
    const dim3 block(1u << (max_t - kRegBits));
    const dim3 g(static_cast<unsigned>(std::min<uint64_t>(grid, total_tiles)));
    if (!c64)
        hipLaunchKernelGGL(k_fused_batch<double>, g, block, (size_t(1) << max_t) * sizeof(double2), st, entries, nentries,
                           total_tiles);
    else
        hipLaunchKernelGGL(k_fused_batch<float>, g, block, (size_t(1) << max_t) * sizeof(float2), st, entries, nentries,
                           total_tiles);
}

}  // namespace ev::fused
