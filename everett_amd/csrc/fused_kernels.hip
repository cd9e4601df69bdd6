// The fused statevector pass kernel for gfx950 (descriptors in fused.hpp).
//
// A tile is 2^t amplitudes handled by one workgroup of W = 2^(t-4) threads,
// 16 amplitudes per thread.  A tile arrives straight in its first stage's
// register mapping and leaves straight from its last stage's; only the
// exchanges between stages go through LDS (64 KiB at most, XOR-swizzled so
// the stage mappings read and write conflict-free).  Ops act on 16 registers
// with compile-time register indices (no scratch).  Matrix-vector products
// are chained fused multiply-adds (one rounding per step); results agree with
// the reference's per-product rounding to ~1e-15 relative per gate.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "common.hpp"
#include "device.hpp"
#include "fused.hpp"
#include "fused_kernels.hpp"

namespace ev::fused {

namespace {

template <class V> using Regs = V[kRegs];

// Op bodies.  Every op reads the thread's 16 amplitudes from `in` and
// writes all 16 to `out` (the op loop alternates two register arrays).
// Writing out of place lets every op kind's code leave its results in the
// same registers, so the loop carries both arrays without copies; in-place
// updates made the register allocator give each op kind's results other
// registers than the loop head's and copy the whole array back after every
// op.  Each output component is one fused multiply-add chain (one rounding
// per step) whose last step adds the replaced input's own term.

// (o0, o1) <- (m0 a0 + m1 a1, m2 a0 + m3 a1).  Per output component one
// chain: o0.re = m0.re a0.re + (-m0.im a0.im + (m1.re a1.re + -(m1.im a1.im)))
// and likewise.
template <class V>
__device__ __forceinline__ void mix2(const V &a0, const V &a1, V &o0, V &o1, V m0, V m1, V m2, V m3)
{
    const auto p0x = dfma(-m0.y, a0.y, dfma(m1.x, a1.x, -(m1.y * a1.y)));
    const auto p0y = dfma(m0.y, a0.x, dfma(m1.x, a1.y, m1.y * a1.x));
    const auto p1x = dfma(m2.x, a0.x, dfma(-m2.y, a0.y, -(m3.y * a1.y)));
    const auto p1y = dfma(m2.x, a0.y, dfma(m2.y, a0.x, m3.y * a1.x));
    o0 = mk<V>(dfma(m0.x, a0.x, p0x), dfma(m0.x, a0.y, p0y));
    o1 = mk<V>(dfma(m3.x, a1.x, p1x), dfma(m3.x, a1.y, p1y));
}

// o_r <- sum_c m[4r + c] a_c for a group of four registers.
template <class V>
__device__ __forceinline__ void mix4(const V &a0, const V &a1, const V &a2, const V &a3, V &o0, V &o1, V &o2, V &o3,
                                     const V *m)
{
    const V x[4] = {a0, a1, a2, a3};
    V y[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const V *row = m + 4 * r;
        decltype(V::x) ax = -(row[r].y * x[r].y), ay = row[r].y * x[r].x;
#pragma unroll
        for (int c = 3; c >= 0; --c) {
            if (c == r) continue;
            ax = dfma(row[c].x, x[c].x, dfma(-row[c].y, x[c].y, ax));
            ay = dfma(row[c].x, x[c].y, dfma(row[c].y, x[c].x, ay));
        }
        y[r] = mk<V>(dfma(row[r].x, x[r].x, ax), dfma(row[r].x, x[r].y, ay));
    }
    o0 = y[0];
    o1 = y[1];
    o2 = y[2];
    o3 = y[3];
}

// Workgroup barrier for the LDS exchanges between stages.  It orders LDS
// only: __syncthreads() would also wait for this thread's outstanding global
// stores and loads (a full workgroup fence), which would serialise a tile's
// HBM write-back and the next tile's gather behind every barrier.
__device__ __forceinline__ void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Per-component value select (a select of the two values, never of their
// addresses, so register arrays stay in registers).
template <class V> __device__ __forceinline__ V pick(bool c, V a, V b)
{
    return mk<V>(c ? a.x : b.x, c ? a.y : b.y);
}

// Identity where the lane's global condition fails: conditional ops then run
// unconditionally (an exact multiply by 1), with no per-register selects.
template <class V> __device__ __forceinline__ V one_unless(bool tok, V m) { return pick(tok, m, mk<V>(1, 0)); }
template <class V> __device__ __forceinline__ V zero_unless(bool tok, V m) { return pick(tok, m, mk<V>(0, 0)); }

// Register conditions: (j & creg) == creg, decided per register by a
// uniform branch; registers an op leaves alone are copied.
template <class V, int B>
__device__ __forceinline__ void op_u1(const Regs<V> &in, Regs<V> &out, V m0, V m1, V m2, V m3, uint32_t creg)
{
#pragma unroll
    for (int j = 0; j < kRegs; ++j) {
        if (j & (1 << B)) continue;
        const int j1 = j | (1 << B);
        if ((uint32_t(j) & creg) == creg) {
            mix2(in[j], in[j1], out[j], out[j1], m0, m1, m2, m3);
        } else {
            out[j] = in[j];
            out[j1] = in[j1];
        }
    }
}

// A register value the compiler cannot trace back to its array element (an
// empty asm, no instruction): a select between two such values stays a
// select of values.  Selecting between the array elements themselves is
// turned into a load through a selected address, which moves both register
// arrays to scratch memory.
template <class V> __device__ __forceinline__ V opaque(V a)
{
    asm("" : "+v"(a.x), "+v"(a.y));
    return a;
}

// Multi-controlled X: exchange the pair where the register condition holds
// and the lane's global condition tok holds.
template <class V, int B> __device__ __forceinline__ void op_cx(const Regs<V> &in, Regs<V> &out, uint32_t creg, bool tok)
{
#pragma unroll
    for (int j = 0; j < kRegs; ++j) {
        if (j & (1 << B)) continue;
        const int j1 = j | (1 << B);
        if ((uint32_t(j) & creg) == creg) {
            const V a = opaque(in[j]), b = opaque(in[j1]);
            out[j] = pick(tok, b, a);
            out[j1] = pick(tok, a, b);
        } else {
            out[j] = in[j];
            out[j1] = in[j1];
        }
    }
}

template <class V, int B> __device__ __forceinline__ void op_x(const Regs<V> &in, Regs<V> &out, uint32_t creg)
{
#pragma unroll
    for (int j = 0; j < kRegs; ++j) {
        if (j & (1 << B)) continue;
        const int j1 = j | (1 << B);
        const bool c = (uint32_t(j) & creg) == creg;
        const V a = opaque(in[j]), b = opaque(in[j1]);
        out[j] = pick(c, b, a);
        out[j1] = pick(c, a, b);
    }
}

// a <- d a with the reference's complex product (src/complex.rs:115-125).
template <class V, int B>
__device__ __forceinline__ void op_diagr(const Regs<V> &in, Regs<V> &out, V d0, V d1, bool d0_one, uint32_t creg)
{
#pragma unroll
    for (int j = 0; j < kRegs; ++j) {
        if ((uint32_t(j) & creg) != creg) out[j] = in[j];
        else if (j & (1 << B)) out[j] = cmul(d1, in[j]);
        else if (!d0_one) out[j] = cmul(d0, in[j]);
        else out[j] = in[j];
    }
}

template <class V> __device__ __forceinline__ void op_diagt(const Regs<V> &in, Regs<V> &out, V d, uint32_t creg)
{
#pragma unroll
    for (int j = 0; j < kRegs; ++j) {
        if ((uint32_t(j) & creg) == creg) out[j] = cmul(d, in[j]);
        else out[j] = in[j];
    }
}

// Dense 4x4 in the basis |b_hi b_lo> (b_hi = MSB).
template <class V, int BH, int BL> __device__ __forceinline__ void op_u2(const Regs<V> &in, Regs<V> &out, const V *m)
{
#pragma unroll
    for (int j = 0; j < kRegs; ++j) {
        if (j & ((1 << BH) | (1 << BL))) continue;
        const int a = j, b = j | (1 << BL), c = j | (1 << BH), d = j | (1 << BH) | (1 << BL);
        mix4(in[a], in[b], in[c], in[d], out[a], out[b], out[c], out[d], m);
    }
}

// Sparse 4x4 (kU2X): s = m00 m03 m11 m22 m30 m33 of a gate in the basis
// |b_hi b_lo> whose other entries are zero, so |00> and |11> mix and |01>
// and |10> only scale: 6 multiply-adds per amplitude instead of 16.
template <class V> __device__ __forceinline__ V dot2(V m0, V x0, V m1, V x1)
{
    decltype(V::x) ax = -(m0.y * x0.y), ay = m0.y * x0.x;
    ax = dfma(m1.x, x1.x, dfma(-m1.y, x1.y, ax));
    ay = dfma(m1.x, x1.y, dfma(m1.y, x1.x, ay));
    return mk<V>(dfma(m0.x, x0.x, ax), dfma(m0.x, x0.y, ay));
}

template <class V, int BH, int BL>
__device__ __forceinline__ void op_u2x(const Regs<V> &in, Regs<V> &out, V s0, V s1, V s2, V s3, V s4, V s5)
{
#pragma unroll
    for (int j = 0; j < kRegs; ++j) {
        if (j & ((1 << BH) | (1 << BL))) continue;
        const int a = j, b = j | (1 << BL), c = j | (1 << BH), d = j | (1 << BH) | (1 << BL);
        out[a] = dot2(s0, in[a], s1, in[d]);
        out[b] = cmul(s2, in[b]);
        out[c] = cmul(s3, in[c]);
        out[d] = dot2(s4, in[a], s5, in[d]);
    }
}

// Complex64 dense 4x4 on the matrix cores: v_mfma_f32_4x4x1_16b_f32 runs 16
// independent 4x4 outer products, block b = lanes 4b..4b+3, where lane 4b+j
// holds column j of B (one float) and receives column j of D (four floats),
// and lane 4b+i holds row i of A.  With A = column k of the gate's real 8x8
// form (lane 4b+i holding row i) and B = input component k of the lane's own
// group, D accumulates the lane's own four output components: each lane
// multiplies its own groups, no amplitude crosses lanes.  Eight rank-1 steps
// per output half (re: Re M x re - Im M x im, im: Im M x re + Re M x im), 64
// MFMAs per op per wave at twice the unpacked VALU's FP32 rate.
typedef float EvF4 __attribute__((ext_vector_type(4)));

// The lane's row of a complex64 gate (the A operands), loaded ahead of the op.
struct U2Row {
    float ar[4], ai[4];
};

template <class V, int BH, int BL>
__device__ __forceinline__ void op_u2_mfma(const Regs<V> &in, Regs<V> &out, const U2Row &r)
{
    constexpr int kG = kRegs / 4;
    int g[kG][4];
    {
        int n = 0;
#pragma unroll
        for (int j = 0; j < kRegs; ++j) {
            if (j & ((1 << BH) | (1 << BL))) continue;
            g[n][0] = j;
            g[n][1] = j | (1 << BL);
            g[n][2] = j | (1 << BH);
            g[n][3] = j | (1 << BH) | (1 << BL);
            ++n;
        }
    }
    // k outer, groups inner: eight independent accumulation chains in flight
    EvF4 ore[kG], oim[kG];
#pragma unroll
    for (int q = 0; q < kG; ++q) {
        ore[q] = EvF4{0.f, 0.f, 0.f, 0.f};
        oim[q] = EvF4{0.f, 0.f, 0.f, 0.f};
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
#pragma unroll
        for (int q = 0; q < kG; ++q) {
            ore[q] = __builtin_amdgcn_mfma_f32_4x4x1f32(r.ar[k], in[g[q][k]].x, ore[q], 0, 0, 0);
            oim[q] = __builtin_amdgcn_mfma_f32_4x4x1f32(r.ai[k], in[g[q][k]].x, oim[q], 0, 0, 0);
        }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const float an = -r.ai[k];
#pragma unroll
        for (int q = 0; q < kG; ++q) {
            ore[q] = __builtin_amdgcn_mfma_f32_4x4x1f32(an, in[g[q][k]].y, ore[q], 0, 0, 0);
            oim[q] = __builtin_amdgcn_mfma_f32_4x4x1f32(r.ar[k], in[g[q][k]].y, oim[q], 0, 0, 0);
        }
    }
#pragma unroll
    for (int q = 0; q < kG; ++q)
#pragma unroll
        for (int i = 0; i < 4; ++i) out[g[q][i]] = mk<V>(ore[q][i], oim[q][i]);
}

template <class V, int BH, int BL> __device__ __forceinline__ void op_swap(const Regs<V> &in, Regs<V> &out)
{
#pragma unroll
    for (int j = 0; j < kRegs; ++j) {
        if (j & ((1 << BH) | (1 << BL))) continue;
        const int a = j | (1 << BL), b = j | (1 << BH);
        out[j] = in[j];
        out[a] = in[b];
        out[b] = in[a];
        out[a | b] = in[a | b];
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

// Descriptors are device memory that the host fills before a launch and no
// kernel writes, so kernels read them through the constant address space:
// uniform reads become scalar loads, which the compiler may keep in flight
// across the amplitude stores.
#define EV_CONST __attribute__((address_space(4)))
template <class T> using cptr = const EV_CONST T *;
template <class T> __device__ __forceinline__ cptr<T> as_const(const T *p) { return (cptr<T>)(p); }

// The header fields of one op descriptor, held in scalar registers.  The op
// loop fetches both ops of a pair up front; an op's inline coefficients are
// read by the arm that uses them (inline_m), since holding both ops' 2x2
// matrices in scalar registers as well spills them to vector lanes.
template <class V> struct OpView {
    uint32_t code, creg;
    uint64_t cglob;
    uint32_t coef, tq;
    cptr<DevOp> p;
};

template <class V> __device__ __forceinline__ OpView<V> fetch_op(cptr<DevOp> p)
{
    OpView<V> o;
    o.p = p;
    o.code = p->code;
    o.creg = p->creg;
    o.cglob = p->cglob;
    o.coef = p->coef;
    o.tq = p->tq;
    return o;
}

// Inline coefficient k of an op (a 2x2 matrix or two diagonal entries).
template <class V> __device__ __forceinline__ V inline_m(const OpView<V> &op, int k)
{
    using F = decltype(V::x);
    const cptr<F> f = (cptr<F>)(op.p->m);
    return mk<V>(f[2 * k], f[2 * k + 1]);
}

template <class V> __device__ __forceinline__ void load_row(cptr<V> m, V *out)
{
    using F = decltype(V::x);
    const cptr<F> f = (cptr<F>)(m);
#pragma unroll
    for (int k = 0; k < 16; ++k) out[k] = mk<V>(f[2 * k], f[2 * k + 1]);
}

// M (complex64 only): dense 4x4 gates run on the matrix cores (op_u2_mfma)
// instead of VALU fused multiply-add chains.
template <class V, bool M> constexpr bool kMfmaU2 = M && std::is_same_v<decltype(V::x), float>;

// The lane's row (lane mod 4) of a kU2 gate for op_u2_mfma; nothing otherwise.
template <class V, bool M> __device__ __forceinline__ U2Row u2_row(const OpView<V> &op, cptr<V> coef)
{
    U2Row r{};
    if constexpr (kMfmaU2<V, M>) {
        if ((op.code & 0xff) == kU2) {
            const cptr<float> m = (cptr<float>)(coef + op.coef + 4 * (threadIdx.x & 3u));
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                r.ar[k] = m[2 * k];
                r.ai[k] = m[2 * k + 1];
            }
        }
    }
    return r;
}

// X: the pass may hold kU2X ops (density-matrix passes); other passes run a
// kernel without that arm.
template <class V, bool M, bool X = false>
__device__ __forceinline__ void run_op(const OpView<V> &op, const Regs<V> &in, Regs<V> &out, uint64_t gfix,
                                       cptr<V> coef, const U2Row &row)
{
    const uint32_t kind = op.code & 0xff, b0 = (op.code >> 8) & 0xff, b1 = (op.code >> 16) & 0xff;
    const uint32_t flags = op.code >> 24;
    // the lane's global condition, evaluated in the arms that have one (a
    // dense 4x4 never does)
    const auto gok = [&] { return (gfix & op.cglob) == op.cglob; };
    switch (kind) {
    case kU1: EV_CASE_B(op_u1, in, out, inline_m(op, 0), inline_m(op, 1), inline_m(op, 2), inline_m(op, 3), 0u); break;
    case kCU1: {
        // a lane whose global condition fails multiplies by the identity (exact)
        const bool tok = gok();
        const V m0 = one_unless(tok, inline_m(op, 0)), m1 = zero_unless(tok, inline_m(op, 1));
        const V m2 = zero_unless(tok, inline_m(op, 2)), m3 = one_unless(tok, inline_m(op, 3));
        EV_CASE_B(op_u1, in, out, m0, m1, m2, m3, op.creg);
        break;
    }
    case kCX:
        if (op.cglob == 0) {
            EV_CASE_B(op_x, in, out, op.creg);
        } else {
            const bool tok = gok();
            EV_CASE_B(op_cx, in, out, op.creg, tok);
        }
        break;
    case kDiagR: {
        const bool one = (flags & kD0IsOne) != 0, tok = gok();
        const V d0 = one_unless(tok, inline_m(op, 0)), d1 = one_unless(tok, inline_m(op, 1));
        EV_CASE_B(op_diagr, in, out, d0, d1, one, op.creg);
        break;
    }
    case kDiagT: {
        const bool tok = gok();
        const V d = one_unless(tok, pick(((gfix >> op.tq) & 1) != 0, inline_m(op, 1), inline_m(op, 0)));
        op_diagt<V>(in, out, d, op.creg);
        break;
    }
    case kU2: {
        if constexpr (kMfmaU2<V, M>) {
            EV_CASE_PAIR(op_u2_mfma, in, out, row);
        } else {
            V m[16];
            load_row<V>(coef + op.coef, m);
            EV_CASE_PAIR(op_u2, in, out, m);
        }
        break;
    }
    case kPhase: {
        // the lane's factor: product of the factors whose mask it meets
        const cptr<uint64_t> masks = (cptr<uint64_t>)(op.p->m);
        using F = decltype(V::x);
        const cptr<F> ff = (cptr<F>)(coef + op.coef);
        V f = pick((gfix & masks[0]) == masks[0], mk<V>(ff[0], ff[1]), mk<V>(1, 0));
        for (uint32_t i = 1; i < b1; ++i) {
            const uint64_t mk_i = masks[i];
            const V g = cmul(f, mk<V>(ff[2 * i], ff[2 * i + 1]));
            f = pick((gfix & mk_i) == mk_i, g, f);
        }
        if (b0 == kPhaseAll) {
            op_diagt<V>(in, out, f, op.creg);
        } else {
            EV_CASE_B(op_diagr, in, out, mk<V>(1, 0), f, true, op.creg);
        }
        break;
    }
    case (X ? uint32_t(kU2X) : 0x100u): {  // no such arm (kind < 0x100) unless X
        // the six coefficients go by value: passed as a local array they
        // left this kernel with 464 bytes of scratch per lane
        using F = decltype(V::x);
        const cptr<F> f = (cptr<F>)(coef + op.coef);
        const auto e = [&](int k) { return mk<V>(f[2 * k], f[2 * k + 1]); };
        EV_CASE_PAIR(op_u2x, in, out, e(0), e(1), e(2), e(3), e(4), e(5));
        break;
    }
    default: EV_CASE_PAIR(op_swap, in, out); break;
    }
}

// One tile's place in its pass: the state, the pass program, and the global
// index bits of the tile (the tile index deposited into the non-tile qubits).
template <class R> struct TileRef {
    Amp<R> *psi;
    cptr<Stage> stages;
    cptr<DevOp> ops;
    cptr<Amp<R>> coef;
    uint64_t base;
    int t, nstages;
    uint32_t swz_a, swz_b;  // the pass's LDS slot swizzle
};

// Inserts a zero bit at every tile qubit (ascending), which spreads the tile
// index over the non-tile qubits.
__device__ __forceinline__ uint64_t tile_base(cptr<Pass> pass, uint64_t T)
{
    uint64_t b = T;
    for (int i = 0; i < pass->t; ++i) b = insert0(b, pass->q[i]);
    return b;
}

template <class R>
__device__ __forceinline__ TileRef<R> tile_ref(Amp<R> *psi, cptr<Pass> pass, cptr<Stage> stages, cptr<DevOp> ops,
                                               cptr<Amp<R>> coef, uint64_t T)
{
    TileRef<R> r;
    r.psi = psi;
    r.stages = stages + pass->stage_begin;
    r.ops = ops;
    r.coef = coef;
    r.base = tile_base(pass, T);
    r.t = pass->t;
    r.nstages = pass->nstages;
    r.swz_a = pass->swz_a;
    r.swz_b = pass->swz_b;
    return r;
}

// A thread's share of a stage mapping: its thread-id bits placed at the
// stage's non-register positions, as an LDS slot (the local index under the
// pass's swizzle) and as global index bits (tile base included).
struct Lane {
    uint32_t sb;
    uint64_t gfix;
};

// Branch-free over every thread-id bit a tile can have: bits at and above
// np are masked off (threads beyond a narrower tile alias into it), and the
// planner zero-fills the unused entries, so the descriptor reads are two
// wide scalar loads with no wait inside a loop.
__device__ __forceinline__ Lane lane_of(cptr<Stage> st, uint32_t tid, uint64_t base, uint32_t swz_a, uint32_t swz_b,
                                       bool with_global = true)
{
    const uint32_t tm = tid & ((1u << st->np) - 1u);
    uint32_t lb = 0;
    uint64_t g = base;
#pragma unroll
    for (int i = 0; i < kMaxTile - kRegBits; ++i) lb |= ((tm >> i) & 1u) << st->p[i];
    if (with_global) {
#pragma unroll
        for (int i = 0; i < kMaxTile - kRegBits; ++i) g |= uint64_t((tm >> i) & 1u) << st->gpos[i];
    }
    return {swz(lb, swz_a, swz_b), g};
}

// Global gather of a tile in its first stage's mapping.  The planner keeps
// that stage's register bits off the lowest local positions, so 2^3 (c128)
// or 2^4 (c64) consecutive lanes read one contiguous 128-byte run.  A
// register's offset goff[j] and the lane's global bits occupy disjoint
// positions, so each access is the lane's address plus a uniform offset (one
// 64-bit add instead of an OR of both halves and a scaled add; measured
// 0.4% faster passes at 30 qubits complex128).
template <class R>
__device__ __forceinline__ void load_tile(const TileRef<R> &tr, uint32_t tid, Regs<Amp<R>> &v)
{
    const cptr<Stage> st = tr.stages;
    const Lane ln = lane_of(st, tid, tr.base, tr.swz_a, tr.swz_b);
    const Amp<R> *lp = tr.psi + ln.gfix;
#pragma unroll
    for (int j = 0; j < kRegs; ++j) v[j] = lp[st->goff[j]];
}

// The stage's ops alternate between v and w, two at a time (the next op's
// descriptor loads are in flight while the current op computes); an odd
// last op is followed by one copy back to v.  (Exchanging straight from w
// instead keeps both arrays live across the exchange, which spills.)
template <class V, bool M, bool X = false>
__device__ __forceinline__ void run_stage_ops(cptr<Stage> st, Regs<V> &v, Regs<V> &w, uint64_t gfix, cptr<DevOp> ops,
                                              cptr<V> coef)
{
    const int ob = __builtin_amdgcn_readfirstlane(st->op_begin), oe = __builtin_amdgcn_readfirstlane(st->op_end);
    int o = ob;
    for (; o + 1 < oe; o += 2) {
        const OpView<V> a = fetch_op<V>(ops + o), b = fetch_op<V>(ops + o + 1);
        // matrix-core gates: both rows are in flight before the first op runs
        const U2Row ra = u2_row<V, M>(a, coef), rb = u2_row<V, M>(b, coef);
        run_op<V, M, X>(a, v, w, gfix, coef, ra);
        run_op<V, M, X>(b, w, v, gfix, coef, rb);
    }
    if (o < oe) {
        const OpView<V> a = fetch_op<V>(ops + o);
        run_op<V, M, X>(a, v, w, gfix, coef, u2_row<V, M>(a, coef));
#pragma unroll
        for (int j = 0; j < kRegs; ++j) v[j] = w[j];
    }
}

// Exchanges move the real parts, then the imaginary parts, through one LDS
// buffer of 2^t reals: half the LDS a whole-amplitude buffer takes.  That
// keeps the register limit on resident waves: 12 per CU at complex128
// (instead of 10 at 10-qubit tiles and 8 at 12-qubit tiles) and 24 at
// complex64 with 12-qubit tiles (instead of 20).  Measured on MI355X at 30
// qubits: complex128 +19% and complex64 +6% gates/s.
template <class R> constexpr bool kSplitExchange = true;
template <class R> constexpr size_t kLdsSlotBytes = kSplitExchange<R> ? sizeof(R) : sizeof(Amp<R>);

// LDS byte addresses as 32-bit integers (local address space).
template <class T> using lptr = __attribute__((address_space(3))) T *;
__device__ __forceinline__ uint32_t lds_addr(void *p)
{
    return static_cast<uint32_t>(reinterpret_cast<uintptr_t>((lptr<unsigned char>)(p)));
}
template <class T> __device__ __forceinline__ lptr<T> lds_ptr(uint32_t a) { return (lptr<T>)(uintptr_t(a)); }

// LDS exchange between two stages: this thread's registers leave through
// slots ws ^ wroff[j] and the next stage's arrive from slots rs ^ rroff[j].
// The slot maps are linear, so byte addresses are the XOR of the shifted
// thread and register parts; each is computed once and serves both halves
// of a split exchange.
template <class R>
__device__ __forceinline__ void exchange(Regs<Amp<R>> &v, void *lds, bool active,
                                         uint32_t ws, cptr<uint32_t> wroff, uint32_t rs, cptr<uint32_t> rroff)
{
    using V = Amp<R>;
    constexpr uint32_t sh = kLdsSlotBytes<R> == 4 ? 2 : kLdsSlotBytes<R> == 8 ? 3 : 4;
    static_assert((1u << sh) == kLdsSlotBytes<R>, "LDS slot size");
    // The buffer is the kernel's only LDS allocation, so its base address is
    // 0 and base + (x ^ y) = (base + x) ^ y for every slot offset: the base
    // and the thread's part are added once, and each register costs one XOR
    // with its (uniform) shifted register part.
    uint32_t wb = lds_addr(lds) + (ws << sh), rb = lds_addr(lds) + (rs << sh);
    asm("" : "+v"(wb), "+v"(rb));  // keeps the shifts from being folded back into every address
    uint32_t wa[kRegs], ra[kRegs];
#pragma unroll
    for (int j = 0; j < kRegs; ++j) {
        wa[j] = wb ^ (wroff[j] << sh);
        ra[j] = rb ^ (rroff[j] << sh);
    }
    if constexpr (kSplitExchange<R>) {
        // every real part is in LDS before the first read, so the reads land
        // in v directly
        if (active) {
#pragma unroll
            for (int j = 0; j < kRegs; ++j) *lds_ptr<R>(wa[j]) = v[j].x;
        }
        lds_barrier();
        if (active) {
#pragma unroll
            for (int j = 0; j < kRegs; ++j) v[j].x = *lds_ptr<R>(ra[j]);
        }
        lds_barrier();  // every real part read before the imaginary parts overwrite them
        if (active) {
#pragma unroll
            for (int j = 0; j < kRegs; ++j) *lds_ptr<R>(wa[j]) = v[j].y;
        }
        lds_barrier();
        if (active) {
#pragma unroll
            for (int j = 0; j < kRegs; ++j) v[j].y = *lds_ptr<R>(ra[j]);
        }
    } else {
        if (active) {
#pragma unroll
            for (int j = 0; j < kRegs; ++j) *lds_ptr<V>(wa[j]) = v[j];
        }
        lds_barrier();
        if (active) {
#pragma unroll
            for (int j = 0; j < kRegs; ++j) v[j] = *lds_ptr<V>(ra[j]);
        }
    }
}

// Runs one tile whose first-stage amplitudes are already in v: stages with
// LDS exchanges between them, then the global scatter in the last stage's
// mapping.  Threads beyond the tile (a batch launch sizes workgroups for its
// widest tile) run the register arithmetic on don't-care values but skip
// every memory access, and keep every barrier; the op loop thus stays out of
// divergent control flow and its descriptor reads stay scalar.
template <class R, bool M, bool X = false>
__device__ __forceinline__ void compute_tile(const TileRef<R> &tr, uint32_t tid, Regs<Amp<R>> &v, void *lds)
{
    using V = Amp<R>;
    const bool active = tid < (1u << (tr.t - kRegBits));
    const int nst = tr.nstages;
    Regs<V> w;  // the op loop's second register array
    Lane ln = lane_of(tr.stages, tid, tr.base, tr.swz_a, tr.swz_b);
    for (int s = 0; s < nst; ++s) {
        const cptr<Stage> st = tr.stages + s;
        run_stage_ops<V, M, X>(st, v, w, ln.gfix, tr.ops, tr.coef);
        if (s + 1 < nst) {
            const cptr<Stage> nx = st + 1;
            // the global index bits only where an op of the stage tests them or the stage scatters
            const Lane nl = lane_of(nx, tid, tr.base, tr.swz_a, tr.swz_b, nx->gneed != 0);
            uint32_t rb = nl.sb;
            if (nx->relabel) {  // the exchange applies the relabelling CNOTs / SWAPs before the next stage
                const uint32_t tm = tid & ((1u << nx->np) - 1u);
                rb = 0;
#pragma unroll
                for (int i = 0; i < kMaxTile - kRegBits; ++i)  // bit i of tm as an all-ones / zero mask
                    rb ^= uint32_t(int32_t(tm << (31 - i)) >> 31) & nx->rimg[i];
#pragma unroll
                for (int k = 0; k < kMaxAffine; ++k)  // CNOTs controlled by the tile index (unused terms are 0)
                    rb ^= (0u - uint32_t((tr.base >> nx->abit[k]) & 1u)) & nx->aoff[k];
            }
            // A relabelled read into this stage took other threads' slots:
            // every thread must be done reading before any overwrites them.
            // (Otherwise each thread writes exactly the slots it read.)
            if (s > 0 && st->relabel) lds_barrier();
            exchange<R>(v, lds, active, ln.sb, st->roff, rb, nx->rroff);
            ln = nl;
        } else if (active) {
#pragma unroll
            for (int j = 0; j < kRegs; ++j) (tr.psi + ln.gfix)[st->goff[j]] = v[j];  // as in load_tile
        }
    }
}

// One workgroup per tile: gather the tile in its first stage's mapping, run
// the stages, scatter it.  Workgroups are short-lived and the hardware
// dispatcher keeps every CU full; a persistent grid that loops over tiles
// measured 1.7x slower on MI355X (any CU slot it fails to place delays a
// whole strip of tiles).
template <class R, bool M, bool X = false>
__device__ __forceinline__ void one_tile(const TileRef<R> &tr, void *lds)
{
    using V = Amp<R>;
    const uint32_t tid = threadIdx.x;
    Regs<V> v;
    load_tile<R>(tr, tid, v);  // threads beyond the tile read an in-tile alias (their bits above t-4 are ignored)
    compute_tile<R, M, X>(tr, tid, v, lds);
}

// Register budget: 168 VGPRs (3 waves per SIMD; the two c128 register
// arrays alone take 128; the c128 kernel takes 152, 160 with the sparse
// 4x4 arm).  With the split exchange a tile of 2^t amplitudes takes 2^t
// reals of LDS, so registers, not LDS, bound the resident waves.  (The
// complex64 kernel takes 77 VGPRs, 6 waves; with matrix-core gates, whose
// row operands and accumulators live in VGPRs too, 116, 4 waves.)
template <class R> constexpr int kWavesPerSimd = 3;

// XCD-aware tile order: dispatch places workgroup b on XCD b mod 8, so XCD x
// gets the contiguous tile range [x * tiles/8, (x+1) * tiles/8), and tiles
// that neighbour in memory share an L2.
template <class R, bool M, bool X>
__global__ __launch_bounds__(kMaxThreads) __attribute__((amdgpu_waves_per_eu(kWavesPerSimd<R>))) void k_fused_pass(
    Amp<R> *__restrict__ psi, const Pass *__restrict__ pass, const Stage *__restrict__ stages,
    const DevOp *__restrict__ ops, const Amp<R> *__restrict__ coef, uint32_t per, uint32_t xcds)
{
    extern __shared__ unsigned char lds_raw[];
    const uint64_t T = uint64_t(blockIdx.x % xcds) * per + blockIdx.x / xcds;
    one_tile<R, M, X>(tile_ref<R>(psi, as_const(pass), as_const(stages), as_const(ops), as_const(coef), T), lds_raw);
}

// XCD-aware order for a run of any length: workgroup b runs on XCD b mod 8,
// and XCD x takes the contiguous run of tiles that starts at
// x * (total / 8) + min(x, total mod 8) (the first total mod 8 XCDs take one
// tile more), so tiles that neighbour in memory share an L2.
__device__ __forceinline__ uint32_t xcd_linear(uint32_t b, uint32_t total)
{
    const uint32_t x = b & 7u, q = total >> 3, rm = total & 7u;
    return x * q + min(x, rm) + (b >> 3);
}

// Batch kernel: workgroup b belongs to the entry whose run of workgroups
// holds it (entries sorted by tile_begin, found by binary search); an entry
// covers the tiles [tile_off, tile_off + count) of its pass.  Within the run
// the tiles go to XCDs in contiguous ranges (xcd_linear over the run; the
// run's workgroups with one residue mod 8 share an XCD), while the entries
// stay interleaved over every XCD, so entries of unequal cost per tile still
// load all eight evenly.
template <class R>
__global__ __launch_bounds__(kMaxThreads) __attribute__((amdgpu_waves_per_eu(kWavesPerSimd<R>))) void k_fused_batch(
    const BatchEntry *__restrict__ entries, int nentries)
{
    extern __shared__ unsigned char lds_raw[];
    const cptr<BatchEntry> ent = as_const(entries);
    const uint64_t lin = blockIdx.x;
    int lo = 0, hi = nentries - 1;  // last entry with tile_begin <= lin
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (ent[mid].tile_begin <= lin) lo = mid; else hi = mid - 1;
    }
    const cptr<BatchEntry> e = ent + lo;
    const uint64_t end = lo + 1 < nentries ? ent[lo + 1].tile_begin : uint64_t(gridDim.x);
    const uint32_t local = xcd_linear(uint32_t(lin - e->tile_begin), uint32_t(end - e->tile_begin));
    one_tile<R, false>(tile_ref<R>(static_cast<Amp<R> *>(e->psi), as_const(e->pass), as_const(e->stages), as_const(e->ops),
                            as_const(static_cast<const Amp<R> *>(e->coef)), local + e->tile_off),
                lds_raw);
}

template <class R, bool M, bool X>
void launch_pass_t(void *psi, int n, int t, const Pass *pass, const Stage *stages, const DevOp *ops, const void *coef,
                   hipStream_t st)
{
    const uint64_t tiles = uint64_t(1) << (n - t);
    const unsigned block = 1u << (t - kRegBits);  // <= kMaxThreads = the kernel's launch bound
    const size_t lds = (size_t(1) << t) * kLdsSlotBytes<R>;
    const uint32_t xcds = (tiles % 8 == 0) ? 8u : 1u;
    hipLaunchKernelGGL((k_fused_pass<R, M, X>), dim3(unsigned(tiles)), dim3(block), lds, st, static_cast<Amp<R> *>(psi), pass,
                       stages, ops, static_cast<const Amp<R> *>(coef), uint32_t(tiles / xcds), xcds);
}

template <class R>
void launch_batch_t(const BatchEntry *entries, int nentries, uint64_t total_tiles, int max_t, hipStream_t st)
{
    const unsigned block = 1u << (max_t - kRegBits);
    const size_t lds = (size_t(1) << max_t) * kLdsSlotBytes<R>;
    hipLaunchKernelGGL(k_fused_batch<R>, dim3(unsigned(total_tiles)), dim3(block), lds, st, entries, nentries);
}

}  // namespace

void launch_pass(void *psi, bool c64, int n, int t, const Pass *pass, const Stage *stages, const DevOp *ops,
                 const void *coef, bool mfma_u2, bool sparse_u2, hipStream_t st)
{
    EV_REQUIRE(t >= kMinTile && t <= kMaxTile && t <= n, EV_E_ARG, "fused pass tile width out of range");
    EV_REQUIRE(n - t < 32, EV_E_LIMIT, "fused pass has more than 2^32 - 1 tiles");
    EV_REQUIRE(c64 || !mfma_u2, EV_E_ARG, "matrix-core dense gates are a complex64 option");
    EV_REQUIRE(!c64 || !sparse_u2, EV_E_ARG, "sparse 4x4 ops are a complex128 (density-matrix) option");
    if (!c64) {
        if (sparse_u2) launch_pass_t<double, false, true>(psi, n, t, pass, stages, ops, coef, st);
        else launch_pass_t<double, false, false>(psi, n, t, pass, stages, ops, coef, st);
    } else if (mfma_u2) {
        launch_pass_t<float, true, false>(psi, n, t, pass, stages, ops, coef, st);
    } else {
        launch_pass_t<float, false, false>(psi, n, t, pass, stages, ops, coef, st);
    }
}

void launch_batch(bool c64, const BatchEntry *entries, int nentries, uint64_t total_tiles, int max_t, hipStream_t st)
{
    EV_REQUIRE(max_t >= kMinTile && max_t <= kMaxTile, EV_E_ARG, "batch tile width out of range");
    EV_REQUIRE(total_tiles < (uint64_t(1) << 32), EV_E_LIMIT, "batch round has more than 2^32 - 1 tiles");
    if (nentries == 0 || total_tiles == 0) return;
    if (!c64) launch_batch_t<double>(entries, nentries, total_tiles, max_t, st);
    else launch_batch_t<float>(entries, nentries, total_tiles, max_t, st);
}

}  // namespace ev::fused
