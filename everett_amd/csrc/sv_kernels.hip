// Per-gate statevector kernels for gfx950 (one HBM pass per gate), FP64
// reductions, collapse, precision conversion and Born-rule sampling.
//
// Layout: amplitudes are contiguous {re, im} pairs (double2 for c128, float2
// for c64), index bit q = qubit q.  Every kernel is a grid-stride loop over a
// bijective index map (bit insertion, src/kernel.rs:37-66), so each amplitude
// is read and written by exactly one thread: consecutive threads take
// consecutive counters, which keeps both halves of every pair coalesced.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "circuit.hpp"
#include "device.hpp"
#include "kernels.hpp"

namespace ev::kern {

namespace {

// Streaming kernels launch one thread per item (the grid covers the work;
// the grid-stride loops below then run once per thread).  Measured on
// MI355X over 4 GiB: a one-qubit pair update at 30 qubits averages 5.56 TB/s
// over the target positions with one item per thread against 5.24 with a
// 2^16-block grid striding over four items per thread; a plain copy 5.68
// against 5.22.
// A dispatch holds fewer than 2^32 work-items (the HSA packet counts its grid
// in work-items, 32 bits wide), so grids stop at 2^31 threads and states
// beyond that many items stride.
constexpr int kBlock = 256;
constexpr uint64_t kMaxGrid = (uint64_t(1) << 31) / kBlock;

inline dim3 grid_for(uint64_t work)
{
    uint64_t b = (work + kBlock - 1) / kBlock;
    if (b < 1) b = 1;
    if (b > kMaxGrid) b = kMaxGrid;
    return dim3(static_cast<unsigned>(b));
}
// The same for blocks of `bs` threads (a power of two dividing the work).
inline dim3 grid_blocks(uint64_t work, uint64_t bs)
{
    uint64_t b = work / bs;
    const uint64_t cap = (uint64_t(1) << 31) / bs;
    if (b < 1) b = 1;
    if (b > cap) b = cap;
    return dim3(static_cast<unsigned>(b));
}

#define GRID_STRIDE(i, count)                                                               \
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < (count);         \
         i += uint64_t(gridDim.x) * blockDim.x)

// The update kernels below touch every item once, so their amplitude loads
// and stores are non-temporal (`nt`: no cache residency to keep), and their
// blocks take the work in XCD-aware order: dispatch places block b on XCD
// b mod 8, so block b works on range (b mod 8) * (grid / 8) + b / 8 and each
// XCD streams one contiguous eighth of the state.  Measured on MI355X, a
// one-qubit pair update at 30 qubits c128 over the 30 target positions:
// 69.2% of 8 TB/s plain, 73.6% with non-temporal loads and stores (either
// alone: no gain), 75.9% with both and the XCD order; a plain copy
// 6.23 TB/s, 6.56 non-temporal.
__device__ __forceinline__ uint64_t xcd_block()
{
    const uint32_t nb = gridDim.x, b = blockIdx.x;
    return (nb % 8 == 0) ? uint64_t(b % 8) * (nb / 8) + b / 8 : uint64_t(b);
}
#define STREAM(i, count)                                                                    \
    for (uint64_t i = xcd_block() * blockDim.x + threadIdx.x; i < (count);                  \
         i += uint64_t(gridDim.x) * blockDim.x)

template <class V> __device__ __forceinline__ V ldnt(const V *p)
{
    V v;
    v.x = __builtin_nontemporal_load(&p->x);
    v.y = __builtin_nontemporal_load(&p->y);
    return v;
}
template <class V> __device__ __forceinline__ void stnt(V *p, V v)
{
    __builtin_nontemporal_store(v.x, &p->x);
    __builtin_nontemporal_store(v.y, &p->y);
}

// Kernels that touch only some of the amplitudes (CNOT, CZ, SWAP, phases,
// controlled gates) move runs of 2^(lowest touched bit) consecutive items.
// Runs shorter than a 128-byte line leave the rest of the line to the cache:
// non-temporal stores of part-lines measured up to 2x slower (CNOT on qubits
// 0 and 1 at 30 qubits: 12.8 against 6.7 ms), so those launches keep plain
// accesses (NT = false).
template <bool NT, class V> __device__ __forceinline__ V ld(const V *p) { return NT ? ldnt(p) : *p; }
template <bool NT, class V> __device__ __forceinline__ void st(V *p, V v)
{
    if constexpr (NT) stnt(p, v);
    else *p = v;
}
template <class R> constexpr int kLineBits = std::is_same_v<R, double> ? 3 : 4;  // log2(128 B / amplitude)

// The one-qubit kernels stride over their work in blocks of kU x blockDim
// items: a thread takes items i0, i0 + S, ..., i0 + (kU - 1) S (S =
// blockDim, so each of the kU steps is one coalesced run) and loads every
// item of its group before storing any.  With a grid that covers the work,
// one item per thread measured fastest (kU = 2: -2%, kU = 4: -4% on the
// pair update at 30 qubits, MI355X), so kU is 1.
constexpr int kU = 1;
#define GRID_STRIDE_GROUPS(i0, S, count)                                                        \
    for (uint64_t i0 = xcd_block() * (kU * blockDim.x) + threadIdx.x, S = blockDim.x;          \
         i0 < (count); i0 += uint64_t(gridDim.x) * (kU * blockDim.x))

template <class R> M2<R> to_m2(const c128 m[4])
{
    M2<R> o;
    for (int i = 0; i < 4; ++i) o.m[i] = mk<Amp<R>>(static_cast<R>(m[i].re), static_cast<R>(m[i].im));
    return o;
}
template <class R> M4<R> to_m4(const c128 m[16])
{
    M4<R> o;
    for (int i = 0; i < 16; ++i) o.m[i] = mk<Amp<R>>(static_cast<R>(m[i].re), static_cast<R>(m[i].im));
    return o;
}

// ------------------------------------------------------------------ kernels

template <class R> __global__ void k_init_basis(Amp<R> *psi, uint64_t len, uint64_t basis)
{
    STREAM(i, len) stnt(psi + i, mk<Amp<R>>(R(i == basis ? 1 : 0), R(0)));
}

// kernel.rs:100-119: new0 = m00 a0 + m01 a1, new1 = m10 a0 + m11 a1.
template <class R> __global__ void k_apply1_dense(Amp<R> *psi, uint64_t npairs, int k, M2<R> g)
{
    const uint64_t bit = uint64_t(1) << k;
    GRID_STRIDE_GROUPS(i0, S, npairs)
    {
        uint64_t j[kU];
        Amp<R> a0[kU], a1[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            j[u] = insert0(i0 + u * S, k);
            if (i0 + u * S < npairs) {
                a0[u] = ldnt(psi + j[u]);
                a1[u] = ldnt(psi + (j[u] | bit));
            }
        }
#pragma unroll
        for (int u = 0; u < kU; ++u)
            if (i0 + u * S < npairs) {
                stnt(psi + j[u], cadd(cmul(g.m[0], a0[u]), cmul(g.m[1], a1[u])));
                stnt(psi + (j[u] | bit), cadd(cmul(g.m[2], a0[u]), cmul(g.m[3], a1[u])));
            }
    }
}

// Workgroup barrier ordering LDS only (a full __syncthreads fence would also
// wait for the thread's outstanding global stores).
__device__ __forceinline__ void lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// The amplitude that lane (this lane XOR m) of the wave holds, m < 64: a
// cross-lane shuffle, no memory.
template <class V> __device__ __forceinline__ V lane_xor(V v, int m)
{
    V o;
    o.x = __shfl_xor(v.x, m, 64);
    o.y = __shfl_xor(v.y, m, 64);
    return o;
}

// Target qubit below the wave (k < 6): a thread owns one amplitude x and
// takes its partner x ^ 2^k from lane (lane ^ 2^k) of its wave with a
// shuffle, so the launch is one sequential stream in which every load and
// store moves whole lines.  One thread per pair instead touches each line
// with two half-used accesses below the 128-byte line (69.6-71.0% of 8 TB/s
// on positions 0-2 at 30 qubits c128) and reads two streams 2^k apart above
// it (77.5%); this form measured 82.1%.  The thread computes only its own
// output, in the reference's operand order (new0 from a0 = its own, new1
// from a0 = the partner's).  The grid-stride bound is uniform across a wave
// because len is a multiple of 64 and every wave covers 64 consecutive,
// 64-aligned amplitudes.
constexpr int kLaneQubits = 6;
template <class R> __global__ void k_apply1_dense_lane(Amp<R> *psi, uint64_t len, int k, M2<R> g)
{
    const int m = 1 << k;
    STREAM(x, len)
    {
        const Amp<R> a = ldnt(psi + x);
        const Amp<R> p = lane_xor(a, m);
        const Amp<R> o = ((x >> k) & 1) ? cadd(cmul(g.m[2], p), cmul(g.m[3], a)) : cadd(cmul(g.m[0], a), cmul(g.m[1], p));
        stnt(psi + x, o);
    }
}

// Target qubit above the wave but inside a workgroup of BS threads (6 <= k <
// log2 BS): the same one sequential stream, with the partner taken through
// LDS (each thread writes its amplitude, reads lane tid ^ 2^k after an
// LDS-only barrier, and a second barrier keeps the next round's writes
// behind every read).  BS divides len, so every thread of a block runs the
// same rounds and meets every barrier.
template <class R, int BS>
__global__ __launch_bounds__(BS) void k_apply1_dense_block(Amp<R> *psi, uint64_t len, int k, M2<R> g)
{
    __shared__ Amp<R> buf[BS];
    const uint32_t m = 1u << k, t = threadIdx.x;
    for (uint64_t x = xcd_block() * BS + t; x < len; x += uint64_t(gridDim.x) * BS) {
        const Amp<R> a = ldnt(psi + x);
        buf[t] = a;
        lds_sync();
        const Amp<R> p = buf[t ^ m];
        lds_sync();
        const Amp<R> o = ((x >> k) & 1) ? cadd(cmul(g.m[2], p), cmul(g.m[3], a)) : cadd(cmul(g.m[0], a), cmul(g.m[1], p));
        stnt(psi + x, o);
    }
}

// Diagonal gate, any target: with m01 = m10 = 0 the reference's sums reduce
// exactly to m00*a0 and m11*a1 (the zero products contribute signed zeros
// only), so a thread owns one amplitude and needs no partner, and the launch
// is one sequential stream for every target qubit.
template <class R> __global__ void k_apply1_diag_lane(Amp<R> *psi, uint64_t len, int k, Amp<R> d0, Amp<R> d1)
{
    STREAM(x, len)
    {
        const Amp<R> a = ldnt(psi + x);
        stnt(psi + x, cmul(((x >> k) & 1) ? d1 : d0, a));
    }
}

// Phase-type diagonal (m00 == 1 exactly): 1*a0 is a0 bit for bit, so only the
// bit-k-set half moves through HBM.
template <class R, bool NT> __global__ void k_apply1_phase(Amp<R> *psi, uint64_t npairs, int k, Amp<R> d1)
{
    const uint64_t bit = uint64_t(1) << k;
    GRID_STRIDE_GROUPS(i0, S, npairs)
    {
        uint64_t j[kU];
        Amp<R> a1[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            j[u] = insert0(i0 + u * S, k) | bit;
            if (i0 + u * S < npairs) a1[u] = ld<NT>(psi + j[u]);
        }
#pragma unroll
        for (int u = 0; u < kU; ++u)
            if (i0 + u * S < npairs) st<NT>(psi + j[u], cmul(d1, a1[u]));
    }
}

__device__ __forceinline__ uint64_t quad_base(uint64_t i, int lo, int hi) { return insert0(insert0(i, lo), hi); }

// kernel.rs:248-260: swap i10 <-> i11 (a = control, b = target).
template <class R, bool NT> __global__ void k_cnot(Amp<R> *psi, uint64_t ngroups, int lo, int hi, uint64_t ba, uint64_t bb)
{
    STREAM(i, ngroups)
    {
        const uint64_t base = quad_base(i, lo, hi);
        const Amp<R> t = ld<NT>(psi + (base | ba)), u = ld<NT>(psi + (base | ba | bb));
        st<NT>(psi + (base | ba), u);
        st<NT>(psi + (base | ba | bb), t);
    }
}

// kernel.rs:277-289: swap i01 <-> i10.
template <class R, bool NT> __global__ void k_swap(Amp<R> *psi, uint64_t ngroups, int lo, int hi, uint64_t ba, uint64_t bb)
{
    STREAM(i, ngroups)
    {
        const uint64_t base = quad_base(i, lo, hi);
        const Amp<R> t = ld<NT>(psi + (base | bb)), u = ld<NT>(psi + (base | ba));
        st<NT>(psi + (base | bb), u);
        st<NT>(psi + (base | ba), t);
    }
}

// kernel.rs:263-274: negate i11.
template <class R, bool NT> __global__ void k_cz(Amp<R> *psi, uint64_t ngroups, int lo, int hi, uint64_t both)
{
    STREAM(i, ngroups)
    {
        const uint64_t j = quad_base(i, lo, hi) | both;
        const Amp<R> v = ld<NT>(psi + j);
        st<NT>(psi + j, mk<Amp<R>>(-v.x, -v.y));
    }
}

// kernel.rs:214-243: dense 4x4 on [i00, i01, i10, i11], sums left to right.
template <class R>
__global__ void k_apply2_dense(Amp<R> *psi, uint64_t ngroups, int lo, int hi, uint64_t ba, uint64_t bb, M4<R> g)
{
    STREAM(i, ngroups)
    {
        const uint64_t base = quad_base(i, lo, hi);
        const uint64_t idx[4] = {base, base | bb, base | ba, base | ba | bb};
        const Amp<R> a[4] = {ldnt(psi + idx[0]), ldnt(psi + idx[1]), ldnt(psi + idx[2]), ldnt(psi + idx[3])};
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const Amp<R> *row = g.m + 4 * r;
            stnt(psi + idx[r],
                 cadd(cadd(cadd(cmul(row[0], a[0]), cmul(row[1], a[1])), cmul(row[2], a[2])), cmul(row[3], a[3])));
        }
    }
}

struct SwapPairs {
    int count;
    int a[kMaxSwapPairs], b[kMaxSwapPairs];
};

// kernel.rs:277-289 for several disjoint pairs at once: pi(x) exchanges the
// bits a[k] and b[k] of x for every pair; pi is an involution, so the
// amplitudes at x and pi(x) trade places (the thread with x < pi(x) moves
// both; fixed points stay).  A pair's bits are never the lowest ones the
// caller keeps contiguous, so runs of consecutive x map to runs of pi(x).
template <class R> __global__ void k_swap_bits(Amp<R> *psi, uint64_t len, SwapPairs sp)
{
    STREAM(x, len)
    {
        uint64_t y = x;
        for (int k = 0; k < sp.count; ++k) {
            const uint64_t d = ((x >> sp.a[k]) ^ (x >> sp.b[k])) & 1;
            y ^= (d << sp.a[k]) | (d << sp.b[k]);
        }
        if (x < y) {
            const Amp<R> u = ldnt(psi + x), v = ldnt(psi + y);
            stnt(psi + x, v);
            stnt(psi + y, u);
        }
    }
}

struct CtrlArgs {
    uint64_t pos_mask;   // every control and the target (repeated controls are one bit)
    uint64_t ctrl_mask;  // all control bits
    uint64_t tbit;
};

// kernel.rs:332-358, enumerating only the pairs whose controls are all set:
// insert zeros at every control and target position (ascending, so each
// insertion sees the bits already placed), then set the controls.  Any
// number of controls: the positions are the set bits of a uniform mask.
template <class R, bool NT> __global__ void k_ctrl1(Amp<R> *psi, uint64_t nwork, CtrlArgs a, M2<R> g)
{
    STREAM(i, nwork)
    {
        uint64_t x = i;
        for (uint64_t m = a.pos_mask; m; m &= m - 1) x = insert0(x, __builtin_ctzll(m));
        const uint64_t i0 = x | a.ctrl_mask, i1 = i0 | a.tbit;
        const Amp<R> a0 = ld<NT>(psi + i0), a1 = ld<NT>(psi + i1);
        st<NT>(psi + i0, cadd(cmul(g.m[0], a0), cmul(g.m[1], a1)));
        st<NT>(psi + i1, cadd(cmul(g.m[2], a0), cmul(g.m[3], a1)));
    }
}

// ---------------------------------------------------------------- reductions

__device__ __forceinline__ double nsq(double2 a) { return __builtin_fma(a.x, a.x, a.y * a.y); }  // complex.rs:80-82
__device__ __forceinline__ double nsq(float2 a)
{
    const double re = a.x, im = a.y;
    return __builtin_fma(re, re, im * im);
}
__device__ __forceinline__ double2 widen(double2 a) { return a; }
__device__ __forceinline__ double2 widen(float2 a) { return make_double2(a.x, a.y); }

// Block tree-sum of NV doubles per thread into LDS, thread 0 writes partial.
template <int NV> __device__ void block_sum(double (&v)[NV], double *partials)
{
    __shared__ double sh[NV][kBlock];
    for (int c = 0; c < NV; ++c) sh[c][threadIdx.x] = v[c];
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int c = 0; c < NV; ++c) sh[c][threadIdx.x] += sh[c][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int c = 0; c < NV; ++c) partials[NV * blockIdx.x + c] = sh[c][0];
}

template <class R> __global__ void k_norm_partial(const Amp<R> *psi, uint64_t len, double *partials)
{
    double v[1] = {0.0};
    GRID_STRIDE(i, len) v[0] += nsq(psi[i]);
    block_sum<1>(v, partials);
}

template <class R> __global__ void k_p1_partial(const Amp<R> *psi, uint64_t half, int k, double *partials)
{
    double v[1] = {0.0};
    const uint64_t bit = uint64_t(1) << k;
    GRID_STRIDE(i, half) v[0] += nsq(psi[insert0(i, k) | bit]);
    block_sum<1>(v, partials);
}

// state.rs:130-138: sum_j conj(other_j) * self_j in f64.
template <class R> __global__ void k_overlap_partial(const Amp<R> *s, const Amp<R> *o, uint64_t len, double *partials)
{
    double v[2] = {0.0, 0.0};
    GRID_STRIDE(i, len)
    {
        const double2 a = widen(s[i]), b = widen(o[i]);
        const double2 p = cmul(make_double2(b.x, -b.y), a);
        v[0] += p.x;
        v[1] += p.y;
    }
    block_sum<2>(v, partials);
}

// state.rs:163-185: r00 = sum |a_j|^2 (bit clear), r11 (bit set), r01 = sum a_j conj(a_j|bit).
template <class R> __global__ void k_bloch_partial(const Amp<R> *psi, uint64_t half, int k, double *partials)
{
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    const uint64_t bit = uint64_t(1) << k;
    GRID_STRIDE(i, half)
    {
        const uint64_t j0 = insert0(i, k);
        const double2 a = widen(psi[j0]), b = widen(psi[j0 | bit]);
        v[0] += nsq(a);
        v[1] += nsq(b);
        const double2 p = cmul(a, make_double2(b.x, -b.y));
        v[2] += p.x;
        v[3] += p.y;
    }
    block_sum<4>(v, partials);
}

template <int NV> __global__ void k_finalize(const double *partials, int nblocks, double *out)
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

template <class R> __global__ void k_collapse(Amp<R> *psi, uint64_t len, int k, int outcome, double scale)
{
    STREAM(j, len)
    {
        const bool keep = (((j >> k) & 1) != 0) == (outcome != 0);
        const Amp<R> a = ldnt(psi + j);
        stnt(psi + j, keep ? mk<Amp<R>>(R(a.x * scale), R(a.y * scale)) : mk<Amp<R>>(R(0), R(0)));
    }
}

template <class R> __global__ void k_scale(Amp<R> *psi, uint64_t len, double s)
{
    STREAM(j, len)
    {
        const Amp<R> a = ldnt(psi + j);
        stnt(psi + j, mk<Amp<R>>(R(a.x * s), R(a.y * s)));
    }
}

// src/state.rs:103-111 with the norm read from device memory: scale by
// 1 / sqrt(<psi|psi>) when it is positive (IEEE sqrt and division, the
// values the host formula gives).
template <class R> __global__ void k_scale_by_norm(Amp<R> *psi, uint64_t len, const double *norm_sqr)
{
    const double x = *norm_sqr;
    if (!(x > 0.0)) return;
    const double s = 1.0 / __builtin_sqrt(x);
    STREAM(j, len)
    {
        const Amp<R> a = ldnt(psi + j);
        stnt(psi + j, mk<Amp<R>>(R(a.x * s), R(a.y * s)));
    }
}

__global__ void k_c128_to_c64(const double2 *src, float2 *dst, uint64_t len)
{
    STREAM(j, len)
    {
        const double2 a = ldnt(src + j);
        stnt(dst + j, make_float2(float(a.x), float(a.y)));
    }
}
__global__ void k_c64_to_c128(const float2 *src, double2 *dst, uint64_t len)
{
    STREAM(j, len)
    {
        const float2 a = ldnt(src + j);
        stnt(dst + j, make_double2(a.x, a.y));
    }
}

// ----------------------------------------------------------------- sampling
// The f64 CDF is defined chunk-wise: chunk c covers indices [c*C, (c+1)*C);
// its total is the sequential sum of |amp|^2 over the chunk; prefix[c] is the
// sequential sum of totals up to and including c.  A shot with uniform u
// targets u * prefix[last] and returns the first index whose running sum
// (prefix[c-1] + sequential in-chunk sum) exceeds the target.

// Chunks of 2^cl amplitudes, cl = n below 10 qubits, else max(10, n - 16):
// at most 2^16 chunks, so the sequential prefix over the chunk totals (one
// dependent f64 add per chunk) stays short at every size, while the chunks
// stay small enough for the per-chunk sums to run in parallel.
constexpr int kChunkLog = 10;
constexpr int kMaxPrefixLog = 16;
inline int chunk_log(int n) { return n < kChunkLog ? n : std::max(kChunkLog, n - kMaxPrefixLog); }

template <class R> __global__ void k_chunk_totals(const Amp<R> *psi, uint64_t nchunks, uint64_t csize, double *tot)
{
    GRID_STRIDE(c, nchunks)
    {
        double s = 0.0;
        const Amp<R> *p = psi + c * csize;
        for (uint64_t j = 0; j < csize; ++j) s += nsq(p[j]);
        tot[c] = s;
    }
}

// The same sums for chunks of at least 64 amplitudes, with coalesced reads:
// a one-wave workgroup takes 64 chunks; per slab of 64 amplitudes it reads
// each chunk's slab across the lanes into LDS (|a|^2 per entry), then lane k
// adds chunk k's 64 weights in index order, so every chunk's sum is formed
// in the same order as above.
template <class R>
__global__ void k_chunk_totals_wide(const Amp<R> *psi, uint64_t nchunks, uint64_t csize, double *tot)
{
    __shared__ double w[64][65];  // [offset in slab][chunk of the group], padded against bank conflicts
    const unsigned lane = threadIdx.x;
    for (uint64_t c0 = uint64_t(blockIdx.x) * 64; c0 < nchunks; c0 += uint64_t(gridDim.x) * 64) {
        double s = 0.0;
        for (uint64_t j0 = 0; j0 < csize; j0 += 64) {
            for (int k = 0; k < 64; ++k) {
                const uint64_t c = c0 + k;
                w[lane][k] = c < nchunks ? nsq(psi[c * csize + j0 + lane]) : 0.0;
            }
            __syncthreads();
            for (int j = 0; j < 64; ++j) s += w[j][lane];
            __syncthreads();
        }
        if (c0 + lane < nchunks) tot[c0 + lane] = s;
    }
}

// Sequential inclusive prefix over the chunk totals, so the CDF has one
// fixed summation order: run = ((t0 + t1) + t2) + ...  One wave reads 64
// totals at a time (one coalesced load, instead of one dependent load per
// chunk) and every lane adds them into the same running sum in chunk order
// (each lane holds the same value; lane k keeps the sum after chunk k).
__global__ void k_prefix(double *tot, uint64_t nchunks)
{
    if (blockIdx.x != 0) return;
    const unsigned lane = threadIdx.x;  // launched with one 64-lane wave
    double run = 0.0;
    for (uint64_t base = 0; base < nchunks; base += 64) {
        const uint64_t c = base + lane;
        const double v = c < nchunks ? tot[c] : 0.0;
        const uint64_t vb = __double_as_longlong(v);
        const uint32_t lo = uint32_t(vb), hi = uint32_t(vb >> 32);
        double mine = 0.0;
#pragma unroll
        for (int k = 0; k < 64; ++k) {  // past the last chunk this adds +0, which leaves run unchanged
            const uint64_t b = uint64_t(uint32_t(__builtin_amdgcn_readlane(int(hi), k))) << 32 |
                               uint32_t(__builtin_amdgcn_readlane(int(lo), k));
            run += __longlong_as_double(int64_t(b));
            if (int(lane) == k) mine = run;
        }
        if (c < nchunks) tot[c] = mine;
    }
}

template <class R>
__global__ void k_sample(const Amp<R> *psi, uint64_t nchunks, uint64_t csize, const double *prefix, const double *u,
                         uint64_t shots, uint64_t *out)
{
    GRID_STRIDE(s, shots)
    {
        const double target = u[s] * prefix[nchunks - 1];
        uint64_t lo = 0, hi = nchunks - 1;  // first c with prefix[c] > target (last chunk if none)
        while (lo < hi) {
            const uint64_t mid = (lo + hi) / 2;
            if (prefix[mid] > target) hi = mid; else lo = mid + 1;
        }
        double run = lo > 0 ? prefix[lo - 1] : 0.0;
        uint64_t pick = lo * csize + csize - 1, last_nz = ~uint64_t(0);
        bool found = false;
        for (uint64_t j = 0; j < csize; ++j) {
            const double w = nsq(psi[lo * csize + j]);
            if (w > 0.0) last_nz = lo * csize + j;
            run += w;
            if (run > target) {
                pick = lo * csize + j;
                found = true;
                break;
            }
        }
        if (!found && last_nz != ~uint64_t(0)) pick = last_nz;
        out[s] = pick;
    }
}

template <class R> Amp<R> *A(void *p) { return static_cast<Amp<R> *>(p); }
template <class R> const Amp<R> *A(const void *p) { return static_cast<const Amp<R> *>(p); }

template <class R> void apply1_t(void *psi, int n, int k, const c128 m[4], hipStream_t st)
{
    const uint64_t npairs = uint64_t(1) << (n - 1);
    const M2<R> g = to_m2<R>(m);
    switch (classify1(m)) {
    case K1::Phase:
        if (k >= kLineBits<R>)
            hipLaunchKernelGGL((k_apply1_phase<R, true>), grid_for(npairs), dim3(kBlock), 0, st, A<R>(psi), npairs, k,
                               g.m[3]);
        else
            hipLaunchKernelGGL((k_apply1_phase<R, false>), grid_for(npairs), dim3(kBlock), 0, st, A<R>(psi), npairs,
                               k, g.m[3]);
        break;
    case K1::Diag:
        hipLaunchKernelGGL(k_apply1_diag_lane<R>, grid_for(2 * npairs), dim3(kBlock), 0, st, A<R>(psi), 2 * npairs, k,
                           g.m[0], g.m[3]);
        break;
    case K1::Dense:
        if (k < kLaneQubits && n >= kLaneQubits)
            hipLaunchKernelGGL(k_apply1_dense_lane<R>, grid_for(2 * npairs), dim3(kBlock), 0, st, A<R>(psi), 2 * npairs, k,
                               g);
        else if (k < 8 && n >= 8)
            hipLaunchKernelGGL((k_apply1_dense_block<R, 256>), grid_for(2 * npairs), dim3(256), 0, st, A<R>(psi),
                               2 * npairs, k, g);
        else if (k < 10 && n >= 10)
            hipLaunchKernelGGL((k_apply1_dense_block<R, 1024>), grid_blocks(2 * npairs, 1024), dim3(1024), 0, st,
                               A<R>(psi), 2 * npairs, k, g);
        else
            hipLaunchKernelGGL(k_apply1_dense<R>, grid_for(npairs), dim3(kBlock), 0, st, A<R>(psi), npairs, k, g);
        break;
    }
}

template <class R> void apply2_t(void *psi, int n, int a, int b, const c128 m[16], hipStream_t st)
{
    const uint64_t ng = uint64_t(1) << (n - 2);
    const int lo = std::min(a, b), hi = std::max(a, b);
    const uint64_t ba = uint64_t(1) << a, bb = uint64_t(1) << b;
    switch (classify2(m)) {
    case K2::Cnot:
        if (lo >= kLineBits<R>)
            hipLaunchKernelGGL((k_cnot<R, true>), grid_for(ng), dim3(kBlock), 0, st, A<R>(psi), ng, lo, hi, ba, bb);
        else
            hipLaunchKernelGGL((k_cnot<R, false>), grid_for(ng), dim3(kBlock), 0, st, A<R>(psi), ng, lo, hi, ba, bb);
        break;
    case K2::Cz:
        if (lo >= kLineBits<R>)
            hipLaunchKernelGGL((k_cz<R, true>), grid_for(ng), dim3(kBlock), 0, st, A<R>(psi), ng, lo, hi, ba | bb);
        else
            hipLaunchKernelGGL((k_cz<R, false>), grid_for(ng), dim3(kBlock), 0, st, A<R>(psi), ng, lo, hi, ba | bb);
        break;
    case K2::Swap:
        if (lo >= kLineBits<R>)
            hipLaunchKernelGGL((k_swap<R, true>), grid_for(ng), dim3(kBlock), 0, st, A<R>(psi), ng, lo, hi, ba, bb);
        else
            hipLaunchKernelGGL((k_swap<R, false>), grid_for(ng), dim3(kBlock), 0, st, A<R>(psi), ng, lo, hi, ba, bb);
        break;
    case K2::Dense:
        hipLaunchKernelGGL(k_apply2_dense<R>, grid_for(ng), dim3(kBlock), 0, st, A<R>(psi), ng, lo, hi, ba, bb,
                           to_m4<R>(m));
        break;
    }
}

template <class R>
void ctrl1_t(void *psi, int n, const uint64_t *controls, int nc, int target, const c128 m[4], hipStream_t st)
{
    EV_REQUIRE(n >= 1 && n <= 63 && target >= 0 && target < n, EV_E_ARG, "controlled gate: target out of range");
    CtrlArgs a{};
    a.tbit = uint64_t(1) << target;
    for (int c = 0; c < nc; ++c) {
        EV_REQUIRE(controls[c] < uint64_t(n) && int(controls[c]) != target, EV_E_ARG,
                   "controlled gate: control out of range or equal to the target");
        a.ctrl_mask |= uint64_t(1) << controls[c];
    }
    a.pos_mask = a.ctrl_mask | a.tbit;
    const uint64_t nwork = uint64_t(1) << (n - __builtin_popcountll(a.pos_mask));
    if (__builtin_ctzll(a.pos_mask) >= kLineBits<R>)
        hipLaunchKernelGGL((k_ctrl1<R, true>), grid_for(nwork), dim3(kBlock), 0, st, A<R>(psi), nwork, a, to_m2<R>(m));
    else
        hipLaunchKernelGGL((k_ctrl1<R, false>), grid_for(nwork), dim3(kBlock), 0, st, A<R>(psi), nwork, a,
                           to_m2<R>(m));
}

inline dim3 rgrid(uint64_t work)
{
    uint64_t b = (work + kBlock - 1) / kBlock;
    return dim3(static_cast<unsigned>(std::max<uint64_t>(1, std::min<uint64_t>(b, kReduceBlocks))));
}

// Streaming copy, the bandwidth reference: one 16-byte item per thread,
// non-temporal, XCD-aware block order (the update kernels' access pattern).
__global__ void k_copy(const double2 *__restrict__ src, double2 *__restrict__ dst, uint64_t n2)
{
    STREAM(i, n2) stnt(dst + i, ldnt(src + i));
}

}  // namespace

void swap_bits(void *psi, Prec p, int n, const int *a, const int *b, int npairs, hipStream_t st)
{
    EV_REQUIRE(npairs >= 0 && npairs <= kMaxSwapPairs, EV_E_ARG, "too many swap pairs");
    SwapPairs sp{};
    sp.count = npairs;
    for (int k = 0; k < npairs; ++k) {
        sp.a[k] = a[k];
        sp.b[k] = b[k];
    }
    const uint64_t len = uint64_t(1) << n;
    if (p == C128) hipLaunchKernelGGL(k_swap_bits<double>, grid_for(len), dim3(kBlock), 0, st, A<double>(psi), len, sp);
    else hipLaunchKernelGGL(k_swap_bits<float>, grid_for(len), dim3(kBlock), 0, st, A<float>(psi), len, sp);
}

void copy(const void *src, void *dst, uint64_t bytes, hipStream_t st)
{
    const uint64_t n2 = bytes / sizeof(double2);
    hipLaunchKernelGGL(k_copy, grid_for(n2), dim3(kBlock), 0, st, static_cast<const double2 *>(src),
                       static_cast<double2 *>(dst), n2);
}

void init_basis(void *psi, Prec p, int n, uint64_t basis, hipStream_t st)
{
    const uint64_t len = uint64_t(1) << n;
    if (p == C128) hipLaunchKernelGGL(k_init_basis<double>, grid_for(len), dim3(kBlock), 0, st, A<double>(psi), len, basis);
    else hipLaunchKernelGGL(k_init_basis<float>, grid_for(len), dim3(kBlock), 0, st, A<float>(psi), len, basis);
}

void apply1(void *psi, Prec p, int n, int k, const c128 m[4], hipStream_t st)
{
    if (p == C128) apply1_t<double>(psi, n, k, m, st);
    else apply1_t<float>(psi, n, k, m, st);
}

void apply2(void *psi, Prec p, int n, int a, int b, const c128 m[16], hipStream_t st)
{
    if (p == C128) apply2_t<double>(psi, n, a, b, m, st);
    else apply2_t<float>(psi, n, a, b, m, st);
}

void ctrl1(void *psi, Prec p, int n, const uint64_t *controls, int nc, int target, const c128 m[4], hipStream_t st)
{
    if (p == C128) ctrl1_t<double>(psi, n, controls, nc, target, m, st);
    else ctrl1_t<float>(psi, n, controls, nc, target, m, st);
}

void reduce_norm(const void *psi, Prec p, int n, double *partials, double *out, hipStream_t st)
{
    const uint64_t len = uint64_t(1) << n;
    const dim3 g = rgrid(len);
    if (p == C128) hipLaunchKernelGGL(k_norm_partial<double>, g, dim3(kBlock), 0, st, A<double>(psi), len, partials);
    else hipLaunchKernelGGL(k_norm_partial<float>, g, dim3(kBlock), 0, st, A<float>(psi), len, partials);
    hipLaunchKernelGGL(k_finalize<1>, dim3(1), dim3(1024), 0, st, partials, int(g.x), out);
}

void reduce_p1(const void *psi, Prec p, int n, int k, double *partials, double *out, hipStream_t st)
{
    const uint64_t half = uint64_t(1) << (n - 1);
    const dim3 g = rgrid(half);
    if (p == C128) hipLaunchKernelGGL(k_p1_partial<double>, g, dim3(kBlock), 0, st, A<double>(psi), half, k, partials);
    else hipLaunchKernelGGL(k_p1_partial<float>, g, dim3(kBlock), 0, st, A<float>(psi), half, k, partials);
    hipLaunchKernelGGL(k_finalize<1>, dim3(1), dim3(1024), 0, st, partials, int(g.x), out);
}

void reduce_overlap(const void *s, const void *o, Prec p, int n, double *partials, double *out, hipStream_t st)
{
    const uint64_t len = uint64_t(1) << n;
    const dim3 g = rgrid(len);
    if (p == C128)
        hipLaunchKernelGGL(k_overlap_partial<double>, g, dim3(kBlock), 0, st, A<double>(s), A<double>(o), len, partials);
    else hipLaunchKernelGGL(k_overlap_partial<float>, g, dim3(kBlock), 0, st, A<float>(s), A<float>(o), len, partials);
    hipLaunchKernelGGL(k_finalize<2>, dim3(1), dim3(1024), 0, st, partials, int(g.x), out);
}

void reduce_bloch(const void *psi, Prec p, int n, int k, double *partials, double *out, hipStream_t st)
{
    const uint64_t half = uint64_t(1) << (n - 1);
    const dim3 g = rgrid(half);
    if (p == C128) hipLaunchKernelGGL(k_bloch_partial<double>, g, dim3(kBlock), 0, st, A<double>(psi), half, k, partials);
    else hipLaunchKernelGGL(k_bloch_partial<float>, g, dim3(kBlock), 0, st, A<float>(psi), half, k, partials);
    hipLaunchKernelGGL(k_finalize<4>, dim3(1), dim3(1024), 0, st, partials, int(g.x), out);
}

void collapse(void *psi, Prec p, int n, int k, int outcome, double scale, hipStream_t st)
{
    const uint64_t len = uint64_t(1) << n;
    if (p == C128)
        hipLaunchKernelGGL(k_collapse<double>, grid_for(len), dim3(kBlock), 0, st, A<double>(psi), len, k, outcome, scale);
    else hipLaunchKernelGGL(k_collapse<float>, grid_for(len), dim3(kBlock), 0, st, A<float>(psi), len, k, outcome, scale);
}

void scale(void *psi, Prec p, int n, double s, hipStream_t st)
{
    const uint64_t len = uint64_t(1) << n;
    if (p == C128) hipLaunchKernelGGL(k_scale<double>, grid_for(len), dim3(kBlock), 0, st, A<double>(psi), len, s);
    else hipLaunchKernelGGL(k_scale<float>, grid_for(len), dim3(kBlock), 0, st, A<float>(psi), len, s);
}

void normalize(void *psi, Prec p, int n, double *partials, double *norm_slot, hipStream_t st)
{
    reduce_norm(psi, p, n, partials, norm_slot, st);
    const uint64_t len = uint64_t(1) << n;
    if (p == C128)
        hipLaunchKernelGGL(k_scale_by_norm<double>, grid_for(len), dim3(kBlock), 0, st, A<double>(psi), len, norm_slot);
    else
        hipLaunchKernelGGL(k_scale_by_norm<float>, grid_for(len), dim3(kBlock), 0, st, A<float>(psi), len, norm_slot);
}

void c128_to_c64(const void *src, void *dst, uint64_t len, hipStream_t st)
{
    hipLaunchKernelGGL(k_c128_to_c64, grid_for(len), dim3(kBlock), 0, st, static_cast<const double2 *>(src),
                       static_cast<float2 *>(dst), len);
}

void c64_to_c128(const void *src, void *dst, uint64_t len, hipStream_t st)
{
    hipLaunchKernelGGL(k_c64_to_c128, grid_for(len), dim3(kBlock), 0, st, static_cast<const float2 *>(src),
                       static_cast<double2 *>(dst), len);
}

size_t sample_scratch_doubles(int n)
{
    const int cl = chunk_log(n);
    return size_t(1) << (n - cl);
}

void sample(const void *psi, Prec p, int n, const double *u, uint64_t shots, double *scratch, uint64_t *out,
            hipStream_t st)
{
    const int cl = chunk_log(n);
    const uint64_t csize = uint64_t(1) << cl, nchunks = uint64_t(1) << (n - cl);
    if (csize >= 64) {
        const dim3 g(unsigned(std::min<uint64_t>((nchunks + 63) / 64, kMaxGrid)));
        if (p == C128)
            hipLaunchKernelGGL(k_chunk_totals_wide<double>, g, dim3(64), 0, st, A<double>(psi), nchunks, csize, scratch);
        else
            hipLaunchKernelGGL(k_chunk_totals_wide<float>, g, dim3(64), 0, st, A<float>(psi), nchunks, csize, scratch);
    } else if (p == C128) {
        hipLaunchKernelGGL(k_chunk_totals<double>, grid_for(nchunks), dim3(kBlock), 0, st, A<double>(psi), nchunks, csize,
                           scratch);
    } else {
        hipLaunchKernelGGL(k_chunk_totals<float>, grid_for(nchunks), dim3(kBlock), 0, st, A<float>(psi), nchunks, csize,
                           scratch);
    }
    hipLaunchKernelGGL(k_prefix, dim3(1), dim3(64), 0, st, scratch, nchunks);
    if (shots == 0) return;
    if (p == C128)
        hipLaunchKernelGGL(k_sample<double>, grid_for(shots), dim3(kBlock), 0, st, A<double>(psi), nchunks, csize, scratch,
                           u, shots, out);
    else
        hipLaunchKernelGGL(k_sample<float>, grid_for(shots), dim3(kBlock), 0, st, A<float>(psi), nchunks, csize, scratch, u,
                           shots, out);
}

// ------------------------------------------------ sampling without allocation
//
// A fixed per-GPU scratch for ev_sv_sample_noalloc: the chunk totals (at most
// 2^16: chunk_log keeps n - chunk_log(n) <= 16), the shots' uniforms and the
// sampled indices.  Each GPU the library runs on holds its own copy until
// the process ends.  Host code never names the array except to look up its
// device address (once per device, that device current); every kernel
// argument and copy uses that address.
namespace {
constexpr size_t kNoAllocChunks = size_t(1) << kMaxPrefixLog;
constexpr size_t kScratchDoubles = kNoAllocChunks + 2 * size_t(kNoAllocMaxShots);
static __device__ __align__(256) double scratch[kScratchDoubles];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};
}  // namespace

double *noalloc_scratch(int device)
{
    EV_REQUIRE(device >= 0 && device < kMaxDevices, EV_E_LIMIT, "device index beyond this build's table of 64");
    int cur = -1;
    EV_HIP(hipGetDevice(&cur));
    EV_REQUIRE(cur == device, EV_E_STATE, "the state's device must be the current device");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *ptr = nullptr;
        EV_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        g_scratch_ptr[device] = static_cast<double *>(ptr);
    }
    return g_scratch_ptr[device];
}

void sample_noalloc(const void *psi, Prec p, int n, uint64_t shots, uint64_t *host_u_then_out, hipStream_t st,
                    int device)
{
    EV_REQUIRE(shots <= kNoAllocMaxShots, EV_E_LIMIT,
               "ev_sv_sample_noalloc takes at most " + std::to_string(kNoAllocMaxShots) + " shots per call");
    EV_REQUIRE(sample_scratch_doubles(n) <= kNoAllocChunks, EV_E_LIMIT, "state too large for the sampling scratch");
    double *base = noalloc_scratch(device);
    double *totals = base, *d_u = base + kNoAllocChunks;
    uint64_t *d_out = reinterpret_cast<uint64_t *>(d_u + kNoAllocMaxShots);
    // Every scratch word a call reads is written earlier in the same call
    // (all chunk totals, then the prefix in place, the uniforms by the copy,
    // every shot's index by k_sample), so nothing is cleared first.
    if (shots) EV_HIP(hipMemcpyAsync(d_u, host_u_then_out, shots * sizeof(double), hipMemcpyHostToDevice, st));
    sample(psi, p, n, d_u, shots, totals, d_out, st);
    if (shots) EV_HIP(hipMemcpyAsync(host_u_then_out, d_out, shots * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    EV_HIP(hipStreamSynchronize(st));
}

}  // namespace ev::kern
