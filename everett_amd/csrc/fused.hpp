// Descriptors of a fused statevector pass, shared by the planner (host) and
// the pass kernel (device).
//
// A pass picks t "tile" qubits Q (always including the lowest qubits, so each
// tile row is a contiguous run of amplitudes).  Workgroup w owns the 2^t
// amplitudes whose non-tile bits spell w; it streams them from HBM into LDS,
// runs the pass program, and streams them back: one HBM read and one write of
// the whole state per pass, whatever the number of gates in it.
//
// The program is a list of stages.  In a stage each of the 2^(t-4) threads
// holds 16 amplitudes in registers, whose local index bits at the stage's 4
// "register" positions vary and whose other bits spell the thread id; the
// stage's ops act on those registers (targets must be register bits, controls
// and diagonal factors may depend on any bit of the global index).  Between
// stages the tile goes back to LDS and is re-read with the next mapping.
#pragma once

#include <cstddef>
#include <cstdint>

#if defined(__HIPCC__)
#define EV_HOST_DEVICE __host__ __device__
#else
#define EV_HOST_DEVICE
#endif

namespace ev::fused {

constexpr int kRegBits = 4;
constexpr int kRegs = 1 << kRegBits;  // amplitudes per thread per stage
constexpr int kMaxTile = 13;          // 2^13 amplitudes: 512 threads x 16 (64 KiB of LDS at c128, 32 KiB at c64)
constexpr int kMaxThreads = 1 << (kMaxTile - kRegBits);
constexpr int kTileSlots = 12;        // per-thread-bit table entries in a Stage (t - 4 <= kTileSlots)
static_assert(kMaxTile - kRegBits <= kTileSlots, "Stage tables hold every thread-id bit");
constexpr int kMinTile = 8;
constexpr int kMaxAffine = 8;         // relabelling terms conditioned on tile-index bits, per stage

enum OpKind : uint8_t {
    kU1 = 0,     // dense 2x2 on register bit b0
    kCU1 = 1,    // dense 2x2 on register bit b0 where the control condition holds
    kCX = 2,     // swap register pairs across bit b0 where the condition holds (CNOT, multi-controlled X)
    kDiagR = 3,  // diag(d0, d1) on register bit b0 where the condition holds
    kDiagT = 4,  // diag(d0, d1) on global qubit tq (not a register bit) where the condition holds
    kU2 = 5,     // dense 4x4 on register bits (b0 = MSB of the gate basis, b1 = LSB), b0 > b1
    kSwap = 6,   // exchange register bits b0 > b1
    kPhase = 7,  // multiply registers whose bit b0 is set (b0 = 4: every register) and that meet
                 // creg by the lane's product of b1 conditioned phase factors (see DevOp)
    kU2X = 8,    // 4x4 on register bits b0 > b1 (as kU2) whose only nonzero entries are m00 m03
                 // m11 m22 m30 m33, six pool entries in that order: |00> and |11> mix, |01> and
                 // |10> scale.  Density-matrix passes only (the superoperator of a depolarizing or
                 // amplitude-damping channel on a (row, column) bit pair has this form)
};

constexpr uint32_t kPhaseAll = 4;     // kPhase b0: the factor applies to all 16 registers
constexpr uint32_t kMaxPhaseTerms = 8;

enum OpFlags : uint8_t { kD0IsOne = 1 };

// Every field is a full dword so the kernel reads descriptors with scalar
// (SMEM) loads; code = kind | b0 << 8 | b1 << 16 | flags << 24.  One-qubit
// and diagonal ops carry their coefficients inline (one fetch per op); kU2
// indexes the pass's coefficient pool.  In c64 runs the inline coefficients
// are stored as floats in the first half of m.
//
// kPhase carries b1 terms: term i multiplies by factor coef[coef + i] (pool)
// the lanes whose global index has every bit of mask i set, where mask i is
// the 64-bit pattern stored in m[i] (not converted for c64).  The lane's
// factor is the product of its terms' factors.
struct DevOp {
    uint32_t code;
    uint32_t creg;   // condition on the register index j: (j & creg) == creg
    uint64_t cglob;  // condition on the thread's fixed global index bits
    uint32_t coef;   // kU2: offset (complex entries) into the coefficient pool
    uint32_t tq;     // kDiagT target qubit
    uint32_t pad[2];
    double m[8];     // kU1/kCU1: m00..m11 {re, im}; kDiagR/kDiagT: d0, d1
};
static_assert(sizeof(DevOp) == 96, "DevOp is fetched as 96 bytes");

EV_HOST_DEVICE inline uint32_t op_code(uint32_t kind, uint32_t b0, uint32_t b1, uint32_t flags)
{
    return kind | (b0 << 8) | (b1 << 16) | (flags << 24);
}

// All fields are dwords (or wider) so the kernel reads them with scalar loads.
//
// relabel != 0: CNOTs and SWAPs between tile qubits that the planner moved to
// the exchange before this stage are applied by the exchange itself.  They
// permute local indices linearly over GF(2), so the amplitude this stage
// wants at local index l sits in the LDS slot of A(l), for the composed
// inverse permutation A; the slot is the XOR of rimg[i] over the thread's set
// id bits and rroff[j] (swizzled images of the stage's bits under A).  A
// CNOT controlled by a qubit outside the tile flips its target only in the
// tiles whose index has that bit set, so A may also XOR the constant
// aoff[k] into every slot of the tiles whose global base has bit abit[k].
struct Stage {
    int32_t op_begin, op_end;
    int32_t np;               // thread-id bits (t - 4)
    int32_t relabel;          // the LDS read goes through rimg / rroff
    uint32_t roff[kRegs];     // swizzled LDS slot offset of register j
    uint64_t goff[kRegs];     // global index offset of register j
    int32_t p[kTileSlots];     // thread-id bit i -> local position p[i]
    int32_t gpos[kTileSlots];  // thread-id bit i -> global qubit
    uint32_t rimg[kTileSlots]; // relabel: swizzled A(1 << p[i])
    uint32_t rroff[kRegs];    // relabel: swizzled A(register offset of j)
    int32_t naff;             // relabel: affine terms
    uint32_t abit[kMaxAffine];
    uint32_t aoff[kMaxAffine];
    int32_t gneed;            // an op reads the lane's global index bits, or the stage scatters
};

struct Pass {
    int32_t t;
    int32_t nstages;
    int32_t stage_begin;
    int32_t ncomp;            // n - t
    int32_t q[16];            // local bit -> global qubit (ascending)
    int32_t comp[40];         // tile-index bit -> global qubit (the non-tile qubits, ascending)
    int32_t swz_a, swz_b;     // the pass's LDS slot swizzle (swz below)
};

// Rewrites the inline coefficients of ops for a c64 run (floats in the
// first half of m).
inline void inline_coef_to_c64(DevOp *ops, size_t n)
{
    for (size_t i = 0; i < n; ++i) {
        if ((ops[i].code & 0xff) == kPhase) continue;  // m holds term masks
        float f[8];
        for (int k = 0; k < 8; ++k) f[k] = float(ops[i].m[k]);
        for (int k = 0; k < 8; ++k) reinterpret_cast<float *>(ops[i].m)[k] = f[k];
    }
}

// LDS slot of local index l in a pass's exchanges: the low 6 bits are XORed
// with the bits a and b places above them (b = 31 folds nothing for tiles of
// at most 16 bits).  Each output bit depends only on itself and higher bits,
// so the map is a bijection of [0, 2^t) for every t; it is linear over XOR,
// so swz(a ^ b) = swz(a) ^ swz(b).  The planner picks (a, b) per pass so the
// 32 lanes that share an LDS cycle (8-byte slots, complex128) or all 64
// (4-byte slots, complex64) hit distinct banks in as many exchanges as it can.
constexpr int kSwzMask = 63;
constexpr int kSwzNone = 31;
EV_HOST_DEVICE inline uint32_t swz(uint32_t l, uint32_t a, uint32_t b)
{
    return l ^ ((l >> a) & kSwzMask) ^ ((l >> b) & kSwzMask);
}

}  // namespace ev::fused
