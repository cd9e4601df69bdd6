// Host mirror of the reference's circuit front end: gate constructors
// (src/gate.rs), the RNG (src/rng.rs), the instruction set (src/op.rs) and the
// Circuit container with its validation rules (src/circuit.rs).
#pragma once

#include <array>
#include <cstdint>
#include <vector>

#include "common.hpp"

namespace ev {

// ----------------------------------------------------------------- rng.rs
// SplitMix64-seeded xoshiro256++ (src/rng.rs:9-56).  Host only.
struct Rng {
    uint64_t s[4];
    static Rng seed_from_u64(uint64_t seed);  // rng.rs:18-30
    uint64_t next_u64();                      // rng.rs:34-48
    double next_f64();                        // rng.rs:52-56
};

// ---------------------------------------------------------------- gate.rs
using Mat2 = std::array<c128, 4>;   // row-major [m00, m01, m10, m11]
using Mat4 = std::array<c128, 16>;  // row-major, basis |q_a q_b>, a = MSB

namespace gate {
Mat2 id();
Mat2 x();
Mat2 y();
Mat2 z();
Mat2 h();
Mat2 s();
Mat2 t();
Mat2 rx(double theta);
Mat2 ry(double theta);
Mat2 rz(double theta);
Mat2 phase(double lambda);
Mat2 compose(const Mat2 &a, const Mat2 &b);  // a * b (gate.rs:153-162)
Mat4 cnot();
Mat4 cz();
Mat4 swap();
}  // namespace gate

// ------------------------------------------------------------------ op.rs
enum class OpKind : uint8_t { Apply1 = 1, Apply2 = 2, Controlled = 3, Measure = 4, If = 5 };

// One instruction.  An If record guards the record that follows it (nested
// IfClassic is a run of If records).  Matrices and control lists live in the
// owning Circuit's pools.
struct Op {
    OpKind kind;
    uint64_t q0;        // Apply1 target | Apply2 a | Controlled target | Measure qubit | If bit
    uint64_t q1;        // Apply2 b | Measure into
    uint32_t ctrl_off;  // Controlled: offset into ctrl_pool
    uint32_t nctrl;
    uint64_t mat_off;   // offset into mat_pool (4 or 16 entries)
};

class Circuit {
  public:
    Circuit(uint64_t num_qubits = 0, uint64_t num_classical = 0) : nq_(num_qubits), ncl_(num_classical) {}

    uint64_t num_qubits() const { return nq_; }
    uint64_t num_classical() const { return ncl_; }
    const std::vector<Op> &ops() const { return ops_; }
    size_t size() const { return ops_.size(); }

    const c128 *mat(const Op &op) const { return mat_pool_.data() + op.mat_off; }
    const uint64_t *ctrls(const Op &op) const { return ctrl_pool_.data() + op.ctrl_off; }

    void apply1(const Mat2 &m, uint64_t target);
    void apply2(const Mat4 &m, uint64_t a, uint64_t b);
    void controlled(const uint64_t *controls, uint32_t nc, const Mat2 &m, uint64_t target);
    void measure(uint64_t qubit, uint64_t into);
    void if_bit(uint64_t bit);  // guards the next record
    // Appends the op that starts at ops[i] (a controlled op with more than
    // EV_MAX_CONTROLS controls continues in EV_OP_CONTROLS records) and
    // returns the index of the record after it.
    size_t append(const ev_op *ops, size_t count, size_t i);
    void compose(const Circuit &sub);  // circuit.rs:124-127
    // ev_op records op i takes (1, or more for a controlled op with more than
    // EV_MAX_CONTROLS controls), and all ops together.
    size_t records(size_t i) const;
    size_t record_count() const;
    // Writes op i as records(i) records from out[0]; returns that count.
    size_t to_ev(size_t i, ev_op *out) const;
    // Register sizes set after the ops (the OpenQASM reader learns them from
    // declarations that may follow the statements).
    void set_registers(uint64_t num_qubits, uint64_t num_classical)
    {
        nq_ = num_qubits;
        ncl_ = num_classical;
    }

    // circuit.rs:245-306; throws Error with the reference's first failure.
    void validate() const;

  private:
    size_t validate_at(size_t i) const;  // returns the index after record i's op
    uint64_t nq_, ncl_;
    std::vector<Op> ops_;
    std::vector<uint64_t> ctrl_pool_;
    std::vector<c128> mat_pool_;
};

// Builders for the reference's algorithm circuits and the benchmark workloads.
Circuit gen_qft(uint64_t n);   // algorithms/qft.rs:54-62
Circuit gen_iqft(uint64_t n);  // algorithms/qft.rs:81-91
Circuit gen_brickwall(uint64_t n, uint64_t depth, Rng &rng);
Circuit gen_dm_random(uint64_t n, uint64_t depth, Rng &rng);

// Kernel classes of the reference's dispatch.  One-qubit gates with
// m01 == m10 == 0 are diagonal; a diagonal gate with m00 == 1 exactly is a
// phase (touches only the bit-set half).  Two-qubit gates are matched by exact
// equality against CNOT / CZ / SWAP (src/kernel.rs:201-212).
enum class K1 { Dense, Diag, Phase };
enum class K2 { Cnot, Cz, Swap, Dense };
K1 classify1(const c128 m[4]);
K2 classify2(const c128 m[16]);

// Kraus operators of a channel (src/noise.rs:179-234), 4 entries each.
std::vector<Mat2> channel_kraus(const ev_channel &ch);

}  // namespace ev
