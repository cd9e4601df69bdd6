// Superoperator algebra and primitives of the device density-matrix backend.
#pragma once

#include <vector>

#include "engine.hpp"

namespace ev {

// A NoiseModel (src/noise.rs:31-40) with its Kraus lists expanded once and the
// channel-only superoperators precomputed (the reference re-expands them on
// every gate, density_matrix.rs:122-123).
struct DmNoise {
    std::vector<Mat2> k1, k2;
    Mat4 s1{}, s2{};
    double readout_error = 0.0;
};

DmNoise make_dm_noise(const ev_noise_model *noise);

// A (x) conj(A) on the |row col> pair of bits (row = MSB).
Mat4 superop_of(const Mat2 &A);
// sum_k (K_k U) (x) conj(K_k U), or U (x) conj(U) without Kraus operators.
Mat4 gate_channel_superop(const Mat2 &U, const std::vector<Mat2> &kraus);
Mat4 channel_superop(const std::vector<Mat2> &kraus);
Mat4 conj4(const c128 *m);
Mat2 conj2(const c128 *m);

// Sum of Re rho[j][j] over j with (j & mask) == want.
double dm_diag_sum(ev_dm *d, uint64_t mask, uint64_t want);
void dm_partial_trace(ev_dm *d, int q, c128 out[4]);
void dm_renormalize(ev_dm *d);
bool dm_measure(ev_dm *d, int q, Rng &rng, double readout_error);
// Consecutive Measure ops from index i as one pass over rho; returns the
// index after them.
size_t dm_measure_run(ev_dm *d, const Circuit &c, size_t i, Rng &rng, double readout_error, uint8_t *classical);

// backend/mod.rs:51-72 over the density-matrix primitives for record i, one
// launch per (super)operator; returns the index after the op.
size_t dm_plain_op(ev_dm *d, const Circuit &c, size_t i, const DmNoise &nz, Rng &rng, uint8_t *classical, bool execute,
                   RunStats &st);

// Fused execution of a density-matrix circuit (planner.cpp).
void execute_planned_dm(ev_dm *d, const Circuit &c, const DmNoise &nz, Rng &rng, uint8_t *classical,
                        const ev_run_options &opt, RunStats &st);

}  // namespace ev
