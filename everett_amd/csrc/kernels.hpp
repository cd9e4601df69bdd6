// Host-callable launchers of the per-gate statevector kernels (one HBM pass
// per gate) and the FP64 reductions.  Implemented in sv_kernels.hip.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>

#include "common.hpp"

namespace ev::kern {

enum Prec : int { C128 = 0, C64 = 1 };
inline size_t amp_bytes(Prec p) { return p == C128 ? 16 : 8; }

constexpr int kReduceBlocks = 1024;  // fixed grid => deterministic reduction order

// psi := |basis> (a fill kernel, no memset, so callers may capture it in a graph).
void init_basis(void *psi, Prec p, int n, uint64_t basis, hipStream_t st);

// src/kernel.rs:76-183.  Diagonal gates touch only what changes: a gate with
// m01 == m10 == 0 multiplies in place; one that also has m00 == 1 (phase,
// S, T, Z) reads and writes only the bit-k-set half.
void apply1(void *psi, Prec p, int n, int k, const c128 m[4], hipStream_t st);
// src/kernel.rs:192-289: exact-equality dispatch to CNOT / CZ / SWAP loops,
// else the dense 4x4 product.
void apply2(void *psi, Prec p, int n, int a, int b, const c128 m[16], hipStream_t st);
// src/kernel.rs:299-437: visits only the 2^(n-1-c) pairs whose controls are set.
void ctrl1(void *psi, Prec p, int n, const uint64_t *controls, int nc, int target, const c128 m[4],
           hipStream_t st);

// FP64 reductions; `partials` holds 2*kReduceBlocks doubles, the result lands
// in out[0] (out[0..1] for the complex overlap) on the device.
void reduce_norm(const void *psi, Prec p, int n, double *partials, double *out, hipStream_t st);
void reduce_p1(const void *psi, Prec p, int n, int k, double *partials, double *out, hipStream_t st);
void reduce_overlap(const void *self_, const void *other, Prec p, int n, double *partials, double *out,
                    hipStream_t st);
// Bloch-vector sums (src/state.rs:163-185): out = {r00, r11, re r01, im r01}.
void reduce_bloch(const void *psi, Prec p, int n, int k, double *partials, double *out, hipStream_t st);

// src/measure.rs:40-48: keep amplitudes with bit k == outcome scaled by `scale`,
// zero the rest.
void collapse(void *psi, Prec p, int n, int k, int outcome, double scale, hipStream_t st);
// src/state.rs:103-111: multiply every amplitude by s.
void scale(void *psi, Prec p, int n, double s, hipStream_t st);
// src/state.rs:103-111 on the device: <psi|psi> into *norm_slot (partials as
// for reduce_norm), then psi /= sqrt of it when positive; no host round trip.
void normalize(void *psi, Prec p, int n, double *partials, double *norm_slot, hipStream_t st);

// Streaming copy of `bytes` (a multiple of 32), the HBM bandwidth reference.
void copy(const void *src, void *dst, uint64_t bytes, hipStream_t st);

// Measured multiply-add rates in TFLOP/s (rate_kernels.hip): vector ALU
// alone, matrix cores alone, both at equal flops.  `scratch` holds
// compute_rates_scratch_bytes(); synchronises `st`.
struct ComputeRates {
    double valu_tflops = 0, mfma_tflops = 0, mixed_tflops = 0;
};
size_t compute_rates_scratch_bytes();
ComputeRates compute_rates(bool f32, void *scratch, hipStream_t st, hipEvent_t e0, hipEvent_t e1);

// SWAPs of the qubit pairs (a[k], b[k]) (pairwise disjoint) on a 2^n state,
// in place in one pass: amplitude x trades places with pi(x), pi exchanging
// the bits of every pair.
constexpr int kMaxSwapPairs = 17;
void swap_bits(void *psi, Prec p, int n, const int *a, const int *b, int npairs, hipStream_t st);

// Precision conversion between a c128 staging buffer and c64 storage.
void c128_to_c64(const void *src, void *dst, uint64_t len, hipStream_t st);
void c64_to_c128(const void *src, void *dst, uint64_t len, hipStream_t st);

// Born-rule sampling support: inclusive f64 prefix sums of |amp|^2 over
// chunks, and a per-shot inverse-CDF search.  `scan` must hold 2^n doubles
// plus chunk totals; see sv_kernels.hip.
size_t sample_scratch_doubles(int n);
// Sampling through a fixed per-GPU scratch (no device allocation): the host
// buffer holds the shots' uniforms (as doubles) on entry and their basis
// indices on return; at most kNoAllocMaxShots shots; synchronises `st`.
constexpr uint64_t kNoAllocMaxShots = uint64_t(1) << 16;
void sample_noalloc(const void *psi, Prec p, int n, uint64_t shots, uint64_t *host_u_then_out, hipStream_t st,
                    int device);
void sample(const void *psi, Prec p, int n, const double *u, uint64_t shots, double *scratch, uint64_t *out,
            hipStream_t st);

}  // namespace ev::kern
