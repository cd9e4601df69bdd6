// Host-callable launchers of the density-matrix reductions (dm_kernels.hip).
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>

#include "kernels.hpp"

namespace ev::kern {

// out[0] = sum of Re rho[j][j] over j with (j & mask) == want (trace: mask 0).
void dm_diag_sum(const void *rho, int n, uint64_t mask, uint64_t want, double *partials, double *out, hipStream_t st);
// Zero rows and columns whose qubit-q bit differs from outcome.
void dm_collapse(void *rho, int n, int q, int outcome, hipStream_t st);
// The collapse and the renormalisation by s = 1 / trace in one pass: kept
// entries become rho * s, the others 0 (read only where kept).
void dm_collapse_scale(void *rho, int n, int q, int outcome, double s, hipStream_t st);
// A block of consecutive measurements: the scales of the measurements that
// renormalised, in order.
constexpr int kMaxBlockMeasures = 32;
struct DmScales {
    int count;
    double s[kMaxBlockMeasures];
};
// diag[j] = Re rho[j][j] (2^n doubles).
void dm_extract_diag(const void *rho, int n, double *diag, hipStream_t st);
// dm_diag_sum over the diagonal copy (the same partial sums, bit for bit).
void diag_vec_sum(const double *diag, int n, uint64_t mask, uint64_t want, double *partials, double *out, hipStream_t st);
// One measurement's collapse (and scale by s) on the diagonal copy.
void diag_collapse(double *diag, int n, int q, int outcome, double s, bool scaled, hipStream_t st);
// Every collapse of a block and the scales in order, in one pass over rho
// (none: an outcome had probability 0 after an earlier one; all entries 0).
void dm_collapse_block(void *rho, int n, uint64_t keep_mask, uint64_t keep_val, bool none, const DmScales &sc,
                       hipStream_t st);
// out[0..8) = reduced 2x2 matrix of qubit q, {re, im} per entry, row-major.
void dm_partial_trace(const void *rho, int n, int q, double *partials, double *out, hipStream_t st);

}  // namespace ev::kern
