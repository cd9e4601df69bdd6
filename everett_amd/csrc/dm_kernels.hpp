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
// out[0..8) = reduced 2x2 matrix of qubit q, {re, im} per entry, row-major.
void dm_partial_trace(const void *rho, int n, int q, double *partials, double *out, hipStream_t st);

}  // namespace ev::kern
