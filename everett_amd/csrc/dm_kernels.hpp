// Host-callable launchers of the density-matrix reductions (dm_kernels.hip).
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>

#include "kernels.hpp"

namespace ev::kern {

// out[0] = sum of Re rho[j][j] over j with (j & mask) == mask (trace: mask 0).
void dm_diag_sum(const void *rho, int n, uint64_t mask, double *partials, double *out, hipStream_t st);
// Zero rows and columns whose qubit-q bit differs from outcome.
void dm_collapse(void *rho, int n, int q, int outcome, hipStream_t st);
// out[0..8) = reduced 2x2 matrix of qubit q, {re, im} per entry, row-major.
void dm_partial_trace(const void *rho, int n, int q, double *partials, double *out, hipStream_t st);

}  // namespace ev::kern
