// Launchers over a batch arena (batch_kernels.hip).  The arena packs the
// states back to back; state i starts at amplitude aoff[i] (a multiple of
// kChunk) and owns chunks [chunk_begin[i], chunk_begin[i+1]).
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>

namespace ev::batchk {

constexpr uint64_t kChunk = 256;  // amplitudes per reduction chunk (<= the smallest state)

// Every state := |0..0>.
void init(void *arena, bool c64, uint64_t total_amps, const uint64_t *d_aoff, int count, hipStream_t st);
// out[i] = <psi_i|psi_i> in f64; part holds total_amps / kChunk doubles.
void norms(const void *arena, bool c64, uint64_t total_amps, const uint64_t *d_chunk_begin, int count, double *part,
           double *out, hipStream_t st);
// out[5i..5i+5) = {re <ref|x>, im <ref|x>, <x|x>, <ref|ref>, max |x - ref|} per
// state, against a c128 arena with the same layout; part holds 5 doubles per chunk.
void compare(const void *arena, bool c64, const void *ref_arena, uint64_t total_amps, const uint64_t *d_chunk_begin,
             int count, double *part, double *out, hipStream_t st);
// d_scale[i] = 1 / sqrt(d_nrm[i]) (0 for a zero norm), correctly rounded.
void inv_sqrt(const double *d_nrm, int count, double *d_scale, hipStream_t st);
// Multiplies state i by d_scale[i].
void scale(void *arena, bool c64, uint64_t total_amps, const uint64_t *d_chunk_begin, int count, const double *d_scale,
           hipStream_t st);

}  // namespace ev::batchk
