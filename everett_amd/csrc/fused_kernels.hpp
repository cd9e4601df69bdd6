// Launcher of the fused statevector pass kernel (fused_kernels.hip).
#pragma once

#include <hip/hip_runtime_api.h>

#include "fused.hpp"

namespace ev::fused {

// One launch = one pass over a 2^n state with tile width t: 2^(n-t)
// workgroups of 2^(t-4) threads, one tile of 2^t amplitudes each.
// Descriptor pointers are device memory; coef holds double2 (c128) or float2
// (c64) entries.  mfma_u2 (c64 only) runs dense 4x4 gates on the matrix
// cores; sparse_u2 (c128 only) must be set when the pass holds kU2X ops.
void launch_pass(void *psi, bool c64, int n, int t, const Pass *pass, const Stage *stages, const DevOp *ops,
                 const void *coef, bool mfma_u2, bool sparse_u2, hipStream_t st);

// One entry of a batched round: a run of consecutive tiles of one state's next
// pass, from tile tile_off on.  tile_begin is the prefix sum of the tile
// counts of the entries before it, so linear tile b of the launch is tile
// b - tile_begin + tile_off of the entry with the last tile_begin <= b.
struct BatchEntry {
    void *psi;
    const Pass *pass;
    const Stage *stages;
    const DevOp *ops;
    const void *coef;
    uint64_t tile_begin;
    uint64_t tile_off;
};

// One launch over every entry of a round: total_tiles workgroups of
// 2^(max_t-4) threads.  Entry e owns the workgroups [tile_begin, next
// entry's tile_begin) and runs one tile of its pass on each, the tiles
// dealt out in an XCD-aware order within that run.
void launch_batch(bool c64, const BatchEntry *entries, int nentries, uint64_t total_tiles, int max_t, hipStream_t st);

}  // namespace ev::fused
