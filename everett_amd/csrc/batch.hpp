// The batch handle behind ev_batch_* (batch.cpp).
#pragma once

#include <hip/hip_runtime_api.h>

#include <vector>

#include "circuit.hpp"
#include "common.hpp"
#include "kernels.hpp"

struct ev_batch {
    ev::kern::Prec prec = ev::kern::C128;
    hipStream_t stream = nullptr;
    std::vector<int> n;             // qubits of state i
    std::vector<uint64_t> aoff;     // amplitude offset of state i in the arena (count + 1 entries)
    std::vector<uint64_t> gates;    // logical gates of circuit i
    uint64_t total_amps = 0;
    void *arena = nullptr;          // all states, back to back
    std::vector<void *> descs;      // per group: pass descriptors and round tables
    // per group: the pinned staging of its descriptors and its graph, released
    // with the batch (hipHostFree waits for the device, so freeing them inside
    // the run would put that wait and the frees on the run's wall time)
    std::vector<void *> staging;
    std::vector<hipGraph_t> graphs;
    std::vector<hipGraphExec_t> gexecs;
    uint64_t *d_meta = nullptr;     // aoff[] then chunk_begin[] on the device
    double *d_part = nullptr;       // reduction scratch
    double *h_out = nullptr;        // pinned results
    ev_batch_stats stats{};
};

namespace ev {

ev_batch *batch_run(const Circuit *const *cs, size_t count, const ev_run_options &opt);
void batch_free(ev_batch *b);
void batch_norms(ev_batch *b, double *out);
void batch_compare(ev_batch *b, ev_batch *ref, double *fid, double *maxabs);
void batch_read(ev_batch *b, size_t i, uint64_t offset, c128 *host, uint64_t len);

}  // namespace ev
