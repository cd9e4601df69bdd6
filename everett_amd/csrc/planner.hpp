// Fused execution of statevector circuits: the planner groups gates into
// tile passes (several gates per HBM sweep) and the pass kernel applies them.
#pragma once

#include <utility>
#include <vector>

#include "engine.hpp"
#include "fused.hpp"

namespace ev {

// Runs c on s through fused passes.  Measure and If records end a segment:
// the planned passes before them are flushed, then they run as in the plain
// executor.
void execute_planned(ev_sv *s, const Circuit &c, Rng &rng, uint8_t *classical, const ev_run_options &opt,
                     RunStats &st, Profiler &prof);

// Fused pass configuration: tile width t, low qubits always in the tile, the
// per-pass compute budget in units of one dense 2x2 over the tile, and the
// number of low qubits whose amplitudes fill one 128-byte run.
struct FusedConfig {
    int t, low;
    float budget;
    int coalesce;
    bool mfma_u2;  // complex64: dense 4x4 gates on the matrix cores
    bool sparse_u2;  // lower 4x4s of the kU2X form to kU2X (density-matrix passes)
};
FusedConfig fused_config(int n, bool c64, const ev_run_options &opt);

// A planned gate segment: passes, their stages, the device ops and the c128
// coefficient pool they index.
struct Plan {
    std::vector<fused::Pass> passes;
    std::vector<fused::Stage> stages;
    std::vector<fused::DevOp> devops;
    std::vector<c128> coef;
    // Disjoint qubit pairs whose SWAPs end the segment and run after the
    // passes as one bit-permutation pass (empty unless the plan was made with
    // the tail peeled off).
    std::vector<std::pair<int, int>> swap_tail;
};

// Plans every gate of a measurement-free circuit for an n-qubit state.  When
// the tile covers the whole state (t == n) the budget is lifted: the circuit
// is one pass.  With peel_swaps, a trailing network of disjoint SWAPs wider
// than a tile is taken out of the passes into swap_tail, as a single run
// executes it; batched and served runs keep every gate in passes.
Plan plan_circuit(const Circuit &c, int n, const FusedConfig &cfg, bool peel_swaps = false);

// Plans a measurement-free circuit without executing it (introspection).
struct PlanBlob {
    std::vector<unsigned char> passes, stages, ops;
    std::vector<c128> coef;
    std::vector<std::pair<int, int>> swap_tail;
    uint32_t np = 0, ns = 0, no = 0, t = 0;
};
PlanBlob debug_plan(const Circuit &c, const ev_run_options &opt);

// Algorithmic floating-point operations per state amplitude of all of a
// plan's device ops (each pass applies its ops to every amplitude once).
double plan_flops_per_amp(const Plan &P);

}  // namespace ev
