// Fused execution of statevector circuits: the planner groups gates into
// tile passes (several gates per HBM sweep) and the pass kernel applies them.
#pragma once

#include <vector>

#include "engine.hpp"

namespace ev {

// Runs c on s through fused passes.  Measure and If records end a segment:
// the planned passes before them are flushed, then they run as in the plain
// executor.
void execute_planned(ev_sv *s, const Circuit &c, Rng &rng, uint8_t *classical, const ev_run_options &opt,
                     RunStats &st, Profiler &prof);

// Plans a measurement-free circuit without executing it (introspection).
struct PlanBlob {
    std::vector<unsigned char> passes, stages, ops;
    std::vector<c128> coef;
    uint32_t np = 0, ns = 0, no = 0, t = 0;
};
PlanBlob debug_plan(const Circuit &c, const ev_run_options &opt);

}  // namespace ev
