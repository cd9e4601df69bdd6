// Fused execution of statevector circuits: the planner groups gates into
// tile passes (several gates per HBM sweep) and the pass kernel applies them.
#pragma once

#include "engine.hpp"

namespace ev {

// Runs c on s through fused passes.  Measure and If records end a segment:
// the planned passes before them are flushed, then they run as in the plain
// executor.
void execute_planned(ev_sv *s, const Circuit &c, Rng &rng, uint8_t *classical, const ev_run_options &opt,
                     RunStats &st, Profiler &prof);

}  // namespace ev
