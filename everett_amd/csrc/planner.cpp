// Fused pass planner and executor (statevector and density matrix).
#include "planner.hpp"

#include "dm_ops.hpp"

namespace ev {

void execute_planned(ev_sv *, const Circuit &, Rng &, uint8_t *, const ev_run_options &, RunStats &, Profiler &)
{
    throw Error(EV_E_STATE, "fused execution is not available in this build; set ev_run_options.plain = 1");
}

void execute_planned_dm(ev_dm *, const Circuit &, const DmNoise &, Rng &, uint8_t *, const ev_run_options &, RunStats &)
{
    throw Error(EV_E_STATE, "fused execution is not available in this build; set ev_run_options.plain = 1");
}

}  // namespace ev
