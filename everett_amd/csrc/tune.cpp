// Autotuner for the fused pass parameters (ev_autotune in include/everett.h).
//
// For one state size it times complete fused runs of a synthetic brick-wall
// circuit followed by a QFT (HIP events around the run on the state's
// stream) over a grid: the three widest tile widths, 2..5 contiguous low
// qubits per tile and three gate budgets per pass, then the budget around
// the winner.  The winner is recorded per (n, precision) and picked up by
// fused_config().
#include <hip/hip_runtime_api.h>

#include <map>
#include <mutex>
#include <vector>

#include "engine.hpp"
#include "planner.hpp"
#include "tune.hpp"

namespace ev {

namespace {
std::mutex g_mu;
std::map<std::pair<int, int>, ev_tune_entry> g_table;
}  // namespace

bool tune_lookup(int n, bool c64, ev_tune_entry *out)
{
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_table.find({n, c64 ? 1 : 0});
    if (it == g_table.end()) return false;
    *out = it->second;
    return true;
}

void tune_store(int n, bool c64, const ev_tune_entry &e)
{
    EV_REQUIRE(e.tile_qubits >= uint32_t(fused::kMinTile) && e.tile_qubits <= uint32_t(fused::kMaxTile) &&
                   e.tile_qubits <= uint32_t(n) && e.low_qubits >= 1 && e.low_qubits <= 6 && e.max_gates_per_pass >= 1,
               EV_E_ARG, "tune entry out of range");
    std::lock_guard<std::mutex> lk(g_mu);
    g_table[{n, c64 ? 1 : 0}] = e;
}

ev_tune_entry autotune(int n, bool c64, int depth)
{
    EV_REQUIRE(n >= fused::kMinTile && n <= 32, EV_E_LIMIT, "autotune needs 8..32 qubits");
    Rng rng = Rng::seed_from_u64(0x70E0ull + uint64_t(n));
    // the random-circuit family the benchmarks use: brick-wall layers, then a QFT
    Circuit c = gen_brickwall(uint64_t(n), uint64_t(depth > 0 ? depth : 8), rng);
    c.compose(gen_qft(uint64_t(n)));
    ev_sv *s = sv_create(n, c64 ? kern::C64 : kern::C128);
    ev_tune_entry best{};
    try {
        auto cost = [&](uint32_t t, uint32_t low, uint32_t budget) {
            ev_run_options o{};
            o.precision = c64 ? EV_C64 : EV_C128;
            o.tile_qubits = t;
            o.low_qubits = low;
            o.max_gates_per_pass = budget;
            std::vector<uint8_t> none;
            double best_ms = 1e300;
            for (int rep = 0; rep < 3; ++rep) {  // first run warms up
                sv_run_into(s, c, 0, none.data(), o);
                if (rep) best_ms = std::min(best_ms, last_stats().device_ms);
            }
            return best_ms / double(c.size());
        };
        // 13-qubit tiles stay an explicit option: their 8-wave workgroups
        // leave a CU one complex128 workgroup (registers), and at 30 qubits
        // they measured slower at both precisions
        const uint32_t tmax = uint32_t(std::min(n, 12));
        best.tile_qubits = tmax;
        best.low_qubits = c64 ? 5 : 4;
        best.max_gates_per_pass = 20;
        best.ms_per_gate = cost(best.tile_qubits, best.low_qubits, best.max_gates_per_pass);
        auto try_cfg = [&](uint32_t t, uint32_t low, uint32_t b) {
            if (low >= t) return;
            const double ms = cost(t, low, b);
            if (ms < best.ms_per_gate) best = ev_tune_entry{t, low, b, 0, ms};
        };
        // grid over the widest tiles, contiguous low qubits and gate budgets
        const uint32_t t0 = std::max<uint32_t>(uint32_t(fused::kMinTile), tmax >= 2 ? tmax - 2 : tmax);
        for (uint32_t t = t0; t <= tmax; ++t)
            for (uint32_t low : {2u, 3u, 4u, 5u})
                for (uint32_t b : {24u, 48u, 1000u}) try_cfg(t, low, b);
        const uint32_t b0 = best.max_gates_per_pass;  // then the budget around the winner
        for (uint32_t b : {b0 / 2, b0 + b0 / 2, b0 * 2}) try_cfg(best.tile_qubits, best.low_qubits, std::max(b, 2u));
    } catch (...) {
        sv_destroy(s);
        throw;
    }
    sv_destroy(s);
    tune_store(n, c64, best);
    return best;
}

}  // namespace ev
