// Recorded fused-pass parameters per (qubits, precision) (tune.cpp).
#pragma once

#include "common.hpp"

namespace ev {

bool tune_lookup(int n, bool c64, ev_tune_entry *out);
void tune_store(int n, bool c64, const ev_tune_entry &e);
ev_tune_entry autotune(int n, bool c64, int depth);

}  // namespace ev
