"""The autotuner (ev_autotune): it times the fused path on the device over its
grid of tile widths, low-qubit counts and gate budgets, records the fastest
entry for (n, precision), and the next plan for that size takes it; a run
with the tuned entry still matches the oracle."""
import numpy as np
import pytest

import everett_amd as ev
import oracle
from tests.fused_emulator import get_plan

pytestmark = pytest.mark.gpu


def test_autotune_records_an_entry_the_next_plan_uses():
    n = 16
    before = ev.tune_get(n, "c64")
    try:
        got = ev.autotune(n, "c64", 4)
        assert 8 <= got["tile_qubits"] <= min(13, n) and got["low_qubits"] >= 1 and got["ms_per_gate"] > 0
        e = ev.tune_get(n, "c64")
        assert (e["tile_qubits"], e["low_qubits"], e["max_gates_per_pass"]) == \
            (got["tile_qubits"], got["low_qubits"], got["max_gates_per_pass"])
        c = ev.workloads.config1(n, 6)
        _, _, _, _, t = get_plan(c, precision="c64")
        assert t == got["tile_qubits"]
        want, _ = oracle.sv_run(c.ops(), n)
        s = ev.StateVectorBackend.run(c, precision="c64").state().amplitudes()
        f = abs(np.vdot(want, s)) ** 2 / np.vdot(s, s).real
        assert 1 - f < 1e-6 and np.max(np.abs(s - want)) < 1e-5
    finally:
        if before is not None:
            ev.tune_set(n, "c64", before)
