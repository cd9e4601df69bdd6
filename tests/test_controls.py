"""Controlled ops with any number of controls (src/op.rs:34-36: `controls:
Vec<QubitId>`, no length limit; src/kernel.rs:299-330 folds them into one
mask, so repeated controls are one condition).

The C ABI carries an op's first 16 controls in its EV_OP_CONTROLLED record and
the rest in EV_OP_CONTROLS continuation records.  CPU tests cover the record
encoding (append, export, QASM, malformed input) and the oracle's reading of
it; GPU tests run 17-23-control gates on the plain and fused paths against
the oracle's apply_controlled_1q.
"""
import numpy as np
import pytest

import everett_amd as ev
import oracle
from everett_amd import _ffi
from tests.helpers import gate_set, ops_array, random_state

TOL128 = 1e-10


def _roundtrip(n, controls, target, m):
    c = ev.Circuit(n)
    c.controlled(controls, ev.Gate1(*m), target)
    return c, c.ops()


@pytest.mark.parametrize("nc", [15, 16, 17, 32, 33, 50])
def test_controls_roundtrip_through_records(nc):
    n = 64
    controls = [(3 * i) % 63 + 1 for i in range(nc)]  # repeats once nc > 63 / 3, which the reference allows
    m = gate_set()["u"]
    c, recs = _roundtrip(n, controls, 0, m)
    assert len(recs) == 1 + max(0, (nc - 16 + 15) // 16) == len(c)
    assert c.num_ops() == 1
    assert int(recs[0]["kind"]) == _ffi.OP_CONTROLLED and int(recs[0]["num_controls"]) == nc
    assert all(int(r["kind"]) == _ffi.OP_CONTROLS for r in recs[1:])
    got, nxt = _ffi.decode_controls(recs, 0)
    assert got == controls and nxt == len(recs)
    c.validate()
    # re-appending the exported records reproduces them
    c2 = ev.Circuit(n).append_ops(recs)
    assert np.array_equal(c2.ops(), recs)
    # the QASM writer emits every control
    cx, _ = _roundtrip(n, controls, 0, gate_set()["x"])
    text = cx.to_qasm()
    assert text.count("q[") == nc + 1 and (f"ctrl({nc}) @" in text)


def test_qasm_many_controls_parse_and_export():
    for k in (17, 50):
        n = k + 1
        src = "OPENQASM 3.0;\nqubit[%d] q;\nctrl(%d) @ x %s, q[%d];\n" % (
            n, k, ", ".join(f"q[{i}]" for i in range(k)), k)
        c = ev.Circuit.from_qasm(src)
        recs = c.ops()
        cs, j = _ffi.decode_controls(recs, 0)
        assert cs == list(range(k)) and j == len(recs) and int(recs[0]["target"]) == k
        c.validate()
        assert ev.Circuit.from_qasm(c.to_qasm()).ops().tobytes() == recs.tobytes()


def test_validation_sees_continued_controls():
    c = ev.Circuit(20)
    c.controlled(list(range(1, 17)) + [25], ev.Gate1.x(), 0)  # the 17th control is out of range
    with pytest.raises(ev.QubitOutOfRange):
        c.validate()
    c = ev.Circuit(20)
    c.controlled(list(range(1, 18)) + [0], ev.Gate1.x(), 0)  # the target as the 18th control
    with pytest.raises(ev.DuplicateQubit):
        c.validate()


def test_malformed_continuations_are_rejected():
    recs = _ffi.controlled_records(0, gate_set()["x"], list(range(1, 21)))
    with pytest.raises(ev.EverettError):
        ev.Circuit(24).append_ops(recs[:1])  # head promises 20 controls, carries 16
    with pytest.raises(ev.EverettError):
        ev.Circuit(24).append_ops(recs[1:])  # continuation without a head
    bad = recs.copy()
    bad[1]["num_controls"] = 16  # more than the head's total
    with pytest.raises(ev.EverettError):
        ev.Circuit(24).append_ops(bad)
    ok = ev.Circuit(24)
    ok.append_ops(recs).append_ops(recs)
    assert ok.num_ops() == 2 and len(ok) == 4


def test_oracle_reads_continued_controls():
    n, target = 20, 7
    controls = [q for q in range(n) if q != target][:19]
    rng = np.random.default_rng(3)
    for name in ("x", "u", "phase"):
        g = gate_set()[name]
        prep = ops_array([(1, q, 0, oracle.gate1("ry", 2.6 + 0.01 * q), None) for q in range(n)])
        ops = np.concatenate([prep, ops_array([(3, target, 0, g, controls)])])
        assert len(ops) == n + 2
        got, _ = oracle.sv_run(ops, n)
        want, _ = oracle.sv_run(prep, n)
        before = want.copy()
        oracle.apply_controlled_1q(want, controls, target, g)
        assert np.array_equal(got, want), name
        assert np.max(np.abs(want - before)) > 1e-3  # the gate acted on a populated subspace
    del rng


# ------------------------------------------------------------------ GPU

def _prep_circuit(n, concentrate):
    """Ry rotations that put most weight on |1> for the qubits in
    `concentrate` (so a many-control gate acts on a populated subspace),
    generic rotations elsewhere, and CZs (diagonal, so the weights stay)."""
    c = ev.Circuit(n)
    r = ev.Rng.seed_from_u64(n)
    for q in range(n):
        th = 2.8 + 0.3 * r.next_f64() if q in concentrate else 6.28 * r.next_f64()
        c.ry(q, th).rz(q, 6.28 * r.next_f64())
    for q in range(0, n - 1, 3):
        c.cz(q, q + 1)
    return c


@pytest.mark.gpu
@pytest.mark.parametrize("n,nc", [(18, 17), (24, 20), (24, 23)])
def test_many_control_kernel_exact(n, nc):
    target = n // 3
    controls = [q for q in range(n) if q != target][:nc]
    c = _prep_circuit(n, set(controls))
    psi, _ = oracle.sv_run(c.ops(), n)
    for name in ("x", "u", "phase"):
        g = gate_set()[name]
        want = psi.copy()
        oracle.apply_controlled_1q(want, controls, target, g)
        got = psi.copy()
        ev.kernel.apply_controlled_1q(got, controls, target, ev.Gate1(*g))
        assert float(np.max(np.abs(got - want))) == 0.0, (name, n, nc)
        assert np.max(np.abs(want - psi)) > 1e-4


@pytest.mark.gpu
@pytest.mark.parametrize("n,nc", [(18, 17), (22, 20)])
def test_many_control_circuits_plain_and_fused(n, nc):
    target = 1
    controls = [q for q in range(n) if q != target][:nc]
    c = _prep_circuit(n, set(controls))
    gs = gate_set()
    c.controlled(controls, ev.Gate1(*gs["u"]), target)
    c.controlled(list(reversed(controls)) + controls[:3], ev.Gate1.x(), target)  # repeats: one condition
    c.controlled(controls[1:], ev.Gate1(*gs["phase"]), controls[0])
    c.h(target)
    c.controlled(controls, ev.Gate1(*gs["rz"]), target)
    want, _ = oracle.sv_run(c.ops(), n)
    plain = ev.StateVectorBackend.run(c, plain=True).state().amplitudes()
    assert float(np.max(np.abs(plain - want))) == 0.0
    for tile in (0, 10):
        fused = ev.StateVectorBackend.run(c, tile_qubits=tile).state().amplitudes()
        assert np.max(np.abs(fused - want)) < TOL128, tile


@pytest.mark.gpu
def test_many_control_device_state_entry_point():
    n, target = 20, 19
    controls = list(range(19))
    c = _prep_circuit(n, set(controls))
    psi, _ = oracle.sv_run(c.ops(), n)
    s = ev.State.from_amplitudes(psi)
    g = gate_set()["u"]
    s.apply_controlled_1q(controls, target, ev.Gate1(*g))
    want = psi.copy()
    oracle.apply_controlled_1q(want, controls, target, g)
    assert float(np.max(np.abs(s.amplitudes() - want))) == 0.0
