"""Batched execution (many uneven statevectors, one launch per round) against
the oracle and against single-circuit runs."""
import numpy as np
import pytest

import everett_amd as ev
import oracle
from tests.test_planner import mixed_circuit

pytestmark = pytest.mark.gpu


def small_batch():
    cs = []
    for i in range(24):
        rng = ev.Rng.seed_from_u64(500 + i)
        n = 8 + rng.next_u64() % 7
        cs.append(ev.workloads.brickwall(n, 3 + rng.next_u64() % 8, rng) if i % 3 else mixed_circuit(n, 120, i))
    return cs


def test_batch_c128_matches_oracle():
    cs = small_batch()
    b = ev.Batch.run(cs, precision="c128", use_graph=False)
    assert len(b) == len(cs)
    for i, c in enumerate(cs):
        n = c.num_qubits()
        want, _ = oracle.sv_run(c.ops(), n)
        assert b.num_qubits(i) == n
        assert np.max(np.abs(b.amplitudes(i) - want)) < 1e-10, i
    st = b.stats()
    assert st["circuits"] == len(cs) and st["logical_gates"] == sum(len(c) for c in cs)


def test_batch_graph_equals_stream_launches():
    cs = small_batch()
    a = ev.Batch.run(cs, precision="c128", use_graph=False)
    g = ev.Batch.run(cs, precision="c128", use_graph=True)
    assert g.stats()["graph"] == 1
    for i in range(len(cs)):
        assert np.array_equal(a.amplitudes(i), g.amplitudes(i))


def test_batch_c64_fidelity_and_renormalisation():
    cs = small_batch()
    ref = ev.Batch.run(cs, precision="c128")
    x = ev.Batch.run(cs, precision="c64", renormalize=True)
    f, m = x.compare(ref)
    assert np.all(1 - f < 1e-6) and np.all(m < 1e-5)
    assert np.max(np.abs(x.norms() - 1.0)) < 1e-6
    assert np.max(np.abs(ref.norms() - 1.0)) < 1e-12


def test_batch_uneven_sizes_like_config3():
    cs = ev.workloads.config3(48)
    ref = ev.Batch.run(cs, precision="c128")
    x = ev.Batch.run(cs, precision="c64")
    f, m = x.compare(ref)
    assert np.all(1 - f < 1e-6) and np.all(m < 1e-5)
    # the c128 batch equals single-circuit fused runs
    for i in (0, 7, 31):
        single = ev.StateVectorBackend.run(cs[i]).state().amplitudes()
        assert np.max(np.abs(ref.amplitudes(i) - single)) < 1e-12
    # and a small one against the oracle
    i = int(np.argmin([c.num_qubits() for c in cs]))
    want, _ = oracle.sv_run(cs[i].ops(), cs[i].num_qubits())
    assert np.max(np.abs(ref.amplitudes(i) - want)) < 1e-10
    # the batch's algorithmic flops are the single runs' (same plans, summed per pass)
    flops = 0.0
    for c in cs:
        ev.StateVectorBackend.run(c, precision="c64")
        flops += ev.last_run_stats().pass_flops
    assert x.stats()["pass_flops"] == pytest.approx(flops, rel=1e-12)


def test_batch_rejects_measurement():
    c = ev.Circuit.with_classical(8, 1)
    c.h(0).measure(0, 0)
    with pytest.raises(ev.InvalidArgument):
        ev.Batch.run([c])


def test_batch_pipelined_groups_equal_single_runs():
    # 9 states of 26 qubits (2^29+ amplitudes in all) run as several pipelined
    # groups; small states sit between them so groups split unevenly
    sizes = [26, 10, 26, 26, 12, 26, 26, 26, 11, 26, 26, 26]
    cs = [ev.workloads.brickwall(n, 6, ev.Rng.seed_from_u64(700 + i)) for i, n in enumerate(sizes)]
    b = ev.Batch.run(cs, precision="c128", use_graph=True)
    st = b.stats()
    assert st["groups"] >= 2 and st["graph"] == 1 and st["circuits"] == len(cs)  # one captured graph per group
    assert np.max(np.abs(b.norms() - 1.0)) < 1e-12
    eager = ev.Batch.run(cs, precision="c128", use_graph=False)
    assert eager.stats()["graph"] == 0
    for i, c in enumerate(cs):  # each group's graph ran once: bit-identical to the stream launches
        n = c.num_qubits()
        w = min(1 << 16, 1 << n)
        for off in (0, (1 << n) - w):
            assert np.array_equal(b.amplitudes(i, off, w), eager.amplitudes(i, off, w)), (i, off)
    del eager
    for i, c in enumerate(cs):
        n = c.num_qubits()
        w = min(1 << 14, 1 << n)
        s = ev.StateVectorBackend.run(c).state()
        for off in (0, (1 << n) // 2 - w // 2, (1 << n) - w):
            assert np.max(np.abs(b.amplitudes(i, off, w) - s.amplitudes_range(off, w))) < 1e-12, (i, off)
        if n <= 12:
            want, _ = oracle.sv_run(c.ops(), n)
            assert np.max(np.abs(b.amplitudes(i) - want)) < 1e-10, i


def test_batch_mixed_tile_widths_match_oracle():
    # 13-qubit tiles: workgroups of 512 threads; states below 13 qubits run
    # narrower tiles in the same launches, their extra threads idle
    cs = small_batch()[:8] + [ev.workloads.config1(n, 6) for n in (13, 14, 15)]
    b = ev.Batch.run(cs, precision="c128", use_graph=False, tile_qubits=13)
    for i, c in enumerate(cs):
        n = c.num_qubits()
        want, _ = oracle.sv_run(c.ops(), n)
        assert np.max(np.abs(b.amplitudes(i) - want)) < 1e-10, i
