"""Fused pass planner on the CPU: structural bounds of every plan and an
emulation of the pass kernel's indexing against the oracle (no GPU)."""
import math

import numpy as np
import pytest

import everett_amd as ev
import oracle
from tests.fused_emulator import check_plan, emulate, get_plan
from tests.helpers import random_unitary


def mixed_circuit(n, nops, seed):
    rng = np.random.default_rng(seed)
    c = ev.Circuit(n)
    for _ in range(nops):
        r = int(rng.integers(0, 12))
        q = int(rng.integers(0, n))
        b = int(rng.integers(0, n - 1))
        b = b + 1 if b >= q else b
        if r < 3:
            c.gate1(ev.Gate1(*random_unitary(rng, 2).reshape(4)), q)
        elif r == 3:
            c.rz(q, float(rng.uniform(0, 6.28)))
        elif r == 4:
            c.cnot(q, b)
        elif r == 5:
            c.cz(q, b)
        elif r == 6:
            c.swap(q, b)
        elif r == 7:
            c.gate2(ev.Gate2(random_unitary(rng, 4)), q, b)
        elif r == 8:
            others = [x for x in range(n) if x != q]
            cs = [int(x) for x in rng.choice(others, size=int(rng.integers(1, 4)), replace=False)]
            c.controlled(cs, ev.Gate1(*random_unitary(rng, 2).reshape(4)), q)
        elif r == 9:
            c.controlled([b], ev.Gate1.phase(float(rng.uniform(0, 6.28))), q)
        elif r == 10:
            others = [x for x in range(n) if x != q]
            cs = [int(x) for x in rng.choice(others, size=2, replace=False)]
            c.controlled(cs, ev.Gate1.x(), q)
        else:
            c.h(q)
    return c


def phase_circuit(n, rounds, seed):
    """Hadamard layers between runs of one- and two-control phases on random qubits."""
    rng = np.random.default_rng(seed)
    c = ev.Circuit(n)
    for _ in range(rounds):
        for q in rng.choice(n, size=n // 2, replace=False):
            c.h(int(q))
        for _ in range(3 * n):
            qs = [int(x) for x in rng.choice(n, size=3, replace=False)]
            ctrls = qs[:1] if rng.random() < 0.7 else qs[:2]
            c.controlled(ctrls, ev.Gate1.phase(float(rng.uniform(0, 6.28))), qs[2])
    return c


CASES = [
    ("brick10+qft", lambda: ev.workloads.config1(10, 6), {}),
    ("brick12", lambda: ev.workloads.config0(12, 8), {"tile_qubits": 9}),
    ("brick11-budget1", lambda: ev.workloads.config0(11, 4), {"max_gates_per_pass": 1, "tile_qubits": 8}),
    ("brick11-budget-big", lambda: ev.workloads.config0(11, 6), {"max_gates_per_pass": 1000, "tile_qubits": 8}),
    ("mixed10", lambda: mixed_circuit(10, 120, 1), {"tile_qubits": 8}),
    ("mixed11", lambda: mixed_circuit(11, 200, 2), {"tile_qubits": 10}),
    ("mixed9-c64", lambda: mixed_circuit(9, 150, 3), {"precision": "c64", "tile_qubits": 9}),
    ("qft12", lambda: ev.algorithms.qft(12), {"tile_qubits": 10}),
    ("mixed13", lambda: mixed_circuit(13, 160, 4), {"tile_qubits": 12}),
    ("mixed15_tile13", lambda: mixed_circuit(15, 220, 5), {"tile_qubits": 13}),
    ("phases14-t10", lambda: phase_circuit(14, 4, 5), {"tile_qubits": 10}),
    ("phases13-t9-c64", lambda: phase_circuit(13, 3, 6), {"tile_qubits": 9, "precision": "c64"}),
    ("qft15-t10", lambda: ev.algorithms.qft(15), {"tile_qubits": 10}),
]


@pytest.mark.parametrize("name,make,opts", CASES, ids=[c[0] for c in CASES])
def test_plan_emulation_matches_oracle(name, make, opts):
    c = make()
    n = c.num_qubits()
    passes, stages, ops, coef, t, tail = get_plan(c, with_tail=True, **opts)
    check_plan(n, passes, stages, ops, coef, coalesce=4 if opts.get("precision") == "c64" else 3)
    rng = np.random.default_rng(n)
    psi = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
    psi /= np.linalg.norm(psi)
    want = psi.copy()
    oracle.sv_apply_ops(c.ops(), want)
    got = emulate(psi.copy(), n, passes, stages, ops, coef, tail)
    assert np.max(np.abs(got - want)) < 1e-12


@pytest.mark.parametrize("n,tile", [(16, 10), (18, 9), (20, 12)])
def test_qft_swap_tail_is_peeled_and_emulates(n, tile):
    # the QFT's bit reversal (n/2 disjoint SWAPs wider than a tile) leaves the passes and runs as one
    # permutation after them, in the plan a single run executes; the emulated plan equals the oracle
    c = ev.workloads.config1(n, 2)
    passes, stages, ops, coef, t, tail = get_plan(c, tile_qubits=tile, with_tail=True)
    assert len(tail) >= 2 and all(min(a, b) >= 3 for a, b in tail)
    qs = [q for pr in tail for q in pr]
    assert len(set(qs)) == len(qs)
    check_plan(n, passes, stages, ops, coef)
    rng = np.random.default_rng(n)
    psi = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
    psi /= np.linalg.norm(psi)
    want = psi.copy()
    oracle.sv_apply_ops(c.ops(), want)
    assert np.max(np.abs(emulate(psi.copy(), n, passes, stages, ops, coef, tail) - want)) < 1e-12


def test_plan_structure_at_30_qubits():
    c = ev.workloads.config1()
    passes, stages, ops, coef, t, tail = get_plan(c, with_tail=True)
    check_plan(30, passes, stages, ops, coef)
    tuned = ev.tune_get(30, "c128")
    assert t == (tuned["tile_qubits"] if tuned else 12)
    # fusion must beat one sweep per gate by a wide margin
    assert len(passes) < len(c) / 20
    # the QFT's bit reversal above the 3 coalesced qubits is one permutation pass after the fused ones
    assert sorted(tail) == sorted((29 - i, i) for i in range(3, 15))


def test_plan_structure_c64_30_qubits():
    c = ev.workloads.config1()
    passes, stages, ops, coef, t = get_plan(c, precision="c64", tile_qubits=12)
    check_plan(30, passes, stages, ops, coef, coalesce=4)
    assert t == 12


def test_tile_option_bounds():
    c = ev.workloads.config0(16, 2)
    with pytest.raises(ev.InvalidArgument):
        get_plan(c, tile_qubits=14)  # 2^14 amplitudes need 1024 threads, above the kernel's launch bound
    with pytest.raises(ev.InvalidArgument):
        get_plan(c, tile_qubits=7)
    assert get_plan(c, tile_qubits=13)[4] == 13


def test_controlled_phases_merge_into_phase_ops():
    # qft(12): 66 controlled phases; within a stage those sharing a register
    # bit (or acting on thread bits only) become one lane-factor op each
    from tests.fused_emulator import PHASE, PHASE_ALL, decode
    c = ev.algorithms.qft(12)
    passes, stages, ops, coef, t = get_plan(c, tile_qubits=10)
    kinds = [decode(o)[0] for o in ops]
    assert kinds.count(PHASE) > 0
    assert sum(decode(o)[2] for o in ops if decode(o)[0] == PHASE) > 2 * kinds.count(PHASE)
    assert len(ops) < 66
    _, _, ops1, _, _ = get_plan(mixed_circuit(10, 120, 1), tile_qubits=8)
    assert any(decode(o)[0] == PHASE and decode(o)[1] == PHASE_ALL for o in ops1)


def test_tile_cnots_become_exchange_relabellings():
    # a CNOT between tile qubits that does not fit a stage's registers is
    # applied by the next exchange's read tables, never as a lane-conditioned op
    from tests.fused_emulator import CX, decode
    c = ev.workloads.config0(12, 8)
    passes, stages, ops, coef, t = get_plan(c, tile_qubits=9)
    assert sum(int(st["relabel"]) for st in stages) > 0
    for P in passes:
        tile = set(int(q) for q in P["q"][: int(P["t"])])
        for st in stages[int(P["stage_begin"]): int(P["stage_begin"]) + int(P["nstages"])]:
            for o in ops[st["op_begin"]: st["op_end"]]:
                if decode(o)[0] == CX and int(o["cglob"]):
                    assert not any((int(o["cglob"]) >> q) & 1 for q in tile)


def _bank_degree(vecs, bits):
    """Lanes per LDS bank when lane bit i adds vecs[i] to the slot (mod 2^bits)."""
    m = (1 << bits) - 1
    vals = set()
    for x in range(1 << bits):
        v = 0
        for i in range(bits):
            if (x >> i) & 1:
                v ^= vecs[i]
        vals.add(v & m)
    return (1 << bits) // len(vals)


@pytest.mark.parametrize("precision,bits", [("c128", 5), ("c64", 6)])
def test_pass_swizzle_keeps_exchanges_conflict_free(precision, bits):
    # the planner picks each pass's LDS swizzle so the 32 lanes of an 8-byte
    # exchange access (complex128) and the 64 of a 4-byte one (complex64) hit
    # distinct banks
    from tests.fused_emulator import swz

    c = ev.workloads.config1(30, 40)
    passes, stages, ops, coef, t = get_plan(c, precision=precision, tile_qubits=12, low_qubits=3,
                                            max_gates_per_pass=1000)
    total = free5 = free6 = 0
    for P in passes:
        sa, sb = int(P["swz_a"]), int(P["swz_b"])
        b0, ns = int(P["stage_begin"]), int(P["nstages"])
        for k in range(ns - 1):
            w, r = stages[b0 + k], stages[b0 + k + 1]
            wl = [int(swz(1 << int(w["p"][i]), sa, sb)) for i in range(6)]
            if int(r["relabel"]):
                rl = [int(r["rimg"][i]) for i in range(6)]
            else:
                rl = [int(swz(1 << int(r["p"][i]), sa, sb)) for i in range(6)]
            total += 1
            free5 += _bank_degree(wl, 5) == 1 and _bank_degree(rl, 5) == 1
            free6 += _bank_degree(wl, 6) == 1 and _bank_degree(rl, 6) == 1
    assert total > 100
    assert (free5 if bits == 5 else free6) >= 0.85 * total, (total, free5, free6)
