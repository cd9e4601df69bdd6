"""Pins the CPU oracle (oracle/refcpu.c) before anything is compared with it.

Each check is chosen so that a plausible restatement mistake (a dropped term,
a wrong sign, a transposed operand, a wrong bit position) fails it:
published RNG vectors, the reference's own inline known answers, brute-force
Kronecker products, closed forms and physical invariants.
"""
import math

import numpy as np
import pytest

import oracle
from tests.helpers import (full_1q, full_controlled, full_op, gate_set, kats, ops_array, random_rho, random_state,
                           random_unitary, sample_state)

TOL = 1e-12


# ------------------------------------------------------------------ complex

def test_complex_known_answers():
    k = kats()
    for case in k["complex_mul"]:
        got = oracle.cmul(complex(*case["a"]), complex(*case["b"]))
        assert got == complex(*case["expect"]), case["src"]
    for case in k["norm_sqr"]:
        assert oracle.norm_sqr(complex(*case["a"])) == case["expect"]


def test_complex_mul_rounding_is_mul_add():
    # (a+bi)(c+di) real part = fma(a, c, -(b*d)): pick values where fused and
    # unfused rounding differ, so a missing fma fails.
    a = complex(1.0 + 2.0**-30, 3.0)
    b = complex(1.0 - 2.0**-30, 1.0 / 3.0)
    got = oracle.cmul(a, b)
    exp_re = math.fma(a.real, b.real, -(a.imag * b.imag)) if hasattr(math, "fma") else None
    if exp_re is not None:
        assert got.real == exp_re
    # the conj product identity of complex.rs:221-226
    z = complex(-2.0, 5.0)
    p = oracle.cmul(z, z.conjugate())
    assert abs(p.real - oracle.norm_sqr(z)) < 1e-15 and abs(p.imag) < 1e-15


# ---------------------------------------------------------------------- rng

def test_xoshiro256pp_published_vector():
    v = kats()["xoshiro256pp_state_1234"]
    r = oracle.Rng(state=v["state"])
    assert [r.next_u64() for _ in range(len(v["expect"]))] == v["expect"]


def test_splitmix64_seed0_published_words():
    r = oracle.Rng(0)
    assert list(r.state) == [int(h, 16) for h in kats()["splitmix64_seed0"]["expect"]]


def test_rng_properties():  # src/rng.rs:59-96
    a, b = oracle.Rng(42), oracle.Rng(42)
    assert [a.next_u64() for _ in range(100)] == [b.next_u64() for _ in range(100)]
    assert oracle.Rng(1).next_u64() != oracle.Rng(2).next_u64()
    r = oracle.Rng(7)
    xs = [r.next_f64() for _ in range(10000)]
    assert all(0.0 <= x < 1.0 for x in xs)
    r = oracle.Rng(123)
    assert abs(np.mean([r.next_f64() for _ in range(100000)]) - 0.5) < 0.01


def test_next_f64_is_top_53_bits():
    r1, r2 = oracle.Rng(99), oracle.Rng(99)
    for _ in range(50):
        assert r1.next_f64() == (r2.next_u64() >> 11) * 2.0**-53


# ---------------------------------------------------------------- index maps

@pytest.mark.parametrize("q", range(1, 9))
def test_index_pair_bijection(q):  # kernel.rs:444-487 (Kani claims), exhaustively
    n = 1 << q
    for k in range(q):
        seen = set()
        for i in range(n // 2):
            i0, i1 = oracle.index_pair(k, i)
            assert i0 < n and i1 < n and i1 == i0 | (1 << k) and i0 & (1 << k) == 0
            seen.update((i0, i1))
        assert len(seen) == n


@pytest.mark.parametrize("q", range(2, 7))
def test_index_quad_partition(q):  # kernel.rs:489-512, exhaustively
    n = 1 << q
    for a in range(q):
        for b in range(q):
            if a == b:
                continue
            seen = []
            for i in range(n // 4):
                idx = oracle.index_quad(a, b, i)
                assert idx[1] == idx[0] | (1 << b) and idx[2] == idx[0] | (1 << a)
                seen += idx
            assert sorted(seen) == list(range(n))


# ------------------------------------------------------------------- kernels

@pytest.mark.parametrize("n", [1, 2, 3, 5, 7])
def test_apply_1q_matches_kronecker(n):
    rng = np.random.default_rng(n)
    for name, g in gate_set().items():
        for k in range(n):
            psi = random_state(rng, n)
            want = full_1q(n, k, g) @ psi
            got = psi.copy()
            oracle.apply_1q(got, k, g)
            assert np.max(np.abs(got - want)) < TOL, (name, n, k)


@pytest.mark.parametrize("n", [2, 3, 5, 6])
def test_apply_2q_matches_kronecker(n):
    rng = np.random.default_rng(100 + n)
    cx = np.eye(4)[[0, 1, 3, 2]]
    cz = np.diag([1, 1, 1, -1])
    sw = np.eye(4)[[0, 2, 1, 3]]
    gates = {"cnot": cx, "cz": cz, "swap": sw, "dense": random_unitary(rng, 4)}
    for name, g in gates.items():
        for a in range(n):
            for b in range(n):
                if a == b:
                    continue
                psi = random_state(rng, n)
                want = full_op(n, [a, b], g) @ psi
                got = psi.copy()
                oracle.apply_2q(got, a, b, g)
                assert np.max(np.abs(got - want)) < TOL, (name, a, b)


def test_specialized_2q_on_sample_state():  # kernel.rs:637-672
    cx = np.eye(4)[[0, 1, 3, 2]]
    cz = np.diag([1, 1, 1, -1])
    sw = np.eye(4)[[0, 2, 1, 3]]
    for n in range(2, 8):
        for a in range(n):
            for b in range(n):
                if a == b:
                    continue
                for g in (cx, cz, sw):
                    base = sample_state(n)
                    got = base.copy()
                    oracle.apply_2q(got, a, b, g)
                    # permutations and sign flips are exact
                    assert np.array_equal(got, full_op(n, [a, b], g).real @ base)


@pytest.mark.parametrize("n", [2, 3, 5])
def test_controlled_matches_kronecker(n):
    rng = np.random.default_rng(200 + n)
    for name, g in gate_set().items():
        for target in range(n):
            others = [q for q in range(n) if q != target]
            for nc in range(0, min(3, len(others)) + 1):
                controls = list(rng.choice(others, size=nc, replace=False)) if nc else []
                psi = random_state(rng, n)
                want = full_controlled(n, controls, target, g) @ psi
                got = psi.copy()
                oracle.apply_controlled_1q(got, controls, target, g)
                assert np.max(np.abs(got - want)) < TOL, (name, target, controls)


def test_known_answer_states():
    k = kats()
    a = np.array([1, 0], dtype=np.complex128)
    oracle.apply_1q(a, 0, [0, 1, 1, 0])
    assert np.array_equal(a, np.array([0, 1], dtype=np.complex128))  # kernel.rs:541-547
    c = k["cnot_10_to_11"]
    s = np.zeros(4, dtype=np.complex128)
    s[c["input_basis"]] = 1
    oracle.apply_2q(s, c["control"], c["target"], np.eye(4)[[0, 1, 3, 2]])
    assert abs(s[c["expect_basis"]] - 1) < 1e-12
    st = np.array([complex(*x) for x in k["controlled_x_matches_cnot"]["state"]])
    v1, v2 = st.copy(), st.copy()
    oracle.apply_2q(v1, 1, 0, np.eye(4)[[0, 1, 3, 2]])
    oracle.apply_controlled_1q(v2, [1], 0, [0, 1, 1, 0])
    assert np.max(np.abs(v1 - v2)) < 1e-12


# -------------------------------------------------------------- measurement

def test_measure_zero_state():  # measure.rs:59-67
    s = np.array([1, 0], dtype=np.complex128)
    assert oracle.measure_qubit(s, 0, oracle.Rng(0)) is False
    assert abs(abs(s[0]) ** 2 - 1) < 1e-12


def test_measure_collapses_and_renormalizes():  # measure.rs:69-79
    s = np.array([1, 0], dtype=np.complex128)
    oracle.apply_1q(s, 0, gate_set()["h"])
    out = oracle.measure_qubit(s, 0, oracle.Rng(99))
    assert abs(oracle.state_norm_sqr(s) - 1) < 1e-12
    assert abs(abs(s[int(out)]) ** 2 - 1) < 1e-12


def test_measure_statistics_and_repeat():  # measure.rs:81-106
    r = oracle.Rng(2024)
    ones = 0
    for _ in range(4000):
        s = np.array([1, 0], dtype=np.complex128)
        oracle.apply_1q(s, 0, gate_set()["h"])
        ones += oracle.measure_qubit(s, 0, r)
    assert abs(ones / 4000 - 0.5) < 0.05
    s = np.zeros(4, dtype=np.complex128)
    s[0] = 1
    oracle.apply_1q(s, 0, gate_set()["h"])
    r = oracle.Rng(5)
    assert oracle.measure_qubit(s, 0, r) == oracle.measure_qubit(s, 0, r)


def test_measure_formula_against_numpy():
    rng = np.random.default_rng(3)
    for n in (1, 3, 6):
        for k in range(n):
            psi = random_state(rng, n)
            r1, r2 = oracle.Rng(11 * n + k), oracle.Rng(11 * n + k)
            u = r2.next_f64()
            p1 = sum(abs(psi[j]) ** 2 for j in range(1 << n) if (j >> k) & 1)
            s = psi.copy()
            out = oracle.measure_qubit(s, k, r1)
            assert out == (u < p1)
            keep = np.array([((j >> k) & 1) == out for j in range(1 << n)])
            want = np.where(keep, psi, 0) / math.sqrt(p1 if out else 1 - p1)
            assert np.max(np.abs(s - want)) < 1e-12


# ------------------------------------------------------------ circuits, qft

@pytest.mark.parametrize("n", range(1, 8))
def test_qft_closed_form(n):  # tests/algorithms_qft.rs:52-74, every basis input j
    N = 1 << n
    for j in sorted({0, 1, N - 1, N // 2, (5 * N) // 7}):
        prep = [(1, q, 0, [0, 1, 1, 0], None) for q in range(n) if (j >> q) & 1]
        ops = np.concatenate([ops_array(prep), oracle.gen_qft(n)]) if prep else oracle.gen_qft(n)
        amps, _ = oracle.sv_run(ops, n)
        k = np.arange(N)
        want = np.exp(2j * np.pi * j * k / N) / math.sqrt(N)
        assert np.max(np.abs(amps - want)) < 1e-12, (n, j)


def test_qft_structure():  # qft.rs:94-105
    ops = oracle.gen_qft(4)
    kinds = list(ops["kind"])
    assert kinds == [1, 3, 3, 3, 1, 3, 3, 1, 3, 1, 2, 2]
    # the first controlled phase is R_2 = phase(TAU/4) with control 2 on target 3
    assert ops[1]["target"] == 3 and ops[1]["controls"][0] == 2
    assert abs(ops[1]["m"][3] - 1j) < 1e-15


def test_brickwall_structure():
    r = oracle.Rng(20)
    ops = oracle.gen_brickwall(20, 20, r)
    assert len(ops) == 1390 and int((ops["kind"] == 1).sum()) == 1200 and int((ops["kind"] == 2).sum()) == 190
    r2 = oracle.Rng(20)
    tx = 2 * math.pi * r2.next_f64()
    assert ops[0]["target"] == 0 and ops[0]["m"][0] == complex(math.cos(tx / 2), 0)
    # layer 1 CNOTs start at qubit 1
    layer1 = ops[20 * 3 + 10 + 20 * 3: 20 * 3 + 10 + 20 * 3 + 9]
    assert list(layer1["target"]) == list(range(1, 19, 2))


def test_if_classic_and_measure_drive():  # backend/mod.rs:51-72
    # X(0); measure 0 -> c0 (=1); if c0: X(1); if c1 (=0): X(2)
    ops = ops_array([
        (1, 0, 0, [0, 1, 1, 0], None),
        (4, 0, 0, None, None),
        (5, 0, 0, None, None), (1, 1, 0, [0, 1, 1, 0], None),
        (5, 1, 0, None, None), (1, 2, 0, [0, 1, 1, 0], None),
    ])
    amps, bits = oracle.sv_run(ops, 3, num_classical=2)
    assert bits == [True, False]
    assert abs(amps[0b011] - 1) < 1e-12


# ------------------------------------------------------------ density matrix

def test_dm_unitary_matches_brute_force():  # density.rs:164-191
    rng = np.random.default_rng(11)
    for n in (1, 2, 3, 4):
        for t in range(n):
            g = random_unitary(rng, 2)
            rho = random_rho(rng, n)
            U = full_1q(n, t, g)
            want = U @ rho @ U.conj().T
            got = rho.reshape(-1).copy()
            oracle.dm_apply_unitary(got, n, g, [t])
            assert np.max(np.abs(got.reshape(rho.shape) - want)) < 1e-12
        if n >= 2:
            for a in range(n):
                for b in range(n):
                    if a == b:
                        continue
                    g = random_unitary(rng, 4)
                    rho = random_rho(rng, n)
                    U = full_op(n, [a, b], g)
                    want = U @ rho @ U.conj().T
                    got = rho.reshape(-1).copy()
                    oracle.dm_apply_unitary(got, n, g, [b, a])  # density_matrix.rs:132
                    assert np.max(np.abs(got.reshape(rho.shape) - want)) < 1e-12


@pytest.mark.parametrize("kind,p", [("depolarizing", 0.1), ("amplitude_damping", 0.3), ("phase_damping", 0.4),
                                    ("bit_flip", 0.2), ("dephasing", 0.3)])
def test_dm_channel_matches_brute_force(kind, p):  # density.rs:198-216
    rng = np.random.default_rng(12)
    ks = oracle.channel_kraus(kind, p)
    # completeness sum K^dag K = I (noise.rs:241-310)
    acc = sum(k.reshape(2, 2).conj().T @ k.reshape(2, 2) for k in ks)
    assert np.max(np.abs(acc - np.eye(2))) < 1e-12
    for n in (1, 3):
        for t in range(n):
            rho = random_rho(rng, n)
            want = sum(full_1q(n, t, k) @ rho @ full_1q(n, t, k).conj().T for k in ks)
            got = rho.reshape(-1).copy()
            oracle.dm_apply_channel(got, n, ks, t)
            assert np.max(np.abs(got.reshape(rho.shape) - want)) < 1e-12


def test_kraus_tables():  # noise.rs:186-231 entries
    p = 0.3
    d = oracle.channel_kraus("depolarizing", p)
    a, b = math.sqrt(1 - p), math.sqrt(p / 3)
    assert np.allclose(d[0], [a, 0, 0, a]) and np.allclose(d[2], [0, -1j * b, 1j * b, 0])
    ad = oracle.channel_kraus("amplitude_damping", p)
    assert np.allclose(ad[1], [0, math.sqrt(p), 0, 0])
    assert len(oracle.channel_kraus(None, 0.0)) == 0


def test_dm_backend_semantics():  # density_matrix.rs:261-382
    h = gate_set()["h"]
    cx = np.eye(4)[[0, 1, 3, 2]]
    bell = ops_array([(1, 0, 0, h, None), (2, 0, 1, cx, None)])
    sv, _ = oracle.sv_run(bell, 2)
    rho, _ = oracle.dm_run(bell, 2)
    assert np.max(np.abs(np.diag(rho.reshape(4, 4)).real - np.abs(sv) ** 2)) < 1e-12
    assert abs(oracle.dm_purity(rho, 2) - 1) < 1e-10
    rho, _ = oracle.dm_run(bell, 2, {"kind_1q": "depolarizing", "p_1q": 0.1, "kind_2q": "depolarizing", "p_2q": 0.1})
    assert abs(oracle.dm_trace(rho, 2) - 1) < 1e-10 and oracle.dm_purity(rho, 2) < 1
    x1 = ops_array([(1, 0, 0, [0, 1, 1, 0], None)])
    rho, _ = oracle.dm_run(x1, 1, {"kind_1q": "amplitude_damping", "p_1q": 0.99})
    assert rho[0].real > 0.9
    h1 = ops_array([(1, 0, 0, h, None)])
    rho, _ = oracle.dm_run(h1, 1, {"kind_1q": "dephasing", "p_1q": 0.5})
    assert abs(rho[1]) < 1e-12 and abs(rho[2]) < 1e-12
    hm = ops_array([(1, 0, 0, h, None), (4, 0, 0, None, None)])
    rho, _ = oracle.dm_run(hm, 1, num_classical=1)
    assert abs(oracle.dm_purity(rho, 1) - 1) < 1e-10


def test_dm_controlled_matches_brute_force():  # density_matrix.rs:140-153, 212-247
    rng = np.random.default_rng(13)
    n = 3
    g = random_unitary(rng, 2)
    ops = ops_array([(1, 0, 0, gate_set()["h"], None), (1, 1, 0, gate_set()["h"], None),
                     (3, 2, 0, g, [0, 1])])
    rho, _ = oracle.dm_run(ops, n)
    psi = np.zeros(8, dtype=np.complex128)
    psi[0] = 1
    psi = full_controlled(n, [0, 1], 2, g) @ full_1q(n, 1, gate_set()["h"]) @ full_1q(n, 0, gate_set()["h"]) @ psi
    assert np.max(np.abs(rho.reshape(8, 8) - np.outer(psi, psi.conj()))) < 1e-12


def test_dm_measure_readout_draw_order():  # density_matrix.rs:155-169
    # |1> measured with readout error 1.0 - eps: the recorded bit flips when the
    # second draw is below readout_error; it is drawn only when readout_error > 0.
    x1 = ops_array([(1, 0, 0, [0, 1, 1, 0], None), (4, 0, 0, None, None)])
    _, bits = oracle.dm_run(x1, 1, {"readout_error": 0.999999}, seed=3, num_classical=1)
    r = oracle.Rng(3)
    r.next_f64()
    assert bits == [not (r.next_f64() < 0.999999)]
    _, bits = oracle.dm_run(x1, 1, {}, seed=3, num_classical=1)
    assert bits == [True]


# ------------------------------------------------------------------ sampler

def test_sampler_semantics():
    rng = np.random.default_rng(5)
    for n in (1, 4, 11, 12):
        psi = random_state(rng, n)
        p = np.abs(psi) ** 2
        cdf = np.cumsum(p)
        u = np.array([0.0, 0.25, 0.5, 0.999, 1e-9])
        got = oracle.sample(psi, n, u)
        for ui, gi in zip(u, got):
            # first index whose cumulative weight exceeds u * total
            want = int(np.searchsorted(cdf, ui * cdf[-1], side="right"))
            assert abs(int(gi) - want) <= 1 and (cdf[int(gi)] >= ui * cdf[-1] - 1e-12)
    psi = np.zeros(8, dtype=np.complex128)
    psi[5] = 1
    assert list(oracle.sample(psi, 3, np.array([0.0, 0.5, 0.9999999]))) == [5, 5, 5]


def test_sampler_distribution_chi2():
    rng = np.random.default_rng(6)
    n = 5
    psi = random_state(rng, n)
    r = oracle.Rng(77)
    u = np.array([r.next_f64() for _ in range(40000)])
    counts = np.bincount(oracle.sample(psi, n, u).astype(np.int64), minlength=32)
    expect = 40000 * np.abs(psi) ** 2
    chi2 = float(np.sum((counts - expect) ** 2 / expect))
    assert chi2 < 80  # 31 dof; p ~ 1e-6 at this bound
