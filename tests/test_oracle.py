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


# ------------------------------------------------------------------- gates
#
# Textbook definitions (src/gate.rs:62-136 doc comments: R_a(t) = exp(-i t a/2)
# for the Pauli a) computed independently with scipy's matrix exponential,
# the reference's own identity tests (src/gate.rs:297-360), and compose
# against numpy's matrix product.  A sign slip in a rotation, a swapped
# sin/cos, a transposed constructor or a reversed compose fails one of them.

PAULI = {
    "x": np.array([[0, 1], [1, 0]], dtype=np.complex128),
    "y": np.array([[0, -1j], [1j, 0]], dtype=np.complex128),
    "z": np.array([[1, 0], [0, -1]], dtype=np.complex128),
}
ANGLES = [0.0, 1e-9, 0.3, -0.7, math.pi / 2, math.pi, 2.1, -math.pi, 3 * math.pi / 2, 2 * math.pi, 5.5, -11.0]


def _m2(v):
    return np.asarray(v, dtype=np.complex128).reshape(2, 2)


@pytest.mark.parametrize("axis", ["x", "y", "z"])
def test_rotations_match_matrix_exponential(axis):
    from scipy.linalg import expm

    for th in ANGLES + list(np.linspace(-4 * math.pi, 4 * math.pi, 37)):
        want = expm(-0.5j * th * PAULI[axis])
        got = _m2(oracle.gate1("r" + axis, th))
        assert np.max(np.abs(got - want)) < 2e-15, (axis, th)


def test_fixed_gates_match_textbook():
    s = math.sqrt(0.5)  # correctly rounded 1/sqrt(2), the value of FRAC_1_SQRT_2 (gate.rs:81-85)
    assert np.array_equal(_m2(oracle.gate1("x")), PAULI["x"])
    assert np.array_equal(_m2(oracle.gate1("y")), PAULI["y"])
    assert np.array_equal(_m2(oracle.gate1("z")), PAULI["z"])
    assert np.array_equal(_m2(oracle.gate1("h")), np.array([[s, s], [s, -s]], dtype=np.complex128))
    assert np.array_equal(_m2(oracle.gate1("s")), np.diag([1, 1j]))
    assert np.max(np.abs(_m2(oracle.gate1("t")) - np.diag([1, np.exp(1j * math.pi / 4)]))) < 2e-16
    for lam in ANGLES:
        assert np.max(np.abs(_m2(oracle.gate1("phase", lam)) - np.diag([1, np.exp(1j * lam)]))) < 2e-16


def test_two_qubit_gates_match_textbook():
    # rows/columns |q_a q_b> = |00>, |01>, |10>, |11> (src/gate.rs:186-232)
    perm = lambda p: np.eye(4, dtype=np.complex128)[p]  # noqa: E731  row r has its 1 in column p[r]
    assert np.array_equal(oracle.gate2("cnot").reshape(4, 4), perm([0, 1, 3, 2]))
    assert np.array_equal(oracle.gate2("swap").reshape(4, 4), perm([0, 2, 1, 3]))
    assert np.array_equal(oracle.gate2("cz").reshape(4, 4), np.diag([1, 1, 1, -1]).astype(np.complex128))
    # CNOT acting on a product state |a>|b> flips b exactly when a is set
    for a in range(2):
        for b in range(2):
            e = np.zeros(4)
            e[2 * a + b] = 1
            out = oracle.gate2("cnot").reshape(4, 4) @ e
            assert out[2 * a + (b ^ a)] == 1


def test_compose_is_matrix_product_in_order():
    rng = np.random.default_rng(5)
    for _ in range(20):
        a = rng.standard_normal((2, 2)) + 1j * rng.standard_normal((2, 2))
        b = rng.standard_normal((2, 2)) + 1j * rng.standard_normal((2, 2))
        got = _m2(oracle.gate_compose(a, b))
        assert np.max(np.abs(got - a @ b)) < 1e-14
        assert np.max(np.abs(a @ b - b @ a)) > 1e-3  # the order is observable
        assert np.array_equal(_m2(oracle.gate_adjoint(a)), a.conj().T)


def _is_identity(m, eps=1e-12):  # gate.rs:166-170
    return np.max(np.abs(_m2(m) - np.eye(2))) <= eps


def _is_unitary(m, eps=1e-12):  # gate.rs:174-177
    return _is_identity(oracle.gate_compose(oracle.gate_adjoint(m), m), eps)


def test_reference_gate_identities():  # src/gate.rs:297-360, restated on the oracle
    g = {k: oracle.gate1(k) for k in ("x", "y", "z", "h", "s", "t")}
    for name, m in g.items():  # paulis_are_unitary
        assert _is_unitary(m), name
    for th in ANGLES:
        for r in ("rx", "ry", "rz", "phase"):
            assert _is_unitary(oracle.gate1(r, th)), (r, th)
    assert _is_identity(oracle.gate_compose(g["x"], g["x"]))  # pauli_x_is_its_own_inverse
    zd = oracle.gate_adjoint(g["z"])
    s2 = oracle.gate_compose(g["s"], g["s"])
    assert _is_identity(oracle.gate_compose(s2, zd))  # s_squared_is_z
    t4 = oracle.gate_compose(oracle.gate_compose(oracle.gate_compose(g["t"], g["t"]), g["t"]), g["t"])
    assert _is_identity(oracle.gate_compose(t4, zd))  # t_to_the_fourth_is_z
    assert not _is_identity(oracle.gate_compose(oracle.gate_compose(g["t"], g["t"]), zd))  # T^2 = S, not Z
    rz = oracle.gate1("rz", math.pi)
    adjusted = np.array([oracle.cmul(complex(v), 1j) for v in rz])  # rz_pi_matches_z_up_to_global_phase
    assert _is_identity(oracle.gate_compose(adjusted, zd))
    for name in ("cnot", "cz", "swap"):  # two_qubit_gates_are_unitary
        u = oracle.gate2(name).reshape(4, 4)
        assert np.max(np.abs(u.conj().T @ u - np.eye(4))) <= 1e-12
    cx = oracle.gate_controlled(g["x"])  # controlled_x_is_cnot
    assert np.array_equal(cx, oracle.gate2("cnot"))
    assert np.array_equal(oracle.gate_controlled(g["z"]), oracle.gate2("cz"))  # controlled_z_matches_cz
    u = random_unitary(np.random.default_rng(3), 2)
    want = np.eye(4, dtype=np.complex128)
    want[2:, 2:] = u
    assert np.array_equal(oracle.gate_controlled(u).reshape(4, 4), want)


# ------------------------------------------------------------------- state

def test_overlap_is_conjugate_linear_in_other():  # state.rs:130-138: <other|self> = sum conj(other_j) self_j
    rng = np.random.default_rng(11)
    for n in (1, 3, 6):
        a = random_state(rng, n) * 1.7
        b = random_state(rng, n) * (0.4 - 0.9j)
        got = oracle.state_overlap(a, b)
        assert abs(got - np.vdot(b, a)) < 1e-13
        assert abs(np.vdot(b, a) - np.vdot(a, b)) > 1e-3  # the conjugated operand is observable


def test_fidelity_ignores_global_phase():  # state.rs:208-217
    a = np.zeros(4, dtype=np.complex128)
    a[0] = 1
    b = a * np.exp(0.7j)
    assert abs(abs(oracle.state_overlap(a, b)) ** 2 - 1.0) < 1e-12
    rng = np.random.default_rng(4)
    psi = random_state(rng, 5)
    assert abs(abs(oracle.state_overlap(psi, psi * np.exp(-2.3j))) ** 2 - 1.0) < 1e-12


def test_normalize_matches_numpy():  # state.rs:103-111 and normalize_fixes_drift (state.rs:227-233)
    s = np.array([2.0, 2.0j], dtype=np.complex128)
    oracle.state_normalize(s)
    assert abs(oracle.state_norm_sqr(s) - 1.0) < 1e-12
    assert np.max(np.abs(s - np.array([1, 1j]) / math.sqrt(2))) < 1e-15
    rng = np.random.default_rng(8)
    for n in (1, 4, 9):
        v = (rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)) * 3.1
        want = v / np.linalg.norm(v)
        assert abs(oracle.state_norm_sqr(v) - np.vdot(v, v).real) < 1e-10
        oracle.state_normalize(v)
        assert np.max(np.abs(v - want)) < 1e-15
    z = np.zeros(8, dtype=np.complex128)  # norm 0: left alone (the `if norm > 0.0` guard)
    oracle.state_normalize(z)
    assert not np.any(z)


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


def test_dm_random_structure():
    """gen_dm_random against the draw order written for it (Fisher-Yates
    i = n-1..1 with j = next_u64() % (i+1); per pair r = next_u64() % 6, r == 5 a
    CNOT, else kind r then kind next_u64() % 5; a leftover qubit draws one kind;
    Rx/Rz draw TAU * next_f64()), re-derived here from the stream."""
    n, depth = 7, 4
    ops = oracle.gen_dm_random(n, depth, oracle.Rng(13))
    r = oracle.Rng(13)
    want = []

    def one(kind, q):
        if kind in (3, 4):
            th = 2 * math.pi * r.next_f64()
            want.append((1, q, oracle.gate1("rx" if kind == 3 else "rz", th)))
        else:
            want.append((1, q, oracle.gate1(("h", "s", "t")[kind])))

    for _ in range(depth):
        perm = list(range(n))
        for i in range(n - 1, 0, -1):
            j = r.next_u64() % (i + 1)
            perm[i], perm[j] = perm[j], perm[i]
        p = 0
        while p + 1 < n:
            k = r.next_u64() % 6
            if k == 5:
                want.append((2, perm[p], perm[p + 1]))
            else:
                one(k, perm[p])
                one(r.next_u64() % 5, perm[p + 1])
            p += 2
        if p < n:
            one(r.next_u64() % 5, perm[p])
    assert len(ops) == len(want)
    for o, w in zip(ops, want):
        assert int(o["kind"]) == w[0] and int(o["target"]) == w[1]
        if w[0] == 1:
            assert np.array_equal(o["m"][:4], w[2])
        else:
            assert int(o["b"]) == w[2] and np.array_equal(o["m"], oracle.gate2("cnot"))


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


@pytest.mark.parametrize("n,seed", [(2, 1), (4, 2), (6, 3), (8, 4)])
def test_dm_superop_run_matches_the_literal_loops(n, seed):
    # ref_dm_run_superop (gates and channels as passes over the 2n-qubit vector) against ref_dm_run
    # (the reference's loops): dm_random circuits plus multi-controlled gates and mid-circuit
    # measurements, with no noise, named channels, and a custom Kraus set with readout error
    rng = np.random.default_rng(seed)
    ops = list(oracle.gen_dm_random(n, 6, oracle.Rng(seed)))
    extra = []
    if n >= 3:
        g = random_unitary(rng, 2).reshape(-1)
        extra.append((3, n - 1, 0, g, list(range(n - 1))))
        extra.append((4, 1, 0, None, None))
        extra.append((3, 0, 0, gate_set()["x"], [1, 2]))
    arr = np.concatenate([np.asarray(ops), ops_array(extra)]) if extra else np.asarray(ops)
    arr = np.concatenate([arr, ops_array([(4, q, q, None, None) for q in range(n)][: max(1, n // 2)])])
    ks = np.array([random_unitary(rng, 2).reshape(-1) * np.sqrt(0.5) for _ in range(2)])
    for noise in (None, {"kind_1q": "depolarizing", "p_1q": 0.03, "kind_2q": "amplitude_damping", "p_2q": 0.05},
                  {"custom_1q": ks, "custom_2q": ks, "readout_error": 0.2}):
        for s_ in (0, 5):
            a, ba = oracle.dm_run(arr, n, noise, s_, num_classical=n)
            b, bb = oracle.dm_run(arr, n, noise, s_, num_classical=n, superop=True)
            assert ba == bb, (noise, s_)
            assert np.max(np.abs(a - b)) < 1e-12, (noise, s_)


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
