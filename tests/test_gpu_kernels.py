"""Per-gate HIP kernels against the oracle, through the C ABI.

The per-gate kernels evaluate the reference's exact arithmetic (same complex
product with mul_add, same sum order), so for complex128 they must agree with
the oracle exactly: max-abs difference 0 (signed zeros compare equal).
Reductions run in a different order on the device, so they are held to 1e-12
relative.
"""
import numpy as np
import pytest

import everett_amd as ev
import oracle
from tests.helpers import gate_set, random_state, random_unitary, sample_state

pytestmark = pytest.mark.gpu

CX = np.eye(4, dtype=np.complex128)[[0, 1, 3, 2]]
CZ = np.diag([1, 1, 1, -1]).astype(np.complex128)
SW = np.eye(4, dtype=np.complex128)[[0, 2, 1, 3]]


class G:
    def __init__(self, m):
        self.m = np.asarray(m, dtype=np.complex128).reshape(-1)


def exact(a, b):
    return float(np.max(np.abs(a - b))) == 0.0


@pytest.mark.parametrize("n", [1, 2, 3, 6, 10, 14])
def test_kernel_apply_1q_exact(n):
    rng = np.random.default_rng(n)
    for name, g in gate_set().items():
        for k in range(n):
            psi = random_state(rng, n)
            want = psi.copy()
            oracle.apply_1q(want, k, g)
            got = psi.copy()
            ev.kernel.apply_1q(got, k, G(g))
            assert exact(got, want), (name, n, k)


@pytest.mark.parametrize("n", [2, 3, 5, 8])
def test_kernel_apply_2q_exact(n):
    rng = np.random.default_rng(50 + n)
    for name, g in {"cnot": CX, "cz": CZ, "swap": SW, "dense": random_unitary(rng, 4)}.items():
        for a in range(n):
            for b in range(n):
                if a == b:
                    continue
                psi = sample_state(n) if name != "dense" else random_state(rng, n)
                want = psi.copy()
                oracle.apply_2q(want, a, b, g)
                got = psi.copy()
                ev.kernel.apply_2q(got, a, b, G(g))
                assert exact(got, want), (name, a, b)


@pytest.mark.parametrize("n", [2, 4, 7, 12])
def test_kernel_controlled_exact(n):
    rng = np.random.default_rng(80 + n)
    for name, g in gate_set().items():
        for target in range(n):
            others = [q for q in range(n) if q != target]
            for nc in range(0, min(4, len(others)) + 1):
                controls = [int(x) for x in rng.choice(others, size=nc, replace=False)] if nc else []
                psi = random_state(rng, n)
                want = psi.copy()
                oracle.apply_controlled_1q(want, controls, target, g)
                got = psi.copy()
                ev.kernel.apply_controlled_1q(got, controls, target, G(g))
                assert exact(got, want), (name, target, controls)


def test_kernel_preconditions_are_errors():
    a = np.zeros(8, dtype=np.complex128)
    with pytest.raises(ev.QubitOutOfRange):
        ev.kernel.apply_1q(a, 3, ev.Gate1.h())
    with pytest.raises(ev.DuplicateQubit):
        ev.kernel.apply_2q(a, 1, 1, ev.Gate2.cnot())
    with pytest.raises(ev.DuplicateQubit):
        ev.kernel.apply_controlled_1q(a, [0, 2], 2, ev.Gate1.x())
    with pytest.raises(ev.DimensionMismatch):
        ev.kernel.apply_1q(np.zeros(6, dtype=np.complex128), 0, ev.Gate1.h())


def test_device_state_ops_and_observables():
    rng = np.random.default_rng(9)
    n = 11
    psi = random_state(rng, n)
    s = ev.State.from_amplitudes(psi)
    ref = psi.copy()
    gs = gate_set()
    for i, (name, g) in enumerate(gs.items()):
        s.apply_1q(i % n, G(g))
        oracle.apply_1q(ref, i % n, g)
    s.apply_2q(3, 7, G(CX))
    oracle.apply_2q(ref, 3, 7, CX)
    s.apply_controlled_1q([1, 9], 4, G(gs["u"]))
    oracle.apply_controlled_1q(ref, [1, 9], 4, gs["u"])
    assert exact(s.amplitudes(), ref)
    assert abs(s.norm_sqr() - oracle.state_norm_sqr(ref)) < 1e-12
    other = ev.State.from_amplitudes(psi)
    assert abs(s.overlap(other) - oracle.state_overlap(ref, psi)) < 1e-12
    assert abs(s.fidelity(other) - abs(oracle.state_overlap(ref, psi)) ** 2) < 1e-12
    for b in (0, 5, 2047):
        assert s.probability(b) == oracle.norm_sqr(ref[b])
    x, y, z = s.bloch_vector(2)
    bit = 1 << 2
    r00 = sum(abs(ref[j]) ** 2 for j in range(1 << n) if not j & bit)
    r11 = sum(abs(ref[j]) ** 2 for j in range(1 << n) if j & bit)
    r01 = sum(ref[j] * np.conj(ref[j | bit]) for j in range(1 << n) if not j & bit)
    assert abs(x - 2 * r01.real) < 1e-12 and abs(y - 2 * r01.imag) < 1e-12 and abs(z - (r00 - r11)) < 1e-12
    c = s.clone()
    c.normalize()
    assert abs(c.norm_sqr() - 1.0) < 1e-14


def test_measure_matches_oracle():
    rng = np.random.default_rng(10)
    for n in (1, 4, 9):
        for k in range(n):
            psi = random_state(rng, n)
            s = ev.State.from_amplitudes(psi)
            ref = psi.copy()
            r1, r2 = ev.Rng.seed_from_u64(1000 * n + k), oracle.Rng(1000 * n + k)
            out = ev.measure_qubit(s, k, r1)
            assert out == oracle.measure_qubit(ref, k, r2)
            assert np.max(np.abs(s.amplitudes() - ref)) < 1e-13
            assert r1.state == r2.state


@pytest.mark.parametrize("n", [1, 3, 10, 13, 27])
def test_sampling_matches_oracle_exactly(n):
    # from 27 qubits on the chunks grow past 2^10 amplitudes (at most 2^16 chunks)
    rng = np.random.default_rng(20 + n)
    psi = random_state(rng, n)
    s = ev.State.from_amplitudes(psi)
    got = s.sample(1024, seed=123)
    r = oracle.Rng(123)
    u = np.array([r.next_f64() for _ in range(1024)])
    assert np.array_equal(got, oracle.sample(psi, n, u))


@pytest.mark.parametrize("n", [1, 5, 12])
def test_c64_kernels_within_bound(n):
    rng = np.random.default_rng(30 + n)
    psi = random_state(rng, n)
    s = ev.State.from_amplitudes(psi, precision="c64")
    ref = psi.copy()
    for i, (name, g) in enumerate(gate_set().items()):
        s.apply_1q(i % n, G(g))
        oracle.apply_1q(ref, i % n, g)
    if n >= 2:
        s.apply_2q(0, n - 1, G(CX))
        oracle.apply_2q(ref, 0, n - 1, CX)
    got = s.amplitudes()
    assert np.max(np.abs(got - ref)) < 1e-6
    assert abs(s.norm_sqr() - 1.0) < 1e-6
