"""Closed-form checks of feed-forward and entangling circuits on the device,
restating the reference's algorithm tests on the public API:

- teleportation (src/algorithms/teleport.rs:45-79; tests/algorithms_teleport.rs:12-91):
  qubit 2's Bloch vector equals the message's for every seed, because the
  feed-forward corrections make all four measurement branches agree;
- Bell and GHZ states (src/algorithms/prep.rs:23-60; tests/algorithms_prep.rs:6-65).

Teleportation also runs inside a 12-qubit register (the message on qubit 9,
the target on qubit 3, the other qubits in a brick-wall state of their own),
so Measure/IfClassic go through the fused path's segments, and on the
density-matrix backend.
"""
import math

import numpy as np
import pytest

import everett_amd as ev

pytestmark = pytest.mark.gpu


def expected_bloch(alpha: complex, beta: complex):  # tests/algorithms_teleport.rs:19-27
    r01 = alpha * beta.conjugate()
    return np.array([2 * r01.real, 2 * r01.imag, abs(alpha) ** 2 - abs(beta) ** 2])


def prep_gate(alpha: complex, beta: complex):  # teleport.rs: Gate1::new(alpha, -conj(beta), beta, conj(alpha))
    return ev.Gate1(alpha, -beta.conjugate(), beta, alpha.conjugate())


def teleport_circuit(alpha, beta, n=3, msg=0, mid=1, dst=2, spectators=()):
    """teleport_state on qubits (msg, mid, dst) of an n-qubit register; the
    spectator qubits get a brick-wall state of their own first."""
    c = ev.Circuit.with_classical(n, 2)
    if spectators:
        r = ev.Rng.seed_from_u64(99)
        for layer in range(3):
            for q in spectators:
                c.rx(q, 6.28 * r.next_f64()).ry(q, 6.28 * r.next_f64())
            for a, b in zip(spectators[layer % 2::2], spectators[layer % 2 + 1::2]):
                c.cnot(a, b)
    c.gate1(prep_gate(alpha, beta), msg)
    c.h(mid).cnot(mid, dst)  # teleport.rs:45-55
    c.cnot(msg, mid).h(msg)
    c.measure(msg, 0).measure(mid, 1)
    c.x_if(1, dst).z_if(0, dst)
    return c


S = 1 / math.sqrt(2)
(SH, CH) = math.sin(0.35), math.cos(0.35)
MESSAGES = [
    (1 + 0j, 0j),
    (0j, 1 + 0j),
    (S + 0j, S + 0j),
    (complex(CH * math.cos(-0.55), CH * math.sin(-0.55)), complex(SH * math.cos(0.55), SH * math.sin(0.55))),
    (S + 0j, S * 1j),
]


@pytest.mark.parametrize("plain", [True, False])
@pytest.mark.parametrize("alpha,beta", MESSAGES)
def test_teleport_bloch_is_the_message_for_every_seed(alpha, beta, plain):
    exp = expected_bloch(alpha, beta)
    c = teleport_circuit(alpha, beta)
    for seed in range(8):  # teleport_is_seed_independent
        ex = ev.StateVectorBackend.run_seeded(c, seed, plain=plain)
        assert np.max(np.abs(np.array(ex.state().bloch_vector(2)) - exp)) < 1e-10, seed


@pytest.mark.parametrize("precision,tol", [("c128", 1e-10), ("c64", 1e-6)])
def test_teleport_inside_a_fused_register(precision, tol):
    n, msg, mid, dst = 12, 9, 6, 3
    spect = [q for q in range(n) if q not in (msg, mid, dst)]
    outcomes = set()
    for alpha, beta in MESSAGES[2:]:
        exp = expected_bloch(alpha, beta)
        c = teleport_circuit(alpha, beta, n, msg, mid, dst, spect)
        for seed in range(8):
            ex = ev.StateVectorBackend.run_seeded(c, seed, precision=precision)
            outcomes.add(tuple(ex.classical()))
            got = np.array(ex.state().bloch_vector(dst))
            assert np.max(np.abs(got - exp)) < tol, (alpha, beta, seed)
    assert len(outcomes) >= 3  # the corrections were exercised on several branches


def test_teleport_on_the_density_matrix_backend():
    # expectation_pauli is Tr(P rho) (density.rs:237-261), so <Y> = -2 Im(alpha conj(beta)), while
    # State::bloch_vector reports y = 2 Im(sum a0 conj(a1)) (state.rs:163-185): the y signs differ
    for alpha, beta in MESSAGES:
        exp = expected_bloch(alpha, beta) * np.array([1.0, -1.0, 1.0])
        c = teleport_circuit(alpha, beta)
        for seed in range(4):
            rho = ev.DensityMatrixBackend.run_with_noise_seeded(c, ev.NoiseModel.ideal(), seed).density_matrix()
            got = np.array([rho.expectation_pauli(2, p) for p in "XYZ"])
            assert np.max(np.abs(got - exp)) < 1e-10, (alpha, beta, seed)


def ghz(n):  # prep.rs:43-60
    c = ev.Circuit(n)
    c.h(0)
    for k in range(n - 1):
        c.cnot(k, k + 1)
    return c


def test_bell_and_small_ghz_probabilities():
    bell = ev.Circuit(2)
    bell.h(0).cnot(0, 1)  # prep.rs:23-27
    s = ev.StateVectorBackend.run(bell).state()
    assert abs(s.probability(0b00) - 0.5) < 1e-12 and abs(s.probability(0b11) - 0.5) < 1e-12
    assert s.probability(0b01) < 1e-12 and s.probability(0b10) < 1e-12
    assert abs(s.norm_sqr() - 1.0) < 1e-12
    for n in (3, 4, 5):
        s = ev.StateVectorBackend.run(ghz(n)).state()
        assert abs(s.probability(0) - 0.5) < 1e-12 and abs(s.probability((1 << n) - 1) - 0.5) < 1e-12
        for b in range(1, (1 << n) - 1):
            assert s.probability(b) < 1e-12
    s = ev.StateVectorBackend.run(ghz(3)).state()
    for q in range(3):  # ghz_qubits_are_maximally_mixed
        assert np.linalg.norm(s.bloch_vector(q)) < 0.01


@pytest.mark.parametrize("n", [16, 26])
def test_large_ghz_on_the_fused_path(n):
    s = ev.StateVectorBackend.run(ghz(n)).state()
    p0, p1 = s.probability(0), s.probability((1 << n) - 1)
    assert abs(p0 - 0.5) < 1e-12 and abs(p1 - 0.5) < 1e-12
    assert abs(s.norm_sqr() - 1.0) < 1e-12  # so every other basis state holds nothing
    for q in (0, n // 2, n - 1):
        assert np.linalg.norm(s.bloch_vector(q)) < 1e-12
