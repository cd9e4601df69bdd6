"""StateVectorBackend / DensityMatrixBackend on the device against the oracle.

Plain (one launch per op) runs of measurement-free circuits evaluate the
reference's arithmetic exactly, so complex128 amplitudes match the oracle
with max-abs 0.  Measurement reduces p1 in a different order, so runs that
measure are held to 1e-12 and must reproduce the classical register exactly.
Density matrices fold gate and channel into one superoperator and are held
to the reference's own 1e-10 (src/backend/density_matrix.rs:369).
"""
import math

import numpy as np
import pytest

import everett_amd as ev
import oracle
from tests.helpers import random_unitary

pytestmark = pytest.mark.gpu


def random_circuit(n, nops, seed, ncl=0, measure=True):
    rng = np.random.default_rng(seed)
    c = ev.Circuit.with_classical(n, ncl)
    gates = [ev.Gate1.h(), ev.Gate1.x(), ev.Gate1.y(), ev.Gate1.z(), ev.Gate1.s(), ev.Gate1.t()]
    for _ in range(nops):
        r = rng.integers(0, 10)
        q = int(rng.integers(0, n))
        if r < 3:
            c.gate1(gates[int(rng.integers(0, len(gates)))], q)
        elif r < 5:
            th = float(rng.uniform(0, 2 * math.pi))
            c.gate1([ev.Gate1.rx, ev.Gate1.ry, ev.Gate1.rz, ev.Gate1.phase][int(rng.integers(0, 4))](th), q)
        elif r < 7 and n >= 2:
            b = int(rng.integers(0, n - 1))
            b = b + 1 if b >= q else b
            g = [ev.Gate2.cnot(), ev.Gate2.cz(), ev.Gate2.swap(), ev.Gate2(random_unitary(rng, 4))][int(rng.integers(0, 4))]
            c.gate2(g, q, b)
        elif r < 8 and n >= 2:
            others = [x for x in range(n) if x != q]
            k = int(rng.integers(1, min(3, len(others)) + 1))
            cs = [int(x) for x in rng.choice(others, size=k, replace=False)]
            c.controlled(cs, ev.Gate1(*random_unitary(rng, 2).reshape(4)), q)
        elif measure and ncl:
            if r == 8:
                c.measure(q, int(rng.integers(0, ncl)))
            else:
                c.x_if(int(rng.integers(0, ncl)), q)
        else:
            c.h(q)
    return c


@pytest.mark.parametrize("n,seed", [(1, 1), (2, 2), (5, 3), (9, 4)])
def test_sv_run_plain_exact_without_measurement(n, seed):
    c = random_circuit(n, 60, seed, measure=False)
    ex = ev.StateVectorBackend.run_seeded(c, seed, plain=True)
    want, _ = oracle.sv_run(c.ops(), n, seed)
    assert float(np.max(np.abs(ex.state().amplitudes() - want))) == 0.0


@pytest.mark.parametrize("n,seed", [(1, 11), (3, 12), (6, 13), (10, 14)])
def test_sv_run_plain_with_measurement(n, seed):
    c = random_circuit(n, 80, seed, ncl=3)
    for s in (0, 7, 99):
        ex = ev.StateVectorBackend.run_seeded(c, s, plain=True)
        want, bits = oracle.sv_run(c.ops(), n, s, num_classical=3)
        assert ex.classical() == bits
        assert np.max(np.abs(ex.state().amplitudes() - want)) < 1e-12


def test_config0_plain_exact():
    c = ev.workloads.config0()
    ex = ev.StateVectorBackend.run(c, plain=True)
    want, _ = oracle.sv_run(c.ops(), 20, 0)
    assert float(np.max(np.abs(ex.state().amplitudes() - want))) == 0.0


@pytest.mark.parametrize("n", [1, 2, 3, 8, 12])
def test_qft_closed_form_on_device(n):  # tests/algorithms_qft.rs:52-74 generalised
    N = 1 << n
    for j in sorted({1, N - 1, N // 3}):
        c = ev.Circuit(n)
        for q in range(n):
            if (j >> q) & 1:
                c.x(q)
        c.compose(ev.algorithms.qft(n))
        got = ev.StateVectorBackend.run(c, plain=True).state().amplitudes()
        want = np.exp(2j * np.pi * j * np.arange(N) / N) / math.sqrt(N)
        assert np.max(np.abs(got - want)) < 1e-12


def test_qft_iqft_identity():  # tests/algorithms_qft.rs:27-50
    n = 4
    for basis in (0, 1, 3, 7, 8, 15):
        c = ev.Circuit(n)
        for q in range(n):
            if basis >> q & 1:
                c.x(q)
        c.compose(ev.algorithms.qft(n)).compose(ev.algorithms.iqft(n))
        amps = ev.StateVectorBackend.run(c, plain=True).state().amplitudes()
        assert abs(abs(amps[basis]) ** 2 - 1) < 1e-12


def test_run_rejects_invalid_circuits():
    c = ev.Circuit(2)
    c.h(5)
    with pytest.raises(ev.QubitOutOfRange):
        ev.StateVectorBackend.run(c, plain=True)
    c, _ = ev.workloads.config2(3, 1)
    with pytest.raises(ev.QubitOutOfRange):
        bad = ev.Circuit(2)
        bad.compose(c)
        ev.DensityMatrixBackend.run(bad, plain=True)


def test_c64_run_fidelity():
    c = ev.workloads.config0(14, 12)
    a = ev.StateVectorBackend.run(c, plain=True).state()
    b = ev.StateVectorBackend.run(c, precision="c64", plain=True).state()
    amps_a, amps_b = a.amplitudes(), b.amplitudes()
    f = abs(np.vdot(amps_a, amps_b)) ** 2
    assert 1 - f < 1e-6 and np.max(np.abs(amps_a - amps_b)) < 1e-5


# ------------------------------------------------------------------ density

NOISES = [
    ("ideal", ev.NoiseModel.ideal(), {}),
    ("depol", ev.NoiseModel.uniform_depolarizing(0.05), {"kind_1q": "depolarizing", "p_1q": 0.05,
                                                          "kind_2q": "depolarizing", "p_2q": 0.05}),
    ("ad", ev.NoiseModel.amplitude_damping(0.2), {"kind_1q": "amplitude_damping", "p_1q": 0.2,
                                                  "kind_2q": "amplitude_damping", "p_2q": 0.2}),
    ("pd", ev.NoiseModel.phase_damping(0.3), {"kind_1q": "phase_damping", "p_1q": 0.3}),
    ("bf", ev.NoiseModel.bit_flip(0.1), {"kind_1q": "bit_flip", "p_1q": 0.1}),
    ("deph", ev.NoiseModel.dephasing(0.25), {"kind_1q": "dephasing", "p_1q": 0.25}),
]


@pytest.mark.parametrize("name,noise,ref_noise", NOISES)
def test_dm_run_plain_matches_oracle(name, noise, ref_noise):
    for n, seed in ((1, 1), (3, 2), (4, 3)):
        c = random_circuit(n, 25, seed, ncl=2)
        ref = dict(ref_noise)
        noise.readout_error = 0.05
        ref["readout_error"] = 0.05
        for s in (0, 5):
            ex = ev.DensityMatrixBackend.run_with_noise_seeded(c, noise, s, plain=True)
            want, bits = oracle.dm_run(c.ops(), n, ref, s, num_classical=2)
            assert ex.classical() == bits, (name, n, s)
            got = ex.density_matrix().data()
            assert np.max(np.abs(got - want)) < 1e-10, (name, n, s)


def test_dm_observables_match_oracle():
    c = random_circuit(4, 30, 21, measure=False)
    noise = ev.NoiseModel.uniform_depolarizing(0.02)
    ref = {"kind_1q": "depolarizing", "p_1q": 0.02, "kind_2q": "depolarizing", "p_2q": 0.02}
    rho = ev.DensityMatrixBackend.run_with_noise(c, noise, plain=True).density_matrix()
    want, _ = oracle.dm_run(c.ops(), 4, ref)
    W = want.reshape(16, 16)
    assert abs(rho.trace() - oracle.dm_trace(want, 4)) < 1e-12
    assert abs(rho.purity() - oracle.dm_purity(want, 4)) < 1e-12
    for b in (0, 7, 15):
        assert abs(rho.probability(b) - W[b, b].real) < 1e-12
    for q in range(4):
        red = np.zeros((2, 2), dtype=np.complex128)
        for i in range(16):
            for j in range(16):
                if (i & ~(1 << q)) == (j & ~(1 << q)):
                    red[(i >> q) & 1, (j >> q) & 1] += W[i, j]
        assert np.max(np.abs(np.array(rho.partial_trace_1q(q)).reshape(2, 2) - red)) < 1e-12
        paulis = {"I": np.eye(2), "X": np.array([[0, 1], [1, 0]]), "Y": np.array([[0, -1j], [1j, 0]]),
                  "Z": np.diag([1, -1])}
        for p, P in paulis.items():
            assert abs(rho.expectation_pauli(q, p) - np.trace(P @ red).real) < 1e-12


def test_config2_small_matches_oracle():
    c, noise = ev.workloads.config2(5, 6)
    ks = np.array([k.m for k in noise.after_1q.ops])
    ref = {"custom_1q": ks, "custom_2q": ks, "readout_error": 0.01}
    for s in (0, 3):
        ex = ev.DensityMatrixBackend.run_with_noise_seeded(c, noise, s, plain=True)
        want, bits = oracle.dm_run(c.ops(), 5, ref, s, num_classical=5)
        assert ex.classical() == bits
        assert np.max(np.abs(ex.density_matrix().data() - want)) < 1e-10
