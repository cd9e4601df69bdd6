"""States past 2^32 amplitudes (64 GiB each): the 64-bit index paths of the
fused passes, the plain gate kernels, the density-matrix channels and
reductions and the range copies, checked through properties that hold at any
size (a round trip back to |0>, a product state's closed form)."""
import gc
import math

import numpy as np
import pytest

import everett_amd as ev
from tests.helpers import adjoint_ops

pytestmark = pytest.mark.gpu

W = 4096


def test_round_trip_at_32_qubits_c128():
    n = 32
    # a random brickwall (the QFT workload stops at 31 qubits, as the reference's angle table does)
    c = ev.workloads.brickwall(n, 8, ev.Rng.seed_from_u64(32))
    inv = ev.Circuit(n)
    inv.append_ops(adjoint_ops(c.ops()))
    c.compose(inv)
    s = ev.StateVectorBackend.run(c).state()
    N = 1 << n
    head = s.amplitudes_range(0, W)
    assert abs(head[0] - 1.0) < 1e-9
    assert np.max(np.abs(head[1:])) < 1e-9
    for lo in ((N // 2) + 12345, N - W):
        assert np.max(np.abs(s.amplitudes_range(lo, W))) < 1e-9, lo
    assert abs(s.norm_sqr() - 1.0) < 1e-10
    del s
    gc.collect()


def product_state_circuit(n):
    # H on every qubit, a phase phi_q on each, then CNOT(control n-1, target 0):
    # amplitude of k is 2^(-n/2) * exp(i * sum of phi_q over the set bits of
    # k' = k ^ (bit n-1 of k)), since the CNOT maps |k'> to |k>
    phis = [0.1 + 0.37 * q for q in range(n)]
    c = ev.Circuit(n)
    for q in range(n):
        c.h(q)
    for q in range(n):
        c.phase(q, phis[q])
    c.cnot(n - 1, 0)
    return c, np.array(phis)


def closed_form(lo, n, phis):
    k = np.arange(lo, lo + W, dtype=np.uint64)
    k = k ^ ((k >> np.uint64(n - 1)) & np.uint64(1))
    bits = ((k[:, None] >> np.arange(n, dtype=np.uint64)[None, :]) & np.uint64(1)).astype(np.float64)
    return np.exp(1j * (bits @ phis)) / math.sqrt(2.0 ** n)


@pytest.mark.parametrize("plain", [False, True])
def test_product_state_closed_form_at_33_qubits_c64(plain):
    n = 33
    c, phis = product_state_circuit(n)
    s = ev.StateVectorBackend.run(c, precision="c64", plain=plain).state()
    N = 1 << n
    amp = 1.0 / math.sqrt(2.0 ** n)
    for lo in (0, (1 << 32) - W // 2, N // 2 + 987654321, N - W):
        got = s.amplitudes_range(lo, W)
        want = closed_form(lo, n, phis)
        # relative to the amplitude's magnitude: 66 single-precision products
        assert np.max(np.abs(got - want)) < 1e-5 * amp, (plain, lo)
    assert abs(s.norm_sqr() - 1.0) < 1e-4
    del s
    gc.collect()


@pytest.mark.parametrize("plain", [False, True])
def test_depolarized_product_state_at_16_qubit_density_matrix(plain):
    # rho of 16 qubits is a 32-qubit vector (64 GiB).  H then a phase on every
    # qubit, a depolarizing channel (probability p) after each gate: each
    # qubit's Bloch vector is s^2 (cos phi, sin phi, 0) with s = 1 - 4p/3, so
    # the purity is ((1 + s^4) / 2)^n and every basis state has probability 2^-n
    n, p = 16, 0.03
    phis = [0.1 + 0.37 * q for q in range(n)]
    c = ev.Circuit(n)
    for q in range(n):
        c.h(q)
    for q in range(n):
        c.phase(q, phis[q])
    rho = ev.DensityMatrixBackend.run_with_noise_seeded(
        c, ev.NoiseModel.uniform_depolarizing(p), 0, plain=plain).density_matrix()
    s2 = (1 - 4 * p / 3) ** 2
    assert abs(rho.trace() - 1.0) < 1e-10
    assert abs(rho.purity() - ((1 + s2 * s2) / 2) ** n) < 1e-10
    for q in (0, 9, 15):
        assert abs(rho.expectation_pauli(q, "X") - s2 * math.cos(phis[q])) < 1e-10, q
        assert abs(rho.expectation_pauli(q, "Y") - s2 * math.sin(phis[q])) < 1e-10, q
        assert abs(rho.expectation_pauli(q, "Z")) < 1e-10, q
    for basis in (0, 0xBEEF, 0xFFFF):
        assert abs(rho.probability(basis) - 2.0 ** -n) < 1e-12, basis
    del rho
    gc.collect()
