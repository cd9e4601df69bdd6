"""Fused passes (planner + pass kernel) on the device against the oracle.

The fused path multiplies consecutive one-qubit gates in f64 and evaluates
gates in a different association order than the reference, so complex128
results are held to 1e-10 max-abs (the f64 parity bound) and complex64 to
1e-5 max-abs with 1 - fidelity <= 1e-6.  At the full 30-qubit size the checks
are size-independent: a circuit followed by its inverse returns |0..0>, the
norm is preserved, and QFT of a basis state matches its closed form.
"""
import math

import numpy as np
import pytest

import everett_amd as ev
import oracle
from tests.helpers import adjoint_ops, random_unitary
from tests.test_planner import mixed_circuit

pytestmark = pytest.mark.gpu

TOL128 = 1e-10


@pytest.mark.parametrize("n,depth,tile", [(8, 6, 0), (12, 10, 0), (13, 8, 9), (16, 12, 0), (15, 10, 13), (17, 12, 13)])
def test_fused_brickwall_qft_matches_oracle(n, depth, tile):
    c = ev.workloads.config1(n, depth)
    want, _ = oracle.sv_run(c.ops(), n)
    got = ev.StateVectorBackend.run(c, tile_qubits=tile).state().amplitudes()
    assert np.max(np.abs(got - want)) < TOL128


@pytest.mark.parametrize("n,seed,budget", [(9, 1, 0), (10, 2, 3), (12, 3, 0), (14, 4, 50)])
def test_fused_mixed_circuits_match_oracle(n, seed, budget):
    c = mixed_circuit(n, 300, seed)
    want, _ = oracle.sv_run(c.ops(), n)
    got = ev.StateVectorBackend.run(c, max_gates_per_pass=budget).state().amplitudes()
    assert np.max(np.abs(got - want)) < TOL128


def test_fused_config0_matches_oracle():
    c = ev.workloads.config0()
    want, _ = oracle.sv_run(c.ops(), 20)
    got = ev.StateVectorBackend.run(c).state().amplitudes()
    assert np.max(np.abs(got - want)) < TOL128


def test_fused_with_measurement_segments():
    n = 10
    c = ev.Circuit.with_classical(n, 3)
    c.compose(ev.workloads.config0(n, 3))
    c.measure(3, 0).measure(7, 1)
    c.x_if(0, 2).z_if(1, 5)
    c.compose(ev.workloads.brickwall(n, 2, ev.Rng.seed_from_u64(5)))
    c.measure(0, 2)
    for seed in (0, 1, 2, 3):
        ex = ev.StateVectorBackend.run_seeded(c, seed)
        want, bits = oracle.sv_run(c.ops(), n, seed, num_classical=3)
        assert ex.classical() == bits
        assert np.max(np.abs(ex.state().amplitudes() - want)) < TOL128


@pytest.mark.parametrize("n,tile", [(12, 0), (18, 0), (18, 13)])
def test_fused_c64_fidelity(n, tile):
    c = ev.workloads.config1(n, 20)
    want, _ = oracle.sv_run(c.ops(), n)
    s = ev.StateVectorBackend.run(c, precision="c64", tile_qubits=tile).state()
    assert abs(s.norm_sqr() - 1.0) < 1e-6  # renormalised in f64 at the end of the run
    got = s.amplitudes()
    fid = abs(np.vdot(want, got)) ** 2 / (np.vdot(got, got).real * np.vdot(want, want).real)
    assert np.max(np.abs(got - want)) < 1e-5 and 1 - fid < 1e-6


@pytest.mark.parametrize("n,tile,seed", [(12, 0, 0), (18, 0, 0), (18, 13, 0), (15, 0, 3)])
def test_fused_c64_matrix_core_gates(n, tile, seed):
    # dense 4x4 gates on v_mfma_f32_4x4x1_16b_f32 (the mfma_u2 run option): same bounds as the VALU path,
    # on brick-walls (dense 4x4 blocks) and on mixed circuits (every op kind beside them)
    c = ev.workloads.config1(n, 20) if seed == 0 else mixed_circuit(n, 300, seed)
    want, _ = oracle.sv_run(c.ops(), n)
    s = ev.StateVectorBackend.run(c, precision="c64", tile_qubits=tile, mfma_u2=True).state()
    got = s.amplitudes()
    fid = abs(np.vdot(want, got)) ** 2 / (np.vdot(got, got).real * np.vdot(want, want).real)
    assert np.max(np.abs(got - want)) < 1e-5 and 1 - fid < 1e-6
    valu = ev.StateVectorBackend.run(c, precision="c64", tile_qubits=tile).state().amplitudes()
    assert np.max(np.abs(got - valu)) < 1e-5


def test_pass_descriptor_ring_wraps_and_grows():
    # passes stream to the device through a per-state ring (1 MiB below 20 qubits): ~1,000 passes of
    # ~2.5 KB each wrap it twice in one run; one pass of ~1.4 MB outgrows it after a smaller run used it
    s = ev.State.zero(14)
    c = mixed_circuit(14, 3000, 11)
    ev.run_into(s, c, 0, tile_qubits=10, max_gates_per_pass=2)
    want, _ = oracle.sv_run(c.ops(), 14)
    assert np.max(np.abs(s.amplitudes() - want)) < TOL128
    s = ev.State.zero(12)
    small, big = mixed_circuit(12, 200, 12), mixed_circuit(12, 12000, 13)
    for circ in (small, big, small):
        ev.run_into(s, circ, 0)
        want, _ = oracle.sv_run(circ.ops(), 12)
        assert np.max(np.abs(s.amplitudes() - want)) < 1e-9


def test_matrix_core_gates_are_a_complex64_option():
    with pytest.raises(ev.EverettError):
        ev.StateVectorBackend.run(ev.workloads.config1(12, 4), precision="c128", mfma_u2=True)


@pytest.mark.parametrize("n", [24, 28])
def test_fused_matches_plain_at_scale(n):
    # the plain path equals the oracle exactly (test_gpu_backends), so it is the reference here
    c = ev.workloads.config1(n, 40)
    a = ev.StateVectorBackend.run(c).state()
    b = ev.StateVectorBackend.run(c, plain=True).state()
    off = (1 << n) // 3
    for lo in (0, off, (1 << n) - 65536):
        x = a.amplitudes_range(lo, 65536)
        y = b.amplitudes_range(lo, 65536)
        assert np.max(np.abs(x - y)) < TOL128
    assert abs(a.fidelity(b) - 1.0) < 1e-10


def test_config1_round_trip_at_30_qubits():
    c = ev.workloads.config1(30, 40)
    inv = ev.Circuit(30)
    inv.append_ops(adjoint_ops(c.ops()))
    c.compose(inv)
    s = ev.StateVectorBackend.run(c).state()
    head = s.amplitudes_range(0, 4096)
    assert abs(head[0] - 1.0) < 1e-9
    assert np.max(np.abs(head[1:])) < 1e-9
    assert abs(s.norm_sqr() - 1.0) < 1e-10


def test_qft_basis_state_closed_form_at_30_qubits():
    n, j = 30, 0x2B5A3C1
    c = ev.Circuit(n)
    for q in range(n):
        if (j >> q) & 1:
            c.x(q)
    c.compose(ev.algorithms.qft(n))
    s = ev.StateVectorBackend.run(c).state()
    N = 1 << n
    for lo in (0, 123456789, N - 4096):
        k = np.arange(lo, lo + 4096, dtype=np.float64)
        want = np.exp(2j * np.pi * ((j * k) % N) / N) / math.sqrt(N)
        got = s.amplitudes_range(lo, 4096)
        assert np.max(np.abs(got - want)) < 1e-12


@pytest.mark.parametrize("name,noise,ref", [
    ("ideal", ev.NoiseModel.ideal(), {}),
    ("depol", ev.NoiseModel.uniform_depolarizing(0.05),
     {"kind_1q": "depolarizing", "p_1q": 0.05, "kind_2q": "depolarizing", "p_2q": 0.05}),
    ("ad", ev.NoiseModel.amplitude_damping(0.1),
     {"kind_1q": "amplitude_damping", "p_1q": 0.1, "kind_2q": "amplitude_damping", "p_2q": 0.1}),
])
def test_fused_dm_matches_oracle(name, noise, ref):
    from tests.test_gpu_backends import random_circuit

    for n, seed in ((4, 1), (5, 2)):
        c = random_circuit(n, 40, seed, ncl=2)
        for s in (0, 4):
            ex = ev.DensityMatrixBackend.run_with_noise_seeded(c, noise, s)
            want, bits = oracle.dm_run(c.ops(), n, ref, s, num_classical=2)
            assert ex.classical() == bits, (name, n, s)
            assert np.max(np.abs(ex.density_matrix().data() - want)) < 1e-10


def test_fused_config2_reduced_matches_oracle():
    c, noise = ev.workloads.config2(6, 8)
    ks = np.array([k.m for k in noise.after_1q.ops])
    ref = {"custom_1q": ks, "custom_2q": ks, "readout_error": 0.01}
    ex = ev.DensityMatrixBackend.run_with_noise_seeded(c, noise, 7)
    want, bits = oracle.dm_run(c.ops(), 6, ref, 7, num_classical=6)
    assert ex.classical() == bits
    assert np.max(np.abs(ex.density_matrix().data() - want)) < 1e-10


def test_fused_dm_measurement_runs_match_oracle():
    # consecutive measurements run as one pass over rho (a device copy of the
    # diagonal gives each one its probabilities); runs here repeat a qubit,
    # follow gates, and feed a conditional, with readout error on
    n = 5
    c = ev.Circuit.with_classical(n, 9)
    c.compose(ev.workloads.brickwall(n, 3, ev.Rng.seed_from_u64(11)))
    c.measure(0, 0).measure(3, 1).measure(0, 2).measure(4, 3)
    c.x_if(1, 2)
    c.compose(ev.workloads.brickwall(n, 2, ev.Rng.seed_from_u64(12)))
    for k, q in enumerate((1, 2, 3, 4, 2)):
        c.measure(q, 4 + k)
    noise = ev.NoiseModel.uniform_depolarizing(0.05)
    noise.readout_error = 0.2
    ref = {"kind_1q": "depolarizing", "p_1q": 0.05, "kind_2q": "depolarizing", "p_2q": 0.05, "readout_error": 0.2}
    for s in range(6):
        ex = ev.DensityMatrixBackend.run_with_noise_seeded(c, noise, s)
        want, bits = oracle.dm_run(c.ops(), n, ref, s, num_classical=9)
        assert ex.classical() == bits, s
        assert np.max(np.abs(ex.density_matrix().data() - want)) < 1e-10, s
        plain = ev.DensityMatrixBackend.run_with_noise_seeded(c, noise, s, plain=True)
        assert plain.classical() == bits, s


@pytest.mark.parametrize("precision,tile", [("c128", 12), ("c64", 12), ("c128", 13), ("c64", 13)])
def test_fused_relabelled_exchanges_are_deterministic(precision, tile):
    # tile 12 = four waves per workgroup, tile 13 = eight.  CNOTs between tile
    # qubits, and CNOTs controlled from outside the tile, become relabellings
    # of the LDS exchange whose reads take other threads' slots: a missing
    # barrier would show as run-to-run differences and oracle mismatches here.
    n = 18
    c = ev.workloads.config1(n, 30)
    runs = [ev.StateVectorBackend.run(c, precision=precision, tile_qubits=tile).state().amplitudes() for _ in range(4)]
    for r in runs[1:]:
        assert np.array_equal(r, runs[0])
    want, _ = oracle.sv_run(c.ops(), n)
    if precision == "c128":
        assert np.max(np.abs(runs[0] - want)) < TOL128
    else:
        assert np.max(np.abs(runs[0] - want)) < 1e-5


def test_run_stats_count_algorithmic_pass_flops():
    # one dense 4x4 on a 10-qubit state is one pass of 30 flops per amplitude
    # (a 4-term complex dot product: 4 complex products of 6, 3 sums of 2)
    c = ev.Circuit(10)
    c.gate2(ev.Gate2(random_unitary(np.random.default_rng(3), 4)), 3, 7)
    ev.StateVectorBackend.run(c)
    rs = ev.last_run_stats()
    assert rs.passes == 1 and rs.pass_flops == 30 * 2**10
    # a CNOT moves amplitudes and does no arithmetic
    c = ev.Circuit(10)
    c.cnot(2, 6)
    ev.StateVectorBackend.run(c)
    assert ev.last_run_stats().pass_flops == 0


def test_trailing_swap_network_runs_as_one_permutation_pass():
    # a segment ending in disjoint SWAPs spread wider than a tile: those above
    # the coalesced low qubits run as one in-place bit-permutation pass, the
    # ones on low qubits stay in the plan; a measurement ends the first segment
    n = 18
    c = ev.Circuit.with_classical(n, 1)
    c.compose(ev.workloads.brickwall(n, 3, ev.Rng.seed_from_u64(21)))
    for i in range(n // 2):
        c.swap(i, n - 1 - i)
    c.measure(5, 0)
    c.compose(ev.workloads.brickwall(n, 2, ev.Rng.seed_from_u64(22)))
    for i in range(2, 8):
        c.swap(i, i + 9)
    for seed in (0, 3):
        ex = ev.StateVectorBackend.run_seeded(c, seed, tile_qubits=12)
        want, bits = oracle.sv_run(c.ops(), n, seed, num_classical=1)
        assert ex.classical() == bits
        assert np.max(np.abs(ex.state().amplitudes() - want)) < TOL128
    # measurement-free, profiled: the permutation pass ran
    c2 = ev.workloads.brickwall(n, 3, ev.Rng.seed_from_u64(23))
    for i in range(n // 2):
        c2.swap(i, n - 1 - i)
    st = ev.State.zero(n, "c128")
    ev.run_into(st, c2, 0, tile_qubits=12, profile=True)
    assert "swap_bits" in {k["name"] for k in ev.last_kernel_stats()}
    want, _ = oracle.sv_run(c2.ops(), n)
    assert np.max(np.abs(st.amplitudes() - want)) < TOL128
