"""The smallest inputs through the C ABI: a zero-qubit state (src/state.rs:39-43:
2^0 = 1 amplitude, |> = 1), one-qubit circuits on the fused path (below one
tile it runs the per-gate kernels), a one-qubit noisy density matrix, and the
stated lower limits of the batched and served paths."""
import numpy as np
import pytest

import everett_amd as ev
import oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("plain", [True, False])
def test_zero_qubit_circuit(plain):
    c = ev.Circuit(0)
    ex = ev.StateVectorBackend.run(c, plain=plain)
    amps = ex.state().amplitudes()
    assert amps.shape == (1,) and amps[0] == 1.0
    assert ex.state().sample(8, 3).tolist() == [0] * 8


@pytest.mark.parametrize("plain", [True, False])
@pytest.mark.parametrize("precision", ["c128", "c64"])
def test_one_qubit_circuit(plain, precision):
    c = ev.Circuit(1)
    c.h(0).rz(0, 0.3).ry(0, 1.1).x(0)
    want, _ = oracle.sv_run(c.ops(), 1, 0)
    got = ev.StateVectorBackend.run(c, plain=plain, precision=precision).state().amplitudes()
    tol = 0.0 if (plain and precision == "c128") else (1e-10 if precision == "c128" else 1e-6)
    assert float(np.max(np.abs(got - want))) <= tol


def test_one_qubit_noisy_density_matrix():
    c = ev.Circuit(1, 1)
    c.h(0).rx(0, 0.7).measure(0, 0)
    noise = ev.NoiseModel.uniform_depolarizing(0.1)
    ref = {"kind_1q": "depolarizing", "p_1q": 0.1, "kind_2q": "depolarizing", "p_2q": 0.1}
    for seed in (0, 5):
        want, bits = oracle.dm_run(c.ops(), 1, ref, seed, 1)
        ex = ev.DensityMatrixBackend.run_with_noise_seeded(c, noise, seed)
        assert ex.classical() == bits
        assert float(np.max(np.abs(ex.density_matrix().data() - want))) <= 1e-12


def test_batched_and_served_lower_limits():
    small = ev.workloads.config0(8, 4)
    x = ev.Batch.run([small, ev.workloads.config0(9, 3)], precision="c64")
    want, _ = oracle.sv_run(small.ops(), 8, 0)
    got = x.amplitudes(0)
    assert float(np.max(np.abs(got - want))) <= 1e-5
    with pytest.raises(ev.EverettError):
        ev.Batch.run([ev.workloads.config0(7, 4)], precision="c64")
    with ev.Server(precision="c64") as srv:
        samples, _ = srv.wait(srv.submit(small, shots=16, seed=9))
        assert len(samples) == 16
        with pytest.raises(ev.EverettError):
            srv.submit(ev.workloads.config0(7, 4), shots=16, seed=9)
