"""The served stream (12-28 qubits, depth 20-200, 1024 shots, complex64) and
the 13-qubit noisy density matrix at their configured sizes.

Serving: every request's samples in the mix equal the same request run alone
(bit for bit: both run the same fused passes on the same complex64
arithmetic), and for the requests of at most 20 qubits the complex64 state
is held against the FP64 oracle (1 - F <= 1e-6, max-abs <= 1e-5).

Density matrix: at 13 qubits the oracle's literal loops take hours, so the
fused run is checked against the plain (one superoperator pass per gate and
channel) run to 1e-10, and the result against what a density matrix must
satisfy: unit trace, Hermiticity, a non-negative diagonal and purity in (0, 1].
Both paths are pinned to the oracle at small sizes in test_gpu_backends.py and
test_gpu_fused.py.  The oracle's superoperator restatement (the same gates and
channels as passes over the 2n-qubit vector, on all host cores; pinned to the
literal loops at n <= 8 in test_oracle.py) reaches 13 qubits in well under a
minute: the fused noisy state is held to it at 1e-10, and the configured run's
classical record to the reference's measurement sequence on its diagonal.
"""
import os

import numpy as np
import pytest

import everett_amd as ev
import oracle

pytestmark = pytest.mark.gpu

# requests 0..19 of the served stream plus request 26 (28 qubits, depth 190):
# 12..28 qubits, depth 20..200
C4_IDS = list(range(20)) + [26]


def test_config4_requests_in_the_mix_equal_alone_and_oracle():
    reqs = [ev.workloads.config4_request(r) for r in C4_IDS]
    sizes = [c.num_qubits() for c, _ in reqs]
    assert max(sizes) == 28 and min(sizes) == 12 and sum(n >= 26 for n in sizes) >= 2
    depth = [(c.num_ops() - 0) / (3 * c.num_qubits() + (c.num_qubits() - 1) / 2) for c, _ in reqs]
    assert max(depth) >= 150
    with ev.Server(precision="c64") as srv:
        ids = [srv.submit(c, shots=1024, seed=s) for c, s in reqs]
        res = [srv.wait(i) for i in ids]
        st = srv.stats()
    assert st["completed"] == len(reqs) and st["failed"] == 0
    oracle.set_threads(16)
    try:
        for (c, seed), (samples, info) in zip(reqs, res):
            n = c.num_qubits()
            assert info["status"] == 0 and len(samples) == 1024
            alone = ev.StateVectorBackend.run(c, precision="c64").state()
            assert np.array_equal(samples, alone.sample(1024, seed)), (n, seed)
            if n <= 20:
                want, _ = oracle.sv_run(c.ops(), n)
                got = alone.amplitudes()
                fid = abs(np.vdot(want, got)) ** 2 / (np.vdot(got, got).real * np.vdot(want, want).real)
                assert 1 - fid <= 1e-6 and np.max(np.abs(got - want)) <= 1e-5, (n, seed, 1 - fid)
            del alone
    finally:
        oracle.set_threads(1)


def _is_density_matrix(a, mixed):
    n = a.shape[0]
    assert abs(np.trace(a).real - 1.0) <= 1e-10 and abs(np.trace(a).imag) <= 1e-12
    assert np.max(np.abs(a - a.conj().T)) <= 1e-12
    d = np.diag(a).real
    assert d.min() >= -1e-15
    # positivity, necessary condition on every 2x2 principal minor: |rho_ij|^2 <= rho_ii rho_jj
    for r0 in range(0, n, 1024):
        blk = np.abs(a[r0:r0 + 1024]) ** 2
        assert np.all(blk <= np.outer(d[r0:r0 + 1024], d) + 1e-13)
    pur = float(np.vdot(a, a).real)
    assert 0.0 < pur <= 1.0 + 1e-12
    if mixed:
        assert pur < 1.0 - 1e-6
    return pur


def test_config2_at_13_qubits_fused_equals_plain_and_is_a_density_matrix():
    # the noisy evolution of the configured circuit (dm_random depth 30, seed 13), before its final
    # measurements erase the mixedness: fused against plain, and the physical constraints
    c, noise = ev.workloads.config2(13, 30)
    body = ev.workloads.dm_random(13, 30, ev.Rng.seed_from_u64(13))
    fused = ev.DensityMatrixBackend.run_with_noise_seeded(body, noise, 13).density_matrix()
    plain = ev.DensityMatrixBackend.run_with_noise_seeded(body, noise, 13, plain=True).density_matrix()
    a, b = fused.matrix(), plain.matrix()
    assert np.max(np.abs(a - b)) <= 1e-10
    del b
    pur = _is_density_matrix(a, mixed=True)
    assert abs(fused.purity() - pur) <= 1e-12 and abs(fused.trace() - 1.0) <= 1e-10
    del a
    # the configured run with its 13 measurements and 1% readout error: the same classical record on
    # both paths, and rho collapsed to one basis projector (readout error flips only the record)
    ef = ev.DensityMatrixBackend.run_with_noise_seeded(c, noise, 13)
    ep = ev.DensityMatrixBackend.run_with_noise_seeded(c, noise, 13, plain=True)
    assert ef.classical() == ep.classical()
    a = ef.density_matrix().matrix()
    assert np.max(np.abs(a - ep.density_matrix().matrix())) <= 1e-10
    x = int(np.argmax(np.diag(a).real))
    assert abs(a[x, x] - 1.0) <= 1e-10
    a[x, x] = 0
    assert np.max(np.abs(a)) <= 1e-12


def _measure_all(diag, n, seed, readout):
    """density_matrix.rs:155-169 on the diagonal of rho, qubits 0..n-1 in order: p1 a sequential sum,
    raw = u < p1, collapse, 1/trace, then the readout draw (only when readout > 0)."""
    d = np.array(diag, dtype=np.float64)
    r = oracle.Rng(seed)
    idx = np.arange(len(d))
    rec = []
    for q in range(n):
        hit = (idx >> q) & 1 == 1
        p1 = float(np.cumsum(d[hit])[-1])
        raw = r.next_f64() < p1
        d[hit != raw] = 0.0
        d = d * (1.0 / float(np.cumsum(d)[-1]))
        rec.append((not raw) if readout > 0 and r.next_f64() < readout else raw)
    return rec


def test_config2_at_13_qubits_against_the_superoperator_oracle():
    n, seed = 13, 13
    c, noise = ev.workloads.config2(n, 30)
    body = ev.workloads.dm_random(n, 30, ev.Rng.seed_from_u64(n))
    ks = np.array([k.m for k in noise.after_1q.ops])
    ref = {"custom_1q": ks, "custom_2q": ks, "readout_error": 0.01}
    oracle.set_threads(min(16, os.cpu_count() or 1))
    try:
        want, _ = oracle.dm_run(body.ops(), n, ref, seed, superop=True)
    finally:
        oracle.set_threads(1)
    got = ev.DensityMatrixBackend.run_with_noise_seeded(body, noise, seed).density_matrix().matrix().reshape(-1)
    assert np.max(np.abs(got - want)) <= 1e-10
    diag = want.reshape(1 << n, 1 << n).diagonal().real.copy()
    del want, got
    ex = ev.DensityMatrixBackend.run_with_noise_seeded(c, noise, seed)
    assert ex.classical() == _measure_all(diag, n, seed, 0.01)
