"""Continuous-batching server: every result in the mix equals the same request
run alone, and the samples follow the oracle's sampler."""
import threading
import time

import numpy as np
import pytest

import everett_amd as ev
import oracle

pytestmark = pytest.mark.gpu


def requests(count, lo=8, hi=15, seed=900):
    out = []
    for r in range(count):
        rng = ev.Rng.seed_from_u64(seed + r)
        n = lo + rng.next_u64() % (hi - lo + 1)
        depth = 2 + rng.next_u64() % 30
        out.append((ev.workloads.brickwall(n, depth, rng), seed + r))
    return out


@pytest.mark.parametrize("precision", ["c128", "c64"])
def test_results_in_the_mix_equal_alone(precision):
    reqs = requests(40)
    with ev.Server(precision=precision) as srv:
        ids = []
        for i, (c, s) in enumerate(reqs):  # staggered: arrivals while others are on the GPU
            ids.append(srv.submit(c, shots=256, seed=s))
            if i % 7 == 0:
                time.sleep(0.002)
        results = [srv.wait(i) for i in ids]
        st = srv.stats()
    assert st["completed"] == len(reqs) and st["failed"] == 0
    assert st["steps"] < sum(r[1]["passes"] for r in results)  # passes of different requests shared launches
    for (c, seed), (samples, info) in zip(reqs, results):
        assert info["status"] == 0 and info["done_ms"] >= info["admit_ms"] >= info["submit_ms"]
        alone = ev.StateVectorBackend.run(c, precision=precision).state()
        assert np.array_equal(samples, alone.sample(256, seed))


def test_samples_match_oracle_sampler_c128():
    reqs = requests(6, 8, 12, seed=77)
    with ev.Server(precision="c128") as srv:
        ids = [srv.submit(c, shots=512, seed=s) for c, s in reqs]
        res = [srv.wait(i)[0] for i in ids]
    for (c, seed), got in zip(reqs, res):
        n = c.num_qubits()
        psi, _ = oracle.sv_run(c.ops(), n)
        r = oracle.Rng(seed)
        u = np.array([r.next_f64() for _ in range(512)])
        want = oracle.sample(psi, n, u)
        # the device state differs from the oracle by <= 1e-10, so a shot may only differ where u
        # lands within 1e-9 of a boundary of the oracle's CDF
        cdf = np.cumsum(np.abs(psi) ** 2)
        for k in np.nonzero(got != want)[0]:
            assert np.min(np.abs(cdf - u[k] * cdf[-1])) < 1e-9, (seed, k)


def test_admission_by_memory_and_concurrent_submitters():
    # budget for two 14-qubit c64 states: admission must queue the rest
    budget = 2 * (1 << 14) * 8
    reqs = requests(12, 13, 14, seed=333)
    with ev.Server(precision="c64", memory_bytes=budget) as srv:
        ids = [None] * len(reqs)

        def worker(k):
            for i in range(k, len(reqs), 3):
                ids[i] = srv.submit(reqs[i][0], shots=64, seed=reqs[i][1])

        ths = [threading.Thread(target=worker, args=(k,)) for k in range(3)]
        for t in ths:
            t.start()
        for t in ths:
            t.join()
        res = [srv.wait(i) for i in ids]
        st = srv.stats()
    assert st["bytes_peak"] <= budget and st["completed"] == len(reqs)
    big = ev.workloads.brickwall(16, 2, ev.Rng.seed_from_u64(1))
    with ev.Server(precision="c64", memory_bytes=budget) as srv:
        with pytest.raises(ev.AdmissionError):
            srv.submit(big, shots=1)


def test_empty_circuit_request_samples_zero_state():
    with ev.Server(precision="c128") as srv:
        rid = srv.submit(ev.Circuit(9), shots=8, seed=1)
        samples, info = srv.wait(rid)
    assert info["status"] == 0 and list(samples) == [0] * 8


def test_requests_in_openqasm_wire_format():
    # requests arrive as OpenQASM 3 text and are parsed on submit: a text request
    # samples exactly like its parsed circuit submitted directly, a text that does
    # not parse is refused before anything is queued, and the text reproduces the
    # original circuit's state to 1e-10 fidelity (angles travel as 17-decimal text)
    reqs = requests(12, 8, 13, seed=4242)
    texts = [c.to_qasm() for c, _ in reqs]
    with ev.Server(precision="c64") as srv:
        ids = [srv.submit_qasm(t, shots=256, seed=s) for t, (c, s) in zip(texts, reqs)]
        ids2 = [srv.submit(ev.Circuit.from_qasm(t), shots=256, seed=s) for t, (c, s) in zip(texts, reqs)]
        a = [srv.wait(i)[0] for i in ids]
        b = [srv.wait(i)[0] for i in ids2]
        with pytest.raises(ev.QasmError) as e:
            srv.submit_qasm("qubit[2] q;\nfoo q[0];\n")
        assert (e.value.line, e.value.col) == (3, 1)
        st = srv.stats()
    assert st["submitted"] == 2 * len(reqs) and st["completed"] == 2 * len(reqs) and st["failed"] == 0
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    for (c, _), t in zip(reqs, texts):
        s1 = ev.StateVectorBackend.run(c).state()
        s2 = ev.StateVectorBackend.run(ev.Circuit.from_qasm(t)).state()
        assert s1.fidelity(s2) > 1 - 1e-10
