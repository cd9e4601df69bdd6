#!/usr/bin/env python3
"""Benchmark: gates/s and HBM GB/s of a 30-qubit complex128 circuit on one MI355X,
plus the batched, served, density-matrix and 20-qubit workloads.

Headline (c1): a 30-qubit brick-wall circuit of depth 40 (Rx, Ry, Rz on every
qubit per layer, angles TAU * Rng::next_f64() from seed 30, then CNOTs on
alternating pairs) followed by QFT(30): 4,660 logical gates.  One step =
reset the HBM-resident state to |0..0> and run the whole circuit; the state is
allocated once, outside the timed region.  `value` is gates/s over the timed
steps; `roofline` is the dominant kernel's algorithmic bytes per launch over
its average launch time (HIP events on the library's stream around every
launch in the timed steps); `cpu_baseline` is the oracle (a single-threaded C
restatement of the reference's loops) on a bounded prefix of the same circuit.

Extra workloads in the same JSON line (see docs/configs.md):
  c0  20-qubit brick-wall depth 20, complex128
  c2  13-qubit noisy density matrix, depth 30, readout error 1%
  c3  4096 uneven statevectors (10..22 qubits), complex64, pipelined groups, one launch per round
  c4  Poisson-arrival serving, 12..28 qubits, 1024 shots each, p50/p99 latency
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import threading
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "gates/sec + HBM GB/s at 30q f64; served circuits/sec with p50/p99 latency"
HBM_PEAK_GBPS = 8000.0
# MI355X vector peaks: FP64 78.6 TFLOP/s; FP32 157.3 TFLOP/s only with packed FP32 math, which this
# build compiles without, so its FP32 vector rate is the unpacked one, equal to FP64's
VALU_PEAK_TFLOPS = {"c128": 78.6, "c64": 78.6}
VALU_BOUND = {"c128": "valu-fp64", "c64": "valu-fp32 (unpacked)"}
# multiply-add rates measured on this GPU at start-up (ev.compute_rates), per precision
MEASURED_RATES: dict = {}
HOST_THREADS = min(16, os.cpu_count() or 1)  # the host share of one GPU on the GPU box


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


# ------------------------------------------------------------------ c1 / c0

def cpu_baseline(circuit, n: int, seconds: float) -> dict:
    """Times the oracle (1 thread) on a prefix of the same circuit at the same n."""
    import numpy as np

    import oracle

    oracle.set_threads(1)
    ops = circuit.ops()
    amps = np.zeros(1 << n, dtype=np.complex128)
    amps[0] = 1.0
    oracle.sv_apply_ops(ops[0:1], amps)  # untimed: pages the state in
    count, i = 0, 1
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds and i < len(ops):
        oracle.sv_apply_ops(ops[i:i + 1], amps)
        i += 1
        count += 1
    dt = time.perf_counter() - t0
    del amps
    return {
        "value": count / dt,
        "unit": "gates/s",
        "cores": 1,
        "kind": "port",
        "sample": f"gates 2..{count + 1} of the same {n}-qubit circuit on a host complex128 state "
                  f"(gate 1 untimed, pages the state in), {dt:.1f} s, oracle/refcpu.c single thread",
    }


def cpu_baseline_all_cores(circuit, n: int, seconds: float) -> dict:
    """The same oracle loops on all host cores (OpenMP over the reduction-free gate loops, so bit
    for bit the 1-thread result): a labelled restatement timing, not the reference's."""
    import numpy as np

    import oracle

    oracle.set_threads(HOST_THREADS)
    ops = circuit.ops()
    amps = np.zeros(1 << n, dtype=np.complex128)
    amps[0] = 1.0
    oracle.sv_apply_ops(ops[0:1], amps)
    count, i = 0, 1
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds and i < len(ops):
        oracle.sv_apply_ops(ops[i:i + 1], amps)
        i += 1
        count += 1
    dt = time.perf_counter() - t0
    oracle.set_threads(1)
    del amps
    return {"value": count / dt, "unit": "gates/s", "cores": HOST_THREADS, "kind": "port",
            "label": f"restatement, {HOST_THREADS} cores -- not the reference",
            "sample": f"gates 2..{count + 1} of the same {n}-qubit circuit, {dt:.1f} s, oracle/refcpu.c with "
                      f"OpenMP on {HOST_THREADS} threads"}


def c0_cpu_comparison(circuit, gpu_gates_per_s: float, seconds: float) -> dict:
    """c0 beside the CPU: the oracle on 1 thread and on all host cores (gate prefixes of the same
    20-qubit circuit, bounded in time), and the GPU's fused result against the whole circuit run by the
    oracle (all host cores: the gate loops have no reductions, so it is the 1-thread result bit for bit)."""
    import numpy as np

    import everett_amd as ev
    import oracle

    n = circuit.num_qubits()
    one = cpu_baseline(circuit, n, seconds)
    allc = cpu_baseline_all_cores(circuit, n, seconds / 2)
    oracle.set_threads(HOST_THREADS)
    t0 = time.perf_counter()
    want, _ = oracle.sv_run(circuit.ops(), n)
    full_s = time.perf_counter() - t0
    oracle.set_threads(1)
    got = ev.StateVectorBackend.run(circuit).state().amplitudes()
    return {"oracle_1_thread": one, "restatement_all_cores": allc,
            "gpu_over_oracle_1_thread": gpu_gates_per_s / one["value"],
            "max_abs_vs_oracle": float(np.max(np.abs(got - want))),
            "oracle_full_run_s": full_s,
            "parity": f"the fused GPU run of all {len(circuit)} gates against the oracle's full run "
                      f"({HOST_THREADS} host threads)"}


def oracle_fidelity(circuits, idxs, state_of) -> dict:
    """Per-circuit fidelity and max-abs of complex64 states against the FP64 oracle (all host cores)."""
    import numpy as np

    import oracle

    oracle.set_threads(HOST_THREADS)
    infid, mabs, sizes = [], [], []
    t0 = time.perf_counter()
    for i in idxs:
        c = circuits[i]
        n = c.num_qubits()
        want, _ = oracle.sv_run(c.ops(), n)
        got = state_of(i)
        f = abs(np.vdot(want, got)) ** 2 / (np.vdot(got, got).real * np.vdot(want, want).real)
        infid.append(1.0 - f)
        mabs.append(float(np.max(np.abs(got - want))))
        sizes.append(n)
    oracle.set_threads(1)
    return {"reference": f"FP64 oracle (oracle/refcpu.c, OpenMP on {HOST_THREADS} host threads)",
            "circuits": len(idxs), "qubits": [min(sizes), max(sizes)] if sizes else None,
            "max_infidelity": float(max(infid)) if infid else None,
            "median_infidelity": float(np.median(infid)) if infid else None,
            "max_abs": float(max(mabs)) if mabs else None, "oracle_s": time.perf_counter() - t0}


def run_sv_workload(circuit, n, precision, steps, warmup, plain, mfma_u2=False):
    import everett_amd as ev

    t = time.perf_counter()
    state = ev.State.zero(n, precision)
    alloc_ms = (time.perf_counter() - t) * 1e3
    for _ in range(warmup):
        ev.run_into(state, circuit, 0, plain=plain, mfma_u2=mfma_u2)
    state.norm_sqr()  # synchronises the handle's stream
    agg: dict = {}
    dev_ms = 0.0
    passes = 0
    pass_flops = 0.0
    t0 = time.perf_counter()
    for _ in range(steps):
        ev.run_into(state, circuit, 0, plain=plain, profile=True, mfma_u2=mfma_u2)
        rs = ev.last_run_stats()
        dev_ms += rs.device_ms
        passes += rs.passes
        pass_flops += rs.pass_flops
        for ks in ev.last_kernel_stats():
            a = agg.setdefault(ks["name"], {"launches": 0, "total_ms": 0.0, "bytes": 0.0})
            a["launches"] += ks["launches"]
            a["total_ms"] += ks["total_ms"]
            a["bytes"] += ks["bytes"]
    norm = state.norm_sqr()
    wall = time.perf_counter() - t0
    del state
    if abs(norm - 1.0) > 1e-6:
        raise SystemExit(f"norm drifted to {norm}")
    return {"wall": wall, "agg": agg, "dev_ms": dev_ms, "alloc_ms": alloc_ms, "passes": passes / max(steps, 1),
            "pass_flops": pass_flops}


def roofline(agg, steps, workload_key, n, precision):
    dom_name, dom = max(agg.items(), key=lambda kv: kv[1]["total_ms"])
    per_launch = dom["bytes"] / dom["launches"]
    avg_ms = dom["total_ms"] / dom["launches"]
    achieved = per_launch / (avg_ms * 1e-3) / 1e9
    traffic = None
    tf = ROOT / "profiles" / "traffic.json"
    if tf.exists():
        entry = json.loads(tf.read_text()).get(f"{dom_name}:{workload_key}:{n}:{precision}")
        if entry:
            traffic = entry.get("hbm_bytes_per_launch")
    return {
        "bound": "hbm",
        "kernel": dom_name,
        "achieved": achieved,
        "peak": HBM_PEAK_GBPS,
        "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBPS,
        "traffic": traffic,
        "bytes_per_launch": per_launch,
        "avg_launch_ms": avg_ms,
        "launches_per_step": dom["launches"] / steps,
    }


def compute_roofline(agg, pass_flops, precision):
    """The fused pass against the FP64 (complex128) or FP32 vector peak: its
    algorithmic flops per launch (each op's complex arithmetic over the
    amplitudes it touches, summed by the library) over its average launch
    time.  The pass is vector-ALU bound once its ops outweigh its HBM
    traffic, which the HBM roofline alone does not show."""
    fp = agg.get("fused_pass")
    if not fp or not fp["launches"] or not pass_flops:
        return None
    per_launch = pass_flops / fp["launches"]
    avg_ms = fp["total_ms"] / fp["launches"]
    key = "c64" if precision == "c64" else "c128"
    peak = VALU_PEAK_TFLOPS[key]
    achieved = per_launch / (avg_ms * 1e-3) / 1e12
    out = {"bound": VALU_BOUND[key], "kernel": "fused_pass",
           "achieved": achieved, "peak": peak, "unit": "TFLOP/s", "frac": achieved / peak,
           "flops_per_launch": per_launch}
    if key in MEASURED_RATES:  # against what multiply-add chains reach on this GPU, as frac_of_copy is for HBM
        att = MEASURED_RATES[key]["valu_tflops"]
        out.update({"attainable": att, "frac_of_attainable": achieved / att})
    return out


def bench_sv(name, circuit, n, desc, args, steps, warmup):
    r = run_sv_workload(circuit, n, args.precision, steps, warmup, args.plain, getattr(args, "mfma_u2", False))
    gates = len(circuit)
    step_bytes = sum(a["bytes"] for a in r["agg"].values()) / steps
    return {
        "gates": gates,
        "gates_per_s": gates * steps / r["wall"],
        "ms_per_step": r["wall"] / steps * 1e3,
        "device_ms_per_step": r["dev_ms"] / steps,
        "passes_per_step": r["passes"],
        "hbm_gbps": step_bytes / (r["wall"] / steps) / 1e9,
        "state_alloc_ms": r["alloc_ms"],
        "roofline": roofline(r["agg"], steps, name, n, args.precision),
        "compute_roofline": compute_roofline(r["agg"], r["pass_flops"], args.precision),
        "kernels": {k: {"launches_per_step": v["launches"] / steps, "ms_per_step": v["total_ms"] / steps,
                        "gbps": (v["bytes"] / (v["total_ms"] * 1e-3) / 1e9) if v["total_ms"] else None}
                    for k, v in sorted(r["agg"].items(), key=lambda kv: -kv[1]["total_ms"])},
        "workload": desc,
    }


# ----------------------------------------------------------------------- c2

def bench_c2(args) -> dict:
    import everett_amd as ev

    c, noise = ev.workloads.config2(13, 30)
    ev.DensityMatrixBackend.run_with_noise_seeded(c, noise, 13)  # warmup
    # a run takes a few milliseconds, so one timing is noisy: the median of five (each allocates rho,
    # runs the circuit and reduces the trace)
    runs = []
    for _ in range(5):
        t0 = time.perf_counter()
        ex = ev.DensityMatrixBackend.run_with_noise_seeded(c, noise, 13)
        rs = ev.last_run_stats()
        tr = ex.density_matrix().trace()
        runs.append((time.perf_counter() - t0, rs.device_ms, ex.classical(), tr))
        del ex
    runs.sort(key=lambda r: r[0])
    wall, dev_ms, classical, tr = runs[len(runs) // 2]
    ms_runs = [round(r[0] * 1e3, 3) for r in runs]
    ev.DensityMatrixBackend.run_with_noise_seeded(c, noise, 13, profile=True)  # per-launch times for the roofline
    agg = {ks["name"]: ks for ks in ev.last_kernel_stats()}
    gates = rs.logical_gates
    pass_bytes = 2 * (4 ** 13) * 16
    return {
        "workload": "13-qubit noisy density matrix: dm_random depth 30 (seed 13), depolarizing 1e-3 then amplitude "
                    "damping 5e-4 after every gate, readout error 1%, all qubits measured",
        "gates": int(gates),
        "gates_per_s": gates / wall,
        "ms": wall * 1e3,
        "ms_runs": ms_runs,
        "device_ms": dev_ms,
        "superop_passes": int(rs.passes),
        "gbps_per_pass": (pass_bytes * rs.passes) / (dev_ms * 1e-3) / 1e9 if dev_ms else None,
        "trace": tr,
        "classical": "".join("1" if b else "0" for b in classical),
        "roofline": roofline(agg, 1, "c2", 26, "c128") if agg else None,
        "vs_oracle": None if args.no_cpu else c2_oracle_parity(c, noise),
    }


def c2_oracle_parity(c, noise) -> dict:
    """The noisy evolution before the final measurements (which collapse rho to a basis projector)
    against the oracle's superoperator restatement on all host cores: the same gates and channels as
    passes over the 26-bit vector, pinned to the oracle's literal loops at n <= 8 by the tests."""
    import numpy as np

    import everett_amd as ev
    import oracle

    n, seed = 13, 13
    body = ev.workloads.dm_random(n, 30, ev.Rng.seed_from_u64(n))
    ks = np.array([k.m for k in noise.after_1q.ops])
    oracle.set_threads(HOST_THREADS)
    t0 = time.perf_counter()
    try:
        want, _ = oracle.dm_run(body.ops(), n, {"custom_1q": ks, "custom_2q": ks, "readout_error": 0.01}, seed,
                                superop=True)
    finally:
        oracle.set_threads(1)
    oracle_s = time.perf_counter() - t0
    got = ev.DensityMatrixBackend.run_with_noise_seeded(body, noise, seed).density_matrix().matrix().reshape(-1)
    return {"max_abs": float(np.max(np.abs(got - want))), "oracle_s": oracle_s,
            "reference": f"oracle superoperator restatement (oracle/refcpu.c, OpenMP on {HOST_THREADS} host threads)",
            "state": "rho after the 30 noisy layers, before the 13 final measurements"}


# ----------------------------------------------------------------------- c3

def bench_c3(args) -> dict:
    import numpy as np

    import everett_amd as ev

    count = args.batch_count
    cs = ev.workloads.config3(count)
    gates = sum(len(c) for c in cs)
    sizes = sorted({c.num_qubits() for c in cs})
    t = time.perf_counter()
    for n in sizes:
        if ev.tune_get(n, "c64") is None:
            ev.autotune(n, "c64", 8)
    tune_s = time.perf_counter() - t
    ev.Batch.run(cs[:64], precision="c64")  # warmup
    t0 = time.perf_counter()
    x = ev.Batch.run(cs, precision="c64", use_graph=True, renormalize=True)
    wall = time.perf_counter() - t0
    st = x.stats()
    rounds_ms = st["device_ms"]
    achieved = st["pass_bytes"] / (rounds_ms * 1e-3) / 1e9
    # FP64 oracle on a stated subset: per qubit count 10..22 the first circuits (in index order)
    # with that count, 6 each up to 17 qubits, 3 at 18-19 and 2 at 20-22
    want = {n: (6 if n <= 17 else 3 if n <= 19 else 2) for n in range(10, 23)}
    sub = []
    for i, c in enumerate(cs):
        n = c.num_qubits()
        if want.get(n, 0) > 0:
            want[n] -= 1
            sub.append(i)
    fid_oracle = oracle_fidelity(cs, sub, lambda i: x.amplitudes(i))
    fid_oracle["subset"] = "per qubit count 10..22 the first circuits in index order: 6 each up to 17 qubits, 3 at 18-19, 2 at 20-22"
    ref = ev.Batch.run(cs, precision="c128", use_graph=True)
    f, m = x.compare(ref)
    infid = 1.0 - f
    del ref, x
    return {
        "workload": f"{count} uneven brick-wall circuits (n 10..22, depth 10..60, seeds 1000+i), complex64 bulk, "
                    "f64 norms and renormalisation, pipelined groups (planned on host threads while the previous group runs), "
                    "one launch per round per group",
        "circuits": count,
        "gates": gates,
        "circuits_per_s": count / wall,
        "gates_per_s": gates / wall,
        "device_circuits_per_s": count / (st["device_ms"] * 1e-3),
        "wall_ms": wall * 1e3,
        "plan_ms": st["plan_ms"],
        "upload_ms": st["upload_ms"],
        "device_ms": st["device_ms"],
        "rounds": st["rounds"],
        "passes": st["passes"],
        "graph": bool(st["graph"]),
        "autotune_s": tune_s,
        "fidelity_vs_f64": fid_oracle,
        "fidelity_vs_gpu_c128": {"circuits": count, "max_infidelity": float(infid.max()),
                                 "median_infidelity": float(np.median(infid)), "max_abs": float(m.max()),
                                 "reference": "the same circuits in complex128 on the GPU (all of them)"},
        "roofline": {"bound": "hbm", "kernel": "fused_batch", "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBPS, "traffic": None, "bytes": st["pass_bytes"],
                     "avg_launch_ms": rounds_ms / max(1, st["rounds"]), "launches": st["rounds"],
                     "note": "algorithmic bytes of all rounds (every pass reads and writes its state once) over the "
                             "HIP-event time of the rounds"},
        "compute_roofline": batch_compute_roofline(st["pass_flops"], rounds_ms, st["rounds"]),
    }


def batch_compute_roofline(flops, rounds_ms, rounds):
    """The batched rounds against the vector ALU's FP32 rate: algorithmic flops of
    every pass (the planner's count per device op) over the HIP-event time of the
    rounds.  The states are at most 22 qubits (32 MB), so the rounds are bound
    by arithmetic, not by HBM."""
    if not rounds_ms or not flops:
        return None
    achieved = flops / (rounds_ms * 1e-3) / 1e12
    peak = VALU_PEAK_TFLOPS["c64"]
    out = {"bound": VALU_BOUND["c64"], "kernel": "fused_batch", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
           "frac": achieved / peak, "flops": flops, "avg_launch_ms": rounds_ms / max(1, rounds), "launches": rounds}
    if "c64" in MEASURED_RATES:
        att = MEASURED_RATES["c64"]["valu_tflops"]
        out.update({"attainable": att, "frac_of_attainable": achieved / att})
    return out


# ----------------------------------------------------------------------- c4

def bench_c4(args) -> dict:
    import numpy as np

    import everett_amd as ev

    R = args.serve_requests
    reqs = [ev.workloads.config4_request(r) for r in range(R)]
    for n in sorted({c.num_qubits() for c, _ in reqs}):
        if ev.tune_get(n, "c64") is None:
            ev.autotune(n, "c64", 8)

    # capacity under a burst of the first requests
    burst = reqs[:min(48, R)]
    with ev.Server(precision="c64") as srv:
        t0 = time.perf_counter()
        ids = [srv.submit(c, shots=1024, seed=s) for c, s in burst]
        for i in ids:
            srv.wait(i)
        cap = len(burst) / (time.perf_counter() - t0)
    rate = args.serve_load * cap
    arrivals = ev.workloads.poisson_arrivals(R, rate)

    with ev.Server(precision="c64") as srv:
        ids = [None] * R
        t0 = time.perf_counter()

        def submitter():
            for k, ((c, s), at) in enumerate(zip(reqs, arrivals)):
                d = at - (time.perf_counter() - t0)
                if d > 0:
                    time.sleep(d)
                ids[k] = srv.submit(c, shots=1024, seed=s)

        th = threading.Thread(target=submitter)
        th.start()
        th.join()
        infos = []
        for k in range(R):
            _, info = srv.wait(ids[k])
            infos.append(info)
        wall = time.perf_counter() - t0
        st = srv.stats()
    lat = np.array([i["done_ms"] - i["submit_ms"] for i in infos])
    span_s = (max(i["done_ms"] for i in infos) - min(i["submit_ms"] for i in infos)) * 1e-3
    served = R / span_s

    # the four largest requests (gates x 2^n) each on an idle server: every latency is at least its request's
    # service time, so the k-th largest latency is at least the k-th largest service time
    work = lambda k: len(reqs[k][0]) * 2.0 ** reqs[k][0].num_qubits()
    biggest = sorted(range(R), key=work, reverse=True)[:4]
    alone = []
    with ev.Server(precision="c64") as srv:
        for k in biggest:
            t2 = time.perf_counter()
            srv.wait(srv.submit(reqs[k][0], shots=1024, seed=reqs[k][1]))
            alone.append({"request": k, "qubits": reqs[k][0].num_qubits(), "gates": len(reqs[k][0]),
                          "arrival_ms": arrivals[k] * 1e3, "alone_ms": (time.perf_counter() - t2) * 1e3,
                          "latency_ms": float(lat[k])})
    big = biggest[0]
    big_alone_ms = alone[0]["alone_ms"]

    # single-request plain-GPU baseline: the same requests one at a time, one launch per gate
    t1 = time.perf_counter()
    done = 0
    for c, s in reqs:
        ex = ev.StateVectorBackend.run(c, precision="c64", plain=True)
        ex.state().sample(1024, s)
        done += 1
        if time.perf_counter() - t1 > args.serve_plain_seconds:
            break
    plain_rate = done / (time.perf_counter() - t1)

    # c64 fidelity against the FP64 oracle: the first 16 requests of at most 20 qubits, run on their own
    # (a served request samples exactly as it does on its own)
    circs = [c for c, _ in reqs]
    sub = [i for i, c in enumerate(circs) if c.num_qubits() <= 20][:16]
    fid = oracle_fidelity(circs, sub, lambda i: ev.StateVectorBackend.run(circs[i], precision="c64").state().amplitudes())
    fid["subset"] = "the first 16 requests (arrival order) with at most 20 qubits, run alone in complex64"
    by_n = {}
    ns = np.array([c.num_qubits() for c in circs])
    for lo, hi in ((12, 15), (16, 19), (20, 23), (24, 28)):
        sel = (ns >= lo) & (ns <= hi)
        if sel.any():
            by_n[f"{lo}-{hi}"] = {"requests": int(sel.sum()), "p50_ms": float(np.percentile(lat[sel], 50)),
                                  "p99_ms": float(np.percentile(lat[sel], 99))}
    return {
        "workload": f"{R} requests, n = round(exp(U[ln 12, ln 28])), depth 20..200 brick-wall, 1024 shots each, "
                    f"Poisson arrivals at {rate:.2f}/s ({args.serve_load:.0%} of the burst capacity), complex64",
        "circuits_per_s": served,
        "p50_latency_ms": float(np.percentile(lat, 50)),
        "p99_latency_ms": float(np.percentile(lat, 99)),
        "offered_rate_per_s": rate,
        "burst_capacity_per_s": cap,
        "single_request_plain_circuits_per_s": plain_rate,
        "single_request_plain_sample": f"first {done} requests in arrival order, one at a time, one launch per gate",
        "speedup_vs_single_plain": served / plain_rate,
        "steps": st["steps"],
        "busy_ms": st["busy_ms"],
        "bytes_peak": st["bytes_peak"],
        "wall_s": wall,
        "latency_by_qubits": by_n,
        "largest_request_alone_ms": big_alone_ms,
        "largest_request": f"request {big}: {reqs[big][0].num_qubits()} qubits, {len(reqs[big][0])} gates, on an idle server",
        "largest_requests": alone,
        "p99_floor_note": "every latency is at least its own service time (largest_requests.alone_ms), and p99 over "
                          f"{R} requests interpolates between the 4th and 3rd largest latencies; requests that "
                          "arrive while a larger one is still running add their service times to each other's",
        "fidelity_vs_f64": fid,
    }


# --------------------------------------------------------------------- main

def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--workload", default="all", choices=["all", "c0", "c1", "c2", "c3", "c4"])
    ap.add_argument("--n", type=int, default=None, help="override the c1 qubit count (profiling runs)")
    ap.add_argument("--precision", default="c128", choices=["c128", "c64"])
    ap.add_argument("--plain", action="store_true", help="one launch per gate instead of fused passes")
    ap.add_argument("--mfma-u2", action="store_true",
                    help="complex64: dense 4x4 gates on the matrix cores instead of VALU FMA chains")
    ap.add_argument("--no-autotune", action="store_true", help="never tune; sizes without a table entry use defaults")
    ap.add_argument("--retune", action="store_true", help="tune the c1 size even when the table has an entry")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batch-count", type=int, default=4096)
    ap.add_argument("--serve-requests", type=int, default=256)
    ap.add_argument("--serve-load", type=float, default=0.7)
    ap.add_argument("--serve-plain-seconds", type=float, default=20.0)
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: the library has no multi-GPU path (the statevector would shard by "
              "its top qubits, one slice per GPU). Run with --gpus 1.", file=sys.stderr)
        return 2

    import everett_amd as ev

    steps = args.steps if args.steps is not None else (1 if args.plain else 5)
    warmup = args.warmup if args.warmup is not None else 1
    wl = args.workload
    n1 = args.n or 30

    out = {"metric": METRIC, "n_gpus": 1, "steps": steps, "warmup": warmup, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64" if args.precision == "c128" else "f32",
           "data": "synthetic (seeded random circuits; no datasets)"}

    copy_gbps = ev.copy_bandwidth(4 << 30, 10)
    out["copy_peak_gbps"] = copy_gbps
    log(f"streaming copy: {copy_gbps:.0f} GB/s")
    for p in ("c128", "c64"):
        MEASURED_RATES[p] = ev.compute_rates(p)
    out["compute_peaks_tflops"] = {
        "c128": MEASURED_RATES["c128"], "c64": MEASURED_RATES["c64"],
        "note": "independent multiply-add chains on this GPU (f64 for c128, f32 for c64): the vector ALU alone, "
                "the matrix cores alone (v_mfma_f64_4x4x4f64 / v_mfma_f32_4x4x1f32), and both at equal flops; "
                "a mixed rate no higher than the faster unit alone means the two share one budget"}
    for p, r in MEASURED_RATES.items():
        log(f"{p} multiply-add rates: vector {r['valu_tflops']:.1f}, matrix {r['mfma_tflops']:.1f}, "
            f"both {r['mixed_tflops']:.1f} TFLOP/s")
    if wl in ("all", "c1"):
        c1 = ev.workloads.config1(n1, 40)
        if not args.plain and not args.no_autotune and (args.retune or ev.tune_get(n1, args.precision) is None):
            t = time.perf_counter()
            tuned = ev.autotune(n1, args.precision, 8)
            log(f"autotune n={n1}: {tuned} ({time.perf_counter() - t:.1f} s)")
        log(f"c1: {len(c1)} gates at n={n1}, {args.precision}, {'plain' if args.plain else 'fused'}")
        r = bench_sv("c1", c1, n1, f"{n1}-qubit brick-wall depth 40 (seed {n1}) + QFT({n1})", args, steps, warmup)
        out.update({
            "value": r["gates_per_s"],
            "unit": "gates/s",
            "ms_per_step": r["ms_per_step"],
            "config": {"workload": r["workload"], "n_qubits": n1, "gates": r["gates"], "precision": args.precision,
                       "path": "plain" if args.plain else "fused", "tune": ev.tune_get(n1, args.precision)},
            "hbm_gbps": r["hbm_gbps"],
            "device_ms_per_step": r["device_ms_per_step"],
            "passes_per_step": r["passes_per_step"],
            "state_alloc_ms": r["state_alloc_ms"],
            "roofline": dict(r["roofline"], frac_of_copy=r["roofline"]["achieved"] / copy_gbps),
            "compute_roofline": r["compute_roofline"],
            "kernels": r["kernels"],
            "cpu_baseline": None,
        })
        cr = r["compute_roofline"]
        log(f"c1: {r['gates_per_s']:.0f} gates/s, {r['ms_per_step']:.1f} ms/step, roofline "
            f"{r['roofline']['achieved']:.0f} GB/s" + (f", {cr['achieved']:.1f} TFLOP/s" if cr else ""))
        if not args.no_cpu:
            log(f"c1 cpu baseline ({args.cpu_seconds:.0f} s sample)")
            out["cpu_baseline"] = cpu_baseline(c1, n1, args.cpu_seconds)
            log(f"c1 cpu restatement on {HOST_THREADS} cores ({args.cpu_seconds / 2:.0f} s sample)")
            out["cpu_restatement_all_cores"] = cpu_baseline_all_cores(c1, n1, args.cpu_seconds / 2)
    extra = {}
    if wl == "all" and args.precision == "c128" and not args.plain:
        c1 = ev.workloads.config1(n1, 40)
        r = bench_sv("c1_c64", c1, n1, f"{n1}-qubit brick-wall depth 40 (seed {n1}) + QFT({n1}), complex64",
                     argparse.Namespace(**dict(vars(args), precision="c64")), steps, warmup)
        extra["c1_c64"] = {k: r[k] for k in ("workload", "gates", "gates_per_s", "ms_per_step", "device_ms_per_step",
                                             "passes_per_step", "roofline", "compute_roofline", "kernels")}
        log(f"c1 c64: {r['gates_per_s']:.0f} gates/s")
        # the same passes with the dense 4x4 gates on the matrix cores (v_mfma_f32_4x4x1_16b_f32)
        m = bench_sv("c1_c64", c1, n1, "", argparse.Namespace(**dict(vars(args), precision="c64", mfma_u2=True)),
                     steps, warmup)
        extra["c1_c64"]["mfma_u2"] = {
            "gates_per_s": m["gates_per_s"], "ms_per_step": m["ms_per_step"],
            "device_ms_per_step": m["device_ms_per_step"], "vs_valu": m["gates_per_s"] / r["gates_per_s"],
            "note": "dense 4x4 gates as 64 v_mfma_f32_4x4x1_16b_f32 per op per wave (twice the unpacked VALU FP32 "
                    "rate on the op alone); the kernel then holds 116 VGPRs (4 waves per SIMD instead of 6)"}
        log(f"c1 c64 with matrix-core 4x4 gates: {m['gates_per_s']:.0f} gates/s")
    if wl in ("all", "c0"):
        c0 = ev.workloads.config0(20, 20)
        r = bench_sv("c0", c0, 20, "20-qubit brick-wall depth 20 (seed 20)", args, 10, 2)
        extra["c0"] = {k: r[k] for k in ("workload", "gates", "gates_per_s", "ms_per_step", "device_ms_per_step",
                                         "passes_per_step")}
        log(f"c0: {r['gates_per_s']:.0f} gates/s")
        if not args.no_cpu:
            extra["c0"]["cpu"] = c0_cpu_comparison(c0, r["gates_per_s"], min(args.cpu_seconds, 10.0))
            log(f"c0 CPU: oracle 1 thread {extra['c0']['cpu']['oracle_1_thread']['value']:.1f} gates/s, "
                f"max-abs vs oracle {extra['c0']['cpu']['max_abs_vs_oracle']:.1e}")
    if wl in ("all", "c2"):
        extra["c2"] = bench_c2(args)
        log(f"c2: {extra['c2']['gates_per_s']:.0f} gates/s")
    if wl in ("all", "c3"):
        extra["c3"] = bench_c3(args)
        log(f"c3: {extra['c3']['circuits_per_s']:.0f} circuits/s")
    if wl in ("all", "c4"):
        extra["c4"] = bench_c4(args)
        log(f"c4: {extra['c4']['circuits_per_s']:.1f} circuits/s, p50 {extra['c4']['p50_latency_ms']:.1f} ms, "
            f"p99 {extra['c4']['p99_latency_ms']:.1f} ms")
    if "value" not in out:  # a single non-headline workload was asked for
        key = next(iter(extra))
        e = extra[key]
        out.update({"value": e.get("circuits_per_s", e.get("gates_per_s")),
                    "unit": "circuits/s" if "circuits_per_s" in e else "gates/s",
                    "ms_per_step": e.get("ms_per_step", e.get("wall_ms", e.get("ms"))),
                    "config": {"workload": e["workload"]}, "roofline": None, "cpu_baseline": None})
    if extra:
        out["workloads"] = extra
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
