#!/usr/bin/env python3
"""Benchmark: gates/s and HBM GB/s of a 30-qubit complex128 random circuit on one MI355X.

Workload (default): a 30-qubit brick-wall circuit of depth 40 (Rx, Ry, Rz on
every qubit per layer, angles TAU * Rng::next_f64() from seed 30, then CNOTs
on alternating pairs) followed by QFT(30): 4,660 logical gates.  One step =
reset the HBM-resident state to |0..0> and run the whole circuit.  The state
is allocated once, outside the timed region.

Prints one JSON line with the rate, the roofline of the dominant kernel
(algorithmic bytes per launch / average launch time, HIP events on the
library's stream over the timed steps) and a CPU baseline: the oracle (a
single-threaded C restatement of the reference's loops) timed on this host
over a bounded prefix of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "gates/sec + HBM GB/s at 30q f64; served circuits/sec with p50/p99 latency"
HBM_PEAK_GBPS = 8000.0


def workload(name: str, n: int | None):
    import everett_amd as ev

    if name == "c1":
        n = n or 30
        c = ev.workloads.config1(n, 40)
        desc = f"{n}-qubit brick-wall depth 40 (seed {n}) + QFT({n})"
    elif name == "c0":
        n = n or 20
        c = ev.workloads.config0(n, 20)
        desc = f"{n}-qubit brick-wall depth 20 (seed {n})"
    else:
        raise SystemExit(f"unknown workload {name}")
    return c, n, desc


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


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--workload", default="c1", choices=["c1", "c0"])
    ap.add_argument("--n", type=int, default=None, help="override the qubit count (profiling runs)")
    ap.add_argument("--precision", default="c128", choices=["c128", "c64"])
    ap.add_argument("--plain", action="store_true", help="one launch per gate instead of fused passes")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: every GPU box here has a single MI355X and this build has no "
              "multi-GPU path (the sharding design is in DESIGN.md). Run with --gpus 1.", file=sys.stderr)
        return 2

    import everett_amd as ev

    steps = args.steps if args.steps is not None else (1 if args.plain else 5)
    warmup = args.warmup if args.warmup is not None else 1

    circuit, n, desc = workload(args.workload, args.n)
    gates = len(circuit)
    print(f"[bench] workload {desc}: {gates} gates, precision {args.precision}, "
          f"{'plain' if args.plain else 'fused'} path", file=sys.stderr, flush=True)

    t = time.perf_counter()
    state = ev.State.zero(n, args.precision)
    alloc_ms = (time.perf_counter() - t) * 1e3

    for w in range(warmup):
        ev.run_into(state, circuit, 0, plain=args.plain)
        print(f"[bench] warmup {w + 1}/{warmup} done", file=sys.stderr, flush=True)

    # Timed region: K steps, HIP events around every kernel launch.
    state.norm_sqr()  # synchronises the handle's stream
    t0 = time.perf_counter()
    agg: dict = {}
    dev_ms = 0.0
    for k in range(steps):
        ev.run_into(state, circuit, 0, plain=args.plain, profile=True)
        dev_ms += ev.last_run_stats().device_ms
        for ks in ev.last_kernel_stats():
            a = agg.setdefault(ks["name"], {"launches": 0, "total_ms": 0.0, "bytes": 0.0})
            a["launches"] += ks["launches"]
            a["total_ms"] += ks["total_ms"]
            a["bytes"] += ks["bytes"]
        print(f"[bench] step {k + 1}/{steps} done", file=sys.stderr, flush=True)
    norm = state.norm_sqr()
    t1 = time.perf_counter()
    wall = t1 - t0
    if abs(norm - 1.0) > 1e-6:
        print(f"[bench] norm drifted to {norm}", file=sys.stderr)
        return 1

    amp_bytes = 16 if args.precision == "c128" else 8
    step_bytes = sum(a["bytes"] for a in agg.values()) / steps
    dom_name, dom = max(agg.items(), key=lambda kv: kv[1]["total_ms"])
    per_launch_bytes = dom["bytes"] / dom["launches"]
    avg_ms = dom["total_ms"] / dom["launches"]
    achieved = per_launch_bytes / (avg_ms * 1e-3) / 1e9
    traffic = None
    tf = ROOT / "profiles" / "traffic.json"
    if tf.exists():
        entry = json.loads(tf.read_text()).get(f"{dom_name}:{args.workload}:{n}:{args.precision}")
        if entry:
            traffic = entry.get("hbm_bytes_per_launch")

    out = {
        "metric": METRIC,
        "value": gates * steps / wall,
        "unit": "gates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": wall / steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "c128" else "f32",
        "data": "synthetic (seeded random brick-wall circuit + QFT; no datasets)",
        "config": {
            "workload": desc,
            "n_qubits": n,
            "gates": gates,
            "precision": args.precision,
            "path": "plain" if args.plain else "fused",
        },
        "hbm_gbps": step_bytes / (wall / steps) / 1e9,
        "device_ms_per_step": dev_ms / steps,
        "state_alloc_ms": alloc_ms,
        "roofline": {
            "bound": "hbm",
            "kernel": dom_name,
            "achieved": achieved,
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBPS,
            "traffic": traffic,
            "bytes_per_launch": per_launch_bytes,
            "avg_launch_ms": avg_ms,
            "launches_per_step": dom["launches"] / steps,
        },
        "kernels": {k: {"launches_per_step": v["launches"] / steps, "ms_per_step": v["total_ms"] / steps,
                        "gbps": (v["bytes"] / (v["total_ms"] * 1e-3) / 1e9) if v["total_ms"] else None}
                    for k, v in sorted(agg.items(), key=lambda kv: -kv[1]["total_ms"])},
        "cpu_baseline": None,
    }
    del state
    if not args.no_cpu:
        print(f"[bench] cpu baseline ({args.cpu_seconds:.0f} s sample)", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(circuit, n, args.cpu_seconds)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
