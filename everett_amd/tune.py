"""Fused-pass tuning tables.

`python -m everett_amd.tune --out FILE [--c128 8-30] [--c64 8-30]` runs the
autotuner (ev_autotune) for each size and writes a JSON table
{"c128": {"30": {...}}, "c64": {...}}.  At import, everett_amd loads
everett_amd/tune_mi355x.json (or the file named by $EVERETT_TUNE_FILE) into
the library, so runs that leave the fused parameters at 0 use it.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

DEFAULT_TABLE = Path(__file__).resolve().parent / "tune_mi355x.json"


def table_path() -> Path:
    return Path(os.environ.get("EVERETT_TUNE_FILE", str(DEFAULT_TABLE)))


def load(path: Path | None = None) -> int:
    """Loads a table into the library; returns the number of entries."""
    from .backend import tune_set

    path = path or table_path()
    if not path.exists():
        return 0
    data = json.loads(path.read_text())
    count = 0
    for prec in ("c128", "c64"):
        for n, e in data.get(prec, {}).items():
            tune_set(int(n), prec, e)
            count += 1
    return count


def _range(spec: str):
    lo, _, hi = spec.partition("-")
    return range(int(lo), int(hi or lo) + 1)


def main() -> int:
    from .backend import autotune

    ap = argparse.ArgumentParser(description="autotune fused-pass parameters")
    ap.add_argument("--out", required=True)
    ap.add_argument("--c128", default="8-30")
    ap.add_argument("--c64", default="8-30")
    ap.add_argument("--depth", type=int, default=8)
    args = ap.parse_args()
    table: dict = {"c128": {}, "c64": {}}
    for prec, spec in (("c128", args.c128), ("c64", args.c64)):
        if not spec:
            continue
        for n in _range(spec):
            t = time.perf_counter()
            e = autotune(n, prec, args.depth)
            table[prec][str(n)] = e
            print(f"[tune] {prec} n={n}: {e} ({time.perf_counter() - t:.1f} s)", file=sys.stderr, flush=True)
            Path(args.out).write_text(json.dumps(table, indent=1, sort_keys=True))
    return 0


if __name__ == "__main__":
    sys.exit(main())
