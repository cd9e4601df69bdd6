"""Summaries of rocprofv3 results (the run_results.db SQLite files it writes).

Run as a script (it needs only the standard library):

  python everett_amd/profiling.py trace --db DB --command CMD --out FILE
      per-kernel dispatch count, total and average duration from a
      `rocprofv3 --kernel-trace --stats` run.

  python everett_amd/profiling.py traffic --fetch DB --write DB --kernel NAME
      --copy-bytes B --algorithmic B --key KEY --json FILE --out FILE
      HBM bytes per launch of one kernel from two `rocprofv3 --pmc` runs
      (FETCH_SIZE, then WRITE_SIZE).  Both counters are scaled by what they
      read for the streaming-copy kernel (k_copy) in the same run, whose
      bytes read and written per launch are known, so the result is in bytes
      of the same kind as the copy's.
"""
from __future__ import annotations

import argparse
import json
import sqlite3
import statistics
import sys
from pathlib import Path


def short_name(name: str) -> str:
    """ev::fused::(anonymous namespace)::k_fused_pass<double>(...) -> k_fused_pass<double>."""
    s = name.replace("(anonymous namespace)::", "").split("(", 1)[0]
    depth, cut = 0, 0
    for i, ch in enumerate(s):  # drop namespace qualifiers outside template arguments
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == ":" and depth == 0 and s[i + 1:i + 2] == ":":
            cut = i + 2
    return s[cut:] or name


def _kernel_rows(db: Path):
    con = sqlite3.connect(str(db))
    rows = con.execute("SELECT name, start, end FROM kernels").fetchall()
    con.close()
    per: dict[str, list[float]] = {}
    for name, s, e in rows:
        per.setdefault(name, []).append((e - s) * 1e-3)  # ns -> us
    return per


def trace(args) -> int:
    per = _kernel_rows(Path(args.db))
    total = sum(sum(v) for v in per.values())
    lines = [f"# {args.title}", "", f"command: {args.command}",
             f"source: rocprofv3 kernel trace ({Path(args.db).name}), {sum(len(v) for v in per.values())} dispatches", ""]
    lines.append(f"{'kernel':<44} {'calls':>7} {'total ms':>11} {'avg us':>11} {'min us':>10} {'max us':>10} {'%':>6}")
    for name, d in sorted(per.items(), key=lambda kv: -sum(kv[1])):
        lines.append(f"{short_name(name)[:44]:<44} {len(d):>7} {sum(d) * 1e-3:>11.3f} {statistics.fmean(d):>11.2f} "
                     f"{min(d):>10.2f} {max(d):>10.2f} {100 * sum(d) / total:>6.2f}")
    text = "\n".join(lines) + "\n"
    Path(args.out).write_text(text)
    print(text, end="")
    return 0


def _counter(db: Path, counter: str) -> dict[str, list[float]]:
    con = sqlite3.connect(str(db))
    rows = con.execute("SELECT dispatch_id, kernel_name, value FROM counters_collection WHERE counter_name = ?",
                       (counter,)).fetchall()
    con.close()
    per: dict[str, list[float]] = {}
    for _, name, value in rows:
        per.setdefault(name, []).append(float(value))
    return per


def _mean_for(per: dict[str, list[float]], pattern: str) -> tuple[float, int]:
    vals = [v for name, vs in per.items() if pattern in name for v in vs]
    if not vals:
        raise SystemExit(f"no dispatches of a kernel matching {pattern!r}")
    return statistics.fmean(vals), len(vals)


def traffic(args) -> int:
    fetch = _counter(Path(args.fetch), "FETCH_SIZE")
    write = _counter(Path(args.write), "WRITE_SIZE")
    k_fetch, nf = _mean_for(fetch, args.kernel)
    k_write, nw = _mean_for(write, args.kernel)
    c_fetch, _ = _mean_for(fetch, args.copy_kernel)
    c_write, _ = _mean_for(write, args.copy_kernel)
    f_fetch = args.copy_bytes / (c_fetch * 1024.0)
    f_write = args.copy_bytes / (c_write * 1024.0)
    fetched = k_fetch * 1024.0 * f_fetch
    written = k_write * 1024.0 * f_write
    total = fetched + written
    entry = {
        "hbm_bytes_per_launch": total,
        "fetch_bytes_per_launch": fetched,
        "write_bytes_per_launch": written,
        "algorithmic_bytes_per_launch": args.algorithmic,
        "traffic_over_algorithmic": total / args.algorithmic,
        "fetch_scale": f_fetch,
        "write_scale": f_write,
        "dispatches": {"fetch": nf, "write": nw},
        "commands": {"fetch": args.fetch_command, "write": args.write_command},
    }
    jf = Path(args.json)
    table = json.loads(jf.read_text()) if jf.exists() else {}
    table[args.key] = entry
    jf.write_text(json.dumps(table, indent=1, sort_keys=True) + "\n")
    lines = [
        f"# {args.title}", "",
        f"fetch run: {args.fetch_command}",
        f"write run: {args.write_command}", "",
        f"kernel matched: {args.kernel} ({nf} dispatches with FETCH_SIZE, {nw} with WRITE_SIZE)",
        f"FETCH_SIZE mean {k_fetch:.6g} KiB/dispatch; copy kernel {c_fetch:.6g} KiB for {args.copy_bytes} B read "
        f"-> scale {f_fetch:.4f}",
        f"WRITE_SIZE mean {k_write:.6g} KiB/dispatch; copy kernel {c_write:.6g} KiB for {args.copy_bytes} B written "
        f"-> scale {f_write:.4f}",
        f"HBM bytes per launch: read {fetched:.6g} + write {written:.6g} = {total:.6g}",
        f"algorithmic bytes per launch: {args.algorithmic:.6g}  (traffic / algorithmic = {total / args.algorithmic:.4f})",
    ]
    text = "\n".join(lines) + "\n"
    Path(args.out).write_text(text)
    print(text, end="")
    return 0


def positions(args) -> int:
    """Average duration and rate of one kernel's launches grouped by position:
    launch i (in start order, the first `limit`) belongs to position
    (i // repeat) % count, e.g. the one-qubit dense launches of a brick-wall
    step, which come `repeat` per qubit in qubit order."""
    con = sqlite3.connect(str(args.db))
    rows = con.execute("SELECT name, start, end FROM kernels ORDER BY start").fetchall()
    con.close()
    d = [(e - s) * 1e-3 for n, s, e in rows if args.kernel in n][: args.limit]
    per: dict[int, list[float]] = {}
    for i, t in enumerate(d):
        per.setdefault((i // args.repeat) % args.count, []).append(t)
    lines = [f"{args.title} ({len(d)} launches of {args.kernel}, {args.bytes:.6g} B each; "
             f"made by: python everett_amd/profiling.py positions --db DB --kernel {args.kernel} "
             f"--repeat {args.repeat} --count {args.count} --limit {args.limit} --bytes {args.bytes:.0f})",
             f"{'pos':>4} {'avg us':>9} {'GB/s':>7} {'of 8 TB/s':>10}"]
    for k in sorted(per):
        m = statistics.fmean(per[k])
        lines.append(f"{k:>4} {m:>9.1f} {args.bytes / (m * 1e-6) / 1e9:>7.0f} {100 * args.bytes / (m * 1e-6) / 8e12:>9.1f}%")
    m = statistics.fmean(d)
    lines.append(f"{'all':>4} {m:>9.1f} {args.bytes / (m * 1e-6) / 1e9:>7.0f} {100 * args.bytes / (m * 1e-6) / 8e12:>9.1f}%")
    text = "\n".join(lines) + "\n"
    if args.append:
        with open(args.append, "a") as f:
            f.write("\n" + text)
    print(text, end="")
    return 0


# Columns of the SQ summary: (header, numerator counter, denominator) where the
# denominator is "waves" (per wave of the dispatch), a counter name,
# "simd-cycles" (a fraction of the device's SIMD time), "simd-waves" (mean
# resident waves per SIMD: SQ_WAVE_CYCLES counts 4-cycle units) or "clock"
# (GHz over the dispatch).  The database sums GRBM_GUI_ACTIVE over the 8 XCDs,
# so one XCD's active cycles are a eighth of it.  A column appears when its
# counters were collected.
SQ_COLUMNS = (
    ("VALU/wave", "SQ_INSTS_VALU", "waves"),
    ("MFMA/wave", "SQ_INSTS_MFMA", "waves"),
    ("LDS/wave", "SQ_INSTS_LDS", "waves"),
    ("bankconf/wave", "SQ_LDS_BANK_CONFLICT", "waves"),
    ("VALU-active", "SQ_ACTIVE_INST_VALU", "SQ_WAVE_CYCLES"),
    ("wait-any", "SQ_WAIT_INST_ANY", "SQ_WAVE_CYCLES"),
    ("wait-lds", "SQ_WAIT_INST_LDS", "SQ_WAVE_CYCLES"),
    ("MFMA-busy", "SQ_VALU_MFMA_BUSY_CYCLES", "simd-cycles"),
    ("waves/SIMD", "SQ_WAVE_CYCLES", "simd-waves"),
    ("GHz", "GRBM_GUI_ACTIVE", "clock"),
)
SIMDS = 1024  # MI355X: 256 CUs x 4 SIMDs
XCDS = 8


def _ratio(c: dict, num: str, den: str, waves: float, us: float = 0.0):
    if num not in c:
        return None
    g = c.get("GRBM_GUI_ACTIVE", 0.0) / XCDS  # one XCD's active cycles
    if den == "waves":
        return c[num] / waves if waves else None
    if den == "simd-cycles":
        return c[num] / (g * SIMDS) if g else None
    if den == "simd-waves":
        return 4.0 * c[num] / (g * SIMDS) if g else None
    if den == "clock":
        return g / (us * 1e3) if us else None
    return c[num] / c[den] if c.get(den) else None


def sq(args) -> int:
    """Per-dispatch SQ counters of one kernel from a `rocprofv3 --pmc SQ_...`
    run, with per-wave figures (waves per dispatch from the kernel's grid and
    workgroup sizes)."""
    con = sqlite3.connect(str(args.db))
    rows = con.execute("SELECT dispatch_id, kernel_name, counter_name, value FROM counters_collection "
                       "ORDER BY dispatch_id").fetchall()
    grid = {d: (gx, wx) for d, gx, wx in con.execute("SELECT dispatch_id, grid_x, workgroup_x FROM kernels")}
    dur = {d: (e - s) * 1e-3 for d, s, e in con.execute("SELECT dispatch_id, start, end FROM kernels")}
    con.close()
    per: dict[int, dict[str, float]] = {}
    for d, name, c, v in rows:
        if args.kernel in name:
            per.setdefault(d, {})[c] = float(v)
    if not per:
        raise SystemExit(f"no dispatches of a kernel matching {args.kernel!r}")
    names = sorted({k for c in per.values() for k in c})
    derived = ("waves", "simd-cycles", "simd-waves", "clock")
    cols = [col for col in SQ_COLUMNS if col[1] in names and (col[2] in derived or col[2] in names)
            and (col[2] not in derived[1:] or "GRBM_GUI_ACTIVE" in names)]
    lines = [f"# {args.title}", "", f"command: {args.command}",
             f"source: rocprofv3 --pmc ({Path(args.db).name}), {len(per)} dispatches of {args.kernel}", "",
             "per dispatch (waves = grid threads / 64; SQ_WAVE_CYCLES and the SQ_ACTIVE_/SQ_WAIT_ counters count in",
             "the SQ's own 4-cycle units, so their ratios are fractions of a wave's life; with G = GRBM_GUI_ACTIVE / "
             f"{XCDS} (one XCD's",
             f"active cycles), MFMA-busy = SQ_VALU_MFMA_BUSY_CYCLES / (G x {SIMDS} SIMDs), waves/SIMD = "
             f"4 SQ_WAVE_CYCLES / (G x {SIMDS}), GHz = G / duration):", "",
             f"{'dispatch':>8} {'us':>9} " + " ".join(f"{h:>13}" for h, _, _ in cols)]
    tot = {k: 0.0 for k in names}
    waves_tot = 0.0
    for d in sorted(per):
        c = per[d]
        gx, _ = grid.get(d, (0, 64))
        waves = gx / 64.0
        waves_tot += waves
        for k in names:
            tot[k] += c.get(k, 0.0)
        cells = []
        for _, num, den in cols:
            r = _ratio(c, num, den, waves, dur.get(d, 0.0))
            cells.append(f"{'-':>13}" if r is None else (f"{r:>13.0f}" if den == "waves" else f"{r:>13.3f}"))
        lines.append(f"{d:>8} {dur.get(d, 0.0):>9.1f} " + " ".join(cells))
    summary = []
    us_tot = sum(dur.get(d, 0.0) for d in per)
    for h, num, den in cols:
        r = _ratio(tot, num, den, waves_tot, us_tot)
        if r is not None:
            summary.append(f"{h} {r:.0f}" if den == "waves" else f"{h} {r:.3f}")
    lines += ["", "all dispatches: " + ", ".join(f"{k} {tot[k]:.4g}" for k in names), "totals: " + ", ".join(summary)]
    text = "\n".join(lines) + "\n"
    Path(args.out).write_text(text)
    print(text, end="")
    return 0


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("trace")
    t.add_argument("--db", required=True)
    t.add_argument("--command", required=True, help="the rocprofv3 command that wrote the db")
    t.add_argument("--title", default="rocprofv3 kernel trace")
    t.add_argument("--out", required=True)
    w = sub.add_parser("traffic")
    w.add_argument("--fetch", required=True)
    w.add_argument("--write", required=True)
    w.add_argument("--fetch-command", required=True)
    w.add_argument("--write-command", required=True)
    w.add_argument("--kernel", required=True, help="substring of the kernel name")
    w.add_argument("--copy-kernel", default="k_copy")
    w.add_argument("--copy-bytes", type=float, required=True, help="bytes read (= written) per copy launch")
    w.add_argument("--algorithmic", type=float, required=True, help="algorithmic bytes per launch of the kernel")
    w.add_argument("--key", required=True)
    w.add_argument("--json", required=True)
    w.add_argument("--title", default="HBM traffic per launch")
    w.add_argument("--out", required=True)
    q = sub.add_parser("sq")
    q.add_argument("--db", required=True)
    q.add_argument("--command", required=True)
    q.add_argument("--kernel", required=True, help="substring of the kernel name")
    q.add_argument("--title", default="SQ counters per dispatch")
    q.add_argument("--out", required=True)
    o = sub.add_parser("positions")
    o.add_argument("--db", required=True)
    o.add_argument("--kernel", required=True)
    o.add_argument("--repeat", type=int, default=1)
    o.add_argument("--count", type=int, required=True)
    o.add_argument("--limit", type=int, default=1 << 30)
    o.add_argument("--bytes", type=float, required=True)
    o.add_argument("--title", default="launches by position")
    o.add_argument("--append", default=None, help="also append the table to this file")
    args = ap.parse_args(argv)
    return {"trace": trace, "traffic": traffic, "sq": sq, "positions": positions}[args.cmd](args)


if __name__ == "__main__":
    sys.exit(main())
