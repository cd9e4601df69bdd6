"""The rocprofv3 summary script (everett_amd/profiling.py) on small synthetic databases."""
from __future__ import annotations

import importlib.util
import json
import sqlite3
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
_spec = importlib.util.spec_from_file_location("ev_profiling", ROOT / "everett_amd" / "profiling.py")
prof = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(prof)


def test_short_names():
    assert prof.short_name("ev::fused::(anonymous namespace)::k_fused_pass<double>(HIP_vector_type<double, 2u>*, "
                           "ev::fused::Pass const*)") == "k_fused_pass<double>"
    assert prof.short_name("ev::kern::(anonymous namespace)::k_copy(HIP_vector_type<double, 4u> const*)") == "k_copy"
    assert prof.short_name("k_finalize<1>") == "k_finalize<1>"
    assert prof.short_name("ns::k<ns::T>(int)") == "k<ns::T>"


def _db(path: Path, kernels=(), counters=()):
    con = sqlite3.connect(str(path))
    con.execute("CREATE TABLE kernels (name TEXT, start INTEGER, end INTEGER)")
    con.execute("CREATE TABLE counters_collection (dispatch_id INTEGER, kernel_name TEXT, counter_name TEXT, "
                "value REAL)")
    con.executemany("INSERT INTO kernels VALUES (?, ?, ?)", kernels)
    con.executemany("INSERT INTO counters_collection VALUES (?, ?, ?, ?)", counters)
    con.commit()
    con.close()


def test_trace_table(tmp_path):
    db = tmp_path / "t.db"
    _db(db, kernels=[("ns::k_a(int)", 0, 2000), ("ns::k_a(int)", 5000, 9000), ("ns::k_b(int)", 0, 1000)])
    out = tmp_path / "t.txt"
    prof.main(["trace", "--db", str(db), "--command", "rocprofv3 ... -- python3 bench.py", "--out", str(out)])
    text = out.read_text()
    row = next(line for line in text.splitlines() if line.startswith("k_a"))
    calls, total_ms, avg_us, mn, mx, pct = row.split()[1:]
    assert int(calls) == 2 and float(total_ms) == pytest.approx(0.006) and float(avg_us) == pytest.approx(3.0)
    assert float(mn) == pytest.approx(2.0) and float(mx) == pytest.approx(4.0)
    assert float(pct) == pytest.approx(100 * 6 / 7, abs=0.01)
    assert "command: rocprofv3 ... -- python3 bench.py" in text


def test_traffic_scaling(tmp_path):
    # copy kernel: 1 MiB read and written per launch; counters read half (fetch) and all (write) of it
    fetch = tmp_path / "f.db"
    write = tmp_path / "w.db"
    _db(fetch, counters=[(1, "k_copy(x)", "FETCH_SIZE", 512.0), (2, "k_pass(x)", "FETCH_SIZE", 1024.0),
                         (3, "k_pass(x)", "FETCH_SIZE", 1024.0)])
    _db(write, counters=[(1, "k_copy(x)", "WRITE_SIZE", 1024.0), (2, "k_pass(x)", "WRITE_SIZE", 2048.0)])
    js = tmp_path / "traffic.json"
    prof.main(["traffic", "--fetch", str(fetch), "--write", str(write), "--fetch-command", "f", "--write-command", "w",
               "--kernel", "k_pass", "--copy-bytes", str(1 << 20), "--algorithmic", str(3 << 20), "--key", "k:x",
               "--json", str(js), "--out", str(tmp_path / "o.txt")])
    e = json.loads(js.read_text())["k:x"]
    assert e["fetch_bytes_per_launch"] == pytest.approx(2 << 20)  # 1024 KiB x scale 2
    assert e["write_bytes_per_launch"] == pytest.approx(2 << 20)  # 2048 KiB x scale 1
    assert e["traffic_over_algorithmic"] == pytest.approx(4 / 3)


def test_sq_summary_per_wave_and_fractions(tmp_path):
    db = tmp_path / "q.db"
    con = sqlite3.connect(str(db))
    con.execute("CREATE TABLE kernels (dispatch_id INTEGER, name TEXT, start INTEGER, end INTEGER, grid_x INTEGER, "
                "workgroup_x INTEGER)")
    con.execute("CREATE TABLE counters_collection (dispatch_id INTEGER, kernel_name TEXT, counter_name TEXT, "
                "value REAL)")
    con.executemany("INSERT INTO kernels VALUES (?, ?, ?, ?, ?, ?)",
                    [(1, "k_fused_pass<float, true>", 0, 2_000_000, 6400, 256), (2, "k_copy", 0, 1000, 64, 64)])
    # GRBM_GUI_ACTIVE summed over 8 XCDs: 16 -> 2 active cycles per XCD over the 2 ms dispatch
    vals = {"SQ_INSTS_VALU": 5000.0, "SQ_INSTS_MFMA": 1000.0, "SQ_WAVE_CYCLES": 400.0, "SQ_ACTIVE_INST_VALU": 100.0,
            "SQ_VALU_MFMA_BUSY_CYCLES": 512.0, "GRBM_GUI_ACTIVE": 16.0}
    con.executemany("INSERT INTO counters_collection VALUES (?, ?, ?, ?)",
                    [(1, "k_fused_pass<float, true>", k, v) for k, v in vals.items()] +
                    [(2, "k_copy", "SQ_INSTS_VALU", 7.0)])
    con.commit()
    con.close()
    out = tmp_path / "q.txt"
    prof.main(["sq", "--db", str(db), "--command", "rocprofv3 --pmc ...", "--kernel", "k_fused_pass",
               "--out", str(out)])
    text = out.read_text()
    row = [l for l in text.splitlines() if l.strip().startswith("1 ")][0].split()
    # 100 waves: VALU 50 per wave, MFMA 10 per wave, VALU-active 0.25, MFMA busy 512 / (2 x 1024) = 0.25,
    # resident waves 4 x 400 / (2 x 1024), clock 2 cycles / 2 ms
    assert row[2:] == ["50", "10", "0.250", "0.250", "0.781", "0.000"]
    assert "MFMA-busy 0.250" in text and "k_copy" not in text
