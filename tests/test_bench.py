"""The bench line's arithmetic and contract, on the CPU: the roofline objects
derived from measured times, and the refusal of more than one GPU."""
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))

import bench  # noqa: E402


def test_batch_compute_roofline(monkeypatch):
    monkeypatch.setattr(bench, "MEASURED_RATES", {})
    r = bench.batch_compute_roofline(16.0e12, 400.0, 200)
    # 16 TFLOP in 0.4 s = 40 TFLOP/s against the unpacked FP32 vector peak
    assert r["achieved"] == pytest.approx(40.0)
    assert r["frac"] == pytest.approx(40.0 / bench.VALU_PEAK_TFLOPS["c64"])
    assert r["avg_launch_ms"] == pytest.approx(2.0) and r["launches"] == 200
    assert r["bound"] == bench.VALU_BOUND["c64"] and "attainable" not in r
    monkeypatch.setattr(bench, "MEASURED_RATES", {"c64": {"valu_tflops": 80.0}})
    r = bench.batch_compute_roofline(16.0e12, 400.0, 200)
    assert r["attainable"] == 80.0 and r["frac_of_attainable"] == pytest.approx(0.5)
    assert bench.batch_compute_roofline(0.0, 400.0, 200) is None
    assert bench.batch_compute_roofline(16.0e12, 0.0, 200) is None


def test_compute_roofline_per_launch(monkeypatch):
    monkeypatch.setattr(bench, "MEASURED_RATES", {"c128": {"valu_tflops": 60.0}})
    agg = {"fused_pass": {"launches": 4, "total_ms": 40.0}}
    r = bench.compute_roofline(agg, 4 * 300e9, "c128")
    # 300 GFLOP per launch in 10 ms = 30 TFLOP/s
    assert r["flops_per_launch"] == pytest.approx(300e9)
    assert r["achieved"] == pytest.approx(30.0)
    assert r["frac"] == pytest.approx(30.0 / bench.VALU_PEAK_TFLOPS["c128"])
    assert r["frac_of_attainable"] == pytest.approx(0.5)
    assert bench.compute_roofline({}, 1e12, "c128") is None


@pytest.mark.parametrize("argv", [["--gpus", "2"], ["--gpus", "8"]])
def test_more_than_one_gpu_is_refused(argv):
    p = subprocess.run([sys.executable, str(ROOT / "bench.py"), *argv], capture_output=True, text=True, timeout=60)
    assert p.returncode == 2
    assert "one GPU" in p.stderr and not p.stdout.strip()
