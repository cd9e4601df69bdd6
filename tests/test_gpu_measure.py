"""The bench's achievable-peak references: the streaming copy (HBM) and the
multiply-add rate probes (vector ALU, matrix cores, both), through the C ABI.
They must return finite rates within MI355X's published peaks."""
import math

import pytest

import everett_amd as ev

pytestmark = pytest.mark.gpu

HBM_PEAK_GBPS = 8000.0
VALU_PEAK = {"c128": 78.6, "c64": 78.6}    # FP64; FP32 without packed math
MFMA_PEAK = {"c128": 78.6, "c64": 157.3}   # dense FP64 / FP32 matrix-core rates


def test_copy_bandwidth_is_a_plausible_hbm_rate():
    gbps = ev.copy_bandwidth(1 << 30, 4)
    assert 1000.0 < gbps <= HBM_PEAK_GBPS


@pytest.mark.parametrize("precision", ["c128", "c64"])
def test_compute_rates_within_published_peaks(precision):
    r = ev.compute_rates(precision)
    print(precision, r)
    for k in ("valu_tflops", "mfma_tflops", "mixed_tflops"):
        assert math.isfinite(r[k]) and r[k] > 5.0, (k, r)
    assert r["valu_tflops"] <= 1.05 * VALU_PEAK[precision]
    assert r["mfma_tflops"] <= 1.05 * MFMA_PEAK[precision]
    assert r["mixed_tflops"] <= 1.05 * (VALU_PEAK[precision] + MFMA_PEAK[precision])


def test_compute_rates_rejects_bad_precision():
    from everett_amd import _ffi

    out = _ffi.EvComputeRates()
    assert _ffi.lib().ev_compute_rates_measure(7, out) != 0
