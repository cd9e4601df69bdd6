"""ev_sv_sample_noalloc (sampling through the library's fixed per-GPU scratch,
no device allocation) against ev_sv_sample, both through the C ABI.

Two calls in a row, the second on a smaller state with fewer shots, must each
equal ev_sv_sample bit for bit: the scratch keeps the previous call's
contents, so a stale chunk total or index would show.  The shot limit is an
error before any device work.
"""
import ctypes as C

import numpy as np
import pytest

import everett_amd as ev
from everett_amd import _ffi

pytestmark = pytest.mark.gpu

MAX_SHOTS = 65536


def _noalloc(state, shots, seed):
    out = np.zeros(max(shots, 1), dtype=np.uint64)
    rc = _ffi.lib().ev_sv_sample_noalloc(state.handle, shots, seed, _ffi.ptr(out))
    return rc, out[:shots]


def _alloc(state, shots, seed):
    out = np.zeros(max(shots, 1), dtype=np.uint64)
    assert _ffi.lib().ev_sv_sample(state.handle, shots, seed, _ffi.ptr(out)) == 0
    return out[:shots]


def _state(n, depth, precision, seed):
    c = ev.workloads.brickwall(n, depth, ev.Rng.seed_from_u64(seed))
    return ev.StateVectorBackend.run(c, precision=precision).state()


@pytest.mark.parametrize("precision", ["c128", "c64"])
def test_noalloc_sampling_equals_allocating_entry_point(precision):
    # a large state first (2^16 chunk totals at 27 qubits), then a smaller one with fewer shots
    cases = [(27 if precision == "c64" else 26, 3, MAX_SHOTS, 11), (14, 6, 777, 12), (9, 4, 0, 13), (9, 4, 5, 14)]
    for n, depth, shots, seed in cases:
        print(f"noalloc case n={n} {precision} shots={shots} seed={seed}", flush=True)
        s = _state(n, depth, precision, seed)
        rc, got = _noalloc(s, shots, seed)
        assert rc == 0, rc
        want = _alloc(s, shots, seed)
        assert np.array_equal(got, want), (n, shots)
        if shots:
            assert int(got.max()) < (1 << n) and len(np.unique(got)) > 1
        del s


def test_noalloc_sampling_rejects_too_many_shots():
    s = _state(10, 2, "c128", 1)
    out = np.zeros(MAX_SHOTS + 1, dtype=np.uint64)
    rc = _ffi.lib().ev_sv_sample_noalloc(s.handle, MAX_SHOTS + 1, 3, _ffi.ptr(out))
    assert rc == _ffi.EV_E_LIMIT
    assert not out.any()  # refused before anything was drawn or launched
    rc = _ffi.lib().ev_sv_sample_noalloc(s.handle, 4, 3, C.c_void_p(0))
    assert rc == _ffi.EV_E_ARG
    rc, got = _noalloc(s, 64, 3)  # the scratch still serves the next call
    assert rc == 0 and np.array_equal(got, _alloc(s, 64, 3))
