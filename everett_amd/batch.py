"""Batched execution of many independent circuits (ev_batch_* in include/everett.h)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _ffi
from .backend import run_options
from .errors import check, raise_for


class Batch:
    """Statevectors of many circuits resident in one HBM arena."""

    __slots__ = ("_h", "_keep")

    def __init__(self, handle, keep=None):
        self._h = handle
        self._keep = keep

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _ffi.lib().ev_batch_destroy(h)
            except Exception:
                pass
            self._h = None

    @staticmethod
    def run(circuits, precision="c64", use_graph: bool = True, renormalize: bool = False, tile_qubits: int = 0,
            max_gates_per_pass: int = 0) -> "Batch":
        """Runs every circuit from |0..0>.  Circuits run in pipelined groups (the host plans the next group
        while the device runs the current one); round r of a group is one launch over the r-th pass of
        every circuit in the group."""
        opt = run_options(precision, False, tile_qubits, max_gates_per_pass)
        opt.use_graph = 1 if use_graph else 0
        opt.renormalize = 1 if renormalize else 0
        arr = (C.c_void_p * len(circuits))(*[c.handle.value for c in circuits])
        h = C.c_void_p()
        err = _ffi.EvError()
        raise_for(_ffi.lib().ev_batch_run(arr, len(circuits), C.byref(opt), C.byref(h), C.byref(err)), err)
        return Batch(h)

    def __len__(self) -> int:
        return int(_ffi.lib().ev_batch_count(self._h))

    def num_qubits(self, i: int) -> int:
        return int(_ffi.lib().ev_batch_num_qubits(self._h, i))

    def amplitudes(self, i: int, offset: int = 0, length: int | None = None) -> np.ndarray:
        if length is None:
            length = (1 << self.num_qubits(i)) - offset
        out = np.empty(length, dtype=np.complex128)
        check(_ffi.lib().ev_batch_read(self._h, i, offset, _ffi.ptr(out), length))
        return out

    def norms(self) -> np.ndarray:
        out = np.zeros(len(self), dtype=np.float64)
        check(_ffi.lib().ev_batch_norms(self._h, _ffi.ptr(out)))
        return out

    def compare(self, ref: "Batch"):
        """Per-circuit (fidelity, max-abs) against a c128 batch of the same circuits."""
        f = np.zeros(len(self), dtype=np.float64)
        m = np.zeros(len(self), dtype=np.float64)
        check(_ffi.lib().ev_batch_compare(self._h, ref._h, _ffi.ptr(f), _ffi.ptr(m)))
        return f, m

    def stats(self) -> dict:
        s = _ffi.EvBatchStats()
        check(_ffi.lib().ev_batch_get_stats(self._h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in s._fields_ if k != "pad"}
