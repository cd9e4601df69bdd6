"""Host-slice kernel entry points with src/kernel.rs's signatures.

Each call copies the numpy complex128 array to the device, applies one gate
with the HIP kernel and copies it back (in place)."""
from __future__ import annotations

import numpy as np

from . import _ffi
from .errors import check


def _amps(a: np.ndarray) -> np.ndarray:
    if not (isinstance(a, np.ndarray) and a.dtype == np.complex128 and a.flags.c_contiguous):
        raise TypeError("amps must be a C-contiguous numpy complex128 array (modified in place)")
    return a


def apply_1q(amps: np.ndarray, k: int, gate) -> None:  # kernel.rs:76
    a = _amps(amps)
    m = np.ascontiguousarray(gate.m, dtype=np.complex128)
    check(_ffi.lib().ev_kernel_apply_1q(_ffi.ptr(a), a.size, k, _ffi.ptr(m)))


def apply_2q(amps: np.ndarray, a_: int, b: int, gate) -> None:  # kernel.rs:192
    a = _amps(amps)
    m = np.ascontiguousarray(gate.m, dtype=np.complex128)
    check(_ffi.lib().ev_kernel_apply_2q(_ffi.ptr(a), a.size, a_, b, _ffi.ptr(m)))


def apply_controlled_1q(amps: np.ndarray, controls, target: int, gate) -> None:  # kernel.rs:299
    a = _amps(amps)
    cs = np.ascontiguousarray(list(controls), dtype=np.uint64)
    m = np.ascontiguousarray(gate.m, dtype=np.complex128)
    check(_ffi.lib().ev_kernel_apply_controlled_1q(_ffi.ptr(a), a.size, _ffi.ptr(cs), cs.size, target, _ffi.ptr(m)))
