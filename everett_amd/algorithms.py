"""Circuit builders of src/algorithms/qft.rs, generated natively."""
from __future__ import annotations

import ctypes as C

from . import _ffi
from .circuit import Circuit
from .errors import check


def qft(n: int) -> Circuit:  # qft.rs:54-62
    h = C.c_void_p()
    check(_ffi.lib().ev_circuit_qft(n, C.byref(h)))
    return Circuit._wrap(h)


def iqft(n: int) -> Circuit:  # qft.rs:81-91
    h = C.c_void_p()
    check(_ffi.lib().ev_circuit_iqft(n, C.byref(h)))
    return Circuit._wrap(h)
