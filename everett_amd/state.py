"""State (src/state.rs) resident in HBM, behind a native ev_sv handle."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _ffi
from .errors import check

PRECISIONS = {"c128": _ffi.EV_C128, "c64": _ffi.EV_C64}


def precision_code(p) -> int:
    if isinstance(p, int):
        return p
    try:
        return PRECISIONS[p]
    except KeyError:
        raise ValueError(f"precision must be 'c128' or 'c64', got {p!r}") from None


class State:
    """A device-resident statevector of 2^n amplitudes."""

    __slots__ = ("_h",)

    def __init__(self, handle):
        self._h = handle

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _ffi.lib().ev_sv_destroy(h)
            except Exception:
                pass
            self._h = None

    @property
    def handle(self):
        return self._h

    @staticmethod
    def zero(n: int, precision="c128") -> "State":  # state.rs:38-44
        h = C.c_void_p()
        check(_ffi.lib().ev_sv_create(n, precision_code(precision), C.byref(h)))
        return State(h)

    @staticmethod
    def from_amplitudes(amps, precision="c128") -> "State":  # state.rs:51-62
        a = np.ascontiguousarray(amps, dtype=np.complex128)
        h = C.c_void_p()
        check(_ffi.lib().ev_sv_from_amplitudes(_ffi.ptr(a), a.size, precision_code(precision), C.byref(h)))
        return State(h)

    def num_qubits(self) -> int:
        return int(_ffi.lib().ev_sv_num_qubits(self._h))

    def dim(self) -> int:
        return 1 << self.num_qubits()

    @property
    def precision(self) -> str:
        return "c64" if _ffi.lib().ev_sv_precision(self._h) == _ffi.EV_C64 else "c128"

    def amplitudes(self) -> np.ndarray:
        """Copies the amplitudes to the host (c64 states are widened)."""
        out = np.empty(self.dim(), dtype=np.complex128)
        check(_ffi.lib().ev_sv_read(self._h, _ffi.ptr(out), out.size))
        return out

    def amplitudes_range(self, offset: int, length: int) -> np.ndarray:
        """Copies amplitudes [offset, offset + length) to the host."""
        out = np.empty(length, dtype=np.complex128)
        check(_ffi.lib().ev_sv_read_range(self._h, offset, _ffi.ptr(out), length))
        return out

    def write(self, amps) -> None:
        a = np.ascontiguousarray(amps, dtype=np.complex128)
        check(_ffi.lib().ev_sv_write(self._h, _ffi.ptr(a), a.size))

    def clone(self) -> "State":
        h = C.c_void_p()
        check(_ffi.lib().ev_sv_copy(self._h, C.byref(h)))
        return State(h)

    # --- gates (src/kernel.rs via the device) ----------------------------
    def apply_1q(self, k: int, gate) -> None:
        m = np.ascontiguousarray(gate.m, dtype=np.complex128)
        check(_ffi.lib().ev_sv_apply_1q(self._h, k, _ffi.ptr(m)))

    def apply_2q(self, a: int, b: int, gate) -> None:
        m = np.ascontiguousarray(gate.m, dtype=np.complex128)
        check(_ffi.lib().ev_sv_apply_2q(self._h, a, b, _ffi.ptr(m)))

    def apply_controlled_1q(self, controls, target: int, gate) -> None:
        cs = np.ascontiguousarray(list(controls), dtype=np.uint64)
        m = np.ascontiguousarray(gate.m, dtype=np.complex128)
        check(_ffi.lib().ev_sv_apply_controlled_1q(self._h, _ffi.ptr(cs), cs.size, target, _ffi.ptr(m)))

    # --- observables (state.rs:94-185, f64 on the device) -----------------
    def norm_sqr(self) -> float:
        out = C.c_double()
        check(_ffi.lib().ev_sv_norm_sqr(self._h, C.byref(out)))
        return out.value

    def normalize(self) -> None:
        check(_ffi.lib().ev_sv_normalize(self._h))

    def probability(self, basis: int) -> float:
        out = C.c_double()
        check(_ffi.lib().ev_sv_probability(self._h, basis, C.byref(out)))
        return out.value

    def overlap(self, other: "State") -> complex:
        out = np.zeros(1, dtype=np.complex128)
        check(_ffi.lib().ev_sv_overlap(self._h, other._h, _ffi.ptr(out)))
        return complex(out[0])

    def fidelity(self, other: "State") -> float:
        out = C.c_double()
        check(_ffi.lib().ev_sv_fidelity(self._h, other._h, C.byref(out)))
        return out.value

    def bloch_vector(self, qubit: int):
        out = np.zeros(3, dtype=np.float64)
        check(_ffi.lib().ev_sv_bloch_vector(self._h, qubit, _ffi.ptr(out)))
        return tuple(float(x) for x in out)

    def sample(self, shots: int, seed: int) -> np.ndarray:
        """Born-rule shots: one Rng::next_f64 per shot, inverse f64 CDF."""
        out = np.zeros(shots, dtype=np.uint64)
        check(_ffi.lib().ev_sv_sample(self._h, shots, seed, _ffi.ptr(out)))
        return out


def measure_qubit(state: State, k: int, rng) -> bool:
    """src/measure.rs:15-51: one draw from rng, collapse and renormalise on the device."""
    u = rng.next_f64()
    out = C.c_int()
    check(_ffi.lib().ev_sv_measure(state.handle, k, u, C.byref(out)))
    return bool(out.value)
