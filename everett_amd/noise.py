"""NoiseModel, Channel and KrausOp (src/noise.rs, src/density.rs:462-463)."""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import _ffi
from .errors import check

_KINDS = {
    "depolarizing": 1,
    "amplitude_damping": 2,
    "phase_damping": 3,
    "bit_flip": 4,
    "dephasing": 5,
    "custom": 6,
}


class KrausOp:
    """A 2x2 Kraus operator, row-major (density.rs:462-463)."""

    __slots__ = ("m",)

    def __init__(self, m):
        self.m = np.asarray(m, dtype=np.complex128).reshape(4).copy()


class Channel:
    """src/noise.rs:160-174."""

    __slots__ = ("kind", "p", "ops")

    def __init__(self, kind: str, p: float = 0.0, ops=None):
        if kind not in _KINDS:
            raise ValueError(f"unknown channel kind {kind!r}")
        self.kind, self.p = kind, float(p)
        self.ops = [o if isinstance(o, KrausOp) else KrausOp(o) for o in (ops or [])]

    @staticmethod
    def Depolarizing(p):
        return Channel("depolarizing", p)

    @staticmethod
    def AmplitudeDamping(g):
        return Channel("amplitude_damping", g)

    @staticmethod
    def PhaseDamping(lam):
        return Channel("phase_damping", lam)

    @staticmethod
    def BitFlip(p):
        return Channel("bit_flip", p)

    @staticmethod
    def Dephasing(p):
        return Channel("dephasing", p)

    @staticmethod
    def Custom(ops):
        return Channel("custom", 0.0, ops)

    def _c(self, keep: list) -> _ffi.EvChannel:
        ch = _ffi.EvChannel()
        ch.kind = _KINDS[self.kind]
        ch.p = self.p
        if self.kind == "custom":
            arr = np.ascontiguousarray(np.concatenate([o.m for o in self.ops]) if self.ops else np.zeros(0), dtype=np.complex128)
            keep.append(arr)
            ch.kraus = arr.ctypes.data if arr.size else None
            ch.num_kraus = len(self.ops)
        return ch

    def kraus_ops(self):
        """noise.rs:179-234, evaluated by the native library."""
        keep: list = []
        ch = self._c(keep)
        n = C.c_size_t()
        check(_ffi.lib().ev_channel_kraus(C.byref(ch), None, 0, C.byref(n)))
        out = np.zeros(4 * max(n.value, 1), dtype=np.complex128)
        check(_ffi.lib().ev_channel_kraus(C.byref(ch), _ffi.ptr(out), n.value, C.byref(n)))
        return [KrausOp(out[4 * i : 4 * i + 4]) for i in range(n.value)]


def _prob(name: str, p: float) -> float:
    if not (0.0 <= p <= 1.0):  # noise.rs:62-65 and siblings panic here
        raise ValueError(f"{name} {p} not in [0, 1]")
    return p


class NoiseModel:
    """src/noise.rs:31-40: after_1q, after_2q (one channel each) and readout_error."""

    __slots__ = ("after_1q", "after_2q", "readout_error")

    def __init__(self, after_1q: Optional[Channel] = None, after_2q: Optional[Channel] = None, readout_error: float = 0.0):
        self.after_1q, self.after_2q, self.readout_error = after_1q, after_2q, float(readout_error)

    @staticmethod
    def ideal():
        return NoiseModel()

    @staticmethod
    def uniform_depolarizing(p):
        ch = Channel.Depolarizing(_prob("depolarizing probability", p))
        return NoiseModel(ch, ch)

    @staticmethod
    def amplitude_damping(g):
        ch = Channel.AmplitudeDamping(_prob("gamma", g))
        return NoiseModel(ch, ch)

    @staticmethod
    def phase_damping(lam):
        return NoiseModel(Channel.PhaseDamping(_prob("lambda", lam)), None)

    @staticmethod
    def bit_flip(p):
        return NoiseModel(Channel.BitFlip(_prob("bit-flip probability", p)), None)

    @staticmethod
    def dephasing(p):
        return NoiseModel(Channel.Dephasing(_prob("dephasing probability", p)), None)

    def _c(self, keep: list) -> _ffi.EvNoiseModel:
        nm = _ffi.EvNoiseModel()
        if self.after_1q is not None:
            nm.after_1q = self.after_1q._c(keep)
        if self.after_2q is not None:
            nm.after_2q = self.after_2q._c(keep)
        nm.readout_error = self.readout_error
        return nm
