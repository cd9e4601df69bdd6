"""Continuous-batching server (ev_server_* in include/everett.h)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _ffi
from .errors import check, raise_for
from .state import precision_code


class Server:
    """Admits circuits by HBM footprint and advances all admitted ones together."""

    def __init__(self, precision="c64", max_active: int = 0, memory_bytes: int = 0, tile_qubits: int = 0,
                 max_gates_per_pass: int = 0):
        cfg = _ffi.EvServerConfig()
        cfg.precision = precision_code(precision)
        cfg.max_active = max_active
        cfg.memory_bytes = memory_bytes
        cfg.tile_qubits = tile_qubits
        cfg.max_gates_per_pass = max_gates_per_pass
        h = C.c_void_p()
        check(_ffi.lib().ev_server_create(C.byref(cfg), C.byref(h)))
        self._h = h
        self._shots = {}

    def close(self):
        if self._h is not None and self._h.value:
            _ffi.lib().ev_server_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def submit(self, circuit, shots: int = 1024, seed: int = 0) -> int:
        rid = C.c_uint64()
        check(_ffi.lib().ev_server_submit(self._h, circuit.handle, shots, seed & 0xFFFFFFFFFFFFFFFF, C.byref(rid)))
        self._shots[rid.value] = shots
        return rid.value

    def submit_qasm(self, src, shots: int = 1024, seed: int = 0) -> int:
        """Submits a request given as OpenQASM 3 text (the serving wire format)."""
        data = src.encode("utf-8") if isinstance(src, str) else bytes(src)
        rid = C.c_uint64()
        err = _ffi.EvError()
        raise_for(_ffi.lib().ev_server_submit_qasm(self._h, data, len(data), shots, seed & 0xFFFFFFFFFFFFFFFF,
                                                     C.byref(rid), C.byref(err)), err)
        self._shots[rid.value] = shots
        return rid.value

    def poll(self, rid: int) -> bool:
        d = C.c_int()
        check(_ffi.lib().ev_server_poll(self._h, rid, C.byref(d)))
        return bool(d.value)

    def wait(self, rid: int):
        """Blocks until request `rid` completes; returns (samples, info dict)."""
        shots = self._shots.pop(rid, 0)
        out = np.zeros(max(shots, 1), dtype=np.uint64)
        info = _ffi.EvRequestInfo()
        check(_ffi.lib().ev_server_wait(self._h, rid, _ffi.ptr(out), C.byref(info)))
        d = {k: getattr(info, k) for k, _ in info._fields_ if k != "pad"}
        return out[:shots], d

    def stats(self) -> dict:
        s = _ffi.EvServerStats()
        check(_ffi.lib().ev_server_get_stats(self._h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in s._fields_}
