"""Circuit (src/circuit.rs) over a native ev_circuit handle.

Builder methods mirror the reference's fluent API and record ops without
validating; validate() (run by every backend) reports the first error in the
reference's order (src/circuit.rs:245-306).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _ffi
from .errors import check, raise_for
from .gate import Gate1, Gate2


class Circuit:
    __slots__ = ("_h",)

    def __init__(self, num_qubits: int, num_classical: int = 0, _handle=None):
        if _handle is not None:
            self._h = _handle
        else:
            h = C.c_void_p()
            check(_ffi.lib().ev_circuit_create(num_qubits, num_classical, C.byref(h)))
            self._h = h

    @staticmethod
    def with_classical(num_qubits: int, num_classical: int) -> "Circuit":
        return Circuit(num_qubits, num_classical)

    @classmethod
    def _wrap(cls, handle) -> "Circuit":
        return cls(0, 0, _handle=handle)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _ffi.lib().ev_circuit_destroy(h)
            except Exception:
                pass
            self._h = None

    @property
    def handle(self):
        return self._h

    # --- registers ------------------------------------------------------
    def num_qubits(self) -> int:
        return int(_ffi.lib().ev_circuit_num_qubits(self._h))

    def num_classical(self) -> int:
        return int(_ffi.lib().ev_circuit_num_classical(self._h))

    def __len__(self) -> int:
        """Number of op records (an IfClassic is its If record plus the guarded op; a
        controlled op with more than 16 controls takes continuation records)."""
        return int(_ffi.lib().ev_circuit_len(self._h))

    def is_empty(self) -> bool:
        return len(self) == 0

    def ops(self) -> np.ndarray:
        """The op records as a structured array of ev_op."""
        n = C.c_size_t()
        check(_ffi.lib().ev_circuit_export(self._h, None, 0, C.byref(n)))
        out = np.zeros(n.value, dtype=_ffi.OP_DTYPE)
        if n.value:
            check(_ffi.lib().ev_circuit_export(self._h, _ffi.ptr(out), n.value, C.byref(n)))
        return out

    def append_ops(self, ops: np.ndarray) -> "Circuit":
        ops = np.ascontiguousarray(ops, dtype=_ffi.OP_DTYPE)
        check(_ffi.lib().ev_circuit_append(self._h, _ffi.ptr(ops), len(ops)))
        return self

    def _push(self, kind, target=0, b=0, m=None):
        rec = np.zeros(1, dtype=_ffi.OP_DTYPE)
        rec["kind"] = kind
        rec["target"] = target
        rec["b"] = b
        if m is not None:
            rec["m"][0, : len(m)] = m
        return self.append_ops(rec)

    # --- generic builders (circuit.rs:91-127) ----------------------------
    def gate1(self, gate: Gate1, target: int) -> "Circuit":
        return self._push(_ffi.OP_APPLY1, target, 0, gate.m)

    def gate2(self, gate: Gate2, a: int, b: int) -> "Circuit":
        return self._push(_ffi.OP_APPLY2, a, b, gate.m)

    def controlled(self, controls, gate: Gate1, target: int) -> "Circuit":
        """Circuit::controlled (src/circuit.rs:104-116): any number of controls."""
        return self.append_ops(_ffi.controlled_records(target, gate.m, controls))

    def compose(self, sub: "Circuit") -> "Circuit":
        check(_ffi.lib().ev_circuit_compose(self._h, sub._h))
        return self

    # --- named gates (circuit.rs:131-195) --------------------------------
    def h(self, q):
        return self.gate1(Gate1.h(), q)

    def x(self, q):
        return self.gate1(Gate1.x(), q)

    def y(self, q):
        return self.gate1(Gate1.y(), q)

    def z(self, q):
        return self.gate1(Gate1.z(), q)

    def s(self, q):
        return self.gate1(Gate1.s(), q)

    def t(self, q):
        return self.gate1(Gate1.t(), q)

    def rx(self, q, theta):
        return self.gate1(Gate1.rx(theta), q)

    def ry(self, q, theta):
        return self.gate1(Gate1.ry(theta), q)

    def rz(self, q, theta):
        return self.gate1(Gate1.rz(theta), q)

    def phase(self, q, lam):
        return self.gate1(Gate1.phase(lam), q)

    def cnot(self, control, target):
        return self.gate2(Gate2.cnot(), control, target)

    def cz(self, a, b):
        return self.gate2(Gate2.cz(), a, b)

    def swap(self, a, b):
        return self.gate2(Gate2.swap(), a, b)

    # --- measurement and classical control (circuit.rs:200-236) ----------
    def measure(self, qubit, into):
        return self._push(_ffi.OP_MEASURE, qubit, into)

    def if_bit(self, bit) -> "Circuit":
        """Guards the next appended op with classical bit `bit` (Op::IfClassic)."""
        return self._push(_ffi.OP_IF, bit)

    def x_if(self, bit, q):
        return self.if_bit(bit).x(q)

    def z_if(self, bit, q):
        return self.if_bit(bit).z(q)

    def num_ops(self) -> int:
        """Number of ops as the reference counts them (Circuit::len, src/circuit.rs:68-71):
        an IfClassic and the op it guards are one op."""
        kinds = self.ops()["kind"]
        return int(np.count_nonzero((kinds != _ffi.OP_IF) & (kinds != _ffi.OP_CONTROLS)))

    # --- OpenQASM 3 -------------------------------------------------------
    @staticmethod
    def from_qasm(src) -> "Circuit":
        """Circuit::from_qasm (src/circuit.rs:324-326): the reference's OpenQASM 3.0
        subset; raises QasmError(line, col, message) on a parse error."""
        data = src.encode("utf-8") if isinstance(src, str) else bytes(src)
        h = C.c_void_p()
        err = _ffi.EvError()
        raise_for(_ffi.lib().ev_circuit_from_qasm(data, len(data), C.byref(h), C.byref(err)), err)
        return Circuit._wrap(h)

    def to_qasm(self) -> str:
        """Circuit::to_qasm (src/circuit.rs:315-317); raises QasmError for a gate
        with no OpenQASM 3 name or a nested IfClassic."""
        n = C.c_size_t()
        err = _ffi.EvError()
        rc = _ffi.lib().ev_circuit_to_qasm(self._h, None, 0, C.byref(n), C.byref(err))
        if rc != _ffi.EV_E_LIMIT:
            raise_for(rc, err)
        buf = C.create_string_buffer(n.value + 1)
        raise_for(_ffi.lib().ev_circuit_to_qasm(self._h, buf, n.value + 1, C.byref(n), C.byref(err)), err)
        return buf.raw[: n.value].decode("utf-8")

    # --- validation ------------------------------------------------------
    def validate(self) -> None:
        err = _ffi.EvError()
        raise_for(_ffi.lib().ev_circuit_validate(self._h, C.byref(err)), err)
