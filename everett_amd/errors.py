"""Error types mirroring src/error.rs:8-53, plus device-side failures."""
from __future__ import annotations

from . import _ffi


class EverettError(Exception):
    """Base class of every error raised by everett_amd."""


class QubitOutOfRange(EverettError):
    """Error::QubitOutOfRange (src/error.rs:12-17)."""

    def __init__(self, qubit: int, num_qubits: int, msg: str = ""):
        super().__init__(msg or f"qubit {qubit} out of range for a {num_qubits}-qubit register")
        self.qubit, self.num_qubits = qubit, num_qubits


class ClassicalBitOutOfRange(EverettError):
    """Error::ClassicalBitOutOfRange (src/error.rs:19-24)."""

    def __init__(self, bit: int, num_classical: int, msg: str = ""):
        super().__init__(msg or f"classical bit {bit} out of range for {num_classical} bits")
        self.bit, self.num_classical = bit, num_classical


class DuplicateQubit(EverettError):
    """Error::DuplicateQubit (src/error.rs:26-29)."""

    def __init__(self, qubit: int, msg: str = ""):
        super().__init__(msg or f"qubit {qubit} used twice")
        self.qubit = qubit


class DimensionMismatch(EverettError):
    """Error::DimensionMismatch (src/error.rs:32-37)."""

    def __init__(self, length: int, expected: int, msg: str = ""):
        super().__init__(msg or f"{length} amplitudes is not a power of two (next: {expected})")
        self.len, self.expected = length, expected


class InvalidArgument(EverettError, ValueError):
    """A malformed argument (the reference panics or debug_asserts here)."""


class DeviceError(EverettError):
    """A HIP runtime failure."""


class OutOfMemory(DeviceError):
    """Device memory exhausted."""


class AdmissionError(EverettError):
    """Serving: the request can never fit the configured device memory."""


class LimitError(EverettError):
    """A documented size limit of this build was exceeded."""


class StateError(EverettError):
    """The handle is not in a state that allows the call."""


class QasmError(EverettError, ValueError):
    """Error::Qasm (src/error.rs:44-52): an OpenQASM 3 parse or emit error at
    line:col (1-based; 0 where none applies)."""

    def __init__(self, line: int, col: int, msg: str = ""):
        super().__init__(msg or f"OpenQASM 3 error at {line}:{col}")
        self.line, self.col = line, col
        prefix = f"OpenQASM 3 error at {line}:{col}: "
        self.message = msg[len(prefix):] if msg.startswith(prefix) else msg


def raise_for(rc: int, err: "_ffi.EvError | None" = None) -> None:
    """Raises the exception matching a C ABI status code (no-op for EV_OK)."""
    if rc == _ffi.EV_OK:
        return
    msg = err.message.decode(errors="replace") if err is not None and err.message else _ffi.last_error()
    a = err.a if err is not None else 0
    b = err.b if err is not None else 0
    if rc == _ffi.EV_E_QUBIT_RANGE:
        raise QubitOutOfRange(a, b, msg)
    if rc == _ffi.EV_E_CBIT_RANGE:
        raise ClassicalBitOutOfRange(a, b, msg)
    if rc == _ffi.EV_E_DUP_QUBIT:
        raise DuplicateQubit(a, msg)
    if rc == _ffi.EV_E_DIM:
        raise DimensionMismatch(a, b, msg)
    if rc == _ffi.EV_E_OOM:
        raise OutOfMemory(msg)
    if rc == _ffi.EV_E_HIP:
        raise DeviceError(msg)
    if rc == _ffi.EV_E_ADMISSION:
        raise AdmissionError(msg)
    if rc == _ffi.EV_E_LIMIT:
        raise LimitError(msg)
    if rc == _ffi.EV_E_STATE:
        raise StateError(msg)
    if rc == _ffi.EV_E_QASM:
        raise QasmError(a, b, msg)
    raise InvalidArgument(msg)


def check(rc: int) -> None:
    raise_for(rc, None)
