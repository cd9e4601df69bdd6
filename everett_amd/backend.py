"""StateVectorBackend and DensityMatrixBackend (src/backend/*.rs) on the device."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _ffi
from .circuit import Circuit
from .errors import check, raise_for
from .noise import NoiseModel
from .state import State, precision_code


def run_options(precision="c128", plain: bool = False, tile_qubits: int = 0, max_gates_per_pass: int = 0,
                profile: bool = False, low_qubits: int = 0, mfma_u2: bool = False):
    o = _ffi.EvRunOptions()
    o.low_qubits = low_qubits
    o.mfma_u2 = 1 if mfma_u2 else 0
    o.precision = precision_code(precision)
    o.plain = 1 if plain else 0
    o.tile_qubits = tile_qubits
    o.max_gates_per_pass = max_gates_per_pass
    o.profile = 1 if profile else 0
    return o


def last_kernel_stats():
    """Per-kernel launches, summed device ms and algorithmic bytes of the last profiled run."""
    n = C.c_size_t()
    check(_ffi.lib().ev_last_run_kernel_stats(None, 0, C.byref(n)))
    arr = (_ffi.EvKernelStat * max(n.value, 1))()
    check(_ffi.lib().ev_last_run_kernel_stats(arr, n.value, C.byref(n)))
    return [
        {"name": arr[i].name.decode(), "launches": int(arr[i].launches), "total_ms": float(arr[i].total_ms),
         "bytes": float(arr[i].bytes)}
        for i in range(n.value)
    ]


def run_into(state: State, circuit: Circuit, seed: int = 0, **opts):
    """Resets `state` to |0..0> and runs `circuit` on it (no new allocation)."""
    opt = run_options(state.precision, **opts)
    ncl = circuit.num_classical()
    bits = np.zeros(max(ncl, 1), dtype=np.uint8)
    err = _ffi.EvError()
    rc = _ffi.lib().ev_sv_run_into(state.handle, circuit.handle, seed & 0xFFFFFFFFFFFFFFFF, C.byref(opt),
                                   _ffi.ptr(bits), ncl, C.byref(err))
    raise_for(rc, err)
    return [bool(b) for b in bits[:ncl]]


@dataclass
class RunStats:
    logical_gates: int
    passes: int
    launches: int
    device_ms: float
    pass_ms: float
    pass_launches: int
    plan_ms: float
    pass_flops: float  # algorithmic flops of the fused pass launches


def last_run_stats() -> RunStats:
    s = _ffi.EvRunStats()
    check(_ffi.lib().ev_last_run_stats(C.byref(s)))
    return RunStats(s.logical_gates, s.passes, s.launches, s.device_ms, s.pass_ms, s.pass_launches, s.plan_ms,
                    s.pass_flops)


class Execution:
    """src/backend/statevector.rs:32-54."""

    __slots__ = ("_state", "_classical")

    def __init__(self, state: State, classical):
        self._state, self._classical = state, classical

    def state(self) -> State:
        return self._state

    def classical(self):
        return list(self._classical)

    def into_state(self) -> State:
        return self._state


class StateVectorBackend:
    """src/backend/statevector.rs:25-107."""

    @staticmethod
    def run(circuit: Circuit, **opts) -> Execution:
        return StateVectorBackend.run_seeded(circuit, 0, **opts)

    @staticmethod
    def run_seeded(circuit: Circuit, seed: int, precision="c128", plain: bool = False, tile_qubits: int = 0,
                   max_gates_per_pass: int = 0, low_qubits: int = 0, mfma_u2: bool = False) -> Execution:
        opt = run_options(precision, plain, tile_qubits, max_gates_per_pass, low_qubits=low_qubits, mfma_u2=mfma_u2)
        ncl = circuit.num_classical()
        bits = np.zeros(max(ncl, 1), dtype=np.uint8)
        h = C.c_void_p()
        err = _ffi.EvError()
        rc = _ffi.lib().ev_sv_run(circuit.handle, seed & 0xFFFFFFFFFFFFFFFF, C.byref(opt), C.byref(h), _ffi.ptr(bits),
                                  ncl, C.byref(err))
        raise_for(rc, err)
        return Execution(State(h), [bool(b) for b in bits[:ncl]])


class DensityMatrix:
    """src/density.rs:31-36, resident on the device."""

    __slots__ = ("_h",)

    def __init__(self, handle):
        self._h = handle

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _ffi.lib().ev_dm_destroy(h)
            except Exception:
                pass
            self._h = None

    @staticmethod
    def zero(n: int) -> "DensityMatrix":
        h = C.c_void_p()
        check(_ffi.lib().ev_dm_create(n, C.byref(h)))
        return DensityMatrix(h)

    def num_qubits(self) -> int:
        return int(_ffi.lib().ev_dm_num_qubits(self._h))

    def dim(self) -> int:
        return 1 << self.num_qubits()

    def data(self) -> np.ndarray:
        """Row-major dim*dim entries (density.rs:101-103), copied to the host."""
        d = self.dim()
        out = np.empty(d * d, dtype=np.complex128)
        check(_ffi.lib().ev_dm_read(self._h, _ffi.ptr(out), out.size))
        return out

    def matrix(self) -> np.ndarray:
        d = self.dim()
        return self.data().reshape(d, d)

    def get(self, i: int, j: int) -> complex:
        return complex(self.data()[i * self.dim() + j])

    def trace(self) -> float:
        out = C.c_double()
        check(_ffi.lib().ev_dm_trace(self._h, C.byref(out)))
        return out.value

    def purity(self) -> float:
        out = C.c_double()
        check(_ffi.lib().ev_dm_purity(self._h, C.byref(out)))
        return out.value

    def probability(self, basis: int) -> float:
        out = C.c_double()
        check(_ffi.lib().ev_dm_probability(self._h, basis, C.byref(out)))
        return out.value

    def expectation_pauli(self, qubit: int, observable: str) -> float:
        out = C.c_double()
        check(_ffi.lib().ev_dm_expectation_pauli(self._h, qubit, observable.encode()[:1], C.byref(out)))
        return out.value

    def partial_trace_1q(self, qubit: int):
        out = np.zeros(4, dtype=np.complex128)
        check(_ffi.lib().ev_dm_partial_trace_1q(self._h, qubit, _ffi.ptr(out)))
        return [complex(x) for x in out]


class DensityMatrixExecution:
    """src/backend/density_matrix.rs:44-69."""

    __slots__ = ("_rho", "_classical")

    def __init__(self, rho: DensityMatrix, classical):
        self._rho, self._classical = rho, classical

    def density_matrix(self) -> DensityMatrix:
        return self._rho

    def classical(self):
        return list(self._classical)

    def into_density_matrix(self) -> DensityMatrix:
        return self._rho


class DensityMatrixBackend:
    """src/backend/density_matrix.rs:37-247."""

    @staticmethod
    def run(circuit: Circuit, **opts) -> DensityMatrixExecution:
        return DensityMatrixBackend.run_with_noise_seeded(circuit, NoiseModel.ideal(), 0, **opts)

    @staticmethod
    def run_with_noise(circuit: Circuit, noise: NoiseModel, **opts) -> DensityMatrixExecution:
        return DensityMatrixBackend.run_with_noise_seeded(circuit, noise, 0, **opts)

    @staticmethod
    def run_with_noise_seeded(circuit: Circuit, noise: NoiseModel, seed: int, plain: bool = False,
                              tile_qubits: int = 0, max_gates_per_pass: int = 0,
                              low_qubits: int = 0, profile: bool = False) -> DensityMatrixExecution:
        opt = run_options("c128", plain, tile_qubits, max_gates_per_pass, low_qubits=low_qubits, profile=profile)
        keep: list = []
        nm = noise._c(keep)
        ncl = circuit.num_classical()
        bits = np.zeros(max(ncl, 1), dtype=np.uint8)
        h = C.c_void_p()
        err = _ffi.EvError()
        rc = _ffi.lib().ev_dm_run(circuit.handle, C.byref(nm), seed & 0xFFFFFFFFFFFFFFFF, C.byref(opt), C.byref(h),
                                  _ffi.ptr(bits), ncl, C.byref(err))
        raise_for(rc, err)
        return DensityMatrixExecution(DensityMatrix(h), [bool(b) for b in bits[:ncl]])


def autotune(n: int, precision="c128", depth: int = 0) -> dict:
    """Times fused-pass parameters for n qubits and records the fastest (see ev_autotune)."""
    e = _ffi.EvTuneEntry()
    check(_ffi.lib().ev_autotune(n, precision_code(precision), depth, C.byref(e)))
    return {"tile_qubits": e.tile_qubits, "low_qubits": e.low_qubits, "max_gates_per_pass": e.max_gates_per_pass,
            "ms_per_gate": e.ms_per_gate}


def tune_get(n: int, precision="c128"):
    e = _ffi.EvTuneEntry()
    found = C.c_int()
    check(_ffi.lib().ev_tune_get(n, precision_code(precision), C.byref(e), C.byref(found)))
    if not found.value:
        return None
    return {"tile_qubits": e.tile_qubits, "low_qubits": e.low_qubits, "max_gates_per_pass": e.max_gates_per_pass,
            "ms_per_gate": e.ms_per_gate}


def tune_set(n: int, precision, entry: dict) -> None:
    e = _ffi.EvTuneEntry()
    e.tile_qubits = entry["tile_qubits"]
    e.low_qubits = entry["low_qubits"]
    e.max_gates_per_pass = entry["max_gates_per_pass"]
    e.ms_per_gate = entry.get("ms_per_gate", 0.0)
    check(_ffi.lib().ev_tune_set(n, precision_code(precision), C.byref(e)))


def compute_rates(precision="c128") -> dict:
    """Measured multiply-add rates of this GPU (TFLOP/s) in the arithmetic of
    `precision` (f64 for c128, f32 for c64): the vector ALU alone, the matrix
    cores alone, and both at equal flops."""
    out = _ffi.EvComputeRates()
    check(_ffi.lib().ev_compute_rates_measure(precision_code(precision), C.byref(out)))
    return {"valu_tflops": out.valu_tflops, "mfma_tflops": out.mfma_tflops, "mixed_tflops": out.mixed_tflops}


def copy_bandwidth(nbytes: int = 4 << 30, reps: int = 10) -> float:
    """Measured streaming-copy bandwidth (read + write GB/s) of this GPU."""
    out = C.c_double()
    check(_ffi.lib().ev_copy_bandwidth(nbytes, reps, C.byref(out)))
    return out.value
