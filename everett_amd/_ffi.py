"""ctypes binding of libeverett.so (the C ABI declared in include/everett.h).

Loading fails loudly when the native library has not been built: there is no
Python or CPU fallback for any simulation path.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

import os

# EVERETT_AMD_LIB selects another build of the same library (the host-sanitized
# one, everett_amd/build.py --sanitize); the default is the in-tree build.
LIB_PATH = Path(os.environ.get("EVERETT_AMD_LIB") or Path(__file__).resolve().parent / "_lib" / "libeverett.so")

MAX_CONTROLS = 16  # controls carried per ev_op record (an op's controls continue in OP_CONTROLS records)

# ev_op (include/everett.h) as a numpy structured dtype; 408 bytes.
OP_DTYPE = np.dtype(
    [
        ("kind", "<u4"),
        ("num_controls", "<u4"),
        ("target", "<u8"),
        ("b", "<u8"),
        ("controls", "<u8", (MAX_CONTROLS,)),
        ("m", "<c16", (16,)),
    ],
    align=True,
)
assert OP_DTYPE.itemsize == 408

OP_APPLY1, OP_APPLY2, OP_CONTROLLED, OP_MEASURE, OP_IF, OP_CONTROLS = 1, 2, 3, 4, 5, 6


def controlled_records(target: int, m, controls, dtype=None) -> np.ndarray:
    """An Op::Controlled as ev_op records: the head carries the total control
    count and the first MAX_CONTROLS controls, OP_CONTROLS records the rest."""
    controls = [int(c) for c in controls]
    extra = max(0, len(controls) - MAX_CONTROLS)
    out = np.zeros(1 + (extra + MAX_CONTROLS - 1) // MAX_CONTROLS, dtype=dtype or OP_DTYPE)
    out[0]["kind"] = OP_CONTROLLED
    out[0]["target"] = target
    mm = np.asarray(m, dtype=np.complex128).reshape(-1)
    out[0]["m"][: mm.size] = mm
    out[0]["num_controls"] = len(controls)
    head = controls[:MAX_CONTROLS]
    out[0]["controls"][: len(head)] = head
    for r in range(1, len(out)):
        part = controls[MAX_CONTROLS * r: MAX_CONTROLS * (r + 1)]
        out[r]["kind"] = OP_CONTROLS
        out[r]["num_controls"] = len(part)
        out[r]["controls"][: len(part)] = part
    return out


def decode_controls(ops: np.ndarray, i: int):
    """The controls of the controlled op whose head record is ops[i], and the
    index of the record after it."""
    total = int(ops[i]["num_controls"])
    cs = [int(x) for x in ops[i]["controls"][: min(total, MAX_CONTROLS)]]
    j = i + 1
    while len(cs) < total:
        k = int(ops[j]["num_controls"])
        cs += [int(x) for x in ops[j]["controls"][:k]]
        j += 1
    return cs, j

EV_C128, EV_C64 = 0, 1

(
    EV_OK,
    EV_E_QUBIT_RANGE,
    EV_E_CBIT_RANGE,
    EV_E_DUP_QUBIT,
    EV_E_DIM,
    EV_E_ARG,
    EV_E_OOM,
    EV_E_HIP,
    EV_E_ADMISSION,
    EV_E_LIMIT,
    EV_E_STATE,
    EV_E_QASM,
) = (0, -1, -2, -3, -4, -5, -6, -7, -8, -9, -10, -11)


class EvError(C.Structure):
    _fields_ = [("code", C.c_int32), ("a", C.c_uint64), ("b", C.c_uint64), ("message", C.c_char * 256)]


class EvRng(C.Structure):
    _fields_ = [("s", C.c_uint64 * 4)]


class EvRunOptions(C.Structure):
    _fields_ = [
        ("precision", C.c_int),
        ("plain", C.c_uint32),
        ("tile_qubits", C.c_uint32),
        ("max_gates_per_pass", C.c_uint32),
        ("profile", C.c_uint32),
        ("use_graph", C.c_uint32),
        ("renormalize", C.c_uint32),
        ("low_qubits", C.c_uint32),
        ("mfma_u2", C.c_uint32),
    ]


class EvChannel(C.Structure):
    _fields_ = [("kind", C.c_int32), ("p", C.c_double), ("kraus", C.c_void_p), ("num_kraus", C.c_size_t)]


class EvNoiseModel(C.Structure):
    _fields_ = [("after_1q", EvChannel), ("after_2q", EvChannel), ("readout_error", C.c_double)]


class EvRunStats(C.Structure):
    _fields_ = [
        ("logical_gates", C.c_uint64),
        ("passes", C.c_uint64),
        ("launches", C.c_uint64),
        ("device_ms", C.c_double),
        ("pass_ms", C.c_double),
        ("pass_launches", C.c_uint64),
        ("plan_ms", C.c_double),
        ("pass_flops", C.c_double),
    ]


class EvKernelStat(C.Structure):
    _fields_ = [("name", C.c_char * 32), ("launches", C.c_uint64), ("total_ms", C.c_double), ("bytes", C.c_double)]


class EvBatchStats(C.Structure):
    _fields_ = [
        ("circuits", C.c_uint64), ("logical_gates", C.c_uint64), ("rounds", C.c_uint64), ("passes", C.c_uint64),
        ("tiles", C.c_uint64), ("plan_ms", C.c_double), ("upload_ms", C.c_double), ("device_ms", C.c_double),
        ("renorm_ms", C.c_double), ("graph", C.c_uint32), ("groups", C.c_uint32), ("pass_bytes", C.c_double),
        ("pass_flops", C.c_double),
    ]


class EvServerConfig(C.Structure):
    _fields_ = [("precision", C.c_int), ("max_active", C.c_uint32), ("memory_bytes", C.c_uint64),
                ("tile_qubits", C.c_uint32), ("max_gates_per_pass", C.c_uint32), ("reserved", C.c_uint32 * 4)]


class EvRequestInfo(C.Structure):
    _fields_ = [("id", C.c_uint64), ("num_qubits", C.c_uint64), ("gates", C.c_uint64), ("passes", C.c_uint64),
                ("shots", C.c_uint64), ("submit_ms", C.c_double), ("admit_ms", C.c_double), ("done_ms", C.c_double),
                ("status", C.c_int32), ("pad", C.c_int32)]


class EvServerStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("failed", C.c_uint64), ("steps", C.c_uint64),
                ("active", C.c_uint64), ("queued", C.c_uint64), ("bytes_in_use", C.c_uint64),
                ("bytes_peak", C.c_uint64), ("busy_ms", C.c_double)]


class EvTuneEntry(C.Structure):
    _fields_ = [("tile_qubits", C.c_uint32), ("low_qubits", C.c_uint32), ("max_gates_per_pass", C.c_uint32),
                ("pad", C.c_uint32), ("ms_per_gate", C.c_double)]


class EvComputeRates(C.Structure):
    _fields_ = [("valu_tflops", C.c_double), ("mfma_tflops", C.c_double), ("mixed_tflops", C.c_double)]


class EvPlanInfo(C.Structure):
    _fields_ = [
        ("num_passes", C.c_uint32), ("num_stages", C.c_uint32), ("num_ops", C.c_uint32), ("num_coef", C.c_uint32),
        ("pass_bytes", C.c_uint32), ("stage_bytes", C.c_uint32), ("op_bytes", C.c_uint32),
        ("tile_qubits", C.c_uint32), ("num_swap_pairs", C.c_uint32),
        ("swap_a", C.c_uint32 * 32), ("swap_b", C.c_uint32 * 32),
    ]


P = C.c_void_p
PP = C.POINTER(C.c_void_p)
U64 = C.c_uint64
SZ = C.c_size_t
I = C.c_int
D = C.c_double

# name -> (restype, argtypes)
PROTOTYPES = {
    "ev_last_error": (C.c_char_p, []),
    "ev_circuit_create": (I, [U64, U64, PP]),
    "ev_circuit_destroy": (None, [P]),
    "ev_circuit_append": (I, [P, P, SZ]),
    "ev_circuit_compose": (I, [P, P]),
    "ev_circuit_len": (SZ, [P]),
    "ev_circuit_num_qubits": (U64, [P]),
    "ev_circuit_num_classical": (U64, [P]),
    "ev_circuit_export": (I, [P, P, SZ, C.POINTER(SZ)]),
    "ev_circuit_validate": (I, [P, C.POINTER(EvError)]),
    "ev_circuit_from_qasm": (I, [C.c_char_p, SZ, PP, C.POINTER(EvError)]),
    "ev_circuit_to_qasm": (I, [P, C.c_char_p, SZ, C.POINTER(SZ), C.POINTER(EvError)]),
    "ev_rng_seed": (None, [C.POINTER(EvRng), U64]),
    "ev_rng_next_u64": (U64, [C.POINTER(EvRng)]),
    "ev_rng_next_f64": (D, [C.POINTER(EvRng)]),
    "ev_circuit_qft": (I, [U64, PP]),
    "ev_circuit_iqft": (I, [U64, PP]),
    "ev_circuit_brickwall": (I, [U64, U64, C.POINTER(EvRng), PP]),
    "ev_circuit_dm_random": (I, [U64, U64, C.POINTER(EvRng), PP]),
    "ev_sv_create": (I, [U64, I, PP]),
    "ev_sv_from_amplitudes": (I, [P, SZ, I, PP]),
    "ev_sv_destroy": (None, [P]),
    "ev_sv_num_qubits": (U64, [P]),
    "ev_sv_precision": (I, [P]),
    "ev_sv_read": (I, [P, P, SZ]),
    "ev_sv_write": (I, [P, P, SZ]),
    "ev_sv_read_range": (I, [P, U64, P, SZ]),
    "ev_sv_copy": (I, [P, PP]),
    "ev_sv_apply_1q": (I, [P, U64, P]),
    "ev_sv_apply_2q": (I, [P, U64, U64, P]),
    "ev_sv_apply_controlled_1q": (I, [P, P, SZ, U64, P]),
    "ev_sv_measure": (I, [P, U64, D, C.POINTER(I)]),
    "ev_sv_norm_sqr": (I, [P, C.POINTER(D)]),
    "ev_sv_normalize": (I, [P]),
    "ev_sv_probability": (I, [P, U64, C.POINTER(D)]),
    "ev_sv_overlap": (I, [P, P, P]),
    "ev_sv_fidelity": (I, [P, P, C.POINTER(D)]),
    "ev_sv_bloch_vector": (I, [P, U64, P]),
    "ev_sv_sample": (I, [P, U64, U64, P]),
    "ev_sv_sample_noalloc": (I, [P, U64, U64, P]),
    "ev_sv_run": (I, [P, U64, C.POINTER(EvRunOptions), PP, P, SZ, C.POINTER(EvError)]),
    "ev_sv_run_into": (I, [P, P, U64, C.POINTER(EvRunOptions), P, SZ, C.POINTER(EvError)]),
    "ev_sv_reset": (I, [P]),
    "ev_kernel_apply_1q": (I, [P, SZ, SZ, P]),
    "ev_kernel_apply_2q": (I, [P, SZ, SZ, SZ, P]),
    "ev_kernel_apply_controlled_1q": (I, [P, SZ, P, SZ, SZ, P]),
    "ev_channel_kraus": (I, [C.POINTER(EvChannel), P, SZ, C.POINTER(SZ)]),
    "ev_dm_create": (I, [U64, PP]),
    "ev_dm_destroy": (None, [P]),
    "ev_dm_num_qubits": (U64, [P]),
    "ev_dm_read": (I, [P, P, SZ]),
    "ev_dm_trace": (I, [P, C.POINTER(D)]),
    "ev_dm_purity": (I, [P, C.POINTER(D)]),
    "ev_dm_probability": (I, [P, U64, C.POINTER(D)]),
    "ev_dm_expectation_pauli": (I, [P, U64, C.c_char, C.POINTER(D)]),
    "ev_dm_partial_trace_1q": (I, [P, U64, P]),
    "ev_dm_run": (I, [P, C.POINTER(EvNoiseModel), U64, C.POINTER(EvRunOptions), PP, P, SZ, C.POINTER(EvError)]),
    "ev_last_run_stats": (I, [C.POINTER(EvRunStats)]),
    "ev_last_run_kernel_stats": (I, [C.POINTER(EvKernelStat), SZ, C.POINTER(SZ)]),
    "ev_debug_plan": (I, [P, C.POINTER(EvRunOptions), P, P, P, P, C.POINTER(EvPlanInfo)]),
    "ev_batch_run": (I, [P, SZ, C.POINTER(EvRunOptions), PP, C.POINTER(EvError)]),
    "ev_batch_destroy": (None, [P]),
    "ev_batch_count": (SZ, [P]),
    "ev_batch_num_qubits": (U64, [P, SZ]),
    "ev_batch_read": (I, [P, SZ, U64, P, SZ]),
    "ev_batch_norms": (I, [P, P]),
    "ev_batch_compare": (I, [P, P, P, P]),
    "ev_batch_get_stats": (I, [P, C.POINTER(EvBatchStats)]),
    "ev_copy_bandwidth": (I, [SZ, C.c_uint32, C.POINTER(D)]),
    "ev_compute_rates_measure": (I, [I, C.POINTER(EvComputeRates)]),
    "ev_autotune": (I, [U64, I, C.c_uint32, C.POINTER(EvTuneEntry)]),
    "ev_tune_get": (I, [U64, I, C.POINTER(EvTuneEntry), C.POINTER(I)]),
    "ev_tune_set": (I, [U64, I, C.POINTER(EvTuneEntry)]),
    "ev_server_create": (I, [C.POINTER(EvServerConfig), PP]),
    "ev_server_destroy": (None, [P]),
    "ev_server_submit": (I, [P, P, U64, U64, C.POINTER(U64)]),
    "ev_server_submit_qasm": (I, [P, C.c_char_p, SZ, U64, U64, C.POINTER(U64), C.POINTER(EvError)]),
    "ev_server_poll": (I, [P, U64, C.POINTER(I)]),
    "ev_server_wait": (I, [P, U64, P, C.POINTER(EvRequestInfo)]),
    "ev_server_get_stats": (I, [P, C.POINTER(EvServerStats)]),
}

_lib = None


def lib():
    """Returns the loaded libeverett.so, loading it on first use."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"everett_amd native library missing at {LIB_PATH}; build it with `python everett_amd/build.py`"
            )
        h = C.CDLL(str(LIB_PATH))
        for name, (res, args) in PROTOTYPES.items():
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _lib = h
    return _lib


def last_error() -> str:
    return lib().ev_last_error().decode(errors="replace")


def ptr(a: np.ndarray) -> C.c_void_p:
    return C.c_void_p(a.ctypes.data)
