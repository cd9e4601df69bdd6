"""everett_amd: an MI355X-native engine with hz2/everett's simulation API.

Circuits are recorded on the host (C++ behind a C ABI); statevectors and
density matrices live in HBM and are evolved by hand-written HIP kernels for
gfx950.  The Python layer is a thin ctypes binding that mirrors the
reference's public names (Circuit, Gate1/Gate2, StateVectorBackend,
DensityMatrixBackend, NoiseModel, Channel, KrausOp, Rng, State).
"""
from . import algorithms, kernel, workloads
from .batch import Batch
from .server import Server
from .backend import (
    DensityMatrix,
    DensityMatrixBackend,
    DensityMatrixExecution,
    Execution,
    RunStats,
    StateVectorBackend,
    autotune,
    compute_rates,
    copy_bandwidth,
    last_kernel_stats,
    last_run_stats,
    run_into,
    tune_get,
    tune_set,
)
from .circuit import Circuit
from .errors import (
    AdmissionError,
    ClassicalBitOutOfRange,
    DeviceError,
    DimensionMismatch,
    DuplicateQubit,
    EverettError,
    InvalidArgument,
    LimitError,
    OutOfMemory,
    QasmError,
    QubitOutOfRange,
    StateError,
)
from .gate import Gate1, Gate2
from .noise import Channel, KrausOp, NoiseModel
from .rng import Rng
from .state import State, measure_qubit

__all__ = [
    "algorithms", "kernel", "workloads", "Batch", "Server", "Circuit", "Gate1", "Gate2", "State", "measure_qubit", "Rng",
    "StateVectorBackend", "Execution", "DensityMatrixBackend", "DensityMatrix", "DensityMatrixExecution",
    "NoiseModel", "Channel", "KrausOp", "RunStats", "last_run_stats", "last_kernel_stats", "run_into", "autotune", "tune_get", "tune_set", "copy_bandwidth", "compute_rates", "EverettError", "QasmError",
    "QubitOutOfRange",
    "ClassicalBitOutOfRange", "DuplicateQubit", "DimensionMismatch", "InvalidArgument", "DeviceError",
    "OutOfMemory", "AdmissionError", "LimitError", "StateError",
]


def _load_tuning() -> None:
    from . import tune

    try:
        tune.load()
    except FileNotFoundError:
        pass


_load_tuning()
