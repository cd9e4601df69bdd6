"""CPU oracle for everett_amd's parity tests (TEST INFRASTRUCTURE ONLY).

A ctypes binding of oracle/build/liboracle.so, the plain-C restatement of the
reference's FP64 algorithms in oracle/refcpu.c.  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg use it, always as the
checker; the product library never loads it.

Parity of the oracle itself is pinned in tests/test_oracle.py (published RNG
vectors, the reference's inline known-answer tests, brute-force Kronecker
products, closed forms).  Functions not pinned there: none.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

import os

HERE = Path(__file__).resolve().parent
# ORACLE_LIB selects the sanitized build (make -C oracle sanitize) for host checks
LIB_PATH = Path(os.environ.get("ORACLE_LIB") or HERE / "build" / "liboracle.so")

OP_DTYPE = np.dtype(
    [
        ("kind", "<u4"),
        ("num_controls", "<u4"),
        ("target", "<u8"),
        ("b", "<u8"),
        ("controls", "<u8", (16,)),
        ("m", "<c16", (16,)),
    ],
    align=True,
)


class RefRng(C.Structure):
    _fields_ = [("s", C.c_uint64 * 4)]


class RefNoise(C.Structure):
    _fields_ = [
        ("kind_1q", C.c_int), ("p_1q", C.c_double), ("custom_1q", C.c_void_p), ("ncustom_1q", C.c_size_t),
        ("kind_2q", C.c_int), ("p_2q", C.c_double), ("custom_2q", C.c_void_p), ("ncustom_2q", C.c_size_t),
        ("readout_error", C.c_double),
    ]


class RC128(C.Structure):
    _fields_ = [("re", C.c_double), ("im", C.c_double)]


_lib = None


def build() -> Path:
    subprocess.run(["make", "-C", str(HERE), "-s"], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        h = C.CDLL(str(LIB_PATH))
        P, SZ, U64, D, I = C.c_void_p, C.c_size_t, C.c_uint64, C.c_double, C.c_int
        R = C.POINTER(RefRng)
        protos = {
            "ref_rng_seed": (None, [R, U64]),
            "ref_rng_next_u64": (U64, [R]),
            "ref_rng_next_f64": (D, [R]),
            "ref_rng_set_state": (None, [R, P]),
            "ref_cmul": (RC128, [RC128, RC128]),
            "ref_norm_sqr": (D, [RC128]),
            "ref_index_pair": (None, [SZ, SZ, P]),
            "ref_index_quad": (None, [SZ, SZ, SZ, P]),
            "ref_set_threads": (None, [I]),
            "ref_apply_1q": (None, [P, SZ, SZ, P]),
            "ref_apply_2q": (None, [P, SZ, SZ, SZ, P]),
            "ref_apply_controlled_1q": (None, [P, SZ, P, SZ, SZ, P]),
            "ref_state_norm_sqr": (D, [P, SZ]),
            "ref_state_normalize": (None, [P, SZ]),
            "ref_state_overlap": (RC128, [P, P, SZ]),
            "ref_measure_qubit": (I, [P, SZ, SZ, R]),
            "ref_sv_run": (I, [P, SZ, SZ, U64, P, P, SZ]),
            "ref_sv_apply_ops": (I, [P, SZ, P, SZ, R, P, SZ]),
            "ref_dm_apply_unitary": (None, [P, SZ, P, P, SZ]),
            "ref_dm_apply_channel": (None, [P, SZ, P, SZ, SZ]),
            "ref_dm_trace": (D, [P, SZ]),
            "ref_dm_purity": (D, [P, SZ]),
            "ref_dm_renormalize": (None, [P, SZ]),
            "ref_channel_kraus": (SZ, [I, D, P, SZ, P]),
            "ref_dm_run": (I, [P, SZ, SZ, C.POINTER(RefNoise), U64, P, P, SZ]),
            "ref_dm_run_superop": (I, [P, SZ, SZ, C.POINTER(RefNoise), U64, P, P, SZ]),
            "ref_gen_qft": (SZ, [SZ, P]),
            "ref_gen_brickwall": (SZ, [SZ, SZ, R, P]),
            "ref_gen_dm_random": (SZ, [SZ, SZ, R, P]),
            "ref_sample": (None, [P, SZ, P, SZ, P]),
            "ref_gate_compose": (None, [P, P, P]),
            "ref_gate_adjoint": (None, [P, P]),
            "ref_gate_controlled": (None, [P, P]),
        }
        for g in ("h", "x", "y", "z", "s", "t", "cnot", "cz", "swap"):
            protos[f"ref_gate_{g}"] = (None, [P])
        for g in ("rx", "ry", "rz", "phase"):
            protos[f"ref_gate_{g}"] = (None, [D, P])
        for name, (res, args) in protos.items():
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _lib = h
    return _lib


def _p(a: np.ndarray):
    return C.c_void_p(a.ctypes.data)


def gate1(name: str, angle: float | None = None) -> np.ndarray:
    """src/gate.rs constructors (restated in refcpu.c): 2x2 row-major as 4 complex."""
    m = np.zeros(4, dtype=np.complex128)
    if name == "id":
        m[0] = m[3] = 1.0  # gate.rs Gate1::id
        return m
    fn = getattr(lib(), f"ref_gate_{name}")
    if angle is None:
        fn(_p(m))
    else:
        fn(float(angle), _p(m))
    return m


def gate2(name: str) -> np.ndarray:
    """Gate2::cnot / cz / swap (src/gate.rs:190-232) as 16 complex, basis |q_a q_b>."""
    m = np.zeros(16, dtype=np.complex128)
    getattr(lib(), f"ref_gate_{name}")(_p(m))
    return m


def gate_compose(a, b) -> np.ndarray:
    """Gate1::compose (src/gate.rs:153-162): a * b, the gate that applies b then a."""
    aa, bb = _mat(a, 4), _mat(b, 4)
    out = np.zeros(4, dtype=np.complex128)
    lib().ref_gate_compose(_p(aa), _p(bb), _p(out))
    return out


def gate_adjoint(a) -> np.ndarray:
    """Gate1::adjoint (src/gate.rs:142-150)."""
    aa = _mat(a, 4)
    out = np.zeros(4, dtype=np.complex128)
    lib().ref_gate_adjoint(_p(aa), _p(out))
    return out


def gate_controlled(u) -> np.ndarray:
    """Gate2::controlled (src/gate.rs:240-247): diag(I, u) in the |q_a q_b> basis."""
    uu = _mat(u, 4)
    out = np.zeros(16, dtype=np.complex128)
    lib().ref_gate_controlled(_p(uu), _p(out))
    return out


def set_threads(n: int) -> None:
    lib().ref_set_threads(n)


class Rng:
    """rng.rs restated (SplitMix64 seeding, xoshiro256++)."""

    def __init__(self, seed: int | None = None, state=None):
        self.r = RefRng()
        if state is not None:
            st = np.asarray(state, dtype=np.uint64)
            lib().ref_rng_set_state(self.r, _p(st))
        else:
            lib().ref_rng_seed(self.r, (seed or 0) & 0xFFFFFFFFFFFFFFFF)

    def next_u64(self) -> int:
        return int(lib().ref_rng_next_u64(self.r))

    def next_f64(self) -> float:
        return float(lib().ref_rng_next_f64(self.r))

    @property
    def state(self):
        return tuple(int(x) for x in self.r.s)


def cmul(a: complex, b: complex) -> complex:
    r = lib().ref_cmul(RC128(a.real, a.imag), RC128(b.real, b.imag))
    return complex(r.re, r.im)


def norm_sqr(a: complex) -> float:
    return float(lib().ref_norm_sqr(RC128(a.real, a.imag)))


def index_pair(k: int, i: int):
    out = np.zeros(2, dtype=np.uint64)
    lib().ref_index_pair(k, i, _p(out))
    return int(out[0]), int(out[1])


def index_quad(a: int, b: int, i: int):
    out = np.zeros(4, dtype=np.uint64)
    lib().ref_index_quad(a, b, i, _p(out))
    return [int(x) for x in out]


def _mat(m, n):
    a = np.ascontiguousarray(np.asarray(m, dtype=np.complex128).reshape(n))
    return a


def apply_1q(amps: np.ndarray, k: int, m) -> None:
    mm = _mat(m, 4)
    lib().ref_apply_1q(_p(amps), amps.size, k, _p(mm))


def apply_2q(amps: np.ndarray, a: int, b: int, m) -> None:
    mm = _mat(m, 16)
    lib().ref_apply_2q(_p(amps), amps.size, a, b, _p(mm))


def apply_controlled_1q(amps: np.ndarray, controls, target: int, m) -> None:
    cs = np.ascontiguousarray(list(controls), dtype=np.uint64)
    mm = _mat(m, 4)
    lib().ref_apply_controlled_1q(_p(amps), amps.size, _p(cs), cs.size, target, _p(mm))


def state_norm_sqr(amps: np.ndarray) -> float:
    return float(lib().ref_state_norm_sqr(_p(amps), amps.size))


def state_normalize(amps: np.ndarray) -> None:
    """State::normalize (src/state.rs:103-111), in place."""
    lib().ref_state_normalize(_p(amps), amps.size)


def state_overlap(a: np.ndarray, other: np.ndarray) -> complex:
    r = lib().ref_state_overlap(_p(a), _p(other), a.size)
    return complex(r.re, r.im)


def measure_qubit(amps: np.ndarray, k: int, rng: Rng) -> bool:
    return bool(lib().ref_measure_qubit(_p(amps), amps.size, k, rng.r))


def sv_run(ops: np.ndarray, n: int, seed: int = 0, num_classical: int = 0):
    ops = np.ascontiguousarray(ops, dtype=OP_DTYPE)
    amps = np.zeros(1 << n, dtype=np.complex128)
    bits = np.zeros(max(num_classical, 1), dtype=np.uint8)
    rc = lib().ref_sv_run(_p(ops), len(ops), n, seed & 0xFFFFFFFFFFFFFFFF, _p(amps), _p(bits), num_classical)
    if rc != 0:
        raise ValueError("malformed op array")
    return amps, [bool(b) for b in bits[:num_classical]]


def sv_apply_ops(ops: np.ndarray, amps: np.ndarray, rng: Rng | None = None, num_classical: int = 0):
    ops = np.ascontiguousarray(ops, dtype=OP_DTYPE)
    rng = rng or Rng(0)
    bits = np.zeros(max(num_classical, 1), dtype=np.uint8)
    rc = lib().ref_sv_apply_ops(_p(ops), len(ops), _p(amps), amps.size, rng.r, _p(bits), num_classical)
    if rc != 0:
        raise ValueError("malformed op array")
    return [bool(b) for b in bits[:num_classical]]


CHANNEL_KINDS = {
    None: 0, "depolarizing": 1, "amplitude_damping": 2, "phase_damping": 3, "bit_flip": 4, "dephasing": 5, "custom": 6,
}


def channel_kraus(kind, p: float = 0.0, custom=None) -> np.ndarray:
    cust = np.ascontiguousarray(np.asarray(custom if custom is not None else np.zeros(0), dtype=np.complex128).reshape(-1))
    out = np.zeros(4 * 64, dtype=np.complex128)
    k = lib().ref_channel_kraus(CHANNEL_KINDS[kind], p, _p(cust) if cust.size else None, cust.size // 4, _p(out))
    return out[: 4 * k].reshape(k, 4)


def dm_apply_unitary(rho: np.ndarray, n: int, u, targets) -> None:
    uu = np.ascontiguousarray(np.asarray(u, dtype=np.complex128).reshape(-1))
    tg = np.ascontiguousarray(list(targets), dtype=np.uint64)
    lib().ref_dm_apply_unitary(_p(rho), n, _p(uu), _p(tg), tg.size)


def dm_apply_channel(rho: np.ndarray, n: int, kraus: np.ndarray, target: int) -> None:
    kk = np.ascontiguousarray(np.asarray(kraus, dtype=np.complex128).reshape(-1))
    lib().ref_dm_apply_channel(_p(rho), n, _p(kk), kk.size // 4, target)


def dm_trace(rho: np.ndarray, n: int) -> float:
    return float(lib().ref_dm_trace(_p(rho), n))


def dm_purity(rho: np.ndarray, n: int) -> float:
    return float(lib().ref_dm_purity(_p(rho), n))


def dm_run(ops: np.ndarray, n: int, noise=None, seed: int = 0, num_classical: int = 0, superop: bool = False):
    """noise: dict(kind_1q, p_1q, custom_1q, kind_2q, p_2q, custom_2q, readout_error).  superop=True runs
    ref_dm_run_superop (every gate and channel as a pass over the 2n-qubit vector; parallel)."""
    ops = np.ascontiguousarray(ops, dtype=OP_DTYPE)
    noise = noise or {}
    keep = []
    nz = RefNoise()

    def chan(kind, p, custom):
        if custom is None:
            return CHANNEL_KINDS[kind], p, None, 0
        arr = np.ascontiguousarray(np.asarray(custom, dtype=np.complex128).reshape(-1))
        keep.append(arr)
        return 6, 0.0, arr.ctypes.data, arr.size // 4

    nz.kind_1q, nz.p_1q, nz.custom_1q, nz.ncustom_1q = chan(noise.get("kind_1q"), noise.get("p_1q", 0.0),
                                                            noise.get("custom_1q"))
    nz.kind_2q, nz.p_2q, nz.custom_2q, nz.ncustom_2q = chan(noise.get("kind_2q"), noise.get("p_2q", 0.0),
                                                            noise.get("custom_2q"))
    nz.readout_error = noise.get("readout_error", 0.0)
    dim = 1 << n
    rho = np.zeros(dim * dim, dtype=np.complex128)
    bits = np.zeros(max(num_classical, 1), dtype=np.uint8)
    fn = lib().ref_dm_run_superop if superop else lib().ref_dm_run
    rc = fn(_p(ops), len(ops), n, C.byref(nz), seed & 0xFFFFFFFFFFFFFFFF, _p(rho), _p(bits), num_classical)
    if rc != 0:
        raise ValueError("malformed op array or noise model")
    return rho, [bool(b) for b in bits[:num_classical]]


def _gen(fn, *args) -> np.ndarray:
    cnt = fn(*args, None)
    out = np.zeros(cnt, dtype=OP_DTYPE)
    return out, cnt


def gen_qft(n: int) -> np.ndarray:
    cnt = lib().ref_gen_qft(n, None)
    out = np.zeros(cnt, dtype=OP_DTYPE)
    lib().ref_gen_qft(n, _p(out))
    return out


def gen_brickwall(n: int, depth: int, rng: Rng) -> np.ndarray:
    probe = Rng(state=rng.state)
    cnt = lib().ref_gen_brickwall(n, depth, probe.r, None)
    out = np.zeros(cnt, dtype=OP_DTYPE)
    lib().ref_gen_brickwall(n, depth, rng.r, _p(out))
    return out


def gen_dm_random(n: int, depth: int, rng: Rng) -> np.ndarray:
    probe = Rng(state=rng.state)
    cnt = lib().ref_gen_dm_random(n, depth, probe.r, None)
    out = np.zeros(cnt, dtype=OP_DTYPE)
    lib().ref_gen_dm_random(n, depth, rng.r, _p(out))
    return out


def sample(amps: np.ndarray, n: int, u: np.ndarray) -> np.ndarray:
    u = np.ascontiguousarray(u, dtype=np.float64)
    out = np.zeros(len(u), dtype=np.uint64)
    lib().ref_sample(_p(amps), n, _p(u), len(u), _p(out))
    return out
