"""Benchmark workloads (synthetic random circuits; definitions in docs/configs.md).

 - brickwall(n, depth, rng): per layer Rx, Ry, Rz on every qubit with angles
   TAU * next_f64(), then CNOTs on (0,1),(2,3),... (even layers) or
   (1,2),(3,4),... (odd layers).
 - dm_random(n, depth, rng): the noisy density-matrix workload.
"""
from __future__ import annotations

import ctypes as C

from . import _ffi
from .algorithms import qft
from .circuit import Circuit
from .errors import check
from .noise import Channel, KrausOp, NoiseModel
from .rng import Rng


def brickwall(n: int, depth: int, rng: Rng) -> Circuit:
    h = C.c_void_p()
    check(_ffi.lib().ev_circuit_brickwall(n, depth, rng._s, C.byref(h)))
    return Circuit._wrap(h)


def dm_random(n: int, depth: int, rng: Rng) -> Circuit:
    h = C.c_void_p()
    check(_ffi.lib().ev_circuit_dm_random(n, depth, rng._s, C.byref(h)))
    return Circuit._wrap(h)


def config0(n: int = 20, depth: int = 20) -> Circuit:
    """Plain SV, 20 qubits: brick-wall depth 20, Rng::seed_from_u64(20)."""
    return brickwall(n, depth, Rng.seed_from_u64(n))


def config1(n: int = 30, depth: int = 40) -> Circuit:
    """Single large SV: brick-wall depth 40 (seed 30) followed by qft(30)."""
    c = brickwall(n, depth, Rng.seed_from_u64(n))
    c.compose(qft(n))
    return c


def composite_noise_kraus(p_dep: float = 1e-3, gamma: float = 5e-4):
    """Kraus set {A_j D_i}: depolarizing(p) first, then amplitude damping(gamma)."""
    import numpy as np

    dep = Channel.Depolarizing(p_dep).kraus_ops()
    ad = Channel.AmplitudeDamping(gamma).kraus_ops()
    ops = []
    for a in ad:
        for d in dep:
            ops.append(KrausOp(a.m.reshape(2, 2) @ d.m.reshape(2, 2)))
    return ops


def config2(n: int = 13, depth: int = 30):
    """Noisy DM: dm_random(n, depth, seed n), composite noise after every gate,
    readout error 1%, all qubits measured at the end.  Returns (circuit, noise)."""
    body = dm_random(n, depth, Rng.seed_from_u64(n))
    c = Circuit.with_classical(n, n)
    c.compose(body)
    for q in range(n):
        c.measure(q, q)
    ch = Channel.Custom(composite_noise_kraus())
    return c, NoiseModel(ch, ch, 0.01)


def config3_circuit(i: int, first_seed: int = 1000) -> Circuit:
    """Batched workload circuit i: Rng::seed_from_u64(first_seed + i) draws
    n = 10 + next_u64() % 13 and depth = 10 + next_u64() % 51, then the
    brick-wall layers continue on the same stream."""
    rng = Rng.seed_from_u64(first_seed + i)
    n = 10 + rng.next_u64() % 13
    depth = 10 + rng.next_u64() % 51
    return brickwall(n, depth, rng)


def config3(count: int = 4096, first_seed: int = 1000):
    return [config3_circuit(i, first_seed) for i in range(count)]


def config4_request(r: int, seed: int = 4000):
    """Serving request r: Rng::seed_from_u64(seed + r) draws
    n = round(exp(ln 12 + (ln 28 - ln 12) * next_f64())) and
    depth = 20 + next_u64() % 181, then brick-wall layers on the same stream.
    Returns (circuit, shots seed)."""
    import math

    rng = Rng.seed_from_u64(seed + r)
    n = int(round(math.exp(math.log(12) + (math.log(28) - math.log(12)) * rng.next_f64())))
    depth = 20 + rng.next_u64() % 181
    return brickwall(n, depth, rng), seed + r


def poisson_arrivals(count: int, rate: float, seed: int = 4000):
    """Arrival times (s) with exponential gaps -ln(1-u)/rate from Rng::seed_from_u64(seed ^ 0xA5A5)."""
    import math

    rng = Rng.seed_from_u64(seed ^ 0xA5A5)
    t, out = 0.0, []
    for _ in range(count):
        t += -math.log(1.0 - rng.next_f64()) / rate
        out.append(t)
    return out
