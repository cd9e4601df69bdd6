"""Test-side brute-force references (textbook linear algebra, numpy).

These are independent of both the oracle and the product: full 2^n x 2^n
operators built from the definition of a gate acting on given qubits
(qubit 0 = least-significant index bit, src/qubit.rs:11-13).
"""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np

GOLDEN = Path(__file__).resolve().parent / "golden"


def kats():
    return json.loads((GOLDEN / "reference_kats.json").read_text())


def full_1q(n: int, k: int, g) -> np.ndarray:
    g = np.asarray(g, dtype=np.complex128).reshape(2, 2)
    return np.kron(np.kron(np.eye(1 << (n - k - 1)), g), np.eye(1 << k))


def full_op(n: int, qubits, g) -> np.ndarray:
    """Full operator of a 2^m x 2^m gate g on `qubits` (qubits[0] = MSB of g's basis)."""
    m = len(qubits)
    g = np.asarray(g, dtype=np.complex128).reshape(1 << m, 1 << m)
    dim = 1 << n
    U = np.zeros((dim, dim), dtype=np.complex128)
    mask = 0
    for q in qubits:
        mask |= 1 << q
    for x in range(dim):
        lx = 0
        for q in qubits:
            lx = (lx << 1) | ((x >> q) & 1)
        base = x & ~mask
        for ly in range(1 << m):
            y = base
            for i, q in enumerate(qubits):
                if (ly >> (m - 1 - i)) & 1:
                    y |= 1 << q
            U[y, x] += g[ly, lx]
    return U


def full_controlled(n: int, controls, target: int, g) -> np.ndarray:
    g = np.asarray(g, dtype=np.complex128).reshape(2, 2)
    dim = 1 << n
    cmask = 0
    for c in controls:
        cmask |= 1 << c
    U = np.zeros((dim, dim), dtype=np.complex128)
    for x in range(dim):
        if x & cmask == cmask:
            t = (x >> target) & 1
            for o in range(2):
                y = (x & ~(1 << target)) | (o << target)
                U[y, x] += g[o, t]
        else:
            U[x, x] = 1
    return U


def random_unitary(rng: np.random.Generator, d: int) -> np.ndarray:
    z = (rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d))) / np.sqrt(2)
    q, r = np.linalg.qr(z)
    ph = np.diag(r) / np.abs(np.diag(r))
    return q * ph


def random_state(rng: np.random.Generator, n: int) -> np.ndarray:
    v = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
    return (v / np.linalg.norm(v)).astype(np.complex128)


def random_rho(rng: np.random.Generator, n: int) -> np.ndarray:
    d = 1 << n
    a = rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d))
    rho = a @ a.conj().T
    return (rho / np.trace(rho)).astype(np.complex128)


def sample_state(n: int) -> np.ndarray:
    """src/kernel.rs:631-635: (0.1 + 0.03 j, -0.2 + 0.017 j) for j = 0..2^n-1."""
    j = np.arange(1 << n, dtype=np.float64)
    return (0.1 + j * 0.03) + 1j * (-0.2 + j * 0.017)


def gate_set():
    """Named gates (src/gate.rs) plus a fixed non-special unitary, as 4-vectors."""
    import math

    s = 0.7071067811865476
    out = {
        "h": [s, s, s, -s],
        "x": [0, 1, 1, 0],
        "y": [0, -1j, 1j, 0],
        "z": [1, 0, 0, -1],
        "s": [1, 0, 0, 1j],
        "t": [1, 0, 0, complex(math.cos(math.pi / 4), math.sin(math.pi / 4))],
        "rx": [math.cos(0.35), -1j * math.sin(0.35), -1j * math.sin(0.35), math.cos(0.35)],
        "ry": [math.cos(-0.65), -math.sin(-0.65), math.sin(-0.65), math.cos(-0.65)],
        "rz": [complex(math.cos(-1.05), math.sin(-1.05)), 0, 0, complex(math.cos(1.05), math.sin(1.05))],
        "phase": [1, 0, 0, complex(math.cos(0.9), math.sin(0.9))],
    }
    u = random_unitary(np.random.default_rng(7), 2)
    out["u"] = list(u.reshape(4))
    return {k: np.asarray(v, dtype=np.complex128) for k, v in out.items()}


def ops_array(records):
    """Builds an ev_op-layout array from (kind, target, b, m, controls) tuples.
    A controlled op with more than 16 controls takes continuation records
    (kind 6, 16 controls each) after its head record."""
    from oracle import OP_DTYPE

    rows = []
    for kind, target, b, m, controls in records:
        controls = list(controls or ())
        extra = max(0, len(controls) - 16)
        recs = np.zeros(1 + (extra + 15) // 16, dtype=OP_DTYPE)
        recs[0]["kind"] = kind
        recs[0]["target"] = target
        recs[0]["b"] = b
        if m is not None:
            mm = np.asarray(m, dtype=np.complex128).reshape(-1)
            recs[0]["m"][: mm.size] = mm
        if controls:
            recs[0]["num_controls"] = len(controls)
            recs[0]["controls"][: min(16, len(controls))] = controls[:16]
            for r in range(1, len(recs)):
                part = controls[16 * r: 16 * (r + 1)]
                recs[r]["kind"] = 6
                recs[r]["num_controls"] = len(part)
                recs[r]["controls"][: len(part)] = part
        rows.append(recs)
    return np.concatenate(rows) if rows else np.zeros(0, dtype=OP_DTYPE)


def adjoint_ops(ops: np.ndarray) -> np.ndarray:
    """The inverse circuit of a measurement-free op list: reversed order, every matrix conjugate-transposed."""
    out = ops[::-1].copy()
    for rec in out:
        k = int(rec["kind"])
        if k in (1, 3):
            m = rec["m"][:4].reshape(2, 2)
            rec["m"][:4] = m.conj().T.reshape(4)
        elif k == 2:
            m = rec["m"].reshape(4, 4)
            rec["m"][:] = m.conj().T.reshape(16)
        else:
            raise ValueError("adjoint_ops takes gate-only circuits")
    return out
