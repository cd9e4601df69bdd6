"""Gate matrices (src/gate.rs): 2x2 and 4x4 complex, row-major.

Constructors evaluate the same libm sin/cos as the reference, so matrices
are bit-identical to Gate1::rx etc.
"""
from __future__ import annotations

import math

import numpy as np

FRAC_1_SQRT_2 = 0.7071067811865476  # std::f64::consts::FRAC_1_SQRT_2
FRAC_PI_4 = math.pi / 4.0


def _expi(theta: float) -> complex:  # src/complex.rs:63-65
    return complex(math.cos(theta), math.sin(theta))


class Gate1:
    """A single-qubit gate: row-major [m00, m01, m10, m11] (src/gate.rs:21-25)."""

    __slots__ = ("m",)

    def __init__(self, m00, m01, m10, m11):
        self.m = np.array([m00, m01, m10, m11], dtype=np.complex128)

    def __eq__(self, other):
        return isinstance(other, Gate1) and bool(np.array_equal(self.m, other.m))

    def __repr__(self):
        return f"Gate1({self.m.tolist()})"

    @staticmethod
    def id():
        return Gate1(1, 0, 0, 1)

    @staticmethod
    def x():
        return Gate1(0, 1, 1, 0)

    @staticmethod
    def y():
        return Gate1(0, -1j, 1j, 0)

    @staticmethod
    def z():
        return Gate1(1, 0, 0, -1)

    @staticmethod
    def h():
        s = FRAC_1_SQRT_2
        return Gate1(s, s, s, -s)

    @staticmethod
    def s():
        return Gate1(1, 0, 0, 1j)

    @staticmethod
    def t():
        return Gate1(1, 0, 0, _expi(FRAC_PI_4))

    @staticmethod
    def rx(theta: float):
        s, c = math.sin(theta / 2.0), math.cos(theta / 2.0)
        return Gate1(complex(c, 0), complex(0, -s), complex(0, -s), complex(c, 0))

    @staticmethod
    def ry(theta: float):
        s, c = math.sin(theta / 2.0), math.cos(theta / 2.0)
        return Gate1(complex(c, 0), complex(-s, 0), complex(s, 0), complex(c, 0))

    @staticmethod
    def rz(theta: float):
        return Gate1(_expi(-theta / 2.0), 0, 0, _expi(theta / 2.0))

    @staticmethod
    def phase(lam: float):
        return Gate1(1, 0, 0, _expi(lam))

    def adjoint(self):  # gate.rs:140-148
        m = self.m
        return Gate1(m[0].conjugate(), m[2].conjugate(), m[1].conjugate(), m[3].conjugate())

    def compose(self, other: "Gate1"):  # gate.rs:153-162 (self * other)
        a = self.m.reshape(2, 2)
        b = other.m.reshape(2, 2)
        r = a @ b
        return Gate1(*r.reshape(4))

    def is_identity(self, eps: float) -> bool:
        return bool(np.all(np.abs(self.m - np.array([1, 0, 0, 1])) <= eps))

    def is_unitary(self, eps: float) -> bool:
        return self.adjoint().compose(self).is_identity(eps)


class Gate2:
    """A two-qubit gate: 16 row-major entries, basis |q_a q_b>, a = MSB (src/gate.rs:31-35)."""

    __slots__ = ("m",)

    def __init__(self, m):
        self.m = np.asarray(m, dtype=np.complex128).reshape(16).copy()

    def __eq__(self, other):
        return isinstance(other, Gate2) and bool(np.array_equal(self.m, other.m))

    @staticmethod
    def cnot():
        m = np.eye(4, dtype=np.complex128)
        m[[2, 3]] = m[[3, 2]]
        return Gate2(m)

    @staticmethod
    def cz():
        return Gate2(np.diag([1, 1, 1, -1]).astype(np.complex128))

    @staticmethod
    def swap():
        m = np.eye(4, dtype=np.complex128)
        m[[1, 2]] = m[[2, 1]]
        return Gate2(m)

    @staticmethod
    def controlled(u: Gate1):  # gate.rs:240-247
        m = np.eye(4, dtype=np.complex128)
        m[2:, 2:] = u.m.reshape(2, 2)
        return Gate2(m)

    def adjoint(self):
        return Gate2(self.m.reshape(4, 4).conj().T)

    def is_unitary(self, eps: float) -> bool:
        u = self.m.reshape(4, 4)
        return bool(np.all(np.abs(u.conj().T @ u - np.eye(4)) <= eps))
