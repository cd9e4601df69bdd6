"""Rng (src/rng.rs:9-56): SplitMix64-seeded xoshiro256++, evaluated by the
native library on the host."""
from __future__ import annotations

from . import _ffi


class Rng:
    __slots__ = ("_s",)

    def __init__(self):
        self._s = _ffi.EvRng()

    @staticmethod
    def seed_from_u64(seed: int) -> "Rng":
        r = Rng()
        _ffi.lib().ev_rng_seed(r._s, seed & 0xFFFFFFFFFFFFFFFF)
        return r

    def next_u64(self) -> int:
        return int(_ffi.lib().ev_rng_next_u64(self._s))

    def next_f64(self) -> float:
        return float(_ffi.lib().ev_rng_next_f64(self._s))

    @property
    def state(self) -> tuple:
        return tuple(int(x) for x in self._s.s)

    def clone(self) -> "Rng":
        r = Rng()
        for i in range(4):
            r._s.s[i] = self._s.s[i]
        return r
