"""Rng (crates/pantometry-core/src/rng.rs:37-152) through the library's host-side generator."""
from __future__ import annotations

import ctypes as C

from ._capi import lib


class Rng:
    """xorshift64* with SplitMix64 seeding; the same stream as the reference's Rng."""

    def __init__(self, seed: int):
        self._s = C.c_uint64(lib().pm_rng_new(seed & 0xFFFFFFFFFFFFFFFF))

    @staticmethod
    def for_index(seed: int, index: int) -> "Rng":
        r = Rng.__new__(Rng)
        r._s = C.c_uint64(lib().pm_rng_for_index(seed & 0xFFFFFFFFFFFFFFFF,
                                                 index & 0xFFFFFFFFFFFFFFFF))
        return r

    def unit(self) -> float:
        return lib().pm_rng_unit(C.byref(self._s))

    def range(self, lo: float, hi: float) -> float:
        return lib().pm_rng_range(C.byref(self._s), lo, hi)

    def gaussian(self) -> float:
        return lib().pm_rng_gaussian(C.byref(self._s))

    def normal(self, mean: float, std_dev: float) -> float:
        return mean + std_dev * self.gaussian()

    def split(self) -> "Rng":
        r = Rng.__new__(Rng)
        r._s = C.c_uint64(lib().pm_rng_split(C.byref(self._s)))
        return r

    def state(self) -> int:
        return self._s.value
