"""An ensemble of independent uniform heat blocks stepped together on the GPU (the batched form of
Ensemble::run over Solid3D samples, crates/pantometry-core/src/ensemble.rs:102-147)."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import List, Sequence, Tuple

import numpy as np

from . import _capi
from ._capi import check, dptr, lib
from .thermal import Substance


@dataclass(frozen=True)
class Member:
    counts: Tuple[int, int, int]
    dx: float
    substance: Substance
    initial: float
    dt: float

    def cells(self) -> int:
        return self.counts[0] * self.counts[1] * self.counts[2]


class Ensemble:
    """Members are independent; each is bit-identical to a standalone Solid3D of its own."""

    def __init__(self, members: Sequence[Member]):
        self.members: List[Member] = list(members)
        arr = (_capi.pm_member * max(len(self.members), 1))()
        for m, spec in enumerate(self.members):
            arr[m] = _capi.pm_member(spec.counts[0], spec.counts[1], spec.counts[2], spec.dx,
                                     spec.substance._c(), spec.initial, spec.dt)
        handle = C.c_void_p()
        check(lib().pm_ensemble_create(arr, len(self.members), C.byref(handle)),
              "pm_ensemble_create")
        self._h = handle

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().pm_ensemble_destroy(h)
            self._h = None

    def set_cells(self, member: int, cells: np.ndarray) -> None:
        c = np.ascontiguousarray(cells, dtype=np.float64).reshape(-1)
        if c.size != self.members[member].cells():
            raise ValueError("one temperature per cell of that member")
        check(lib().pm_ensemble_set_cells(self._h, member, dptr(c)), "pm_ensemble_set_cells")

    def cells(self, member: int) -> np.ndarray:
        nx, ny, nz = self.members[member].counts
        out = np.empty(nx * ny * nz)
        check(lib().pm_ensemble_get_cells(self._h, member, dptr(out)), "pm_ensemble_get_cells")
        return out.reshape(nz, ny, nx)

    def run(self, steps: int) -> float:
        """Steps every member `steps` times in one graph launch; returns device milliseconds."""
        v = _capi.pm_violation()
        ms = C.c_double(0.0)
        check(lib().pm_ensemble_run(self._h, int(steps), C.byref(ms), C.byref(v)),
              "pm_ensemble_run", v)
        return ms.value

    def ledgers(self) -> np.ndarray:
        """Each member's stored heat (the stored entry of Solid3D::ledger)."""
        out = np.empty(max(len(self.members), 1))
        check(lib().pm_ensemble_ledgers(self._h, dptr(out)), "pm_ensemble_ledgers")
        return out[:len(self.members)]

    def set_schedule(self, resident_edge: int, z_chunk: int) -> None:
        check(lib().pm_ensemble_set_schedule(self._h, resident_edge, z_chunk),
              "pm_ensemble_set_schedule")

    def set_tiles(self, policy: int) -> None:
        """Streamed tile shapes: 0 whole rows (default), 1 half as many rows, 2 rows up to 64
        cells then 64 x 8, 3 32-column tiles.  Results are the same under every policy."""
        check(lib().pm_ensemble_set_tiles(self._h, int(policy)), "pm_ensemble_set_tiles")

    def set_steps_per_pass(self, steps: int) -> None:
        """Streamed members advance 2 steps per launch (default) or 1; the results are the same."""
        check(lib().pm_ensemble_set_steps_per_pass(self._h, int(steps)),
              "pm_ensemble_set_steps_per_pass")

    def tune(self) -> str:
        buf = C.create_string_buffer(16384)
        check(lib().pm_ensemble_tune(self._h, buf, len(buf)), "pm_ensemble_tune")
        return buf.value.decode()
