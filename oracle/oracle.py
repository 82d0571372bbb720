"""Python face of the CPU oracle (pm_oracle.c).  TEST INFRASTRUCTURE ONLY: imported by tests/,
the smoke check and the benchmark's CPU-baseline leg, never by the product.

`OracleSolid3D` and `OracleHall` restate Solid3D and Hall (crates/pantometry-thermal/src/solid.rs,
crates/pantometry-acoustic/src/hall.rs) on top of the C functions, with the same method names as
the GPU domains so one test body can drive either.
"""
from __future__ import annotations

import ctypes as C
import math
import sys
import os
import subprocess
from pathlib import Path
from typing import Callable, Optional, Tuple

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"

_D = C.c_double
_DP = C.POINTER(C.c_double)
_S = C.c_size_t
_U64 = C.c_uint64
_I = C.c_int
_IP = C.POINTER(C.c_int)
_U8P = C.POINTER(C.c_ubyte)
_SP = C.POINTER(C.c_size_t)

_SIG = {
    "orc_mix64": (_U64, [_U64]),
    "orc_rng_new": (_U64, [_U64]),
    "orc_rng_for_index": (_U64, [_U64, _U64]),
    "orc_rng_next_u64": (_U64, [C.POINTER(_U64)]),
    "orc_rng_unit": (_D, [C.POINTER(_U64)]),
    "orc_rng_range": (_D, [C.POINTER(_U64), _D, _D]),
    "orc_rng_gaussian": (_D, [C.POINTER(_U64)]),
    "orc_rng_split": (_U64, [C.POINTER(_U64)]),
    "orc_rng_digest": (_U64, [_U64, _I]),
    "orc_heat_capacity": (_D, [_D, _D, _D]),
    "orc_series": (_D, [_D, _D]),
    "orc_resolve": (_D, [_S, _S, _S, _D, _DP, _DP, _DP, _DP, _DP, _DP, _DP]),
    "orc_solid_step": (None, [_S, _S, _S, _DP, _DP, _DP, _DP, _D, _DP, _DP]),
    "orc_solid_step_omp": (None, [_S, _S, _S, _DP, _DP, _DP, _DP, _D, _DP, _DP, _I]),
    "orc_solid_bus_heat": (None, [_S, _DP, _D, _DP]),
    "orc_stored_heat": (_D, [_S, _DP, _DP, _D]),
    "orc_stored_heat_compensated": (_D, [_S, _DP, _DP, _D]),
    "orc_stored_heat_compensated_omp": (_D, [_S, _DP, _DP, _D, _I]),
    "orc_release_mode": (None, [_S, _S, _S, _I, _I, _I, _D, _D, _DP]),
    "orc_mode_amplitude": (_D, [_S, _S, _S, _I, _I, _I, _DP]),
    "orc_mode_amplification": (_D, [_S, _S, _S, _I, _I, _I, _D, _D, _D, _D, _D]),
    "orc_hall_invariant": (_D, [_S, _S, _S, _D, _D, _D, _DP, _DP, _DP, _DP, _DP]),
    "orc_hall_invariant_compensated_omp": (_D, [_S, _S, _S, _D, _D, _D, _DP, _DP, _DP, _DP, _DP,
                                                _I]),
    "orc_hall_sweep": (_D, [_S, _S, _S, _D, _D, _D, _D, _D, _DP, _DP, _DP, _DP, _DP, _S, _D, _D,
                            _I]),
    "orc_hall_mode": (None, [_S, _S, _S, _D, _I, _I, _I, _D, _DP]),
    "orc_workload_hot_spot": (None, [_S, _S, _S, _D, _D, _D, _D, _U64, _DP, _I]),
    "orc_workload_heterogeneous": (None, [_S, _S, _S, _U64, _U64, _I, _DP, _DP, _I]),
    "orc_workload_member": (None, [_U64, _U64, _D, _D, _D, _D, C.POINTER(C.c_longlong), _DP]),
    "orc_loss_from": (_D, [_D, _D, _D, _D, _D]),
    "orc_sample": (_D, [_S, _S, _S, _D, _DP, _D, _D, _D]),
    "orc_hall_sample": (_D, [_S, _S, _S, _D, _DP, _D, _D, _D]),
    "orc_room_invariant": (_D, [_S, _S, _D, _D, _D, _D, _DP, _DP, _DP, _DP]),
    "orc_resolve_void": (_D, [_S, _S, _S, _D, _DP, _DP, _U8P, _DP, _DP, _DP, _DP, _DP]),
    "orc_solid_bus_heat_void": (None, [_S, _DP, _U8P, _D, _DP]),
    "orc_find_gaps": (_S, [_S, _S, _S, _D, _U8P, _DP, _SP, _SP, _DP]),
    "orc_gap_exchange": (None, [_S, _SP, _SP, _DP, _DP, _DP, _D, _DP]),
    "orc_gap_rates": (_D, [_S, _SP, _SP, _DP, _DP, _DP, _D]),
    "orc_sample_void": (_D, [_S, _S, _S, _D, _DP, _U8P, _D, _D, _D]),
    "orc_stored_heat_phase": (_D, [_S, _DP, _DP, _D, _DP, _DP, _DP, _DP, _DP]),
    "orc_bus_heat_phase": (None, [_S, _DP, _U8P, _D, _DP, _DP, _DP, _DP, _DP, _DP]),
    "orc_solid_step_phase": (None, [_S, _S, _S, _D, _DP, _DP, _DP, _DP, _DP, _D, _DP, _DP, _IP,
                                    _DP, _DP, _DP, _DP, _DP, _DP, _U8P, _S, _SP, _SP, _DP, _DP,
                                    _DP, _DP, _DP, _DP, _DP, _DP]),
    "orc_deposit_phase": (None, [_S, _D, _DP, _DP, _DP, _DP, _DP, _DP, _DP]),
    "orc_bar_step": (None, [_S, _D, _D, _D, _DP, _DP, _DP]),
    "orc_bar_stored": (_D, [_S, _D, _DP, _D]),
    "orc_bar_sample": (_D, [_S, _D, _DP, _D]),
    "orc_tube_invariant": (_D, [_S, _D, _D, _D, _D, _DP, _DP, _DP]),
    "orc_tube_sweep": (_D, [_S, _D, _D, _D, _D, _D, _D, _I, _D, _I, _D, _DP, _DP, _DP]),
    "orc_series_with_half_cell": (_D, [_D, _D, _D, _D]),
    "orc_film_flux": (_D, [_S, _S, _S, _D, _S, _S, _S, _D, _IP, _DP, _DP, _DP, _D, _D]),
    "orc_loss_conductance": (_D, [_S, _S, _S, _D, _S, _S, _S, _D, _IP, _DP, _DP, _DP, _D, _D]),
    "orc_worst_rate_now": (_D, [_S, _S, _S, _D, _D, _DP, _DP, _DP, _IP, _DP, _DP, _DP, _DP,
                                _DP]),
    "orc_solid_step_ext": (None, [_S, _S, _S, _D, _DP, _DP, _DP, _DP, _DP, _D, _DP, _DP, _IP,
                                  _DP, _DP, _DP, _DP, _DP, _DP, _DP, _DP]),
}

_lib = None


def build() -> Path:
    """Compiles the oracle with its Makefile (gcc, -O2 -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", str(HERE), "-B"], check=True)
    return LIB


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        h = C.CDLL(str(LIB))
        for name, (res, args) in _SIG.items():
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _lib = h
    return _lib


def threads() -> int:
    n = os.cpu_count() or 1
    return max(1, min(16, n))


def _u8(a: np.ndarray):
    return a.ctypes.data_as(_U8P)


def _sp(a: np.ndarray):
    return a.ctypes.data_as(_SP)


def p(a: np.ndarray):
    return a.ctypes.data_as(_DP)


def _ip(a: np.ndarray):
    return a.ctypes.data_as(_IP)


def _thermal_or(value: float, conductivity: float, missing: float) -> float:
    """A thermal property, or what the reference uses when the substance has no thermal
    properties at all (`thermal.map_or(missing, ...)`; a NaN conductivity marks that here)."""
    return missing if math.isnan(conductivity) else float(value)


# ---- Rng ----------------------------------------------------------------------------------
class OracleRng:
    def __init__(self, seed: int):
        self.s = C.c_uint64(lib().orc_rng_new(seed))

    @staticmethod
    def for_index(seed: int, index: int) -> "OracleRng":
        r = OracleRng.__new__(OracleRng)
        r.s = C.c_uint64(lib().orc_rng_for_index(seed, index))
        return r

    def unit(self) -> float:
        return lib().orc_rng_unit(C.byref(self.s))

    def range(self, lo: float, hi: float) -> float:
        return lib().orc_rng_range(C.byref(self.s), lo, hi)

    def gaussian(self) -> float:
        return lib().orc_rng_gaussian(C.byref(self.s))

    def next_u64(self) -> int:
        return lib().orc_rng_next_u64(C.byref(self.s))


def rng_digest(seed: int, draws: int) -> int:
    return lib().orc_rng_digest(seed, draws)


# ---- Ledger / audit (conserved.rs) ---------------------------------------------------------
def ledger_total(entries) -> Tuple[float, float]:
    """Ledger::with over `entries` (conserved.rs:78-89): (total, scale)."""
    total, scale = 0.0, 0.0
    for e in entries:
        total += e
        scale = max(scale, abs(e))
    return total, scale


def audit_breaks(before: Tuple[float, float], after: Tuple[float, float], tol: float) -> bool:
    """audit_with for one quantity (conserved.rs:335-360)."""
    b, sb = before
    a, sa = after
    scale = max(abs(b), abs(a), sb, sa)
    if scale < 1e-300:
        return False
    return abs(a - b) / scale > tol


# ---- Solid3D -------------------------------------------------------------------------------
class OracleSolid3D:
    """Solid3D's plain branch on the CPU, as the reference computes it."""

    def __init__(self, name: str, density: float, conductivity: float, specific_heat: float,
                 counts: Tuple[int, int, int], dx: float, initial: float,
                 emissivity: float = 0.0, fusion=None):
        self.name_ = name
        self.counts_ = tuple(max(1, int(c)) for c in counts)
        nx, ny, nz = self.counts_
        self.n = nx * ny * nz
        self.dx = float(dx)
        self.reference = float(initial)
        self.cells = np.full(self.n, float(initial))
        self.saved = self.cells.copy()
        self.kcell = np.full(self.n, float(conductivity))
        cap = lib().orc_heat_capacity(density, specific_heat, self.dx)
        self.cap = np.full(self.n, cap)
        self.first = (density, conductivity, specific_heat)
        self.absorbed = 0.0
        self._old = np.empty(self.n)
        # The film's per-cell inputs are the cell's substance's emissivity and conductivity
        # (0 and +inf for a substance that does not say, solid.rs:1193-1195).
        self.emis = np.full(self.n, _thermal_or(emissivity, conductivity, 0.0))
        self.ksub = np.full(self.n, _thermal_or(conductivity, conductivity, math.inf))
        self.on = np.zeros(6, dtype=np.int32)      # exposed faces, Face order
        self.ambient = np.zeros(6)
        self.h = np.zeros(6)
        self.area = np.zeros(6)
        self.source = np.zeros(self.n)             # watts per cell
        self.lost = 0.0
        self.supplied = 0.0
        # Void cells (solid.rs:238, 992-1005) and the gaps between solid cells across them.
        self.voids = np.zeros(self.n, dtype=np.uint8)
        self.gap_a = np.zeros(0, dtype=np.uintp)
        self.gap_b = np.zeros(0, dtype=np.uintp)
        self.gap_coef = np.zeros(0)
        # Phase change (solid.rs:586-680): per-cell solid and liquid conductivity and capacity,
        # melting point and latent joules, and the melted fraction, once any substance melts.
        self.phase = False
        self.two_phase = False
        self._phase_of(np.ones(self.n, dtype=bool), density, conductivity, specific_heat, fusion)
        if self.phase:
            # The fresh build: a cell above its melting point starts liquid (solid.rs:652-665).
            self.melted = np.where(self.cells > self.point, 1.0, 0.0)
            self.saved_melted = self.melted.copy()
        self.resolve()

    def _phase_of(self, m, density, conductivity, specific_heat, fusion):
        """The per-cell phase tables of the cells `m` selects (resolve's props, solid.rs:599-646).
        `fusion` is (melting point K, latent J/kg, liquid (k, c_p) or None) or None."""
        thermal = not math.isnan(conductivity)
        volume = self.dx * self.dx * self.dx
        cs = (density * volume) * specific_heat
        if fusion is not None and not self.phase:
            self.phase = True
            self.k_solid = self.kcell.copy()
            self.k_liquid = self.kcell.copy()
            self.c_solid = self.cap.copy()
            self.c_liquid = self.cap.copy()
            self.point = np.full(self.n, math.inf)
            self.latent = np.zeros(self.n)
            self.melted = np.zeros(self.n)
            self.saved_melted = np.zeros(self.n)
        if fusion is not None and fusion[2] is not None and thermal:
            if fusion[2][0] != conductivity or fusion[2][1] != specific_heat:
                self.two_phase = True
        if not self.phase:
            return
        liquid = fusion[2] if fusion is not None else None
        self.k_solid[m] = conductivity if thermal else math.nan
        self.c_solid[m] = cs
        self.k_liquid[m] = liquid[0] if liquid is not None else self.k_solid[m]
        self.c_liquid[m] = (density * volume) * liquid[1] if liquid is not None else cs
        if fusion is not None and thermal:
            self.point[m] = fusion[0]
            self.latent[m] = (density * volume) * fusion[1]
        else:
            self.point[m] = math.inf
            self.latent[m] = 0.0

    def _phase_args(self):
        return (p(self.melted), p(self.point), p(self.latent), p(self.c_solid), p(self.c_liquid))

    def resolve(self) -> None:
        """Solid3D::resolve (solid.rs:588-784) and find_gaps (solid.rs:791-847)."""
        nx, ny, nz = self.counts_
        if self.phase:
            # Mixed by melt fraction, arithmetically (solid.rs:686-702).
            phi = self.melted
            self.kcell = self.k_solid + phi * (self.k_liquid - self.k_solid)
            self.cap = self.c_solid + phi * (self.c_liquid - self.c_solid)
        self.kcell[self.voids != 0] = 0.0
        self.cap[self.voids != 0] = 0.0
        self.kx = np.empty((nx + 1) * ny * nz)
        self.ky = np.empty(nx * (ny + 1) * nz)
        self.kz = np.empty(nx * ny * (nz + 1))
        self.mob = np.empty(self.n)
        self.face_sum = np.empty(self.n)
        self.worst_rate = lib().orc_resolve_void(nx, ny, nz, self.dx, p(self.kcell), p(self.cap),
                                                 _u8(self.voids), p(self.kx), p(self.ky),
                                                 p(self.kz), p(self.mob), p(self.face_sum))
        a = np.zeros(3 * self.n, dtype=np.uintp)
        b = np.zeros(3 * self.n, dtype=np.uintp)
        coef = np.zeros(3 * self.n)
        count = lib().orc_find_gaps(nx, ny, nz, self.dx, _u8(self.voids), p(self.emis), _sp(a),
                                    _sp(b), p(coef))
        self.gap_a, self.gap_b, self.gap_coef = a[:count].copy(), b[:count].copy(), coef[:count].copy()

    def empty(self, mask: np.ndarray):
        """Solid3D::empty, solid.rs:992-1005: the selected cells hold nothing from now on."""
        m = np.asarray(mask).reshape(-1).astype(bool)
        self.voids[m] = 1
        self.resolve()
        return self

    def is_void(self, i: int, j: int, k: int) -> bool:
        nx, ny, nz = self.counts_
        if i >= nx or j >= ny or k >= nz:
            return False
        return bool(self.voids[i + nx * (j + ny * k)])

    def void_cells(self) -> int:
        return int(np.count_nonzero(self.voids))

    def temperature_at(self, i: int, j: int, k: int) -> float:
        """solid.rs:1323-1335: clamped, NaN for a void cell."""
        nx, ny, nz = self.counts_
        c = min(i, nx - 1) + nx * (min(j, ny - 1) + ny * min(k, nz - 1))
        return math.nan if self.voids[c] else float(self.cells[c])

    def statistics(self) -> Tuple[float, float, float]:
        """(peak, mean, coldest) over the cells that are there (solid.rs:1377-1405)."""
        peak, coldest, total, count = -sys.float_info.max, sys.float_info.max, 0.0, 0
        for c in range(self.n):
            if self.voids[c]:
                continue
            t = float(self.cells[c])
            peak = max(peak, t) if not math.isnan(t) else peak
            coldest = min(coldest, t) if not math.isnan(t) else coldest
            total += t
            count += 1
        return peak, (total / count if count else math.nan), coldest

    def fill(self, density: float, conductivity: float, specific_heat: float, mask: np.ndarray,
             emissivity: float = 0.0, fusion=None):
        """Solid3D::fill, solid.rs:519-543: a void cell stays void; a melted fraction is carried."""
        m = np.asarray(mask).reshape(-1).astype(bool)
        self._phase_of(m, density, conductivity, specific_heat, fusion)
        self.kcell[m] = conductivity
        self.cap[m] = lib().orc_heat_capacity(density, specific_heat, self.dx)
        self.emis[m] = _thermal_or(emissivity, conductivity, 0.0)
        self.ksub[m] = _thermal_or(conductivity, conductivity, math.inf)
        self.resolve()
        return self

    def losing_from(self, face: int, ambient: float, convection: float, area: float):
        """Solid3D::losing_from, solid.rs:1038-1042: the face (0..5, XMin..ZMax) loses to an
        environment; a later call for the same face replaces it."""
        self.on[face] = 1
        self.ambient[face] = ambient
        self.h[face] = convection
        self.area[face] = area
        return self

    def dissipating(self, watts: float, mask: np.ndarray):
        """Solid3D::dissipating, solid.rs:1058-1083 (no void cells): `watts` spread evenly over the
        selected cells, replacing what they generated before."""
        m = np.asarray(mask).reshape(-1).astype(bool) & (self.voids == 0)
        chosen = int(m.sum())
        if chosen:
            self.source[m] = watts / float(chosen)
        return self

    def exposed(self) -> bool:
        return bool(self.on.any())

    def worst_rate_now(self) -> float:
        nx, ny, nz = self.counts_
        if len(self.gap_a) and not math.isnan(self.worst_rate):
            w = self._film_rate()
            return lib().orc_gap_rates(len(self.gap_a), _sp(self.gap_a), _sp(self.gap_b),
                                       p(self.gap_coef), p(self.cells), p(self.cap), w)
        return self._film_rate()

    def _film_rate(self) -> float:
        nx, ny, nz = self.counts_
        return lib().orc_worst_rate_now(nx, ny, nz, self.dx, self.worst_rate, p(self.mob),
                                        p(self.face_sum), p(self.cells), _ip(self.on),
                                        p(self.ambient), p(self.h), p(self.area), p(self.emis),
                                        p(self.ksub))

    def with_conductivity(self, k: np.ndarray):
        self.kcell = np.ascontiguousarray(k, dtype=np.float64).reshape(-1).copy()
        if self.phase:
            self.k_solid = self.kcell.copy()
            self.k_liquid = self.kcell.copy()
        self.resolve()
        return self

    def set_cells(self, cells: np.ndarray) -> None:
        """Every cell at once, as set_temperature would set it (solid.rs:1346-1361)."""
        self.cells = np.ascontiguousarray(cells, dtype=np.float64).reshape(-1).copy()
        if self.phase:
            latent = self.latent > 0.0
            self.melted[latent] = np.where(self.cells[latent] > self.point[latent], 1.0, 0.0)

    def max_stable_dt(self, now: float = 0.0) -> float:
        rate = self.worst_rate_now()
        if math.isnan(rate):
            return math.inf
        return math.inf if rate == 0.0 else 1.0 / rate

    def refuses(self, dt: float) -> Optional[str]:
        ratio = dt * self.worst_rate_now()
        if math.isnan(ratio):
            return "substance has no diffusivity"
        if ratio > 1.0 + 6e-12:
            return "Fourier number"
        return None

    def step(self, dt: float, gained: float = 0.0, threads_: int = 1) -> None:
        why = self.refuses(dt)
        if why:
            raise ValueError(why)
        nx, ny, nz = self.counts_
        if gained != 0.0:
            self.absorbed += gained
            if self.phase:
                lib().orc_bus_heat_phase(self.n, p(self.cap), _u8(self.voids), gained,
                                         p(self.cells), *self._phase_args())
            else:
                lib().orc_solid_bus_heat_void(self.n, p(self.cap), _u8(self.voids), gained,
                                              p(self.cells))
        gaps = len(self.gap_a)
        if self.phase:
            lost, supplied = C.c_double(self.lost), C.c_double(self.supplied)
            lib().orc_solid_step_phase(nx, ny, nz, self.dx, p(self.kx), p(self.ky), p(self.kz),
                                       p(self.mob), p(self.cap), dt, p(self.cells), p(self._old),
                                       _ip(self.on), p(self.ambient), p(self.h), p(self.area),
                                       p(self.emis), p(self.ksub), p(self.source), _u8(self.voids),
                                       gaps, _sp(self.gap_a), _sp(self.gap_b), p(self.gap_coef),
                                       *self._phase_args(), C.byref(lost), C.byref(supplied))
            self.lost, self.supplied = lost.value, supplied.value
            if self.two_phase:
                self.resolve()
        elif self.exposed() or self.source.any() or gaps:
            lost, supplied = C.c_double(self.lost), C.c_double(self.supplied)
            lib().orc_solid_step_ext(nx, ny, nz, self.dx, p(self.kx), p(self.ky), p(self.kz),
                                     p(self.mob), p(self.cap), dt, p(self.cells), p(self._old),
                                     _ip(self.on), p(self.ambient), p(self.h), p(self.area),
                                     p(self.emis), p(self.ksub), p(self.source), C.byref(lost),
                                     C.byref(supplied))
            self.lost, self.supplied = lost.value, supplied.value
            if gaps:
                lib().orc_gap_exchange(gaps, _sp(self.gap_a), _sp(self.gap_b), p(self.gap_coef),
                                       p(self._old), p(self.cap), dt, p(self.cells))
        elif threads_ > 1:
            lib().orc_solid_step_omp(nx, ny, nz, p(self.kx), p(self.ky), p(self.kz), p(self.mob),
                                     dt, p(self.cells), p(self._old), threads_)
        else:
            lib().orc_solid_step(nx, ny, nz, p(self.kx), p(self.ky), p(self.kz), p(self.mob), dt,
                                 p(self.cells), p(self._old))

    def stored_heat(self) -> float:
        if self.phase:
            return lib().orc_stored_heat_phase(self.n, p(self.cap), p(self.cells), self.reference,
                                               *self._phase_args())
        return lib().orc_stored_heat(self.n, p(self.cap), p(self.cells), self.reference)

    def set_temperature(self, i: int, j: int, k: int, t: float) -> None:
        """solid.rs:1346-1361: naming a temperature names the phase of a cell that melts."""
        nx, ny, nz = self.counts_
        if i >= nx or j >= ny or k >= nz:
            return
        c = i + nx * (j + ny * k)
        self.cells[c] = t
        if self.phase and self.latent[c] > 0.0:
            self.melted[c] = 1.0 if t > self.point[c] else 0.0

    def set_melted_fraction(self, i: int, j: int, k: int, fraction: float) -> None:
        """solid.rs:448-466."""
        nx, ny, nz = self.counts_
        if i >= nx or j >= ny or k >= nz or not self.phase:
            return
        c = i + nx * (j + ny * k)
        if self.latent[c] <= 0.0:
            return
        phi = min(max(fraction, 0.0), 1.0)
        self.melted[c] = phi
        point = self.point[c]
        if phi >= 1.0:
            self.cells[c] = max(self.cells[c], point)
        elif phi <= 0.0:
            self.cells[c] = min(self.cells[c], point)
        else:
            self.cells[c] = point

    def melted_fraction_at(self, i: int, j: int, k: int) -> float:
        nx, ny, nz = self.counts_
        if not self.phase:
            return 0.0
        return float(self.melted[min(i, nx - 1) + nx * (min(j, ny - 1) + ny * min(k, nz - 1))])

    def melted_volume(self) -> float:
        """solid.rs:476-478: the left fold of the fractions times the cell volume."""
        if not self.phase:
            return 0.0
        total = 0.0
        for v in self.melted:
            total += float(v)
        return total * (self.dx * self.dx * self.dx)

    def checkpoint(self) -> None:
        self.saved = self.cells.copy()
        if self.phase:
            self.saved_melted = self.melted.copy()
        self.saved_books = (self.lost, self.supplied)

    def restore(self) -> None:
        """solid.rs:1945-1950: the operator is not rebuilt."""
        self.cells = self.saved.copy()
        if self.phase:
            self.melted = self.saved_melted.copy()
        self.lost, self.supplied = getattr(self, "saved_books", (0.0, 0.0))

    def stored_heat_compensated(self, threads_: int = 1) -> float:
        if threads_ > 1:
            return lib().orc_stored_heat_compensated_omp(self.n, p(self.cap), p(self.cells),
                                                         self.reference, threads_)
        return lib().orc_stored_heat_compensated(self.n, p(self.cap), p(self.cells),
                                                 self.reference)

    def ledger(self) -> Tuple[float, float]:
        return ledger_total([self.stored_heat(), self.lost, -self.supplied])

    def release_mode(self, mode, mean: float, amplitude: float) -> None:
        nx, ny, nz = self.counts_
        lib().orc_release_mode(nx, ny, nz, mode[0], mode[1], mode[2], mean, amplitude,
                               p(self.cells))
        self.saved = self.cells.copy()

    def mode_amplitude(self, mode) -> float:
        nx, ny, nz = self.counts_
        return lib().orc_mode_amplitude(nx, ny, nz, mode[0], mode[1], mode[2], p(self.cells))

    def mode_amplification(self, mode, dt: float) -> float:
        nx, ny, nz = self.counts_
        d, k, c = self.first
        return lib().orc_mode_amplification(nx, ny, nz, mode[0], mode[1], mode[2], k, d, c,
                                            self.dx, dt)

    def deposit(self, i: int, j: int, k: int, joules: float) -> None:
        nx, ny, nz = self.counts_
        if i >= nx or j >= ny or k >= nz:
            return
        idx = i + nx * (j + ny * k)
        if self.phase:
            lib().orc_deposit_phase(idx, joules, p(self.cells), p(self.melted), p(self.cap),
                                    p(self.point), p(self.latent), p(self.c_solid),
                                    p(self.c_liquid))
        else:
            self.cells[idx] += joules / self.cap[idx]
        self.absorbed += joules

    def grid(self) -> np.ndarray:
        nx, ny, nz = self.counts_
        return self.cells.reshape(nz, ny, nx)

    def sample(self, x: float, y: float, z: float) -> float:
        """ScalarField::at (solid.rs:2045-2102)."""
        nx, ny, nz = self.counts_
        if self.voids.any():
            return lib().orc_sample_void(nx, ny, nz, self.dx, p(self.cells), _u8(self.voids), x,
                                         y, z)
        return lib().orc_sample(nx, ny, nz, self.dx, p(self.cells), x, y, z)


# ---- Hall ----------------------------------------------------------------------------------
class OracleHall:
    """Hall on the CPU, as the reference computes it (plus the point-source extension)."""

    def __init__(self, width: float, height: float, depth: float, across: int, speed: float,
                 density: float):
        nx = max(across, 2)
        dx = max(width / float(nx - 1), 2.2250738585072014e-308)

        def quantise(length: float) -> int:
            return int(max(_round_half_away(length / dx), 0.0)) + 1

        self.nx, self.ny, self.nz = nx, quantise(height), quantise(depth)
        self.dx, self.speed, self.density = dx, speed, density
        n = self.nx * self.ny * self.nz
        self.p = np.zeros(n)
        self.pp = np.zeros(n)
        self.vx = np.zeros(max((self.nx - 1) * self.ny * self.nz, 1))
        self.vy = np.zeros(max(self.nx * (self.ny - 1) * self.nz, 1))
        self.vz = np.zeros(max(self.nx * self.ny * (self.nz - 1), 1))
        self.staggered = False
        self.offset = 0.0
        self.datum = 0.0
        self.supplied = 0.0
        self.src = None
        self.increments = None
        self.cursor = 0

    def nodes(self):
        return self.nx, self.ny, self.nz

    def invariant(self) -> float:
        return lib().orc_hall_invariant(self.nx, self.ny, self.nz, self.dx, self.speed,
                                        self.density, p(self.p), p(self.pp), p(self.vx),
                                        p(self.vy), p(self.vz))

    def energy(self) -> float:
        return self.invariant() + self.offset

    def invariant_compensated(self, threads_: int = 1) -> float:
        """The invariant's terms, each as `invariant` computes it, summed with compensation:
        the exact sum the reference's left fold approximates (test infrastructure for the
        full-length ledger checks, not a reference function)."""
        return lib().orc_hall_invariant_compensated_omp(
            self.nx, self.ny, self.nz, self.dx, self.speed, self.density, p(self.p), p(self.pp),
            p(self.vx), p(self.vy), p(self.vz), threads_)

    def release(self, pressure: np.ndarray) -> None:
        self.p = np.ascontiguousarray(pressure, dtype=np.float64).reshape(-1).copy()
        self.pp = self.p.copy()
        self.vx[:] = 0.0
        self.vy[:] = 0.0
        self.vz[:] = 0.0
        self.staggered = False
        self.offset = 0.0
        self.supplied = 0.0
        self.cursor = 0
        self.datum = self.invariant()

    def mode(self, mode, amplitude: float) -> np.ndarray:
        out = np.empty(self.nx * self.ny * self.nz)
        lib().orc_hall_mode(self.nx, self.ny, self.nz, self.dx, mode[0], mode[1], mode[2],
                            amplitude, p(out))
        return out

    def set_source(self, node: int, increments: np.ndarray) -> None:
        self.src = node
        self.increments = np.asarray(increments, dtype=np.float64)
        self.cursor = 0

    def max_stable_dt(self) -> float:
        return self.dx / (self.speed * math.sqrt(3.0))

    def sample(self, x: float, y: float, z: float) -> float:
        """ScalarField::at (hall.rs:590-617)."""
        return lib().orc_hall_sample(self.nx, self.ny, self.nz, self.dx, p(self.p), x, y, z)

    def courant(self, dt: float) -> float:
        return self.speed * dt * math.sqrt(3.0) / self.dx

    def step(self, dt: float, threads_: int = 1) -> None:
        if self.courant(dt) > 1.0 + 1e-12:
            raise ValueError("Courant number")
        if dt <= 0.0:
            return
        vh = dt if self.staggered else 0.5 * dt
        starting = not self.staggered
        self.staggered = True
        src, inc, prev = (2**64 - 1), 0.0, 0.0
        if self.src is not None and self.cursor <= len(self.increments):
            src = self.src
            if self.cursor < len(self.increments):
                inc = float(self.increments[self.cursor])
            if self.cursor > 0:
                prev = float(self.increments[self.cursor - 1])
        self.cursor += 1
        added = lib().orc_hall_sweep(self.nx, self.ny, self.nz, self.dx, self.speed, self.density,
                                     dt, vh, p(self.p), p(self.pp), p(self.vx), p(self.vy),
                                     p(self.vz), src, inc, prev, threads_)
        self.supplied += added
        if starting:
            inv = self.invariant() - added
            off = self.datum - inv
            scale = max(abs(self.datum), abs(inv))
            if scale > 0.0 and abs(off) / scale > 0.25:
                raise ValueError("energy (starting the leapfrog)")
            self.offset = off

    def ledger(self) -> Tuple[float, float]:
        entries = [self.energy()]
        if self.supplied != 0.0:
            entries.append(-self.supplied)
        return ledger_total(entries)


class OracleRoom(OracleHall):
    """Room (crates/pantometry-acoustic/src/room.rs) on the CPU: the same staggered scheme as a
    hall one node deep (Room::new's node counts, room.rs:90-120), stepped under the
    two-dimensional Courant limit dx/(c*sqrt 2) (room.rs:336-338, 361-366), its energy the
    Room's own invariant (room.rs:243-281) for a room `depth` metres deep."""

    def __init__(self, width: float, height: float, cells_across: int, depth: float,
                 speed: float, density: float):
        super().__init__(width, height, 0.0, max(cells_across, 3), speed, density)
        nx = max(cells_across, 3)
        dx = width / float(nx - 1)
        ny = max(int(_round_half_away(height / dx)) + 1, 3)
        self.nx, self.ny, self.nz, self.dx = nx, ny, 1, dx
        self.depth_m = float(depth)
        self.p = np.zeros(nx * ny)
        self.pp = np.zeros(nx * ny)
        self.vx = np.zeros((nx - 1) * ny)
        self.vy = np.zeros(nx * (ny - 1))
        self.vz = np.zeros(1)

    def invariant(self) -> float:
        return lib().orc_room_invariant(self.nx, self.ny, self.dx, self.depth_m, self.speed,
                                        self.density, p(self.p), p(self.pp), p(self.vx),
                                        p(self.vy))

    def max_stable_dt(self) -> float:
        if self.speed <= 0.0:
            return math.inf
        return self.dx / (self.speed * math.sqrt(2.0))

    def courant(self, dt: float) -> float:
        return self.speed * dt * math.sqrt(2.0) / self.dx

    def mode2(self, nx_mode: int, ny_mode: int, amplitude: float) -> np.ndarray:
        """Room::released_in_mode's field (room.rs:166-182): ((a*px)*py)."""
        lx, ly = (self.nx - 1) * self.dx, (self.ny - 1) * self.dx
        out = np.empty(self.nx * self.ny)
        for j in range(self.ny):
            for i in range(self.nx):
                px = math.cos(float(nx_mode) * math.pi * (float(i) * self.dx) / lx) if lx > 0 else 1.0
                py = math.cos(float(ny_mode) * math.pi * (float(j) * self.dx) / ly) if ly > 0 else 1.0
                out[i + self.nx * j] = amplitude * px * py
        return out


class OracleBar1D:
    """Bar1D (crates/pantometry-thermal/src/lib.rs:425-660) on the CPU: cells of a uniform
    substance, insulated ends, the bus share into cell 0 and per-face joules into their cells."""

    def __init__(self, cells: int, dx: float, area: float, density: float, conductivity: float,
                 specific_heat: float, initial: float):
        n = max(cells, 2)
        self.n, self.dx, self.area = n, float(dx), float(area)
        self.cells = np.full(n, float(initial))
        self.saved = self.cells.copy()
        self.reference = float(initial)
        self.absorbed = 0.0
        self.alpha = conductivity / (density * specific_heat)
        self.capacity = (density * (self.area * self.dx)) * specific_heat
        self._prev = np.empty(n)

    def fourier_number(self, dt: float) -> float:
        """lib.rs:546-551."""
        return self.alpha * dt / (self.dx * self.dx)

    def max_stable_dt(self) -> float:
        """lib.rs:563-568."""
        return self.dx * self.dx / (2.0 * self.alpha)

    def step(self, dt: float, gained: float = 0.0, joules=None) -> None:
        f = self.fourier_number(dt)
        if not math.isfinite(f):
            raise ValueError("substance has no diffusivity")
        if f > 0.5 + 1e-12:
            raise ValueError("Fourier number")
        self.absorbed += gained
        j = None
        if joules is not None:
            j = np.ascontiguousarray(joules, dtype=np.float64)
        lib().orc_bar_step(self.n, f, self.capacity, gained, p(j) if j is not None else None,
                           p(self.cells), p(self._prev))
        if j is not None:
            total = 0.0
            for v in j:
                total += float(v)
            self.absorbed += total

    def stored(self) -> float:
        return lib().orc_bar_stored(self.n, self.capacity, p(self.cells), self.reference)

    def ledger(self) -> Tuple[float, float]:
        return ledger_total([self.stored()])

    def sample(self, x: float) -> float:
        return lib().orc_bar_sample(self.n, self.dx, p(self.cells), x)


class OracleTube:
    """Tube (crates/pantometry-acoustic/src/lib.rs:136-563) on the CPU.  Ends: 0 closed, 1 open,
    2 matched to an impedance."""

    def __init__(self, length: float, cells: int, area: float, density: float, speed: float,
                 left=(0, 0.0), right=(0, 0.0)):
        n = max(cells, 3)
        self.n = n
        self.dx = length / float(n - 1)
        self.area, self.density, self.speed = float(area), float(density), float(speed)
        self.left, self.right = left, right
        self.p = np.zeros(n)
        self.pp = np.zeros(n)
        self.u = np.zeros(n - 1)
        self.staggered = False
        self.offset = 0.0
        self.datum = 0.0
        self.radiated = 0.0

    def impedance(self) -> float:
        return self.density * self.speed

    def release(self, pressure) -> None:
        self.p = np.ascontiguousarray(pressure, dtype=np.float64).copy()
        self.pp = self.p.copy()
        self.u[:] = 0.0
        self.staggered = False
        self.offset = 0.0
        self.datum = self.invariant()

    def invariant(self) -> float:
        return lib().orc_tube_invariant(self.n, self.dx, self.area, self.speed, self.density,
                                        p(self.p), p(self.pp), p(self.u))

    def energy(self) -> float:
        return self.invariant() + self.offset

    def courant(self, dt: float) -> float:
        return self.speed * dt / self.dx

    def max_stable_dt(self) -> float:
        """lib.rs:486-514."""
        if self.speed <= 0.0:
            return math.inf
        wave = self.dx / self.speed
        rc2 = self.density * self.speed * self.speed

        def drain(end):
            kind, z = end
            if kind == 0:
                return math.inf
            if kind == 1:
                return self.dx * self.impedance() / (2.0 * rc2)
            zz = z if z > 0.0 else self.impedance()
            return self.dx * zz / (2.0 * rc2)

        return min(min(wave, drain(self.left)), drain(self.right))

    def step(self, dt: float) -> float:
        """Domain::step, lib.rs:516-533; returns the joules published this step."""
        if self.courant(dt) > 1.0 + 1e-12:
            raise ValueError("Courant number")
        if dt <= 0.0:
            return 0.0
        vh = dt if self.staggered else 0.5 * dt
        starting = not self.staggered
        self.staggered = True
        absorbed = lib().orc_tube_sweep(self.n, self.dx, self.area, self.speed, self.density, dt,
                                        vh, self.left[0], self.left[1], self.right[0],
                                        self.right[1], p(self.p), p(self.pp), p(self.u))
        published = 0.0
        if absorbed > 0.0:
            self.radiated += absorbed
            published = absorbed
        if starting:
            inv = self.invariant()
            off = self.datum - inv
            scale = max(abs(self.datum), abs(inv))
            if scale > 0.0 and abs(off) / scale <= 0.25:
                self.offset = off
        return published

    def ledger(self) -> Tuple[float, float]:
        return ledger_total([self.energy()])


def _round_half_away(x: float) -> float:
    """Rust's f64::round: halves away from zero (x - floor(x) is exact below 2**52)."""
    a = abs(x)
    f = math.floor(a)
    r = f + 1.0 if a - f >= 0.5 else f
    return math.copysign(r, x)


# ---- workloads -----------------------------------------------------------------------------
def hot_spot_cells(counts, amp=100.0, sigma=6.0, noise=0.1, seed=0x5EED0000, base=300.0):
    nx, ny, nz = counts
    out = np.empty(nx * ny * nz)
    lib().orc_workload_hot_spot(nx, ny, nz, base, amp, sigma, noise, seed, p(out), threads())
    return out


def heterogeneous_fields(counts, spots=8, k_seed=0x5EED0001, spot_seed=0x5EED0002):
    nx, ny, nz = counts
    k = np.empty(nx * ny * nz)
    t = np.empty(nx * ny * nz)
    lib().orc_workload_heterogeneous(nx, ny, nz, k_seed, spot_seed, spots, p(k), p(t), threads())
    return k, t


def member_shape(seed: int, i: int, lo_edge=16.0, hi_edge=96.0, lo_alpha=1e-5, hi_alpha=1e-3):
    edge = C.c_longlong()
    alpha = C.c_double()
    lib().orc_workload_member(seed, i, lo_edge, hi_edge, lo_alpha, hi_alpha, C.byref(edge),
                              C.byref(alpha))
    return edge.value, alpha.value
