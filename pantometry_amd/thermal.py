"""Solid3D on the GPU: the Domain of crates/pantometry-thermal/src/solid.rs, stepped by the
gfx950 library through its C ABI.

Temperatures are kelvin, lengths metres, energies joules (the reference's units are identity
newtypes over SI `f64`, crates/pantometry-units/src/lib.rs:95-122).
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass
from enum import IntEnum
from typing import Callable, List, Optional, Tuple, Union

import numpy as np

from . import _capi
from ._capi import check, dptr, lib
from .core import ENERGY, Exchange, HEAT, Kind, Ledger, Violation

STABLE_FOURIER_3D = 1.0 / 6.0  # solid.rs:84


def celsius(v: float) -> float:
    """Temperature::celsius, pantometry-units/src/lib.rs:729."""
    return v + 273.15


def mm(v: float) -> float:
    """Length::mm, pantometry-units/src/lib.rs:669."""
    return v * 1e-3


@dataclass(frozen=True)
class Substance:
    """The thermal part of Substance (crates/pantometry-core/src/substance.rs).  None for a
    conductivity or specific heat is a substance that does not say."""

    name: str
    density: float
    conductivity: Optional[float]
    specific_heat: Optional[float]
    emissivity: float = 0.0
    fusion: Optional["Fusion"] = None

    @staticmethod
    def aluminium_6061() -> "Substance":
        return Substance("Al 6061", 2.70 * 1e3, 167.0, 896.0, 0.09)  # substance.rs:362-382

    @staticmethod
    def borosilicate_crown() -> "Substance":
        return Substance("N-BK7", 2.51 * 1e3, 1.114, 858.0, 0.90)    # substance.rs:336-357

    @staticmethod
    def ice() -> "Substance":
        """substance.rs:613-650: one-phase ice, melting at 0 C with 333.55 kJ/kg."""
        return Substance("ice", 0.917 * 1e3, 2.22, 2050.0, 0.97,
                         Fusion(273.15, 333.55 * 1e3))

    @staticmethod
    def of(name: str, density: float, conductivity: float, specific_heat: float,
           emissivity: float = 0.0, fusion: Optional["Fusion"] = None) -> "Substance":
        return Substance(name, density, conductivity, specific_heat, emissivity, fusion)

    def with_fusion(self, fusion: Optional["Fusion"]) -> "Substance":
        return Substance(self.name, self.density, self.conductivity, self.specific_heat,
                         self.emissivity, fusion)

    def with_emissivity(self, emissivity: float) -> "Substance":
        """The same substance with another surface finish, clamped to 0..1 (substance.rs:398-403);
        a substance with no thermal properties is returned unchanged."""
        if self.conductivity is None or self.specific_heat is None:
            return self
        return Substance(self.name, self.density, self.conductivity, self.specific_heat,
                         min(max(emissivity, 0.0), 1.0), self.fusion)

    def diffusivity(self) -> Optional[float]:
        """k/(rho*c_p), substance.rs:211-216."""
        if self.conductivity is None or self.specific_heat is None:
            return None
        return self.conductivity / (self.density * self.specific_heat)

    def heat_capacity(self, volume: float) -> Optional[float]:
        """(rho*volume)*c_p, substance.rs:219-222."""
        if self.specific_heat is None:
            return None
        return (self.density * volume) * self.specific_heat

    def _c(self) -> _capi.pm_substance:
        f = self.fusion
        liquid = f.liquid if f is not None else None
        return _capi.pm_substance(
            self.density,
            math.nan if self.conductivity is None else self.conductivity,
            math.nan if self.specific_heat is None else self.specific_heat,
            self.emissivity,
            math.nan if f is None else f.melting_point,
            math.nan if f is None else f.latent_heat,
            math.nan if liquid is None else liquid[0],
            math.nan if liquid is None else liquid[1])


@dataclass(frozen=True)
class Fusion:
    """FusionProps (crates/pantometry-core/src/substance.rs:119-193): the melting point in kelvin,
    the latent heat in J/kg, and the liquid's (conductivity, specific heat) when it differs from
    the solid's (None: the one-phase model)."""

    melting_point: float
    latent_heat: float
    liquid: Optional[Tuple[float, float]] = None

    def with_liquid(self, conductivity: float, specific_heat: float) -> "Fusion":
        return Fusion(self.melting_point, self.latent_heat, (conductivity, specific_heat))


@dataclass(frozen=True)
class GapPatch:
    """One patch of a clearance (solid.rs:289-310): how many facing pairs, the bounding span of
    the facing cells across the gap (m), the separation of the two surfaces (m), and whether
    the facing cells fill their bounding box."""

    pairs: int
    span: Tuple[float, float]
    distance: float
    rectangular: bool

    def view_factor(self) -> float:
        """F12 of two equal, parallel, directly opposed rectangles (solid.rs:312-367), with the
        small-argument series below X = Y = 0.01 and the form's own limits at the ends."""
        if not math.isfinite(self.distance) or math.isnan(self.span[0]) or math.isnan(self.span[1]):
            return 0.0
        if self.distance <= 0.0:
            return 1.0 if self.span[0] > 0.0 and self.span[1] > 0.0 else 0.0
        x, y = self.span[0] / self.distance, self.span[1] / self.distance
        if x <= 0.0 or y <= 0.0:
            return 0.0
        x2, y2 = x * x, y * y
        if x < 0.01 and y < 0.01:
            return min(max(x * y / math.pi * (1.0 - (x2 + y2) / 3.0), 0.0), 1.0)
        rx, ry = math.sqrt(1.0 + x2), math.sqrt(1.0 + y2)
        t = (math.log(math.sqrt((1.0 + x2) * (1.0 + y2) / (1.0 + x2 + y2)))
             + x * ry * math.atan(x / ry) + y * rx * math.atan(y / rx)
             - x * math.atan(x) - y * math.atan(y))
        return min(max(2.0 / (math.pi * x * y) * t, 0.0), 1.0)


def gap_patches_of(pairs, counts: Tuple[int, int, int], dx: float) -> List[GapPatch]:
    """Solid3D::gap_patches (solid.rs:857-940): the gap pairs grouped by axis and separation,
    then into four-connected sheets of facing cells."""
    nx, ny, _ = counts
    strides = ((0, 1), (1, nx), (2, nx * ny))
    sheets: dict = {}
    for a, b, _coef in pairs:
        delta = b - a
        found = None
        for axis, stride in strides:
            limit_ok = delta < nx if axis == 0 else (delta < nx * ny if axis == 1 else True)
            if delta % stride == 0 and delta // stride > 1 and limit_ok:
                found = (axis, stride)
                break
        if found is None:
            continue
        axis, stride = found
        cells = delta // stride - 1
        i, j, k = a % nx, (a // nx) % ny, a // (nx * ny)
        across = (j, k) if axis == 0 else ((i, k) if axis == 1 else (i, j))
        sheets.setdefault((axis, cells), []).append(across)
    out = []
    for (axis, cells) in sorted(sheets):
        face = sorted(sheets[(axis, cells)])
        index = {p: n for n, p in enumerate(face)}
        seen = [False] * len(face)
        for start in range(len(face)):
            if seen[start]:
                continue
            stack, members = [start], []
            seen[start] = True
            while stack:
                n = stack.pop()
                u, v = face[n]
                members.append((u, v))
                for du, dv in ((1, 0), (-1, 0), (0, 1), (0, -1)):
                    m = index.get((u + du, v + dv))
                    if m is not None and not seen[m]:
                        seen[m] = True
                        stack.append(m)
            u0, u1 = min(m[0] for m in members), max(m[0] for m in members)
            v0, v1 = min(m[1] for m in members), max(m[1] for m in members)
            wu, wv = u1 + 1 - u0, v1 + 1 - v0
            out.append(GapPatch(len(members), (wu * dx, wv * dx), cells * dx,
                                len(members) == wu * wv))
    return out


class Face(IntEnum):
    """One outer face of a block, solid.rs:128-150, in the enum's order."""
    XMIN = 0
    XMAX = 1
    YMIN = 2
    YMAX = 3
    ZMIN = 4
    ZMAX = 5


@dataclass(frozen=True)
class Environment:
    """What a face loses heat to (crates/pantometry-thermal/src/lib.rs:38-49): the ambient in
    kelvin, the convective coefficient in W/m^2/K and the whole face's area in m^2."""

    ambient: float
    convection_w_per_m2_k: float
    area: float

    @staticmethod
    def still_air(ambient: float, area: float) -> "Environment":
        """A part in still room air, h = 7 W/m^2/K (lib.rs:52-59)."""
        return Environment(ambient, 7.0, area)

    def _c(self) -> _capi.pm_environment:
        return _capi.pm_environment(self.ambient, self.convection_w_per_m2_k, self.area)


def _phase(m: int, n: int, idx: np.ndarray) -> np.ndarray:
    """cos(m*PI*(i+0.5)/n) through libm, or 1 on a one-cell axis (solid.rs:1481-1487)."""
    if n <= 1:
        return np.ones(len(idx))
    return np.array([math.cos(float(m) * math.pi * (float(i) + 0.5) / float(n)) for i in idx])


class Solid3D:
    """A block of cubic cells conducting in three dimensions (solid.rs:173-287), on the GPU."""

    def __init__(self, name: str, substance: Substance, counts: Tuple[int, int, int], dx: float,
                 initial: float):
        self._name = name
        self._substances = [substance]
        self._dx = float(dx)
        self._initial = float(initial)
        handle = C.c_void_p()
        check(lib().pm_solid3d_create(name.encode(), C.byref(substance._c()), int(counts[0]),
                                      int(counts[1]), int(counts[2]), self._dx, self._initial,
                                      C.byref(handle)), "pm_solid3d_create")
        self._h = handle
        out = (C.c_size_t * 3)()
        check(lib().pm_solid3d_counts(self._h, out), "pm_solid3d_counts")
        self._counts = (out[0], out[1], out[2])

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().pm_solid3d_destroy(h)
            self._h = None

    # ---- construction ------------------------------------------------------------------
    def fill(self, substance: Substance,
             which: Union[Callable[[int, int, int], bool], np.ndarray]) -> "Solid3D":
        """Solid3D::fill, solid.rs:519-543."""
        nx, ny, nz = self._counts
        if callable(which):
            mask = np.zeros((nz, ny, nx), dtype=np.uint8)
            for k in range(nz):
                for j in range(ny):
                    for i in range(nx):
                        mask[k, j, i] = 1 if which(i, j, k) else 0
        else:
            mask = np.ascontiguousarray(np.asarray(which).reshape(nz, ny, nx), dtype=np.uint8)
        if substance not in self._substances:
            self._substances.append(substance)
        check(lib().pm_solid3d_fill(self._h, C.byref(substance._c()),
                                    mask.ctypes.data_as(C.POINTER(C.c_uint8))), "pm_solid3d_fill")
        return self

    def with_conductivity(self, conductivity: np.ndarray) -> "Solid3D":
        """Per-cell conductivity, the capacity unchanged (no reference counterpart; equal to
        one fill per distinct value)."""
        k = np.ascontiguousarray(conductivity, dtype=np.float64).reshape(-1)
        if k.size != self.cell_count():
            raise ValueError("one conductivity per cell")
        check(lib().pm_solid3d_set_conductivity(self._h, dptr(k)), "pm_solid3d_set_conductivity")
        return self

    def substances(self) -> int:
        return len(self._substances)

    # ---- state -------------------------------------------------------------------------
    def name(self) -> str:
        return self._name

    def counts(self) -> Tuple[int, int, int]:
        return self._counts

    def spacing(self) -> float:
        return self._dx

    def cell_count(self) -> int:
        nx, ny, nz = self._counts
        return nx * ny * nz

    def volume(self) -> float:
        nx, ny, nz = self._counts
        return float(nx * ny * nz) * (self._dx * self._dx * self._dx)

    def index(self, i: int, j: int, k: int) -> Optional[int]:
        nx, ny, nz = self._counts
        return i + nx * (j + ny * k) if (i < nx and j < ny and k < nz) else None

    def cells(self) -> np.ndarray:
        """Every temperature, x-fastest, shape (nz, ny, nx)."""
        nx, ny, nz = self._counts
        out = np.empty(nx * ny * nz, dtype=np.float64)
        check(lib().pm_solid3d_get_cells(self._h, dptr(out)), "pm_solid3d_get_cells")
        return out.reshape(nz, ny, nx)

    def set_cells(self, cells: np.ndarray) -> None:
        c = np.ascontiguousarray(cells, dtype=np.float64).reshape(-1)
        if c.size != self.cell_count():
            raise ValueError("one temperature per cell")
        check(lib().pm_solid3d_set_cells(self._h, dptr(c)), "pm_solid3d_set_cells")

    def copy_cells_from(self, other: "Solid3D") -> None:
        """This block's temperatures := `other`'s, device to device (same extents)."""
        check(lib().pm_solid3d_copy_cells(self._h, other._h), "pm_solid3d_copy_cells")

    def relative_l2(self, other: "Solid3D", reference: float) -> float:
        """||T - T_other|| / ||T_other - reference||, computed on the device."""
        out = C.c_double(0.0)
        check(lib().pm_solid3d_relative_l2(self._h, other._h, float(reference), C.byref(out)),
              "pm_solid3d_relative_l2")
        return out.value

    def l2_parts(self, other: "Solid3D", reference: float) -> Tuple[float, float]:
        """(||T - T_other||, ||T_other - reference||), computed on the device."""
        diff, norm = C.c_double(0.0), C.c_double(0.0)
        check(lib().pm_solid3d_l2_parts(self._h, other._h, float(reference), C.byref(diff),
                                        C.byref(norm)), "pm_solid3d_l2_parts")
        return diff.value, norm.value

    def set_temperature(self, i: int, j: int, k: int, t: float) -> None:
        check(lib().pm_solid3d_set_temperature(self._h, i, j, k, float(t)),
              "pm_solid3d_set_temperature")

    def deposit(self, i: int, j: int, k: int, joules: float) -> None:
        check(lib().pm_solid3d_deposit(self._h, i, j, k, float(joules)), "pm_solid3d_deposit")

    def temperature_at(self, i: int, j: int, k: int) -> float:
        """solid.rs:1323-1335: clamped to the block, NaN for a void cell."""
        nx, ny, nz = self._counts
        i, j, k = min(i, nx - 1), min(j, ny - 1), min(k, nz - 1)
        if self.is_void(i, j, k):
            return math.nan
        return float(self.cells()[k, j, i])

    def empty(self, which: Union[Callable[[int, int, int], bool], np.ndarray]) -> "Solid3D":
        """Solid3D::empty, solid.rs:992-1005: the selected cells hold nothing from now on; solid
        cells facing each other across them radiate (solid.rs:791-847, 1866-1878)."""
        mask = self._mask(which)
        check(lib().pm_solid3d_empty(self._h, mask.ctypes.data_as(C.POINTER(C.c_uint8))),
              "pm_solid3d_empty")
        return self

    def set_region_temperature(self, which, t: float) -> None:
        """set_temperature on every selected cell (solid.rs:1346-1361), phase included."""
        mask = self._mask(which)
        check(lib().pm_solid3d_set_region(self._h, mask.ctypes.data_as(C.POINTER(C.c_uint8)),
                                          float(t)), "pm_solid3d_set_region")

    def set_melted_fraction(self, i: int, j: int, k: int, fraction: float) -> None:
        """Solid3D::set_melted_fraction, solid.rs:448-466."""
        check(lib().pm_solid3d_set_melted_fraction(self._h, int(i), int(j), int(k), float(fraction)),
              "pm_solid3d_set_melted_fraction")

    def melted_fractions(self) -> np.ndarray:
        """Every cell's melted fraction, shape (nz, ny, nx)."""
        nx, ny, nz = self._counts
        out = np.empty(nx * ny * nz)
        check(lib().pm_solid3d_melted(self._h, dptr(out), None), "pm_solid3d_melted")
        return out.reshape(nz, ny, nx)

    def melted_fraction_at(self, i: int, j: int, k: int) -> float:
        """solid.rs:427-430, clamped."""
        nx, ny, nz = self._counts
        return float(self.melted_fractions()[min(k, nz - 1), min(j, ny - 1), min(i, nx - 1)])

    def melted_volume(self) -> float:
        """solid.rs:476-478, m^3."""
        v = C.c_double(0.0)
        check(lib().pm_solid3d_melted(self._h, None, C.byref(v)), "pm_solid3d_melted")
        return v.value

    def is_void(self, i: int, j: int, k: int) -> bool:
        return bool(lib().pm_solid3d_is_void(self._h, int(i), int(j), int(k)))

    def void_cells(self) -> int:
        return int(lib().pm_solid3d_void_cells(self._h))

    def gap_pairs(self):
        """The radiating pairs in the reference's order: (cell a, cell b, W/K^4)."""
        n = int(lib().pm_solid3d_gap_pairs(self._h, None, None, None, 0))
        a, b = np.empty(n, dtype=np.int64), np.empty(n, dtype=np.int64)
        coef = np.empty(n)
        lib().pm_solid3d_gap_pairs(self._h, a.ctypes.data_as(C.POINTER(C.c_int64)),
                                   b.ctypes.data_as(C.POINTER(C.c_int64)), dptr(coef), n)
        return list(zip(a.tolist(), b.tolist(), coef.tolist()))

    def gap_patches(self) -> List[GapPatch]:
        """Solid3D::gap_patches (solid.rs:857-940) over the block's radiating pairs."""
        return gap_patches_of(self.gap_pairs(), self._counts, self._dx)

    def _mask(self, which) -> np.ndarray:
        nx, ny, nz = self._counts
        if callable(which):
            mask = np.zeros((nz, ny, nx), dtype=np.uint8)
            for k in range(nz):
                for j in range(ny):
                    for i in range(nx):
                        mask[k, j, i] = 1 if which(i, j, k) else 0
            return mask.reshape(-1)
        return np.ascontiguousarray(np.asarray(which).reshape(-1), dtype=np.uint8)

    def release_mode(self, mode: Tuple[int, int, int], mean: float, amplitude: float) -> None:
        """Solid3D::release_mode, solid.rs:1478-1500: mean + ((amp*px)*py)*pz per cell."""
        nx, ny, nz = self._counts
        px = _phase(mode[0], nx, range(nx))
        py = _phase(mode[1], ny, range(ny))
        pz = _phase(mode[2], nz, range(nz))
        a = amplitude * px
        cells = mean + (a[None, None, :] * py[None, :, None]) * pz[:, None, None]
        self.set_cells(cells)
        # `self.saved.clone_from(&self.cells)`: a restore returns to the released mode.
        self.checkpoint()

    def mode_amplitude(self, mode: Tuple[int, int, int]) -> float:
        """Solid3D::mode_amplitude, solid.rs:1507-1535 (left fold in index order)."""
        nx, ny, nz = self._counts
        px = _phase(mode[0], nx, range(nx))
        py = _phase(mode[1], ny, range(ny))
        pz = _phase(mode[2], nz, range(nz))
        norm = 1.0
        for m, n in zip(mode, self._counts):
            norm = norm * (1.0 if (n <= 1 or m == 0) else 0.5)
        terms = ((self.cells() * px[None, None, :]) * py[None, :, None]) * pz[:, None, None]
        total = float(np.cumsum(terms.reshape(-1))[-1])
        return total / (float(nx * ny * nz) * norm)

    def fourier_number(self, dt: float) -> float:
        """alpha*dt/dx^2 for the constructor's substance, solid.rs:1423-1428."""
        alpha = self._substances[0].diffusivity()
        if alpha is None:
            return math.inf
        return alpha * dt / (self._dx * self._dx)

    def mode_amplification(self, mode: Tuple[int, int, int], dt: float) -> float:
        """1 + F * eigenvalue * dx^2, solid.rs:1455-1474."""
        f = self.fourier_number(dt)
        total = 0.0
        terms = []
        for m, n in zip(mode, self._counts):
            if n <= 1:
                terms.append(0.0)
            else:
                s = math.sin(float(m) * math.pi / (2.0 * float(n)))
                terms.append(-4.0 * s * s)
        total = terms[0] + terms[1] + terms[2]
        return 1.0 + f * total

    def _statistics(self) -> Tuple[float, float, float]:
        out = (C.c_double * 3)()
        check(lib().pm_solid3d_statistics(self._h, out), "pm_solid3d_statistics")
        return out[0], out[1], out[2]

    def mean_temperature(self) -> float:
        """solid.rs:1377-1384: over the cells that are there, NaN when none is."""
        return self._statistics()[1]

    def peak_temperature(self) -> float:
        return self._statistics()[0]

    def coldest_temperature(self) -> float:
        return self._statistics()[2]

    def heat_capacity(self) -> float:
        return float(lib().pm_solid3d_heat_capacity(self._h))

    def absorbed_energy(self) -> float:
        out = (C.c_double * 4)()
        check(lib().pm_solid3d_readings(self._h, out), "pm_solid3d_readings")
        return out[3]

    # ---- Domain ------------------------------------------------------------------------
    def kind(self) -> Kind:
        return Kind.EVOLVING

    def books_balance(self) -> bool:
        return True

    def supports_restore(self) -> bool:
        return True

    def residual(self) -> float:
        return 0.0

    def max_stable_dt(self, now: float = 0.0) -> float:
        return float(lib().pm_solid3d_max_stable_dt(self._h))

    def stability_ratio(self, dt: float) -> float:
        return float(lib().pm_solid3d_stability_ratio(self._h, dt))

    def step(self, t: float, dt: float, bus: Exchange) -> None:
        """Domain::step, solid.rs:1717-1901: refuse first, then take the bus share, then sweep."""
        v = _capi.pm_violation()
        check(lib().pm_solid3d_refuses(self._h, float(dt), C.byref(v)), "pm_solid3d_refuses", v)
        gained = bus.take_share(HEAT, dt)
        check(lib().pm_solid3d_step(self._h, float(dt), float(gained), C.byref(v)),
              "pm_solid3d_step", v)

    def losing_from(self, face: Face, environment: Environment) -> "Solid3D":
        """Solid3D::losing_from (solid.rs:1038-1042): the face loses heat to the environment by
        convection and radiation, inside the sweep; a later call for the face replaces it."""
        env = environment._c()
        check(lib().pm_solid3d_losing_from(self._h, int(face), C.byref(env)),
              "pm_solid3d_losing_from")
        return self

    def dissipating(self, watts: float, where) -> "Solid3D":
        """Solid3D::dissipating (solid.rs:1058-1083): `watts` in total spread evenly over the cells
        `where` selects (a boolean array over the cells, or a function of (i, j, k)), replacing
        what they generated before."""
        nx, ny, nz = self._counts
        if callable(where):
            k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
            mask = np.vectorize(where)(i, j, k)
        else:
            mask = np.asarray(where)
        m = np.ascontiguousarray(mask.reshape(-1), dtype=np.uint8)
        if m.size != nx * ny * nz:
            raise ValueError("dissipating: the selection must cover every cell")
        check(lib().pm_solid3d_dissipating(self._h, float(watts),
                                           m.ctypes.data_as(C.POINTER(C.c_uint8))),
              "pm_solid3d_dissipating")
        return self

    def sample(self, points) -> np.ndarray:
        """ScalarField::at (solid.rs:2045-2102) at each point (x, y, z in metres from the block's
        corner): trilinear between cell centres, kelvin."""
        pts = np.ascontiguousarray(np.asarray(points, dtype=np.float64).reshape(-1, 3))
        out = np.empty(len(pts))
        check(lib().pm_solid3d_sample(self._h, dptr(pts), len(pts), dptr(out)),
              "pm_solid3d_sample")
        return out

    def generated_power(self) -> float:
        out = C.c_double(0.0)
        check(lib().pm_solid3d_generated_power(self._h, C.byref(out)),
              "pm_solid3d_generated_power")
        return out.value

    def generated_energy(self) -> float:
        return self.books()[2]

    def lost_energy(self) -> float:
        return self.books()[1]

    def books(self) -> Tuple[float, float, float]:
        b = _capi.pm_books()
        check(lib().pm_solid3d_ledger(self._h, C.byref(b)), "pm_solid3d_ledger")
        return b.stored, b.lost, b.supplied

    def ledger(self) -> Ledger:
        stored, lost, supplied = self.books()
        return Ledger().with_(ENERGY, stored).with_(ENERGY, lost).with_(ENERGY, -supplied)

    def checkpoint(self) -> None:
        check(lib().pm_solid3d_checkpoint(self._h), "pm_solid3d_checkpoint")

    def restore(self) -> None:
        check(lib().pm_solid3d_restore(self._h), "pm_solid3d_restore")

    def readings(self):
        out = (C.c_double * 4)()
        check(lib().pm_solid3d_readings(self._h, out), "pm_solid3d_readings")
        # Solid3D::readings, solid.rs:1962-2002: "generated" only for a block with a source.
        out_ = [(self._name, "peak", out[0], "C"), (self._name, "mean", out[1], "C"),
                (self._name, "coldest", out[2], "C"), (self._name, "absorbed", out[3], "J")]
        if self.generated_power() != 0.0:
            out_.append((self._name, "generated", self.generated_energy(), "J"))
        # "melted", only for a block that can melt (solid.rs:2003-2010).
        if lib().pm_solid3d_melting_cells(self._h) > 0:
            out_.append((self._name, "melted", self.melted_volume() * 1e9, "mm3"))
        return out_

    # ---- device-resident loop ----------------------------------------------------------
    def advance(self, dt: float, steps: int, tolerance: float = 1e-9, want_totals: bool = True):
        """`steps` steps of `dt` with the audit after each, on the device.  Returns
        (ledger totals after each step, device milliseconds).  Raises Violation on a refused
        step or an audit breach (the state is left as the breaking step made it)."""
        v = _capi.pm_violation()
        totals = np.empty(max(steps, 1), dtype=np.float64) if want_totals else None
        ms = C.c_double(0.0)
        check(lib().pm_solid3d_advance(self._h, float(dt), int(steps), float(tolerance),
                                       dptr(totals) if want_totals else None, C.byref(ms),
                                       C.byref(v)), "pm_solid3d_advance", v)
        return (totals[:steps] if want_totals else None), ms.value

    def set_tile(self, tile_y: int = 0, z_chunk: int = 0, prefetch: int = 0) -> None:
        check(lib().pm_solid3d_set_tile(self._h, tile_y, z_chunk, prefetch), "pm_solid3d_set_tile")

    def set_profiling(self, on: bool) -> None:
        check(lib().pm_solid3d_set_profiling(self._h, 1 if on else 0), "pm_solid3d_set_profiling")

    def kernel_times(self) -> Tuple[float, float, int]:
        """Mean (sweep, audit) device milliseconds per launch over the last profiled advance,
        and the steps one sweep launch advanced."""
        sw, au, k = C.c_double(), C.c_double(), C.c_int()
        check(lib().pm_solid3d_kernel_times(self._h, C.byref(sw), C.byref(au), C.byref(k)),
              "pm_solid3d_kernel_times")
        return sw.value, au.value, k.value

    def set_steps_per_pass(self, steps: int) -> None:
        """1, or 2 (default): advance runs pairs of steps in one pass where the form allows."""
        check(lib().pm_solid3d_set_steps_per_pass(self._h, int(steps)),
              "pm_solid3d_set_steps_per_pass")

    def steps_per_pass(self) -> int:
        return int(lib().pm_solid3d_steps_per_pass(self._h))


class Bar1D:
    """Bar1D (crates/pantometry-thermal/src/lib.rs:425-660): explicit one-dimensional conduction
    with insulated ends, on the GPU.  Temperatures in kelvin at cell centres (i + 1/2) dx."""

    def __init__(self, name: str, substance: Substance, cells: int, dx: float, area: float,
                 initial: float):
        self._name = name
        self._substance = substance
        self._dx = float(dx)
        self._area = float(area)
        handle = C.c_void_p()
        sub = substance._c()
        check(lib().pm_bar1d_create(name.encode(), C.byref(sub), int(cells), self._dx,
                                    self._area, float(initial), C.byref(handle)),
              "pm_bar1d_create")
        self._h = handle
        self._n = int(lib().pm_bar1d_cells(self._h))
        self._face_area: Optional[float] = None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().pm_bar1d_destroy(h)
            self._h = None

    def exposing(self, face_area: float) -> "Bar1D":
        """Bar1D::exposing (lib.rs:483-486): each cell has a face of `face_area` that heat can be
        delivered onto, one value per cell, by `step_with_flux`."""
        self._face_area = float(face_area)
        return self

    # ---- state ---------------------------------------------------------------------------
    def name(self) -> str:
        return self._name

    def cell_count(self) -> int:
        return self._n

    def cells(self) -> np.ndarray:
        out = np.empty(self._n)
        check(lib().pm_bar1d_get_cells(self._h, dptr(out)), "pm_bar1d_get_cells")
        return out

    def set_cells(self, cells) -> None:
        c = np.ascontiguousarray(cells, dtype=np.float64).reshape(-1)
        if c.size != self._n:
            raise ValueError(f"{c.size} temperatures for {self._n} cells")
        check(lib().pm_bar1d_set_cells(self._h, dptr(c)), "pm_bar1d_set_cells")

    def temperature_at(self, index: int) -> float:
        """lib.rs:494-496: clamped to the ends."""
        return float(self.cells()[min(index, self._n - 1)])

    def _readings(self) -> Tuple[float, float, float]:
        out = (C.c_double * 3)()
        check(lib().pm_bar1d_readings(self._h, out), "pm_bar1d_readings")
        return out[0], out[1], out[2]

    def mean_temperature(self) -> float:
        """lib.rs:504-506: the left fold over the cells, over their count."""
        return self._readings()[0]

    def end_to_end(self) -> float:
        c = self.cells()
        return float(c[-1] - c[0])

    def absorbed_energy(self) -> float:
        absorbed = C.c_double()
        check(lib().pm_bar1d_ledger(self._h, None, C.byref(absorbed)), "pm_bar1d_ledger")
        return absorbed.value

    def fourier_number(self, dt: float) -> float:
        return float(lib().pm_bar1d_fourier_number(self._h, float(dt)))

    # ---- Domain ----------------------------------------------------------------------------
    def kind(self) -> Kind:
        return Kind.EVOLVING

    def books_balance(self) -> bool:
        return True

    def supports_restore(self) -> bool:
        return True

    def residual(self) -> float:
        return 0.0

    def max_stable_dt(self, now: float = 0.0) -> float:
        return float(lib().pm_bar1d_max_stable_dt(self._h))

    def step(self, t: float, dt: float, bus: Exchange, joules=None) -> None:
        """Domain::step, lib.rs:571-627: refuse first, then take the bus share into cell 0 and
        any per-cell joules into their cells, then conduct."""
        v = _capi.pm_violation()
        check(lib().pm_bar1d_refuses(self._h, float(dt), C.byref(v)), "pm_bar1d_refuses", v)
        gained = bus.take_share(HEAT, dt)
        j = None
        if joules is not None:
            j = np.ascontiguousarray(joules, dtype=np.float64).reshape(-1)
            if j.size != self._n:
                raise Violation.at(self._name, "flux on a different grid", float(j.size))
        check(lib().pm_bar1d_step(self._h, float(dt), float(gained),
                                  dptr(j) if j is not None else None, C.byref(v)),
              "pm_bar1d_step", v)

    def stored_heat(self) -> float:
        stored = C.c_double()
        check(lib().pm_bar1d_ledger(self._h, C.byref(stored), None), "pm_bar1d_ledger")
        return stored.value

    def ledger(self) -> Ledger:
        """lib.rs:629-631: the heat stored since the start."""
        return Ledger().with_(ENERGY, self.stored_heat())

    def checkpoint(self) -> None:
        check(lib().pm_bar1d_checkpoint(self._h), "pm_bar1d_checkpoint")

    def restore(self) -> None:
        check(lib().pm_bar1d_restore(self._h), "pm_bar1d_restore")

    def readings(self):
        """lib.rs:644-660: mean and peak in celsius, and what it has absorbed."""
        mean, peak, absorbed = self._readings()
        return [(self._name, "mean", mean - 273.15, "C"),
                (self._name, "peak", peak - 273.15, "C"),
                (self._name, "absorbed", absorbed, "J")]

    def sample(self, xs) -> np.ndarray:
        """ScalarField::at (lib.rs:736-751) at each x in metres (or each (x, y, z) point, y and z
        ignored): linear between cell centres, held past the ends."""
        pts = np.asarray(xs, dtype=np.float64)
        x = np.ascontiguousarray(pts[:, 0] if pts.ndim == 2 else pts.reshape(-1))
        out = np.empty(len(x))
        check(lib().pm_bar1d_sample(self._h, dptr(x), len(x), dptr(out)), "pm_bar1d_sample")
        return out

    def advance(self, dt: float, steps: int, tolerance: float = 1e-9):
        """`steps` steps with no bus input, each audited on the device.  Returns (stored heat
        after each step, device milliseconds)."""
        v = _capi.pm_violation()
        totals = np.empty(max(steps, 1))
        ms = C.c_double(0.0)
        check(lib().pm_bar1d_advance(self._h, float(dt), int(steps), float(tolerance), dptr(totals),
                                     C.byref(ms), C.byref(v)), "pm_bar1d_advance", v)
        return totals[:steps], ms.value
