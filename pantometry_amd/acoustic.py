"""Hall and Room on the GPU: the Domains of crates/pantometry-acoustic/src/hall.rs and room.rs,
stepped by the gfx950 library through its C ABI.  Pressures in pascal, lengths in metres, energies in joules."""
from __future__ import annotations

import ctypes as C
import math
from typing import Callable, List, Optional, Tuple, Union

import numpy as np

from . import _capi
from ._capi import check, dptr, lib
from .core import ENERGY, Exchange, Kind, Ledger


class Hall:
    """A rectangular room with rigid walls on a node-centred staggered grid (hall.rs:65-90)."""

    def __init__(self, name: str, width: float, height: float, depth: float, across: int,
                 speed: float, density: float):
        self._name = name
        self._speed = float(speed)
        self._density = float(density)
        handle = C.c_void_p()
        check(lib().pm_hall_create(name.encode(), float(width), float(height), float(depth),
                                   int(across), self._speed, self._density, C.byref(handle)),
              "pm_hall_create")
        self._h = handle
        out = (C.c_size_t * 3)()
        check(lib().pm_hall_nodes(self._h, out), "pm_hall_nodes")
        self._nodes = (out[0], out[1], out[2])
        self._dx = float(lib().pm_hall_spacing(self._h))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().pm_hall_destroy(h)
            self._h = None

    @staticmethod
    def of_air(name: str, width: float, height: float, depth: float, across: int) -> "Hall":
        """343 m/s, 1.204 kg/m^3 (hall.rs:137-153)."""
        return Hall(name, width, height, depth, across, 343.0, 1.204)

    # ---- geometry ----------------------------------------------------------------------
    def name(self) -> str:
        return self._name

    def nodes(self) -> Tuple[int, int, int]:
        return self._nodes

    def spacing(self) -> float:
        return self._dx

    def width(self) -> float:
        return float(self._nodes[0] - 1) * self._dx

    def height(self) -> float:
        return float(self._nodes[1] - 1) * self._dx

    def depth(self) -> float:
        return float(self._nodes[2] - 1) * self._dx

    def node_index(self, i: int, j: int, k: int) -> int:
        nx, ny, _ = self._nodes
        return i + nx * (j + ny * k)

    # ---- release -----------------------------------------------------------------------
    def released_from(self, profile: Union[Callable[[float, float, float], float], np.ndarray]) -> "Hall":
        """Hall::released_from, hall.rs:166-184; the profile sees node positions (i, j, k)*dx."""
        nx, ny, nz = self._nodes
        if callable(profile):
            p = np.empty((nz, ny, nx), dtype=np.float64)
            for k in range(nz):
                for j in range(ny):
                    for i in range(nx):
                        p[k, j, i] = profile(float(i) * self._dx, float(j) * self._dx,
                                             float(k) * self._dx)
        else:
            p = np.ascontiguousarray(profile, dtype=np.float64).reshape(nz, ny, nx)
        p = np.ascontiguousarray(p.reshape(-1))
        check(lib().pm_hall_release(self._h, dptr(p)), "pm_hall_release")
        return self

    def mode_profile(self, mode: Tuple[int, int, int], amplitude: float) -> np.ndarray:
        """The pressure of hall.rs:190-210 at every node: ((a*px)*py)*pz."""
        nx, ny, nz = self._nodes
        lx, ly, lz = self.width(), self.height(), self.depth()

        def phase(n: int, count: int, length: float) -> np.ndarray:
            if length > 0.0:
                return np.array([math.cos(float(n) * math.pi * (float(i) * self._dx) / length)
                                 for i in range(count)])
            return np.ones(count)

        px, py, pz = phase(mode[0], nx, lx), phase(mode[1], ny, ly), phase(mode[2], nz, lz)
        a = amplitude * px
        return (a[None, None, :] * py[None, :, None]) * pz[:, None, None]

    def released_in_mode(self, mode: Tuple[int, int, int], amplitude: float) -> "Hall":
        return self.released_from(self.mode_profile(mode, amplitude))

    def set_source(self, node: int, increments: Optional[np.ndarray]) -> None:
        """A point source: increments[n] Pa added at `node` after step n's pressure sweep, its
        energy booked as supplied (no reference counterpart)."""
        if increments is None or len(increments) == 0:
            check(lib().pm_hall_set_source(self._h, 0, None, 0), "pm_hall_set_source")
            return
        inc = np.ascontiguousarray(increments, dtype=np.float64)
        check(lib().pm_hall_set_source(self._h, int(node), dptr(inc), inc.size),
              "pm_hall_set_source")

    # ---- closed forms ------------------------------------------------------------------
    def mode_frequency(self, mode: Tuple[int, int, int]) -> float:
        """hall.rs:266-279."""
        total = 0.0
        for n, length in zip(mode, (self.width(), self.height(), self.depth())):
            total += (float(n) / length) ** 2 if length > 0.0 else 0.0
        return self._speed / 2.0 * math.sqrt(total)

    def modes_below(self, limit: float) -> List[Tuple[Tuple[int, int, int], float]]:
        """hall.rs:286-316."""
        lengths = (self.width(), self.height(), self.depth())

        def bound(length: float) -> int:
            if length <= 0.0 or self._speed <= 0.0:
                return 0
            return int(math.floor(2.0 * limit * length / self._speed))

        mx, my, mz = (bound(length) for length in lengths)
        found = []
        for a in range(mx + 1):
            for b in range(my + 1):
                for c in range(mz + 1):
                    if a == b == c == 0:
                        continue
                    f = self.mode_frequency((a, b, c))
                    if f <= limit:
                        found.append(((a, b, c), f))
        found.sort(key=lambda x: x[1])
        return found

    def courant(self, dt: float) -> float:
        return float(lib().pm_hall_courant(self._h, float(dt)))

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
        return float(lib().pm_hall_max_stable_dt(self._h))

    def step(self, t: float, dt: float, bus: Exchange) -> None:
        v = _capi.pm_violation()
        check(lib().pm_hall_step(self._h, float(dt), C.byref(v)), "pm_hall_step", v)

    def energy_parts(self) -> Tuple[float, float, float]:
        inv, off, sup = C.c_double(), C.c_double(), C.c_double()
        check(lib().pm_hall_energy(self._h, C.byref(inv), C.byref(off), C.byref(sup)),
              "pm_hall_energy")
        return inv.value, off.value, sup.value

    def energy(self) -> float:
        """Hall::energy, hall.rs:249-251: invariant + startup offset."""
        inv, off, _ = self.energy_parts()
        return inv + off

    def startup_adjustment(self) -> float:
        return self.energy_parts()[1]

    def supplied(self) -> float:
        return self.energy_parts()[2]

    def ledger(self) -> Ledger:
        inv, off, sup = self.energy_parts()
        led = Ledger().with_(ENERGY, inv + off)
        if sup != 0.0:
            led.add(ENERGY, -sup)
        return led

    def checkpoint(self) -> None:
        check(lib().pm_hall_checkpoint(self._h), "pm_hall_checkpoint")

    def restore(self) -> None:
        check(lib().pm_hall_restore(self._h), "pm_hall_restore")

    def fields(self):
        """(p, pressure_prev, vx, vy, vz) as flat x-fastest arrays."""
        nx, ny, nz = self._nodes
        p = np.empty(nx * ny * nz)
        pp = np.empty(nx * ny * nz)
        vx = np.empty(max((nx - 1) * ny * nz, 1))
        vy = np.empty(max(nx * (ny - 1) * nz, 1))
        vz = np.empty(max(nx * ny * (nz - 1), 1))
        check(lib().pm_hall_get_fields(self._h, dptr(p), dptr(pp), dptr(vx), dptr(vy), dptr(vz)),
              "pm_hall_get_fields")
        return (p, pp, vx[:(nx - 1) * ny * nz], vy[:nx * (ny - 1) * nz], vz[:nx * ny * (nz - 1)])

    def pressure_at(self, i: int, j: int, k: int) -> float:
        nx, ny, nz = self._nodes
        p = self.fields()[0]
        return float(p[self.node_index(min(i, nx - 1), min(j, ny - 1), min(k, nz - 1))])

    def set_tile(self, tile_y: int = 0, z_chunk: int = 0) -> None:
        """Rows per 64-wide column tile (8 or 16) and z-planes per workgroup (0: automatic)."""
        check(lib().pm_hall_set_tile(self._h, int(tile_y), int(z_chunk)), "pm_hall_set_tile")

    def set_tile_width(self, tile_x: int = 0) -> None:
        """Width of the paired step's column tiles: 64, 56 or 48 (0: the default, 64)."""
        check(lib().pm_hall_set_tile_width(self._h, int(tile_x)), "pm_hall_set_tile_width")

    def set_steps_per_pass(self, steps: int) -> None:
        """1, or 2 (default): advance runs pairs of leapfrog steps in one pass."""
        check(lib().pm_hall_set_steps_per_pass(self._h, int(steps)), "pm_hall_set_steps_per_pass")

    def peak_pressure(self) -> float:
        return float(lib().pm_hall_peak_pressure(self._h))

    def sample(self, points) -> np.ndarray:
        """ScalarField::at (hall.rs:590-617) at each point (x, y, z in metres from the corner
        node): trilinear between nodes, pascals."""
        pts = np.ascontiguousarray(np.asarray(points, dtype=np.float64).reshape(-1, 3))
        out = np.empty(len(pts))
        check(lib().pm_hall_sample(self._h, dptr(pts), len(pts), dptr(out)), "pm_hall_sample")
        return out

    def readings(self):
        return [(self._name, "peak", self.peak_pressure(), "Pa")]

    def advance(self, dt: float, steps: int, tolerance: float = 1e-9):
        """`steps` steps with the energy audit after each, on the device.  Returns (ledger
        totals after each step, device milliseconds)."""
        v = _capi.pm_violation()
        totals = np.empty(max(steps, 1))
        ms = C.c_double(0.0)
        check(lib().pm_hall_advance(self._h, float(dt), int(steps), float(tolerance), dptr(totals),
                                    C.byref(ms), C.byref(v)), "pm_hall_advance", v)
        return totals[:steps], ms.value


class Room(Hall):
    """Room (crates/pantometry-acoustic/src/room.rs): a two-dimensional rigid-walled room, `depth`
    metres deep, on the hall's kernels as a hall one node deep.  Its own node counts (room.rs:
    90-120), Courant limit dx/(c sqrt 2) (room.rs:361-366) and energy invariant (room.rs:243-281)
    are applied by the library; the fields are those of the hall's sweep with no z faces."""

    def __init__(self, name: str, width: float, height: float, cells_across: int, depth: float,
                 density: float, speed: float):
        self._name = name
        self._speed = float(speed)
        self._density = float(density)
        self._depth = float(depth)
        handle = C.c_void_p()
        check(lib().pm_room_create(name.encode(), float(width), float(height), int(cells_across),
                                   self._depth, self._density, self._speed, C.byref(handle)),
              "pm_room_create")
        self._h = handle
        out = (C.c_size_t * 3)()
        check(lib().pm_hall_nodes(self._h, out), "pm_hall_nodes")
        self._nodes = (out[0], out[1], out[2])
        self._dx = float(lib().pm_hall_spacing(self._h))

    @staticmethod
    def of_air(name: str, width: float, height: float, cells_across: int) -> "Room":
        """Air at 20 C, one metre deep: 1.204 kg/m^3, 343 m/s (room.rs:122-137)."""
        return Room(name, width, height, cells_across, 1.0, 1.204, 343.0)

    def cells(self) -> Tuple[int, int]:
        return self._nodes[0], self._nodes[1]

    def room_depth(self) -> float:
        return self._depth

    def cell_area(self) -> float:
        """room.rs:283-285."""
        return self._dx * self._depth

    def released_from(self, profile) -> "Room":
        """Room::released_from, room.rs:140-160; the profile sees node positions (i, j)*dx."""
        nx, ny = self.cells()
        if callable(profile):
            p = np.empty((ny, nx), dtype=np.float64)
            for j in range(ny):
                for i in range(nx):
                    p[j, i] = profile(float(i) * self._dx, float(j) * self._dx)
        else:
            p = np.ascontiguousarray(profile, dtype=np.float64).reshape(ny, nx)
        p = np.ascontiguousarray(p.reshape(-1))
        check(lib().pm_hall_release(self._h, dptr(p)), "pm_hall_release")
        return self

    def mode_profile(self, mode: Tuple[int, int], amplitude: float) -> np.ndarray:
        """The pressure of room.rs:166-182 at every node: (a*px)*py."""
        m = tuple(mode) + (0,) * (3 - len(tuple(mode)))
        return super().mode_profile((m[0], m[1], 0), amplitude).reshape(self._nodes[1],
                                                                          self._nodes[0])

    def released_in_mode(self, mode: Tuple[int, int], amplitude: float) -> "Room":
        return self.released_from(self.mode_profile(mode, amplitude))

    def set_source(self, node: int, increments: Optional[np.ndarray]) -> None:
        raise TypeError("a room has no source")

    def mode_frequency(self, mode: Tuple[int, int]) -> float:
        """room.rs:291-300: (c/2) sqrt((nx/Lx)^2 + (ny/Ly)^2), zero for a degenerate room."""
        lx, ly = self.width(), self.height()
        if lx <= 0.0 or ly <= 0.0:
            return 0.0
        a = float(mode[0]) / lx
        b = float(mode[1]) / ly
        return self._speed / 2.0 * math.sqrt(a * a + b * b)

    def modes_below(self, limit: float) -> List[Tuple[int, int, float]]:
        """room.rs:306-332."""
        def bound(length: float) -> int:
            return 0 if length <= 0.0 else int(math.floor(2.0 * limit * length / self._speed))

        found = []
        for a in range(bound(self.width()) + 1):
            for b in range(bound(self.height()) + 1):
                if a == 0 and b == 0:
                    continue
                f = self.mode_frequency((a, b))
                if f <= limit:
                    found.append((a, b, f))
        found.sort(key=lambda x: x[2])
        return found

    def pressure_at(self, i: int, j: int, k: int = 0) -> float:
        """room.rs:200-204, clamped to the walls."""
        return super().pressure_at(i, j, 0)

    def sample(self, points) -> np.ndarray:
        """ScalarField::at (room.rs:587-600) at each (x, y) point, or (x, y, z) with z ignored:
        bilinear between nodes, clamped to the walls, pascals."""
        pts = np.asarray(points, dtype=np.float64)
        pts = pts.reshape(-1, pts.shape[-1] if pts.ndim > 1 else 2)
        xyz = np.zeros((len(pts), 3))
        xyz[:, :2] = pts[:, :2]
        return super().sample(xyz)


def impedance(density: float, sound_speed: float) -> float:
    """rho c, crates/pantometry-acoustic/src/lib.rs:69-71."""
    return density * sound_speed


def reflection_coefficient(z_from: float, z_into: float) -> float:
    """(Z2 - Z1)/(Z2 + Z1), lib.rs:86-92; zero for a pair summing to zero."""
    if z_from + z_into == 0.0:
        return 0.0
    return (z_into - z_from) / (z_into + z_from)


def power_transmission(z_from: float, z_into: float) -> float:
    """1 - R^2 clamped to [0, 1], lib.rs:100-103."""
    r = reflection_coefficient(z_from, z_into)
    return min(max(1.0 - r * r, 0.0), 1.0)


class End:
    """What an end of a tube does to a wave (lib.rs:106-116): closed, open, or matched to an
    impedance."""

    CLOSED, OPEN, MATCHED = 0, 1, 2

    def __init__(self, kind: int, z: float = 0.0):
        self.kind, self.z = kind, float(z)

    @staticmethod
    def closed() -> "End":
        return End(End.CLOSED)

    @staticmethod
    def open() -> "End":
        return End(End.OPEN)

    @staticmethod
    def matched(z: float) -> "End":
        return End(End.MATCHED, z)

    def __eq__(self, other) -> bool:
        return isinstance(other, End) and (self.kind, self.z) == (other.kind, other.z)

    def __repr__(self) -> str:
        return ("Closed", "Open", f"Matched({self.z})")[self.kind]


class Tube:
    """Tube (crates/pantometry-acoustic/src/lib.rs:136-563): pressure at the samples, particle
    velocity on the faces between them, leapfrogged, with closed, open or matched ends."""

    def __init__(self, name: str, length: float, cells: int, area: float, density: float,
                 speed: float):
        self._name = name
        self._density, self._speed, self._area = float(density), float(speed), float(area)
        handle = C.c_void_p()
        check(lib().pm_tube_create(name.encode(), float(length), int(cells), self._area,
                                   self._density, self._speed, C.byref(handle)), "pm_tube_create")
        self._h = handle
        self._n = int(lib().pm_tube_cells(self._h))
        self._dx = float(lib().pm_tube_spacing(self._h))
        self._ends = (End.closed(), End.closed())

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().pm_tube_destroy(h)
            self._h = None

    @staticmethod
    def of_air(name: str, length: float, cells: int, area: float) -> "Tube":
        """Air at 20 C: 1.204 kg/m^3, 343 m/s (lib.rs:196-205)."""
        return Tube(name, length, cells, area, 1.204, 343.0)

    def with_ends(self, left: End, right: End) -> "Tube":
        check(lib().pm_tube_set_ends(self._h, left.kind, left.z, right.kind, right.z),
              "pm_tube_set_ends")
        self._ends = (left, right)
        return self

    def released_from(self, profile) -> "Tube":
        """Tube::released_from, lib.rs:226-236; the profile sees sample positions i*dx."""
        if callable(profile):
            p = np.array([profile(float(i) * self._dx) for i in range(self._n)], dtype=np.float64)
        else:
            p = np.ascontiguousarray(profile, dtype=np.float64).reshape(-1)
        check(lib().pm_tube_release(self._h, dptr(p)), "pm_tube_release")
        return self

    # ---- geometry and closed forms ---------------------------------------------------------
    def name(self) -> str:
        return self._name

    def cells(self) -> int:
        return self._n

    def spacing(self) -> float:
        return self._dx

    def length(self) -> float:
        return float(self._n - 1) * self._dx

    def sound_speed(self) -> float:
        return self._speed

    def impedance(self) -> float:
        return impedance(self._density, self._speed)

    def mode_frequency(self, n: int) -> float:
        """lib.rs:340-355: n c/2L for like ends, odd multiples of c/4L otherwise."""
        length = self.length()
        if length <= 0.0 or n == 0:
            return 0.0
        kinds = (self._ends[0].kind, self._ends[1].kind)
        if kinds in ((End.CLOSED, End.CLOSED), (End.OPEN, End.OPEN)):
            return float(n) * self._speed / (2.0 * length)
        return float(2 * n - 1) * self._speed / (4.0 * length)

    def courant(self, dt: float) -> float:
        return float(lib().pm_tube_courant(self._h, float(dt)))

    # ---- state ---------------------------------------------------------------------------
    def fields(self):
        """(pressure, pressure_prev, velocity)."""
        p, pp, u = np.empty(self._n), np.empty(self._n), np.empty(self._n - 1)
        check(lib().pm_tube_get_fields(self._h, dptr(p), dptr(pp), dptr(u)), "pm_tube_get_fields")
        return p, pp, u

    def pressure(self) -> np.ndarray:
        return self.fields()[0]

    def pressure_at(self, cell: int) -> float:
        return float(self.pressure()[min(cell, self._n - 1)])

    def velocity_at(self, face: int) -> float:
        return float(self.fields()[2][min(face, self._n - 2)])

    def peak_pressure(self) -> float:
        peak = 0.0
        for v in self.pressure():
            peak = max(peak, abs(float(v)))
        return peak

    def energy_parts(self) -> Tuple[float, float, float]:
        inv, off, rad = C.c_double(), C.c_double(), C.c_double()
        check(lib().pm_tube_energy(self._h, C.byref(inv), C.byref(off), C.byref(rad)),
              "pm_tube_energy")
        return inv.value, off.value, rad.value

    def energy(self) -> float:
        """Tube::energy, lib.rs:288-290: the invariant plus the startup offset."""
        inv, off, _ = self.energy_parts()
        return inv + off

    def startup_adjustment(self) -> float:
        return self.energy_parts()[1]

    def radiated_energy(self) -> float:
        return self.energy_parts()[2]

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
        return float(lib().pm_tube_max_stable_dt(self._h))

    def step(self, t: float, dt: float, bus: Exchange) -> None:
        """Domain::step, lib.rs:516-533: what the ends absorbed goes onto the bus as energy."""
        v = _capi.pm_violation()
        published = C.c_double(0.0)
        check(lib().pm_tube_step(self._h, float(dt), C.byref(published), C.byref(v)),
              "pm_tube_step", v)
        if published.value > 0.0:
            bus.publish(ENERGY, published.value)

    def ledger(self) -> Ledger:
        return Ledger().with_(ENERGY, self.energy())

    def checkpoint(self) -> None:
        check(lib().pm_tube_checkpoint(self._h), "pm_tube_checkpoint")

    def restore(self) -> None:
        check(lib().pm_tube_restore(self._h), "pm_tube_restore")

    def readings(self):
        return [(self._name, "peak", self.peak_pressure(), "Pa")]

    def advance(self, dt: float, steps: int, tolerance: float = 1e-9):
        """`steps` steps, each audited on the device (energy plus what the ends radiated).
        Returns (books after each step, device milliseconds)."""
        v = _capi.pm_violation()
        totals = np.empty(max(steps, 1))
        ms = C.c_double(0.0)
        check(lib().pm_tube_advance(self._h, float(dt), int(steps), float(tolerance), dptr(totals),
                                    C.byref(ms), C.byref(v)), "pm_tube_advance", v)
        return totals[:steps], ms.value
