"""Scene files for the domains on this backend's path, and the frames a run of one captures.

A scene is the reference's JSON format (crates/pantometry-world/src/lib.rs:55-130, `Scene`):
a title, a coupling schedule, a duration cut into frames, a conservation tolerance, declared
materials and a list of domains.  These kinds of domain are built here, the grids on the GPU:

  block   Solid3D (DomainSpec::Block, lib.rs:2185-2400): cells, cell size, starting temperature,
          bulk material, regions of other materials or of void (later wins) with their own
          starting temperature, a hot spot, cooled faces (CoolingSpec) and dissipating boxes
          (DissipationSpec); materials may melt (FusionProps)
  hall    Hall (DomainSpec::Hall, lib.rs:1786-1805): an air-filled room released in a mode
  room    Room (DomainSpec::Room, lib.rs:1405-1437): released in a mode or as a Gaussian pulse
  bar     Bar1D of aluminium (DomainSpec::Bar, lib.rs:1682-1701)
  heater  Heater (crates/pantometry-world/src/heater.rs): a reserve of joules spent at a fixed
          power onto the bus, the heat every thermal domain takes its share of

Every other kind of domain, and the features outside this path (designed parts from STL,
composite materials), is refused with a message that says so rather than built approximately.  A run (World::run, lib.rs:3204-3218) advances the simulation by
duration/frames per frame and captures each domain's readings and its field sampled over its
extent (scene::capture, crates/pantometry-scene/src/lib.rs:337-404); `to_json` writes the frames
in the viewer's run-file format (crates/pantometry-view/src/data.rs:55-171).
"""
from __future__ import annotations

import json
import math
from dataclasses import dataclass, field
from pathlib import Path
from typing import Dict, List, Optional, Tuple, Union

import numpy as np

from .acoustic import Hall, Room
from .core import ENERGY, HEAT, Exchange, Kind, Ledger, Schedule, Simulation, Tolerances
from .thermal import Bar1D, Environment, Face, Fusion, Solid3D, Substance, celsius, mm

FORMAT = 1  # lib.rs:369, the highest scene revision this reads

# The catalogue a scene can name, Substance::from_name (crates/pantometry-core/src/substance.rs:
# 292-331): density in g/cm^3 (stored as *1e3 kg/m^3), conductivity, specific heat, emissivity,
# and whether the substance freezes (its `fusion` is present).
_CATALOGUE = {
    "aluminium": ("Al 6061", 2.70, 167.0, 896.0, 0.09, False),
    "borosilicate": ("N-BK7", 2.51, 1.114, 858.0, 0.90, False),
    "copper": ("Cu", 8.96, 401.0, 385.0, 0.04, False),
    "electrical_steel": ("electrical steel", 7.65, 25.0, 460.0, 0.30, False),
    "fr4": ("FR-4", 1.85, 0.30, 1100.0, 0.90, False),
    "ice": ("ice", 0.917, 2.22, 2050.0, 0.97, True),
    "pla": ("PLA", 1.24, 0.13, 1800.0, 0.90, False),
    "stainless_304": ("304 stainless", 8.00, 16.2, 500.0, 0.28, False),
    "water": ("water", 0.998, 0.598, 4182.0, 0.96, False),
}
CATALOGUE = tuple(_CATALOGUE)
VOID = "void"

_FACES = {"x-min": Face.XMIN, "x-max": Face.XMAX, "y-min": Face.YMIN, "y-max": Face.YMAX,
          "z-min": Face.ZMIN, "z-max": Face.ZMAX}
_SCHEDULES = {"one-way": Schedule.ONE_WAY, "staggered": Schedule.STAGGERED,
              "multirate": Schedule.MULTIRATE}
_SCENE_KEYS = {"format", "title", "schedule", "domains", "duration_s", "frames",
               "conservation_tolerance", "tolerance_for", "materials", "composites"}
_BLOCK_KEYS = {"kind", "name", "cells", "cell_mm", "initial_c", "material", "regions",
               "hot_spot", "parts", "cooling", "dissipation"}
_HALL_KEYS = {"kind", "name", "width_m", "height_m", "depth_m", "nodes_across", "mode",
              "amplitude_pa"}
_ROOM_KEYS = {"kind", "name", "width_m", "height_m", "cells_across", "release"}
_RELEASE_KEYS = {"mode": {"as", "nx", "ny", "amplitude_pa"},
                 "pulse": {"as", "x_m", "y_m", "radius_m", "amplitude_pa"}}
_BAR_KEYS = {"kind", "name", "length_mm", "cells", "area_mm2", "initial_c", "exposes"}
_HEATER_KEYS = {"kind", "name", "watts", "reserve_j"}
_FUSION_KEYS = {"melting_point", "latent_heat", "liquid"}


class SceneError(ValueError):
    """A scene this backend refuses, with the reason."""


def _unknown(site: str, got: dict, known: set) -> None:
    extra = sorted(set(got) - known)
    if extra:  # serde(deny_unknown_fields)
        raise SceneError(f"{site}: unknown field(s) {', '.join(extra)}")


@dataclass
class Scene:
    title: str
    schedule: Schedule
    duration_s: float
    frames: int
    conservation_tolerance: float
    tolerance_for: Dict[str, float]
    materials: Dict[str, dict]
    domains: List[dict]
    composites: Dict[str, dict] = field(default_factory=dict)


def load(source: Union[str, Path, dict]) -> Scene:
    """A scene from a file path, JSON text or a parsed dict, checked as World::build checks it
    (lib.rs:400-428, 2846-2864)."""
    if isinstance(source, dict):
        raw = source
    else:
        text = str(source)
        p = Path(text)
        raw = json.loads(p.read_text() if not text.lstrip().startswith("{") and p.exists()
                         else text)
    _unknown("scene", raw, _SCENE_KEYS)
    fmt = int(raw.get("format", 1))
    if fmt == 0:
        raise SceneError(f"format 0 is not a version this format ever had; scenes are {FORMAT}")
    if fmt > FORMAT:
        raise SceneError(f"this scene is format {fmt} and this build reads up to {FORMAT}")
    schedule = raw.get("schedule", "multirate")
    if schedule not in _SCHEDULES:
        raise SceneError(f"schedule {schedule!r} is not one of {', '.join(_SCHEDULES)}")
    domains = raw.get("domains") or []
    if not domains:
        raise SceneError("a scene needs at least one domain")
    duration = float(raw["duration_s"])
    if not duration > 0.0:
        raise SceneError(f"duration must be positive, got {duration}")
    frames = int(raw["frames"])
    if frames <= 0:
        raise SceneError("a scene needs at least one frame")
    return Scene(raw["title"], _SCHEDULES[schedule], duration, frames,
                 float(raw.get("conservation_tolerance", 1e-6)),
                 dict(raw.get("tolerance_for", {})), dict(raw.get("materials", {})), domains,
                 dict(raw.get("composites", {})))


class Palette:
    """The substances a scene can name: the catalogue plus its own declarations (lib.rs:1072-
    1150).  A declaration may not shadow a catalogue name, and one nothing asks for is refused."""

    def __init__(self, declared: Dict[str, dict]):
        self.declared: Dict[str, Tuple[Substance, bool]] = {}
        self.asked = set()
        for key, spec in declared.items():
            if key in _CATALOGUE or key == VOID:
                raise SceneError(f"materials.{key}: shadows the catalogue's {key!r}")
            thermal = spec.get("thermal")
            density = float(spec.get("density", math.nan))
            if not (density > 0.0 and math.isfinite(density)):
                raise SceneError(f"materials.{key}: density is {density}")
            fusion = _fusion(f"materials.{key}", spec.get("fusion"))
            if thermal is None:
                sub = Substance(spec.get("name", key), density, None, None, 0.0, fusion)
            else:
                k, cp = float(thermal["conductivity"]), float(thermal["specific_heat"])
                if not (k >= 0.0 and cp > 0.0):
                    raise SceneError(f"materials.{key}: conductivity {k}, specific heat {cp}")
                sub = Substance(spec.get("name", key), density, k, cp,
                                min(max(float(thermal.get("emissivity", 0.0)), 0.0), 1.0), fusion)
            self.declared[key] = (sub, fusion is not None)

    def get(self, site: str, name: str) -> Substance:
        self.asked.add(name)
        if name in self.declared:
            sub, freezes = self.declared[name]
        elif name in _CATALOGUE:
            label, dens, k, cp, eps, freezes = _CATALOGUE[name]
            sub = Substance.ice() if freezes else Substance(label, dens * 1e3, k, cp, eps)
        else:
            raise SceneError(f"{site}: no material {name!r}; the catalogue has "
                             f"{', '.join(CATALOGUE)} and the scene declares "
                             f"{', '.join(self.declared) or 'none'}")
        if sub.conductivity is None or sub.specific_heat is None:
            raise SceneError(f"{site}: {name!r} has no thermal properties")
        return sub

    def unused(self) -> List[str]:
        return sorted(k for k in self.declared if k not in self.asked)


def _fusion(site: str, spec: Optional[dict]) -> Optional[Fusion]:
    """FusionProps as a scene declares it (substance.rs:132-154): a melting point in kelvin, a
    latent heat in J/kg, and optionally the liquid's ThermalProps."""
    if spec is None:
        return None
    _unknown(f"{site}.fusion", spec, _FUSION_KEYS)
    point, latent = float(spec["melting_point"]), float(spec["latent_heat"])
    if not (point > 0.0 and math.isfinite(point) and latent > 0.0 and math.isfinite(latent)):
        raise SceneError(f"{site}.fusion: melting point {point} K, latent heat {latent} J/kg")
    liquid = spec.get("liquid")
    if liquid is None:
        return Fusion(point, latent)
    kl, cl = float(liquid["conductivity"]), float(liquid["specific_heat"])
    if not (kl >= 0.0 and cl > 0.0):
        raise SceneError(f"{site}.fusion.liquid: conductivity {kl}, specific heat {cl}")
    return Fusion(point, latent, (kl, cl))


def _box(counts, lo, hi) -> np.ndarray:
    nx, ny, nz = counts
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    return ((i >= lo[0]) & (i < hi[0]) & (j >= lo[1]) & (j < hi[1]) & (k >= lo[2]) &
            (k < hi[2])).reshape(-1)


def _selects(site: str, spec: dict, cells) -> None:
    empty = any(spec["to"][a] <= spec["from"][a] for a in range(3))
    outside = any(spec["to"][a] > cells[a] for a in range(3))
    if empty or outside:
        raise SceneError(f"{site}: {spec['from']}..{spec['to']} selects no cells of a "
                         f"{cells[0]}x{cells[1]}x{cells[2]} block; `to` is one past the last "
                         f"cell, so a single cell is `to = from + 1`")


def build_block(spec: dict, palette: Palette) -> Solid3D:
    """DomainSpec::Block, in the builder's order (lib.rs:2185-2400)."""
    name = spec["name"]
    _unknown(name, spec, _BLOCK_KEYS)
    cells = tuple(int(c) for c in spec["cells"])
    cell_mm = float(spec["cell_mm"])
    initial_c = float(spec["initial_c"])
    bulk = palette.get(name, spec.get("material") or "aluminium")
    block = Solid3D(name, bulk, cells, mm(cell_mm), celsius(initial_c))
    if spec.get("parts"):
        raise SceneError(f"{name}: designed parts are voxelised from STL meshes, which is not on "
                         f"this backend's path")
    for n, r in enumerate(spec.get("regions", [])):
        site = f"{name}/regions[{n}]"
        _selects(site, r, cells)
        inside = _box(cells, r["from"], r["to"])
        if r["material"] == VOID:
            # A box of nothing: an air gap without a mesh (lib.rs:2276-2285).
            if r.get("initial_c") is not None:
                raise SceneError(f"{site}: a \"void\" region cannot start at a temperature, "
                                 f"because there is nothing there to have one")
            block.empty(inside)
            continue
        block.fill(palette.get(site, r["material"]), inside)
        if r.get("initial_c") is not None:
            c = float(r["initial_c"])
            if not math.isfinite(c):
                raise SceneError(f"{site}: initial_c is {c}")
            block.set_region_temperature(inside, celsius(c))
    seen = set()
    for n, cool in enumerate(spec.get("cooling", [])):
        site = f"{name}/cooling[{n}]"
        area, h, amb = float(cool["area_cm2"]), float(cool["convection_w_per_m2_k"]), \
            float(cool["ambient_c"])
        if not (math.isfinite(area) and area > 0.0):
            raise SceneError(f"{site}: area_cm2 is {area}, and a face with no area loses nothing")
        if not (math.isfinite(h) and h >= 0.0):
            raise SceneError(f"{site}: convection_w_per_m2_k is {h}")
        if not math.isfinite(amb):
            raise SceneError(f"{site}: ambient_c is {amb}")
        if cool["face"] not in _FACES:
            raise SceneError(f"{site}: face {cool['face']!r} is not one of {', '.join(_FACES)}")
        if cool["face"] in seen:
            raise SceneError(f"{site}: {cool['face']} is already cooled by an earlier entry")
        seen.add(cool["face"])
        block.losing_from(_FACES[cool["face"]], Environment(celsius(amb), h, area * 1e-4))
    spot = spec.get("hot_spot")
    if spot is not None:
        i, j, k = spot["at"]
        block.set_temperature(i, j, k, celsius(initial_c + float(spot["above_k"])))
    for n, d in enumerate(spec.get("dissipation", [])):
        site = f"{name}/dissipation[{n}]"
        _selects(site, d, cells)
        watts = float(d["watts"])
        if not math.isfinite(watts):
            raise SceneError(f"{site}: {watts} is not a number of watts")
        block.dissipating(watts, _box(cells, d["from"], d["to"]))
    return block


def build_hall(spec: dict) -> Hall:
    """DomainSpec::Hall (lib.rs:1786-1805): Hall::of_air released in the stated mode."""
    name = spec["name"]
    _unknown(name, spec, _HALL_KEYS)
    hall = Hall.of_air(name, float(spec["width_m"]), float(spec["height_m"]),
                       float(spec["depth_m"]), int(spec["nodes_across"]))
    m = spec["mode"]
    return hall.released_in_mode((int(m[0]), int(m[1]), int(m[2])), float(spec["amplitude_pa"]))


def build_room(spec: dict) -> Room:
    """DomainSpec::Room (lib.rs:1405-1437): Room::of_air released in a mode or as a pulse
    a*exp(-((x-x0)^2 + (y-y0)^2)/r^2), r at least 1e-9."""
    name = spec["name"]
    _unknown(name, spec, _ROOM_KEYS)
    release = spec.get("release") or {"as": "mode", "nx": 1, "ny": 1, "amplitude_pa": 1.0}
    kind = release.get("as")
    if kind not in _RELEASE_KEYS:
        raise SceneError(f"{name}.release: {kind!r} is not one of mode, pulse")
    _unknown(f"{name}.release", release, _RELEASE_KEYS[kind])
    room = Room.of_air(name, float(spec["width_m"]), float(spec["height_m"]),
                       int(spec["cells_across"]))
    if kind == "mode":
        return room.released_in_mode((int(release["nx"]), int(release["ny"])),
                                     float(release["amplitude_pa"]))
    cx, cy = float(release["x_m"]), float(release["y_m"])
    r, a = max(float(release["radius_m"]), 1e-9), float(release["amplitude_pa"])

    def pulse(x, y):
        dx, dy = x - cx, y - cy
        return a * math.exp(-(dx * dx + dy * dy) / (r * r))

    return room.released_from(pulse)


def build_bar(spec: dict) -> Bar1D:
    """DomainSpec::Bar (lib.rs:1682-1701): aluminium, length_mm/cells per cell."""
    name = spec["name"]
    _unknown(name, spec, _BAR_KEYS)
    cells = int(spec["cells"])
    bar = Bar1D(name, Substance.aluminium_6061(), cells, mm(float(spec["length_mm"]) / float(cells)),
                float(spec["area_mm2"]) * 1e-6, celsius(float(spec["initial_c"])))
    exposes = spec.get("exposes")
    if exposes is not None:
        bar.exposing(float(exposes["face_area_mm2"]) * 1e-6)
    return bar


class Heater:
    """Heater (crates/pantometry-world/src/heater.rs:17-104): a reserve of joules spent at a
    fixed power, published on the bus as heat.  Its ledger is the reserve."""

    def __init__(self, name: str, watts: float, reserve: float):
        self._name = name
        self.watts = max(watts, 0.0)
        self.reserve = max(reserve, 0.0)
        self._saved: Optional[float] = None

    def name(self) -> str:
        return self._name

    def kind(self) -> Kind:
        return Kind.QUASI_STATIC

    def max_stable_dt(self, now: float = 0.0) -> float:
        return math.inf

    def books_balance(self) -> bool:
        return False  # the Domain default: a heater makes no claim of its own (sim.rs:152-154)

    def step(self, t: float, dt: float, bus: Exchange) -> None:
        joules = max(min(self.watts * dt, self.reserve), 0.0)
        self.reserve -= joules
        bus.publish(HEAT, joules)

    def ledger(self) -> Ledger:
        return Ledger().with_(ENERGY, self.reserve)

    def checkpoint(self) -> None:
        self._saved = self.reserve

    def restore(self) -> None:
        if self._saved is not None:
            self.reserve = self._saved

    def supports_restore(self) -> bool:
        return True

    def readings(self):
        return [(self._name, "reserve", self.reserve, "J")]


def build_heater(spec: dict) -> Heater:
    _unknown(spec["name"], spec, _HEATER_KEYS)
    return Heater(spec["name"], float(spec["watts"]), float(spec["reserve_j"]))


@dataclass
class Panel:
    name: str
    unit: str
    nx: int
    ny: int
    nz: int
    values: np.ndarray


@dataclass
class Frame:
    time_s: float
    panels: List[Panel] = field(default_factory=list)
    readings: List[Tuple[str, str, float, str]] = field(default_factory=list)


def _extent_points(size: Tuple[float, float, float], samples: Tuple[int, int, int]) -> np.ndarray:
    """The sample points of an extent from 0 to `size` (scene::sample, scene/src/lib.rs:376-404):
    lo + (hi - lo) * f with f = i/(n-1), or the middle for a single sample; x fastest."""
    nx, ny, nz = samples

    def along(i, n):
        return i / (n - 1) if n > 1 else 0.5

    pts = np.empty((nx * ny * nz, 3))
    q = 0
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                f = (along(i, nx), along(j, ny), along(k, nz))
                pts[q] = [0.0 + (size[a] - 0.0) * f[a] for a in range(3)]
                q += 1
    return pts


class World:
    """A built scene: the simulation, its domains and where each is sampled."""

    def __init__(self, scene: Scene):
        self.scene = scene
        tol = Tolerances(scene.conservation_tolerance)
        for channel, t in scene.tolerance_for.items():
            if channel != "energy":
                raise SceneError(f"tolerance_for: {channel!r} is not a channel of these domains")
            tol = tol.with_("energy", float(t))
        self.sim = Simulation(scene.schedule)
        self.sim.conservation_tol = tol
        palette = Palette(scene.materials)
        if scene.composites:
            raise SceneError(f"composites: {', '.join(sorted(scene.composites))} mix substances "
                             f"by volume (crates/pantometry-core/src/mixture.rs), which is not on "
                             f"this backend's path")
        self.extent: Dict[str, Tuple[Tuple[float, float, float], Tuple[int, int, int]]] = {}
        for spec in scene.domains:
            kind = spec.get("kind")
            if kind == "room":
                d = build_room(spec)
                # Placement::field(Extent::rectangle(width, height, nx, ny)), lib.rs:3402-3415
                nx = max(int(spec["cells_across"]), 3)
                w, h = float(spec["width_m"]), float(spec["height_m"])
                ny = max(int(_round_half_away((h / w) * float(nx - 1))) + 1, 3)
                self.extent[d.name()] = ((w, h, 0.0), (nx, ny, 1))
            elif kind == "bar":
                d = build_bar(spec)
                # Placement::field(Extent::line(length, cells)), lib.rs:3416-3418
                self.extent[d.name()] = ((float(spec["length_mm"]) * 1e-3, 0.0, 0.0),
                                         (max(int(spec["cells"]), 2), 1, 1))
            elif kind == "heater":
                d = build_heater(spec)
            elif kind == "block":
                d = build_block(spec, palette)
                cells = tuple(int(c) for c in spec["cells"])
                size = tuple(cells[a] * float(spec["cell_mm"]) * 1e-3 for a in range(3))
                self.extent[d.name()] = (size, cells)
            elif kind == "hall":
                d = build_hall(spec)
                nx, ny, nz = d.nodes()
                dx = d.spacing()
                # Placement::field(Extent::volume(((n-1) dx per axis), nodes)), lib.rs:3461-3482
                size = ((nx - 1) * dx, (ny - 1) * dx, (nz - 1) * dx)
                self.extent[d.name()] = (size, (nx, ny, nz))
            else:
                raise SceneError(f"{spec.get('name', '?')}: a {kind!r} domain is not on this "
                                 f"backend's path (block, hall, room, bar and heater are)")
            self.sim.with_(d)
        unused = palette.unused()
        if unused:
            raise SceneError(f"materials: {', '.join(unused)} declared and never used")
        self._pts = {n: _extent_points(*e) for n, e in self.extent.items()}

    def capture(self) -> Frame:
        """scene::capture (scene/src/lib.rs:337-357): every domain's readings, and each field
        (a block's temperature, a hall's pressure) sampled over its extent through
        ScalarField::at on the device."""
        frame = Frame(self.sim.t)
        for d in self.sim.domains:
            frame.readings.extend(d.readings())
            if d.name() in self._pts:
                (size, (nx, ny, nz)) = self.extent[d.name()]
                unit = "Pa" if isinstance(d, (Hall, Room)) else "K"
                frame.panels.append(Panel(d.name(), unit, nx, ny, nz, d.sample(self._pts[d.name()])))
        return frame

    def advance(self, dt: float):
        return self.sim.advance(dt)

    def run(self) -> List[Frame]:
        """World::run (lib.rs:3204-3218): a frame before the first advance and after each."""
        dt = self.scene.duration_s / float(self.scene.frames)
        frames = [self.capture()]
        for _ in range(self.scene.frames):
            self.advance(dt)
            frames.append(self.capture())
        return frames


def _round_half_away(x: float) -> float:
    """f64::round: halves away from zero."""
    return math.floor(abs(x) + 0.5) * (1.0 if x >= 0.0 else -1.0) if abs(x) < 2.0 ** 52 else x


def build(source: Union[str, Path, dict, Scene]) -> World:
    return World(source if isinstance(source, Scene) else load(source))


def _number(x: float) -> str:
    """Rust's `{:.6e}` (crates/pantometry-view/src/data.rs:149-151): six decimals, an exponent
    with no sign for positive powers and no padding; `null` for a sample that is not finite."""
    if not math.isfinite(x):
        return "null"
    mant, exp = f"{x:.6e}".split("e")
    return f"{mant}e{int(exp)}"


def _quote(s: str) -> str:
    out = []
    for c in s:
        if c == '"':
            out.append('\\"')
        elif c == "\\":
            out.append("\\\\")
        elif ord(c) < 32 or 127 <= ord(c) < 160:
            out.append(" ")
        else:
            out.append(c)
    return '"' + "".join(out) + '"'


def to_json(title: str, frames: List[Frame]) -> str:
    """The run file the viewer reads (pantometry-view data.rs:55-138), field panels only."""
    out = [f'{{\n  "title": {_quote(title)},\n  "frames": [\n']
    for fi, fr in enumerate(frames):
        out.append(f'    {{ "t": {fr.time_s:.6f}, "panels": [')
        for pi, p in enumerate(fr.panels):
            out.append(f'\n      {{ "name": {_quote(p.name)}, "unit": {_quote(p.unit)}, ')
            out.append(f'"kind": "field", "nx": {p.nx}, "ny": {p.ny}, "nz": {p.nz}, "values": '
                       + "[" + ",".join(_number(float(v)) for v in p.values) + "]")
            out.append(" }" if pi + 1 == len(fr.panels) else " },")
        out.append('\n    ], "readings": [')
        for ri, (dom, label, value, unit) in enumerate(fr.readings):
            out.append(f'\n      {{ "domain": {_quote(dom)}, "label": {_quote(label)}, "unit": '
                       f'{_quote(unit)}, "value": {_number(float(value))} }}'
                       + ("" if ri + 1 == len(fr.readings) else ","))
        out.append("\n    ] }\n" if fi + 1 == len(frames) else "\n    ] },\n")
    out.append("  ]\n}\n")
    return "".join(out)
