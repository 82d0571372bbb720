"""Scenes on the GPU: the reference's scene files built into device domains and run frame by
frame, every block field and every captured panel bit-identical to the same scene built and
stepped in the CPU oracle (multirate substeps at each block's own limit), the run file parsing as
JSON, and the refusals that need a built block."""
import json
import math
from pathlib import Path

import numpy as np
import pytest

from tests.conftest import gpu_case

pytestmark = pytest.mark.gpu

SCENES = Path(__file__).resolve().parent / "golden" / "scenes"
FACES = {"x-min": 0, "x-max": 1, "y-min": 2, "y-max": 3, "z-min": 4, "z-max": 5}


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).reshape(-1).view(np.uint64)


def _fusion(sub):
    f = sub.fusion
    return None if f is None else (f.melting_point, f.latent_heat, f.liquid)


def oracle_block(oracle, spec, palette):
    """The builder's steps (crates/pantometry-world/src/lib.rs:2185-2400) on the oracle."""
    from pantometry_amd import scene as S
    cells = tuple(spec["cells"])
    bulk = palette.get("x", spec.get("material") or "aluminium")
    o = oracle.OracleSolid3D(spec["name"], bulk.density, bulk.conductivity, bulk.specific_heat,
                             cells, spec["cell_mm"] * 1e-3, spec["initial_c"] + 273.15,
                             emissivity=bulk.emissivity, fusion=_fusion(bulk))
    nx, ny, _ = cells
    for r in spec.get("regions", []):
        inside = S._box(cells, r["from"], r["to"])
        if r["material"] == "void":
            o.empty(inside)
            continue
        sub = palette.get("x", r["material"])
        o.fill(sub.density, sub.conductivity, sub.specific_heat, inside, emissivity=sub.emissivity,
               fusion=_fusion(sub))
        if r.get("initial_c") is not None:
            for c in np.flatnonzero(inside):
                o.set_temperature(c % nx, (c // nx) % ny, c // (nx * ny), r["initial_c"] + 273.15)
    for c in spec.get("cooling", []):
        o.losing_from(FACES[c["face"]], c["ambient_c"] + 273.15, c["convection_w_per_m2_k"],
                      c["area_cm2"] * 1e-4)
    if spec.get("hot_spot"):
        i, j, k = spec["hot_spot"]["at"]
        nx, ny, _ = cells
        o.cells[i + nx * (j + ny * k)] = spec["initial_c"] + spec["hot_spot"]["above_k"] + 273.15
    for d in spec.get("dissipation", []):
        o.dissipating(d["watts"], S._box(cells, d["from"], d["to"]))
    return o


@pytest.mark.parametrize("name,frames", [("15-a-hot-spot-in-a-block", None),
                                         ("19-a-coating-stops-the-heat", None),
                                         ("24-a-power-module-junction-to-ambient", 2)])
def test_block_scenes_match_the_oracle(oracle, name, frames):
    gpu_case(f"scene {name}")
    from pantometry_amd import scene as S
    from pantometry_amd.core import substeps_for
    sc = S.load(SCENES / f"{name}.json")
    world = S.build(sc)
    palette = S.Palette(sc.materials)
    blocks = {spec["name"]: oracle_block(oracle, spec, palette) for spec in sc.domains}
    dt = sc.duration_s / sc.frames
    count = sc.frames if frames is None else frames
    first = world.capture()
    for n in range(count):
        world.advance(dt)
        for spec in sc.domains:
            o = blocks[spec["name"]]
            m = substeps_for(dt, o.max_stable_dt())
            h = dt / m
            for _ in range(m):
                o.step(h)
            g = world.sim.domain(spec["name"])
            assert np.array_equal(bits(g.cells()), bits(o.cells)), (name, n)
    frame = world.capture()
    for p in frame.panels:
        o = blocks[p.name]
        pts = world._pts[p.name]
        expect = np.array([o.sample(*pt) for pt in pts])
        assert np.array_equal(bits(p.values), bits(expect))
    labels = [r[1] for r in frame.readings]
    assert labels[:4] == ["peak", "mean", "coldest", "absorbed"]
    if name.startswith("24"):
        assert labels[4] == "generated"
        g = world.sim.domain("module")
        assert g.generated_power() == pytest.approx(45.0, rel=1e-15)
        assert g.lost_energy() > 0.0
    doc = json.loads(S.to_json(sc.title, [first, frame]))
    assert len(doc["frames"]) == 2


def test_the_hall_scene_matches_the_oracle(oracle):
    """Scene 16: every frame's pressure field and captured panel bit-identical to the oracle's
    hall released in the same mode and advanced by the same multirate substeps."""
    gpu_case("scene 16 hall")
    from pantometry_amd import scene as S
    from pantometry_amd.core import substeps_for
    sc = S.load(SCENES / "16-a-room-with-a-ceiling.json")
    spec = sc.domains[0]
    world = S.build(sc)
    o = oracle.OracleHall(spec["width_m"], spec["height_m"], spec["depth_m"], spec["nodes_across"],
                          343.0, 1.204)
    o.release(o.mode(tuple(spec["mode"]), spec["amplitude_pa"]))
    dt = sc.duration_s / sc.frames
    for n in range(sc.frames):
        world.advance(dt)
        m = substeps_for(dt, o.max_stable_dt())
        for _ in range(m):
            o.step(dt / m)
        g = world.sim.domain(spec["name"])
        assert np.array_equal(bits(g.fields()[0]), bits(o.p)), n
    frame = world.capture()
    panel = frame.panels[0]
    assert panel.unit == "Pa" and (panel.nx, panel.ny, panel.nz) == (o.nx, o.ny, o.nz)
    expect = np.array([o.sample(*pt) for pt in world._pts[spec["name"]]])
    assert np.array_equal(bits(panel.values), bits(expect))
    assert frame.readings[0][1] == "peak"


def test_scenes_off_the_path_are_refused():
    gpu_case("scene refusals")
    from pantometry_amd import scene as S
    with pytest.raises(S.SceneError, match="composites"):
        S.build(SCENES / "22-wax-in-an-aluminium-matrix.json")
    base = json.loads((SCENES / "15-a-hot-spot-in-a-block.json").read_text())
    with pytest.raises(S.SceneError, match="'orbit' domain"):
        S.build(dict(base, domains=[{"kind": "orbit", "name": "o"}]))
    block = base["domains"][0]
    for change, words in [({"regions": [{"material": "void", "from": [0, 0, 0],
                                         "to": [1, 1, 1], "initial_c": 5.0}]},
                           "cannot start at a temperature"),
                          ({"regions": [{"material": "copper", "from": [0, 0, 0],
                                         "to": [10, 1, 1]}]}, "selects no cells"),
                          ({"parts": [{"stl": "a.stl", "material": "copper"}]}, "STL"),
                          ({"cooling": [{"face": "z-min", "ambient_c": 20.0,
                                         "convection_w_per_m2_k": 10.0, "area_cm2": 1.0},
                                        {"face": "z-min", "ambient_c": 20.0,
                                         "convection_w_per_m2_k": 10.0, "area_cm2": 1.0}]},
                           "already cooled")]:
        bad = dict(base, domains=[dict(block, **change)])
        with pytest.raises(S.SceneError, match=words):
            S.build(bad)


def _march(world, dt, frames, others):
    """Advance the built scene and, beside it, the oracle domains by the same multirate
    substeps, with the heaters' heat taken by share on a bus of the same rules."""
    from pantometry_amd.core import Exchange, HEAT, substeps_for
    bus = Exchange()
    for n in range(frames):
        world.advance(dt)
        bus.covering(dt)
        for kind, o in others:
            if kind == "heater":
                o.step(0.0, dt, bus)
                continue
            m = substeps_for(dt, o.max_stable_dt())
            h = dt / m
            for _ in range(m):
                if kind in ("block", "bar"):
                    o.step(h, bus.take_share(HEAT, h))
                else:
                    o.step(h)
        yield n


@pytest.mark.parametrize("name", ["01-room-mode", "03-room-pulse"])
def test_room_scenes_match_the_oracle(oracle, name):
    """Scenes 01 and 03: the room released in a mode or as a pulse, every frame's pressure field
    and the captured panel bit-identical to the room oracle's."""
    import math
    gpu_case(f"scene {name}")
    from pantometry_amd import scene as S
    sc = S.load(SCENES / f"{name}.json")
    spec = sc.domains[0]
    world = S.build(sc)
    o = oracle.OracleRoom(spec["width_m"], spec["height_m"], spec["cells_across"], 1.0, 343.0,
                          1.204)
    rel = spec["release"]
    if rel["as"] == "mode":
        o.release(o.mode2(rel["nx"], rel["ny"], rel["amplitude_pa"]))
    else:
        r = max(rel["radius_m"], 1e-9)
        o.release(np.array([rel["amplitude_pa"] * math.exp(
            -(((i % o.nx) * o.dx - rel["x_m"]) ** 2 + ((i // o.nx) * o.dx - rel["y_m"]) ** 2)
            / (r * r)) for i in range(o.nx * o.ny)]))
    g = world.sim.domain(spec["name"])
    assert np.array_equal(bits(g.fields()[0]), bits(o.p))
    for n in _march(world, sc.duration_s / sc.frames, sc.frames, [("room", o)]):
        assert np.array_equal(bits(g.fields()[0]), bits(o.p)), n
    frame = world.capture()
    panel = frame.panels[0]
    assert panel.unit == "Pa" and panel.nz == 1
    expect = np.array([o.sample(*pt) for pt in world._pts[spec["name"]]])
    assert np.array_equal(bits(panel.values), bits(expect))
    assert frame.readings == [(spec["name"], "peak", g.peak_pressure(), "Pa")]


def test_the_heater_and_bar_scene_matches_the_oracle(oracle):
    """Scene 04: a heater's reserve spent onto a bar, the bar's cells bit-identical to the bar
    oracle's each frame, the heater's reserve and the panel the same."""
    gpu_case("scene 04 heater and bar")
    from pantometry_amd import scene as S
    sc = S.load(SCENES / "04-heater-and-bar.json")
    world = S.build(sc)
    hs, bs = sc.domains
    heater = S.Heater(hs["name"], hs["watts"], hs["reserve_j"])
    al = S.Palette({}).get("x", "aluminium")
    o = oracle.OracleBar1D(bs["cells"], bs["length_mm"] / bs["cells"] * 1e-3, bs["area_mm2"] * 1e-6,
                           al.density, al.conductivity, al.specific_heat, bs["initial_c"] + 273.15)
    g = world.sim.domain(bs["name"])
    for n in _march(world, sc.duration_s / sc.frames, sc.frames, [("heater", heater), ("bar", o)]):
        assert np.array_equal(bits(g.cells()), bits(o.cells)), n
        assert world.sim.domain(hs["name"]).reserve == heater.reserve
    frame = world.capture()
    panel = [p for p in frame.panels if p.name == bs["name"]][0]
    expect = np.array([o.sample(pt[0]) for pt in world._pts[bs["name"]]])
    assert np.array_equal(bits(panel.values), bits(expect))
    assert g.absorbed_energy() == pytest.approx(hs["reserve_j"], rel=1e-12)


@pytest.mark.parametrize("name,frames", [("20-melting-a-block-of-ice", 4),
                                         ("21-a-wax-thermal-buffer", 3),
                                         ("23-a-part-radiating-to-its-lid", 5)])
def test_melting_and_void_scenes_match_the_oracle(oracle, name, frames):
    """Scenes 20 and 21 (a heater melting ice, and a two-phase wax), and 23 (a part radiating
    across void to its lid): every cell and melted fraction bit-identical to the oracle's."""
    gpu_case(f"scene {name}")
    from pantometry_amd import scene as S
    sc = S.load(SCENES / f"{name}.json")
    world = S.build(sc)
    palette = S.Palette(sc.materials)
    others = []
    for spec in sc.domains:
        if spec["kind"] == "heater":
            others.append(("heater", S.Heater(spec["name"], spec["watts"], spec["reserve_j"])))
        else:
            others.append(("block", oracle_block(oracle, spec, palette)))
    for n in _march(world, sc.duration_s / sc.frames, frames, others):
        for (kind, o), spec in zip(others, sc.domains):
            if kind != "block":
                continue
            g = world.sim.domain(spec["name"])
            assert np.array_equal(bits(g.cells()), bits(o.cells)), (name, n)
            if o.phase:
                assert np.array_equal(bits(g.melted_fractions()), bits(o.melted)), (name, n)
    frame = world.capture()
    for (kind, o), spec in zip(others, sc.domains):
        if kind != "block":
            continue
        panel = [p for p in frame.panels if p.name == spec["name"]][0]
        expect = np.array([o.sample(*pt) for pt in world._pts[spec["name"]]])
        assert np.array_equal(np.isnan(panel.values), np.isnan(expect))
        fin = ~np.isnan(expect)
        assert np.array_equal(bits(panel.values[fin]), bits(expect[fin]))
        labels = [r[1] for r in frame.readings if r[0] == spec["name"]]
        assert ("melted" in labels) == bool(o.phase and (o.latent > 0).any())
