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


def oracle_block(oracle, spec, palette):
    """The builder's steps (crates/pantometry-world/src/lib.rs:2185-2400) on the oracle."""
    from pantometry_amd import scene as S
    cells = tuple(spec["cells"])
    bulk = palette.get("x", spec.get("material") or "aluminium")
    o = oracle.OracleSolid3D(spec["name"], bulk.density, bulk.conductivity, bulk.specific_heat,
                             cells, spec["cell_mm"] * 1e-3, spec["initial_c"] + 273.15,
                             emissivity=bulk.emissivity)
    for r in spec.get("regions", []):
        sub = palette.get("x", r["material"])
        inside = S._box(cells, r["from"], r["to"])
        o.fill(sub.density, sub.conductivity, sub.specific_heat, inside, emissivity=sub.emissivity)
        if r.get("initial_c") is not None:
            o.cells[inside] = r["initial_c"] + 273.15
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
    with pytest.raises(S.SceneError, match="room"):
        S.build(SCENES / "01-room-mode.json")
    with pytest.raises(S.SceneError, match="heater"):
        S.build(SCENES / "20-melting-a-block-of-ice.json")
    base = json.loads((SCENES / "15-a-hot-spot-in-a-block.json").read_text())
    block = base["domains"][0]
    for change, words in [({"regions": [{"material": "void", "from": [0, 0, 0],
                                         "to": [1, 1, 1]}]}, "void"),
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
