"""Scene files and run files on the host: what the loader accepts and refuses (World::build's
checks, crates/pantometry-world/src/lib.rs:400-428 and 2846-2864, the palette of lib.rs:1072-1150)
and the run file's number spelling (Rust's `{:.6e}`, crates/pantometry-view/src/data.rs:140-156).
Scene fixtures are copies of the reference's own scene files (tests/golden/scenes)."""
import json
from pathlib import Path

import pytest

from pantometry_amd import scene as S

SCENES = Path(__file__).resolve().parent / "golden" / "scenes"


def test_the_reference_scenes_load():
    for name in ("01-room-mode", "02-room-higher-mode", "03-room-pulse", "04-heater-and-bar",
                 "15-a-hot-spot-in-a-block", "16-a-room-with-a-ceiling",
                 "19-a-coating-stops-the-heat", "20-melting-a-block-of-ice",
                 "21-a-wax-thermal-buffer", "22-wax-in-an-aluminium-matrix",
                 "23-a-part-radiating-to-its-lid", "24-a-power-module-junction-to-ambient"):
        sc = S.load(SCENES / f"{name}.json")
        assert sc.frames > 0 and sc.duration_s > 0.0


def test_bad_files_are_refused():
    base = json.loads((SCENES / "15-a-hot-spot-in-a-block.json").read_text())
    for change, words in [({"format": 0}, "format 0"), ({"format": 2}, "format 2"),
                          ({"frames": 0}, "frame"), ({"duration_s": -1.0}, "positive"),
                          ({"domains": []}, "domain"), ({"schedule": "iterative"}, "schedule"),
                          ({"speed": 3}, "unknown")]:
        bad = dict(base, **change)
        with pytest.raises(S.SceneError, match=words):
            S.load(bad)


def test_the_palette():
    p = S.Palette({"silicon": {"name": "Si", "density": 2329.0,
                               "thermal": {"conductivity": 148.0, "specific_heat": 705.0,
                                           "expansion": 2.6e-6, "emissivity": 0.6}}})
    si = p.get("x", "silicon")
    assert (si.density, si.conductivity, si.specific_heat, si.emissivity) == (2329.0, 148.0,
                                                                              705.0, 0.6)
    cu = p.get("x", "copper")
    assert (cu.density, cu.conductivity, cu.specific_heat, cu.emissivity) == (8960.0, 401.0,
                                                                              385.0, 0.04)
    ice = p.get("x", "ice")  # Substance::ice, substance.rs:613-650: one-phase, 333.55 kJ/kg
    assert ice.fusion == S.Fusion(273.15, 333.55 * 1e3)
    wax = S.Palette({"wax": {"density": 814.0,
                             "thermal": {"conductivity": 0.358, "specific_heat": 1934.0},
                             "fusion": {"melting_point": 301.3, "latent_heat": 244000.0,
                                        "liquid": {"conductivity": 0.152,
                                                   "specific_heat": 2196.0}}}}).get("x", "wax")
    assert wax.fusion == S.Fusion(301.3, 244000.0, (0.152, 2196.0))
    for bad, words in [({"melting_point": -1.0, "latent_heat": 1.0}, "melting point"),
                       ({"melting_point": 300.0, "latent_heat": 1.0, "solidus": 2.0}, "unknown")]:
        with pytest.raises(S.SceneError, match=words):
            S.Palette({"w": {"density": 1.0, "thermal": {"conductivity": 1.0,
                                                         "specific_heat": 1.0},
                             "fusion": bad}})
    with pytest.raises(S.SceneError, match="no material"):
        p.get("x", "unobtainium")
    with pytest.raises(S.SceneError, match="shadows"):
        S.Palette({"copper": {"density": 1.0, "thermal": {"conductivity": 1.0,
                                                           "specific_heat": 1.0}}})
    assert S.Palette({"a": {"density": 1.0}}).unused() == ["a"]


@pytest.mark.parametrize("x,rust", [(293.15, "2.931500e2"), (0.001, "1.000000e-3"),
                                    (-45.0, "-4.500000e1"), (0.0, "0.000000e0"),
                                    (1.0, "1.000000e0"), (123456789.0, "1.234568e8"),
                                    (float("nan"), "null"), (float("inf"), "null")])
def test_numbers_are_spelled_as_rust_spells_them(x, rust):
    assert S._number(x) == rust


def test_a_run_file_is_json():
    frames = [S.Frame(0.0, [S.Panel("b", "K", 2, 1, 1, [300.0, float("nan")])],
                      [("b", "peak", 26.85, "C")]),
              S.Frame(0.5, [S.Panel("b", "K", 2, 1, 1, [301.0, 302.0])],
                      [("b", "peak", 28.85, "C")])]
    doc = json.loads(S.to_json('a "quoted" title', frames))
    assert doc["title"] == 'a "quoted" title'
    assert [f["t"] for f in doc["frames"]] == [0.0, 0.5]
    assert doc["frames"][0]["panels"][0]["values"] == [300.0, None]
    assert doc["frames"][1]["readings"][0] == {"domain": "b", "label": "peak", "unit": "C",
                                               "value": 28.85}


def test_rooms_bars_and_heaters_parse_as_the_reference_states_them():
    # DomainSpec::Room / Bar / Heater (crates/pantometry-world/src/lib.rs:464-488, 601-617) and
    # Release (lib.rs:1301-1337): unknown fields and release kinds are refused.
    base = json.loads((SCENES / "03-room-pulse.json").read_text())
    room = base["domains"][0]
    for change, words in [({"release": {"as": "spiral"}}, "spiral"),
                          ({"release": {"as": "mode", "nx": 1, "ny": 1, "amplitude_pa": 1.0,
                                        "phase": 0.0}}, "unknown"),
                          ({"depth_m": 1.0}, "unknown")]:
        with pytest.raises(S.SceneError, match=words):
            S.build_room(dict(room, **change))
    with pytest.raises(S.SceneError, match="unknown"):
        S.build_heater({"kind": "heater", "name": "h", "watts": 1.0, "reserve_j": 1.0,
                        "volts": 3.0})
    h = S.Heater("h", 2.0, 3.0)
    from pantometry_amd.core import Exchange, HEAT
    bus = Exchange()
    h.step(0.0, 1.0, bus)
    h.step(0.0, 1.0, bus)  # heater.rs:54-59: min(watts*dt, reserve), never below zero
    assert bus.peek(HEAT) == 3.0 and h.reserve == 0.0
    assert h.readings() == [("h", "reserve", 0.0, "J")]


def test_composites_are_refused_by_name():
    with pytest.raises(S.SceneError, match="composites"):
        S.build(SCENES / "22-wax-in-an-aluminium-matrix.json")
