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
    for name in ("15-a-hot-spot-in-a-block", "16-a-room-with-a-ceiling",
                 "19-a-coating-stops-the-heat", "24-a-power-module-junction-to-ambient"):
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
    with pytest.raises(S.SceneError, match="phase"):
        p.get("x", "ice")
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
