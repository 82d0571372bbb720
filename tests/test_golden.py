"""The committed known-answer fixtures against the CPU oracle and against what the reference's
own tests fix: the RNG digest (rng.rs:295-308), the exact eigen-decay of a mode
(solid_closed_forms.rs:46-72), heat conservation over the run (a_layered_wall.rs:363-416) and
Hall's invariant (hall_closed_forms.rs).  The oracle must replay every fixture bit for bit."""
from __future__ import annotations

import numpy as np
import pytest

from tests.golden_cases import arr, bits, load, num

FX = load()
SOLIDS = ["mode_7x5x3", "scene15", "hetero_16"]


def test_rng_fixture(oracle):
    f = FX["rng"]
    assert int(f["digest"]) == 6_777_642_030_472_145_829          # rng.rs:295-308
    assert oracle.rng_digest(f["seed"], f["draws"]) == int(f["digest"])
    r = oracle.OracleRng(f["seed"])
    assert [format(r.next_u64(), "016x") for _ in range(8)] == f["first"]


def _replay_solid(oracle, f):
    b = oracle.OracleSolid3D("block", f["density"], f["conductivity"], f["specific_heat"],
                             tuple(f["counts"]), f["dx"], f["initial"])
    if "k" in f:
        b.with_conductivity(arr(f["k"]))
    b.set_cells(arr(f["start"]))
    dt = num(f["dt"])
    assert dt == 0.9 * b.max_stable_dt()
    for _ in range(f["steps"]):
        b.step(dt)
    return b


@pytest.mark.parametrize("name", SOLIDS)
def test_solid_fixture_replays_bit_exact(name, oracle):
    f = FX[name]
    b = _replay_solid(oracle, f)
    assert np.array_equal(bits(b.cells), bits(arr(f["final"])))
    assert b.ledger()[0] == num(f["stored_after"])
    before, after = num(f["stored_before"]), num(f["stored_after"])
    # conserved to rounding (a_layered_wall.rs:363-416): every update rounds at the magnitude
    # of the temperature itself, so a few ulps of sum(C |T|) per step bound the drift (a mode
    # about the reference stores almost nothing, so the total cannot set the scale)
    scale = float(np.sum(np.abs(b.cap * arr(f["start"]))))
    assert abs(after - before) <= f["steps"] * 8 * 2.2e-16 * scale


def test_mode_fixture_decays_as_the_closed_form(oracle):
    f = FX["mode_7x5x3"]
    b = oracle.OracleSolid3D("block", f["density"], f["conductivity"], f["specific_heat"],
                             tuple(f["counts"]), f["dx"], f["initial"])
    b.set_cells(arr(f["start"]))
    a0 = b.mode_amplitude(tuple(f["mode"]))
    g = b.mode_amplification(tuple(f["mode"]), num(f["dt"]))
    b.set_cells(arr(f["final"]))
    a = b.mode_amplitude(tuple(f["mode"]))
    assert abs(a - a0 * g ** f["steps"]) <= 1e-12 * abs(a0)
    assert abs(a0 - f["amplitude"]) <= 1e-12


def test_scene15_fixture_matches_the_scene():
    f = FX["scene15"]
    start = arr(f["start"]).reshape(9, 9, 9)
    assert start[4, 4, 4] == 80.0 + 273.15 and start[0, 0, 0] == 20.0 + 273.15
    assert np.count_nonzero(start != 20.0 + 273.15) == 1


def test_hall_fixture_replays_bit_exact(oracle):
    f = FX["hall_13"]
    h = oracle.OracleHall(f["width"], f["width"], f["width"], f["nodes"][0], f["speed"],
                          f["density"])
    assert np.array_equal(bits(h.mode(tuple(f["mode"]), f["amplitude"])), bits(arr(f["start"])))
    h.release(arr(f["start"]))
    dt = num(f["dt"])
    for _ in range(f["steps"]):
        h.step(dt)
    for key, got in (("p", h.p), ("p_prev", h.pp), ("vx", h.vx), ("vy", h.vy), ("vz", h.vz)):
        assert np.array_equal(bits(got), bits(arr(f[key]))), key
    e0, e1 = num(f["energy_before"]), num(f["energy_after"])
    assert abs(e1 - e0) <= 1e-12 * abs(e0)                        # hall_closed_forms.rs
