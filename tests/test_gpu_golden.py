"""The gfx950 path against the committed known-answer fixtures, bit for bit."""
from __future__ import annotations

import numpy as np
import pytest

from tests.conftest import gpu_case
from tests.golden_cases import arr, bits, load, num

pytestmark = pytest.mark.gpu

FX = load()


@pytest.mark.parametrize("name", ["mode_7x5x3", "scene15", "hetero_16"])
def test_solid_fixture_on_the_gpu(name):
    gpu_case(f"golden solid {name}")
    from pantometry_amd import Solid3D, Substance
    f = FX[name]
    sub = Substance.of("fixture", f["density"], f["conductivity"], f["specific_heat"])
    b = Solid3D("block", sub, tuple(f["counts"]), f["dx"], f["initial"])
    if "k" in f:
        b.with_conductivity(arr(f["k"]))
    b.set_cells(arr(f["start"]))
    dt = num(f["dt"])
    assert dt == 0.9 * b.max_stable_dt(0.0)
    totals, _ = b.advance(dt, f["steps"], tolerance=1e-12)
    assert np.array_equal(bits(b.cells()), bits(arr(f["final"])))
    # the audit sums in another order than the oracle's left fold: rounding at the magnitude of
    # C |T| per cell (a mode about the reference stores almost nothing, so not relative to it)
    cap = (f["density"] * f["dx"] ** 3) * f["specific_heat"]
    scale = cap * float(np.sum(np.abs(arr(f["final"]))))
    assert abs(totals[-1] - num(f["stored_after"])) <= 1e-14 * scale


def test_hall_fixture_on_the_gpu():
    gpu_case("golden hall_13")
    from pantometry_amd import Hall
    f = FX["hall_13"]
    h = Hall("hall", f["width"], f["width"], f["width"], f["nodes"][0], f["speed"], f["density"])
    h.released_from(arr(f["start"]).reshape(f["nodes"][2], f["nodes"][1], f["nodes"][0]))
    h.advance(num(f["dt"]), f["steps"], tolerance=1e-12)
    p, pp, vx, vy, vz = h.fields()
    for key, got in (("p", p), ("p_prev", pp), ("vx", vx), ("vy", vy), ("vz", vz)):
        assert np.array_equal(bits(got), bits(arr(f[key]))), key
    assert abs(h.energy() - num(f["energy_after"])) <= 1e-12 * abs(num(f["energy_after"]))
