"""Void cells and radiating gaps on the GPU against the Solid3D oracle: bit-exact cells through
the sweep, the films, the sources, the bus heat and the gap exchange; the same pairs and the
same limit; and the reference's own gap and void tests run on the device
(crates/pantometry-thermal/tests/a_gap_that_radiates.rs, a_grid_with_nothing_in_it.rs,
nothing_is_not_a_temperature.rs)."""
import math

import numpy as np
import pytest

from tests.conftest import gpu_case

pytestmark = pytest.mark.gpu

C0 = 273.15
COPPER = ("Cu", 8.96 * 1000.0, 401.0, 385.0, 0.04)
PLA = ("PLA", 1.24 * 1000.0, 0.13, 1800.0, 0.90)
STEEL = ("steel", 7.85 * 1000.0, 45.0, 490.0, 0.35)


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).reshape(-1).view(np.uint64)


def sub(t, emissivity=None):
    import pantometry_amd as pm
    return pm.Substance.of(t[0], t[1], t[2], t[3], t[4] if emissivity is None else emissivity)


def pair(counts, dx, initial, first=COPPER, emissivity=None):
    import pantometry_amd as pm
    from oracle.oracle import OracleSolid3D
    e = first[4] if emissivity is None else emissivity
    g = pm.Solid3D("b", sub(first, e), counts, dx, initial)
    o = OracleSolid3D("b", first[1], first[2], first[3], counts, dx, initial, e)
    return g, o


def mask_of(counts, pred):
    nx, ny, nz = counts
    m = np.zeros(nx * ny * nz, dtype=np.uint8)
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                m[i + nx * (j + ny * k)] = 1 if pred(i, j, k) else 0
    return m


# 70 x 20 x 70: at least 256 column tiles (64 x 16) x planes, so the sweep marches z columns.
@pytest.mark.parametrize("counts", [(7, 6, 5), (33, 9, 12), (70, 20, 70)], ids=str)
def test_a_block_with_void_gaps_films_and_sources_is_bit_exact(counts):
    gpu_case(f"void parity {counts}")
    import pantometry_amd as pm
    rng = np.random.default_rng(sum(counts))
    n = counts[0] * counts[1] * counts[2]
    g, o = pair(counts, 2e-3, 300.0)
    second = mask_of(counts, lambda i, j, k: (i + 2 * j + k) % 3 == 0)
    g.fill(sub(PLA), second)
    o.fill(PLA[1], PLA[2], PLA[3], second, PLA[4])
    third = mask_of(counts, lambda i, j, k: i == 1)
    g.fill(sub(STEEL), third)
    o.fill(STEEL[1], STEEL[2], STEEL[3], third, STEEL[4])
    start = C0 + rng.uniform(20.0, 700.0, size=n)
    g.set_cells(start)
    o.set_cells(start)
    voids = (mask_of(counts, lambda i, j, k: k == 2 or (i == 3 and j == 2))
             | (rng.random(n) < 0.08).astype(np.uint8))
    g.empty(voids)
    o.empty(voids)
    assert g.void_cells() == o.void_cells()
    want_pairs = list(zip(o.gap_a.tolist(), o.gap_b.tolist(), o.gap_coef.tolist()))
    assert g.gap_pairs() == want_pairs and len(want_pairs) > 10
    env = pm.Environment(290.0, 12.0, 4e-4)
    g.losing_from(pm.Face.XMIN, env)
    o.losing_from(0, 290.0, 12.0, 4e-4)
    g.losing_from(pm.Face.ZMAX, pm.Environment(310.0, 3.0, 1e-3))
    o.losing_from(5, 310.0, 3.0, 1e-3)
    heat = mask_of(counts, lambda i, j, k: i == counts[0] - 1)
    g.dissipating(2.5, heat)
    o.dissipating(2.5, heat)
    assert g.max_stable_dt() == o.max_stable_dt()
    dt = 0.7 * o.max_stable_dt()
    for s in range(5):
        bus = pm.Exchange()
        gained = 0.05 * s
        if gained:
            bus.publish(pm.HEAT, gained)
            bus.covering(dt)
        g.step(0.0, dt, bus)
        o.step(dt, gained)
        assert np.array_equal(bits(g.cells()), bits(o.cells)), f"step {s}"
    stored, lost, supplied = g.books()
    assert supplied == o.supplied
    assert abs(lost - o.lost) <= 1e-12 * max(abs(o.lost), 1e-300)
    g.advance(dt, 40, tolerance=1e-9)
    for _ in range(40):
        o.step(dt)
    assert np.array_equal(bits(g.cells()), bits(o.cells))
    assert g.max_stable_dt() == o.max_stable_dt()
    peak, mean, coldest = o.statistics()
    assert (g.peak_temperature(), g.mean_temperature(), g.coldest_temperature()) == (peak, mean, coldest)
    nx, ny, nz = counts
    for (i, j, k) in ((3, 2, 0), (0, 0, 2), (nx - 1, ny - 1, nz - 1)):
        a, b = g.temperature_at(i, j, k), o.temperature_at(i, j, k)
        assert (math.isnan(a) and math.isnan(b)) or a == b
    pts = rng.uniform(-1e-3, 2e-3 * max(counts) + 1e-3, size=(300, 3))
    got = g.sample(pts)
    want = np.array([o.sample(*p) for p in pts])
    assert np.array_equal(np.isnan(got), np.isnan(want)) and np.isnan(got).any()
    fin = ~np.isnan(want)
    assert np.array_equal(bits(got[fin]), bits(want[fin]))


def _plates(gap, emissivity, hot_c, cold_c):
    import pantometry_amd as pm
    n = 2 + gap
    b = pm.Solid3D("plates", sub(COPPER, emissivity), (1, 1, n), 5e-3, cold_c + C0)
    b.set_temperature(0, 0, 0, hot_c + C0)
    return b.empty(lambda i, j, k: 0 < k < n - 1)


def test_the_gap_tests_on_the_device():
    # a_gap_that_radiates.rs:46-222.
    gpu_case("gap reference tests")
    import pantometry_amd as pm
    sigma = 5.670374419e-8
    e = 0.8
    b = _plates(1, e, 500.0, 20.0)
    hot, cold = b.temperature_at(0, 0, 0), b.temperature_at(0, 0, 2)
    capacity = b.heat_capacity() / 2.0
    b.step(0.0, 1e-4, pm.Exchange())
    moved = abs(cold - b.temperature_at(0, 0, 2)) * capacity
    expect = sigma * 25e-6 * (hot ** 4 - cold ** 4) / (2.0 / e - 1.0) * 1e-4
    assert abs(moved - expect) / expect < 1e-9

    c = _plates(1, 0.9, 600.0, 20.0)
    opening = c.ledger().get(pm.ENERGY) or 0.0
    c.advance(c.max_stable_dt(), 5000, tolerance=1e-9)
    assert c.temperature_at(0, 0, 2) > 30.0 + C0
    closing = c.ledger().get(pm.ENERGY) or 0.0
    assert abs(closing - opening) / max(abs(opening), 1.0) < 1e-12

    # 40000 steps rather than the reference's 400000: within a kelvin of level already.
    level = _plates(1, 0.9, 400.0, 0.0)
    level.advance(level.max_stable_dt(), 40_000, tolerance=1e-9)
    h, k = level.temperature_at(0, 0, 0), level.temperature_at(0, 0, 2)
    assert abs(h - k) < 1.0
    assert abs(0.5 * (h + k) - 0.5 * (400.0 + 2 * C0)) < 1e-6

    def step_once(gap):
        p = _plates(gap, 0.8, 500.0, 20.0)
        before = p.temperature_at(0, 0, 1 + gap)
        p.step(0.0, 1e-4, pm.Exchange())
        return p.temperature_at(0, 0, 1 + gap) - before

    assert abs(step_once(1) - step_once(5)) / step_once(1) < 1e-12

    assert _plates(1, 0.9, 1200.0, 20.0).max_stable_dt() < _plates(1, 0.9, 40.0, 20.0).max_stable_dt()
    hot_pair = _plates(1, 0.9, 1200.0, 20.0)
    dt = hot_pair.max_stable_dt()
    for _ in range(200):
        hot_pair.step(0.0, dt, pm.Exchange())
        assert hot_pair.temperature_at(0, 0, 0) >= hot_pair.temperature_at(0, 0, 2)


def test_the_void_tests_on_the_device():
    # a_grid_with_nothing_in_it.rs:44-226 and nothing_is_not_a_temperature.rs:19-96.
    gpu_case("void reference tests")
    import pantometry_amd as pm

    def bar(emissivity=None):
        b = pm.Solid3D("bar", sub(COPPER, emissivity), (1, 1, 5), 5e-3, 20.0 + C0)
        for k in range(2):
            b.set_temperature(0, 0, k, 300.0 + C0)
        return b

    def march(b):
        dt = b.max_stable_dt()
        b.advance(dt, 4000, tolerance=1e-9)
        return b.temperature_at(0, 0, 4)

    copper = march(bar())
    plastic = march(bar().fill(sub(PLA), lambda i, j, k: k == 2))
    nothing = march(bar().empty(lambda i, j, k: k == 2))
    start = 20.0 + C0
    assert copper > start + 50.0 and plastic > start + 1.0 and nothing > start
    assert plastic < copper and nothing < plastic
    mirrors = bar(0.0).empty(lambda i, j, k: k == 2)
    assert mirrors.gap_pairs() == []
    assert abs(march(mirrors) - start) < 1e-9

    part = pm.Solid3D("part", sub(COPPER), (4, 1, 1), 5e-3, 100.0 + C0)
    part.set_temperature(0, 0, 0, 300.0 + C0)
    solid_mean = part.mean_temperature()
    part.empty(lambda i, j, k: i == 3)
    assert part.is_void(3, 0, 0) and not part.is_void(0, 0, 0) and part.void_cells() == 1
    assert math.isnan(part.temperature_at(3, 0, 0))
    assert part.mean_temperature() > solid_mean
    assert math.isfinite(part.peak_temperature()) and math.isfinite(part.coldest_temperature())

    def share(void):
        b = pm.Solid3D("part", sub(COPPER), (4, 1, 1), 5e-3, 20.0 + C0)
        if void:
            b.empty(lambda i, j, k: i >= 2)
        opening = b.ledger().get(pm.ENERGY) or 0.0
        bus = pm.Exchange()
        bus.publish(pm.HEAT, 1000.0)
        b.step(0.0, 1e-6, bus)
        return b, (b.ledger().get(pm.ENERGY) or 0.0) - opening

    whole, gw = share(False)
    half, gh = share(True)
    assert abs(gw - 1000.0) < 1e-6 and abs(gh - 1000.0) < 1e-6
    rw = whole.mean_temperature() - (20.0 + C0)
    rh = half.mean_temperature() - (20.0 + C0)
    assert abs(rh / rw - 2.0) < 1e-9

    nothing_at_all = pm.Solid3D("nothing", sub(COPPER), (2, 2, 2), 5e-3, 20.0 + C0)
    nothing_at_all.empty(lambda i, j, k: True)
    assert nothing_at_all.void_cells() == 8
    assert math.isinf(nothing_at_all.max_stable_dt())
    assert (nothing_at_all.ledger().get(pm.ENERGY) or 0.0) == 0.0
    nothing_at_all.step(0.0, 1.0, pm.Exchange())
    assert math.isnan(nothing_at_all.mean_temperature())

    gap_bar = pm.Solid3D("bar", sub(COPPER), (4, 1, 1), 10e-3, 20.0 + C0)
    gap_bar.set_temperature(0, 0, 0, 300.0 + C0)
    gap_bar.empty(lambda i, j, k: i in (1, 2))
    at = gap_bar.sample([[mm * 1e-3, 5e-3, 5e-3] for mm in (5.0, 35.0, 15.0, 20.0, 25.0, 9.0, 11.0)])
    assert abs(at[0] - (300.0 + C0)) < 1e-12 and abs(at[1] - (20.0 + C0)) < 1e-12
    assert all(math.isnan(v) for v in at[2:5]) and math.isfinite(at[5]) and math.isnan(at[6])


def test_gap_patches_on_the_device():
    # how_much_of_the_gap_is_an_assumption.rs:93-128: the housing's one 4x4 patch, two cells
    # apart, from the device's pairs.
    gpu_case("gap patches")
    import pantometry_amd as pm
    b = pm.Solid3D("housing", sub(COPPER), (6, 6, 7), 8e-3, 20.0 + C0)
    b.empty(lambda i, j, k: k < 3 and (i in (0, 5) or j in (0, 5)))
    b.empty(lambda i, j, k: k in (3, 4))
    patches = b.gap_patches()
    assert len(patches) == 1
    p = patches[0]
    assert p.pairs == 16 and p.rectangular and abs(p.distance - 0.016) < 1e-12
    assert abs(p.view_factor() - 0.41525) < 5e-5
