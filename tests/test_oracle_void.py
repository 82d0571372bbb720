"""Void cells and radiating gaps in the Solid3D oracle, pinned against the reference's own tests
(crates/pantometry-thermal/tests/a_gap_that_radiates.rs, a_grid_with_nothing_in_it.rs,
nothing_is_not_a_temperature.rs), the parallel-plate closed form, and by-hand restatements of
find_gaps and the exchange."""
import math

import numpy as np
import pytest

from oracle.oracle import OracleSolid3D

SIGMA = 5.670374419e-8
COPPER = (8.96 * 1000.0, 401.0, 385.0, 0.04)
PLA = (1.24 * 1000.0, 0.13, 1800.0, 0.90)
C0 = 273.15


def block(counts, dx, initial_c, substance=COPPER, emissivity=None):
    d, k, c, e = substance
    return OracleSolid3D("b", d, k, c, counts, dx, initial_c + C0,
                         e if emissivity is None else emissivity)


def mask_of(counts, pred):
    nx, ny, nz = counts
    m = np.zeros(nx * ny * nz, dtype=bool)
    for kk in range(nz):
        for j in range(ny):
            for i in range(nx):
                m[i + nx * (j + ny * kk)] = pred(i, j, kk)
    return m


def plates(gap, emissivity, hot_c, cold_c):
    """Two single-cell copper plates with `gap` empty cells between them, hot at k = 0
    (a_gap_that_radiates.rs:21-36)."""
    n = 2 + gap
    b = block((1, 1, n), 5e-3, cold_c, emissivity=emissivity)
    b.cells[0] = hot_c + C0
    return b.empty(mask_of((1, 1, n), lambda i, j, k: 0 < k < n - 1))


def test_one_step_moves_what_the_parallel_plate_form_says():
    # a_gap_that_radiates.rs:46-74.
    e = 0.8
    b = plates(1, e, 500.0, 20.0)
    hot, cold = b.temperature_at(0, 0, 0), b.temperature_at(0, 0, 2)
    capacity = b.cap[0]
    b.step(1e-4)
    moved = abs(cold - b.temperature_at(0, 0, 2)) * capacity
    expect = SIGMA * 25e-6 * (hot ** 4 - cold ** 4) / (2.0 / e - 1.0) * 1e-4
    assert abs(moved - expect) / expect < 1e-9


def test_a_gap_conserves_exactly():
    # a_gap_that_radiates.rs:82-103.
    b = plates(1, 0.9, 600.0, 20.0)
    opening = b.ledger()[0]
    dt = b.max_stable_dt()
    for _ in range(5000):
        b.step(dt)
    assert b.temperature_at(0, 0, 2) > 30.0 + C0
    assert abs(b.ledger()[0] - opening) / max(abs(opening), 1.0) < 1e-12


def test_an_isolated_pair_levels_and_keeps_its_mean():
    # a_gap_that_radiates.rs:111-131, over 40000 steps rather than 400000 (the pair is within a
    # kelvin of level by then; the mean is conserved at every step).
    b = plates(1, 0.9, 400.0, 0.0)
    dt = b.max_stable_dt()
    for _ in range(40000):
        b.step(dt)
    hot, cold = b.temperature_at(0, 0, 0), b.temperature_at(0, 0, 2)
    assert abs(hot - cold) < 1.0
    assert abs(0.5 * (hot + cold) - 0.5 * (400.0 + 2 * C0)) < 1e-6


def test_a_wide_gap_carries_the_same_because_the_sides_are_mirrors():
    # a_gap_that_radiates.rs:150-172.
    def step_once(gap):
        b = plates(gap, 0.8, 500.0, 20.0)
        before = b.temperature_at(0, 0, 1 + gap)
        b.step(1e-4)
        return b.temperature_at(0, 0, 1 + gap) - before

    narrow, wide = step_once(1), step_once(5)
    assert abs(narrow - wide) / narrow < 1e-12


def test_a_hot_pair_gets_a_shorter_limit_and_does_not_overshoot():
    # a_gap_that_radiates.rs:200-222.
    cool = plates(1, 0.9, 40.0, 20.0).max_stable_dt()
    hot = plates(1, 0.9, 1200.0, 20.0).max_stable_dt()
    assert hot < cool
    b = plates(1, 0.9, 1200.0, 20.0)
    dt = b.max_stable_dt()
    for _ in range(200):
        b.step(dt)
        assert b.temperature_at(0, 0, 0) >= b.temperature_at(0, 0, 2)


def _bar(n, hot_cells, substance=COPPER, emissivity=None):
    b = block((1, 1, n), 5e-3, 20.0, substance, emissivity)
    for k in range(hot_cells):
        b.cells[k] = 300.0 + C0
    return b


def _march(b, steps):
    dt = b.max_stable_dt()
    for _ in range(steps):
        b.step(dt)


def test_void_breaks_the_path_and_a_poor_conductor_only_slows_it():
    # a_grid_with_nothing_in_it.rs:44-86.
    solid = _bar(5, 2)
    _march(solid, 4000)
    insulated = _bar(5, 2)
    insulated.fill(PLA[0], PLA[1], PLA[2], mask_of((1, 1, 5), lambda i, j, k: k == 2), PLA[3])
    _march(insulated, 4000)
    gapped = _bar(5, 2).empty(mask_of((1, 1, 5), lambda i, j, k: k == 2))
    _march(gapped, 4000)
    copper, plastic, nothing = (b.temperature_at(0, 0, 4) for b in (solid, insulated, gapped))
    start = 20.0 + C0
    assert copper > start + 50.0
    assert plastic > start + 1.0
    assert nothing > start
    assert plastic < copper and nothing < plastic


def test_a_gap_between_mirrors_carries_nothing():
    # a_grid_with_nothing_in_it.rs:94-120.
    b = _bar(5, 2, emissivity=0.0).empty(mask_of((1, 1, 5), lambda i, j, k: k == 2))
    assert len(b.gap_a) == 0
    _march(b, 4000)
    assert abs(b.temperature_at(0, 0, 4) - (20.0 + C0)) < 1e-9


def test_a_void_cell_has_no_temperature_and_no_vote():
    # a_grid_with_nothing_in_it.rs:127-156.
    b = block((4, 1, 1), 5e-3, 100.0)
    b.cells[0] = 300.0 + C0
    solid_mean = b.statistics()[1]
    b.empty(mask_of((4, 1, 1), lambda i, j, k: i == 3))
    assert b.is_void(3, 0, 0) and not b.is_void(0, 0, 0)
    assert b.void_cells() == 1
    assert math.isnan(b.temperature_at(3, 0, 0))
    peak, mean, coldest = b.statistics()
    assert mean > solid_mean
    assert math.isfinite(peak) and math.isfinite(coldest)


def test_void_takes_no_share_of_the_bus():
    # a_grid_with_nothing_in_it.rs:164-200.
    def run(void):
        b = block((4, 1, 1), 5e-3, 20.0)
        if void:
            b.empty(mask_of((4, 1, 1), lambda i, j, k: i >= 2))
        opening = b.ledger()[0]
        b.step(1e-6, gained=1000.0)
        return b, b.ledger()[0] - opening

    whole, gained_whole = run(False)
    half, gained_half = run(True)
    assert abs(gained_whole - 1000.0) < 1e-6 and abs(gained_half - 1000.0) < 1e-6
    rise_whole = whole.statistics()[1] - (20.0 + C0)
    rise_half = half.statistics()[1] - (20.0 + C0)
    assert abs(rise_half / rise_whole - 2.0) < 1e-9


def test_a_block_of_nothing_is_not_a_division_by_zero():
    # a_grid_with_nothing_in_it.rs:207-226.
    b = block((2, 2, 2), 5e-3, 20.0).empty(np.ones(8, dtype=bool))
    assert b.void_cells() == 8
    assert math.isinf(b.max_stable_dt())
    assert b.ledger()[0] == 0.0
    b.step(1.0)
    assert math.isnan(b.statistics()[1])


def test_a_sample_in_the_gap_has_no_value_and_one_on_the_metal_is_exact():
    # nothing_is_not_a_temperature.rs:19-96.
    b = block((4, 1, 1), 10e-3, 20.0)
    b.cells[0] = 300.0 + C0
    b.empty(mask_of((4, 1, 1), lambda i, j, k: i in (1, 2)))

    def at(mm):
        return b.sample(mm * 1e-3, 5e-3, 5e-3)

    assert abs(at(5.0) - (300.0 + C0)) < 1e-12
    assert abs(at(35.0) - (20.0 + C0)) < 1e-12
    for mm in (15.0, 20.0, 25.0, 11.0):
        assert math.isnan(at(mm))
    assert math.isfinite(at(9.0))
    face = at(10.0)
    assert math.isnan(face) or abs(face - (300.0 + C0)) < 1e-12
    assert math.isnan(b.sample(math.nan, 0.0, 0.0))


def _pairs_by_hand(counts, voids, emis, dx):
    """find_gaps (solid.rs:791-847) restated: every line along x, then y, then z."""
    nx, ny, nz = counts
    out = []

    def walk(line):
        last, between = None, False
        for c in line:
            if voids[c]:
                between = last is not None
                continue
            if last is not None and between and emis[last] > 0.0 and emis[c] > 0.0:
                r = 1.0 / emis[last] + 1.0 / emis[c] - 1.0
                if r > 0.0:
                    out.append((last, c, SIGMA * (dx * dx) / r))
            last, between = c, False

    for k in range(nz):
        for j in range(ny):
            walk([i + nx * (j + ny * k) for i in range(nx)])
    for k in range(nz):
        for i in range(nx):
            walk([i + nx * (j + ny * k) for j in range(ny)])
    for j in range(ny):
        for i in range(nx):
            walk([i + nx * (j + ny * k) for k in range(nz)])
    return out


def test_the_gaps_and_their_exchange_to_the_bit():
    # A ragged 7x6x5 block of two substances with a void slab, a void column and scattered void
    # cells: the oracle's pairs equal find_gaps restated by hand, and one step's exchange
    # (read off the step's input, applied pair by pair after the sweep) equals the restatement
    # in Python floats to the bit.
    rng = np.random.default_rng(17)
    counts = (7, 6, 5)
    nx, ny, nz = counts
    b = block(counts, 2e-3, 20.0)
    b.fill(PLA[0], PLA[1], PLA[2], mask_of(counts, lambda i, j, k: (i + j + k) % 3 == 0), PLA[3])
    b.cells = C0 + rng.uniform(20.0, 900.0, size=b.n)
    voids = mask_of(counts, lambda i, j, k: k == 2 or (i == 3 and j == 2)) | (rng.random(b.n) < 0.1)
    b.empty(voids)
    want = _pairs_by_hand(counts, b.voids, b.emis, b.dx)
    got = list(zip(b.gap_a.tolist(), b.gap_b.tolist(), b.gap_coef.tolist()))
    assert got == want and len(want) > 20
    before = b.cells.copy()
    dt = 0.5 * b.max_stable_dt()
    b.step(dt)
    # The sweep alone, on a copy without gaps, then the exchange by hand.
    plain = block(counts, 2e-3, 20.0)
    plain.fill(PLA[0], PLA[1], PLA[2], mask_of(counts, lambda i, j, k: (i + j + k) % 3 == 0), PLA[3])
    plain.empty(voids)
    plain.gap_a = plain.gap_a[:0]
    plain.gap_b = plain.gap_b[:0]
    plain.gap_coef = plain.gap_coef[:0]
    plain.cells = before.copy()
    plain.on[0] = 0
    plain.step(dt)
    cells = [float(v) for v in plain.cells]
    for a, c, coef in want:
        watts = coef * ((before[a] * before[a]) * (before[a] * before[a])
                        - (before[c] * before[c]) * (before[c] * before[c]))
        if watts == 0.0:
            continue
        joules = watts * dt
        cells[a] += -joules / b.cap[a]
        cells[c] += joules / b.cap[c]
    assert np.array_equal(b.cells, np.array(cells))


def test_the_view_factor_is_the_parallel_rectangle_form():
    # how_much_of_the_gap_is_an_assumption.rs:16-90.
    from pantometry_amd.thermal import GapPatch

    def patch(span, distance):
        return GapPatch(1, (span, span), distance, True)

    assert abs(patch(1.0, 1.0).view_factor() - 0.19982) < 5e-6
    previous = 0.0
    for n in range(1, 7):
        f = patch(10.0 ** n, 1.0).view_factor()
        assert f > previous
        previous = f
    assert previous > 0.999
    worst = 0.0
    for c in (10.0, 30.0, 100.0, 300.0, 1000.0, 10000.0, 1e6):
        off = abs(patch(1.0, c).view_factor() / (1.0 / (math.pi * c * c)) - 1.0)
        assert off <= 2.0 / (3.0 * c * c) * 1.05 + 1e-9
        if c >= 300.0:
            worst = max(worst, off)
    assert worst < 1e-5
    assert patch(0.0, 1.0).view_factor() == 0.0 and patch(1.0, 0.0).view_factor() == 1.0
    assert math.isfinite(patch(1.0, math.nan).view_factor())
    assert math.isfinite(patch(math.nan, 1.0).view_factor())


def test_a_patch_is_a_sheet_and_clearances_stay_separate():
    # how_much_of_the_gap_is_an_assumption.rs:93-180, over the oracle's pairs.
    from pantometry_amd.thermal import gap_patches_of
    counts = (6, 6, 7)
    b = block(counts, 8e-3, 20.0)
    b.empty(mask_of(counts, lambda i, j, k: k < 3 and (i in (0, 5) or j in (0, 5))))
    b.empty(mask_of(counts, lambda i, j, k: k in (3, 4)))
    pairs = list(zip(b.gap_a.tolist(), b.gap_b.tolist(), b.gap_coef.tolist()))
    patches = gap_patches_of(pairs, counts, 8e-3)
    assert len(patches) == 1
    p = patches[0]
    assert p.pairs == 16 and p.rectangular
    assert abs(p.span[0] - 0.032) < 1e-12 and abs(p.span[1] - 0.032) < 1e-12
    assert abs(p.distance - 0.016) < 1e-12
    assert abs(p.view_factor() - 0.41525) < 5e-5
    two = block((5, 1, 3), 5e-3, 20.0).empty(mask_of((5, 1, 3), lambda i, j, k: k == 1 and i in (0, 4)))
    ps = sorted(gap_patches_of(list(zip(two.gap_a.tolist(), two.gap_b.tolist(),
                                        two.gap_coef.tolist())), (5, 1, 3), 5e-3),
                key=lambda q: q.pairs)
    assert len(ps) == 2
    for q in ps:
        assert q.pairs == 1 and abs(q.distance - 0.005) < 1e-12
        assert abs(q.view_factor() - 0.19982) < 5e-6
    solid = block((4, 4, 4), 5e-3, 20.0)
    assert gap_patches_of([], (4, 4, 4), 5e-3) == [] and len(solid.gap_a) == 0
