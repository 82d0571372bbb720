"""The one-dimensional oracles pinned: Bar1D against the reference's own bar tests
(crates/pantometry-thermal/src/lib.rs:1175-1290, 1402-1450) and the discrete cosine mode's exact
decay; Tube against its tests (crates/pantometry-acoustic/src/lib.rs:566-990) and the
three-point Laplacian a step from rest must reproduce."""
import math

import numpy as np
import pytest

from oracle.oracle import OracleBar1D, OracleTube, audit_breaks

# Substance::aluminium_6061 and borosilicate_crown (crates/pantometry-core/src/substance.rs:
# 335-370): density in g/cm3 times 1e3, conductivity, specific heat.
ALUMINIUM = (2.70 * 1000.0, 167.0, 896.0)
GLASS = (2.51 * 1000.0, 1.114, 858.0)
CELSIUS20 = 293.15


def bar(cells, substance=ALUMINIUM, dx=1e-3, area=1e-4, initial=CELSIUS20):
    return OracleBar1D(cells, dx, area, *substance, initial)


def test_explicit_conduction_refuses_an_unstable_step():
    # lib.rs:1178-1204.
    b = bar(20)
    limit = b.max_stable_dt()
    assert abs(limit * 1e3 - 7.2) < 0.2
    assert abs(b.fourier_number(limit) - 0.5) < 1e-12
    b.step(limit)
    with pytest.raises(ValueError, match="Fourier"):
        b.step(limit * 1.5)


def test_two_materials_on_one_grid_need_different_steps():
    # lib.rs:1209-1228.
    glass = bar(10, GLASS).max_stable_dt()
    metal = bar(10).max_stable_dt()
    assert abs(glass - 0.967) < 0.02
    assert abs(metal - 0.0072) < 0.001
    assert glass / metal > 100.0


def test_conduction_conserves_heat_and_flattens_a_gradient():
    # lib.rs:1232-1270.
    b = bar(21)
    b.cells[10] = 60.0 + 273.15
    before = float(np.sum(b.cells))
    spread = float(np.max(b.cells) - np.min(b.cells))
    dt = b.max_stable_dt()
    for _ in range(500):
        b.step(dt * 0.9)
    assert abs(float(np.sum(b.cells)) / before - 1.0) < 1e-12
    assert float(np.max(b.cells) - np.min(b.cells)) < spread / 10.0
    assert float(np.mean(b.cells)) - 273.15 > 21.0


def test_a_cosine_mode_decays_by_the_discrete_amplification_factor():
    # The insulated, cell-centred three-point operator has eigenvectors cos(pi m (i + 1/2)/n)
    # with eigenvalue -4 sin^2(pi m / 2n), so a mode decays by 1 - 4 f sin^2(pi m / 2n) per step.
    # A clamped (not mirrored) end, a wrong factor on the centre or a transposed neighbour all
    # break it.
    n, m, steps = 24, 3, 200
    b = bar(n, dx=2e-3)
    shape = np.array([math.cos(math.pi * m * (i + 0.5) / n) for i in range(n)])
    b.cells = CELSIUS20 + 5.0 * shape
    dt = 0.37 * b.max_stable_dt()
    f = b.fourier_number(dt)
    g = 1.0 - 4.0 * f * math.sin(math.pi * m / (2.0 * n)) ** 2
    for _ in range(steps):
        b.step(dt)
    got = (b.cells - CELSIUS20) / 5.0
    assert np.allclose(got, shape * g ** steps, rtol=0.0, atol=1e-11)


def test_the_sweep_is_bar1d_step_to_the_bit():
    # A restatement of Bar1D::step (lib.rs:584-626) in Python floats on a random bar, with bus
    # heat and per-face joules, agrees with the oracle bit for bit.
    rng = np.random.default_rng(11)
    b = bar(13, dx=5e-4)
    b.cells = CELSIUS20 + rng.uniform(-3.0, 3.0, size=13)
    cells = [float(v) for v in b.cells]
    dt = 0.8 * b.max_stable_dt()
    f = b.fourier_number(dt)
    cap = b.capacity
    for s in range(9):
        gained = 0.01 * s
        joules = rng.uniform(0.0, 0.02, size=13) if s % 3 == 1 else None
        cells[0] += gained / cap
        if joules is not None:
            for i in range(13):
                cells[i] += float(joules[i]) / cap
        prev = list(cells)
        for i in range(13):
            left = prev[max(i - 1, 0)]
            right = prev[min(i + 1, 12)]
            cells[i] = prev[i] + f * (left - 2.0 * prev[i] + right)
        b.step(dt, gained, joules)
        assert np.array_equal(b.cells, np.array(cells))


def test_heat_lands_where_it_is_sent_and_the_books_hold():
    # lib.rs:1277-1355: the same joules, lumped (into cell 0) or resolved onto one face.
    lumped, placed = bar(21), bar(21)
    joules = 2.0
    lumped.step(0.5 * lumped.max_stable_dt(), joules)
    face = np.zeros(21)
    face[10] = joules
    placed.step(0.5 * placed.max_stable_dt(), 0.0, face)
    assert int(np.argmax(lumped.cells)) == 0
    assert int(np.argmax(placed.cells)) == 10
    for b in (lumped, placed):
        assert abs(b.stored() - joules) < 1e-12 * joules
        assert b.absorbed == joules


def test_the_field_samples_the_bar_and_stops_at_its_ends():
    # lib.rs:736-751 and its test at lib.rs:1401-1445.
    b = bar(11, dx=1e-3)
    b.cells = np.arange(11, dtype=np.float64) * 2.0 + 300.0
    for i in range(11):
        assert b.sample((i + 0.5) * 1e-3) == b.cells[i]
    mid = b.sample(3.0e-3)
    assert abs(mid - 0.5 * (b.cells[2] + b.cells[3])) < 1e-12
    assert b.sample(-5.0) == b.cells[0]
    assert b.sample(1.0) == b.cells[10]
    assert b.sample(math.nan) == b.cells[0]


def air_tube(cells, ends=((0, 0.0), (0, 0.0)), length=1.0):
    return OracleTube(length, cells, 1e-4, 1.204, 343.0, *ends)


def test_one_step_from_rest_is_the_three_point_laplacian():
    # lib.rs:572-610.
    n = 33
    t = air_tube(n)
    t.release(np.array([math.cos(math.pi * (i * t.dx)) for i in range(n)]))
    h = t.max_stable_dt() * 0.7
    before = t.p.copy()
    t.step(h)
    scale = float(np.max(np.abs(before))) / (t.dx * t.dx)
    for i in range(n):
        left = before[1] if i == 0 else before[i - 1]
        right = before[n - 2] if i + 1 == n else before[i + 1]
        expected = (left - 2.0 * before[i] + right) / (t.dx * t.dx)
        implied = (t.p[i] - before[i]) / (0.5 * h * h * t.speed * t.speed)
        assert abs(implied - expected) < scale * 1e-12


def test_an_absorbing_end_tightens_the_step_and_a_closed_one_does_not():
    # lib.rs:620-645.
    closed = air_tube(101)
    wave = closed.max_stable_dt()
    assert abs(wave - closed.dx / closed.speed) < 1e-15
    z = closed.impedance()
    assert abs(air_tube(101, ((0, 0.0), (2, z))).max_stable_dt() / (wave / 2.0) - 1.0) < 1e-12
    assert abs(air_tube(101, ((0, 0.0), (2, 4.0 * z))).max_stable_dt() - wave) < 1e-15
    assert abs(air_tube(101, ((1, 0.0), (0, 0.0))).max_stable_dt() - wave / 2.0) < 1e-15


def test_the_courant_limit_is_enforced():
    # lib.rs:697-716.
    t = air_tube(101)
    limit = t.max_stable_dt()
    assert abs(limit * 1e6 - 29.15) < 0.05
    assert abs(t.courant(limit) - 1.0) < 1e-12
    t.step(limit)
    with pytest.raises(ValueError, match="Courant"):
        t.step(limit * 1.01)


def test_a_pulse_travels_at_the_speed_of_sound():
    # lib.rs:721-765.
    cells = 401
    t = air_tube(cells, length=4.0)
    t.release(np.array([100.0 * math.exp(-((i * t.dx - 1.0) / 0.05) ** 2) for i in range(cells)]))
    dt = t.max_stable_dt()
    dx = 4.0 / (cells - 1)
    steps = int(round(0.005 / dt))
    for _ in range(steps):
        t.step(dt)
    start = int(round(1.0 / dx))
    best, best_at = 0.0, start
    for i in range(start + 5, cells):
        if t.p[i] > best:
            best, best_at = t.p[i], i
    distance = (best_at - start) * dx
    assert abs(distance / (343.0 * dt * steps) - 1.0) < 0.02
    assert best > 30.0


@pytest.mark.parametrize("ends", [((0, 0.0), (0, 0.0))])
def test_the_fundamental_oscillates_at_its_predicted_period(ends):
    # lib.rs:812-852: 1/171.5 s for a closed 1 m tube.
    t = air_tube(201, ends)
    t.release(np.array([100.0 * math.cos(math.pi * i * t.dx) for i in range(201)]))
    period = 2.0 * 1.0 / 343.0
    assert abs(period - 1.0 / 171.5) < 1e-6
    dt = t.max_stable_dt()
    steps = int(round(period / dt))
    start = t.p[0]
    for _ in range(steps // 2):
        t.step(dt)
    assert t.p[0] < -0.9 * start
    for _ in range(steps - steps // 2):
        t.step(dt)
    assert abs(t.p[0] / start - 1.0) < 0.05


def _bump(t, centre, width, amplitude):
    return np.array([amplitude * math.exp(-((i * t.dx - centre) / width) ** 2)
                     for i in range(t.n)])


def test_a_closed_tube_keeps_its_energy_to_itself():
    # lib.rs:856-882.
    t = air_tube(201)
    t.release(_bump(t, 0.5, 0.08, 200.0))
    start = t.energy()
    assert start > 0.0
    dt = t.max_stable_dt()
    for _ in range(2000):
        assert t.step(dt) == 0.0
    assert t.radiated == 0.0
    assert abs(t.energy() / start - 1.0) < 1e-9


def test_a_matched_end_absorbs_and_publishes_what_it_took():
    # lib.rs:887-930.
    z = air_tube(3).impedance()
    t = air_tube(201, ((0, 0.0), (2, z)))
    t.release(_bump(t, 0.3, 0.05, 150.0))
    start = t.energy()
    dt = t.max_stable_dt()
    collected = 0.0
    for _ in range(3000):
        collected += t.step(dt)
    assert collected > 0.0
    assert collected == t.radiated
    left = t.energy()
    assert abs(collected / (start - left) - 1.0) < 2e-3
    assert left < start * 0.01


def test_the_domain_runs_under_the_multirate_schedule():
    # lib.rs:933-960: 14.58 us at 5 mm, 69 substeps a millisecond, 21 ms audited at 1e-9.
    t = air_tube(201)
    t.release(_bump(t, 0.5, 0.08, 200.0))
    limit = t.max_stable_dt()
    assert abs(limit * 1e6 - 14.58) < 0.05
    n = int(math.ceil(1e-3 / limit - 1e-12))
    assert n == 69
    before = t.ledger()
    for _ in range(21 * n):
        t.step(1e-3 / n)
        after = t.ledger()
        assert not audit_breaks(before, after, 1e-9)
        before = after


def test_degenerate_tubes_are_handled():
    # lib.rs:963-990.
    assert air_tube(1).n == 3
    frozen = OracleTube(1.0, 11, 1e-4, 1.0, 0.0)
    assert not math.isfinite(frozen.max_stable_dt())
    frozen.step(1.0)
    assert float(np.max(np.abs(frozen.p))) == 0.0
    quiet = air_tube(21)
    quiet.release(np.full(21, 5.0))
    before = quiet.p[3]
    quiet.step(0.0)
    assert quiet.p[3] == before


def test_the_tube_sweep_is_tube_advance_to_the_bit():
    # A restatement of Tube::advance (lib.rs:389-446) in Python floats, open left end and a
    # matched right end, agrees with the oracle bit for bit, absorbed joules included.
    rng = np.random.default_rng(5)
    z = 1.7 * 1.204 * 343.0
    t = air_tube(17, ((1, 0.0), (2, z)))
    p0 = _bump(t, 0.4, 0.2, 50.0) + rng.uniform(-1.0, 1.0, size=17)
    t.release(p0)
    p, u = [float(v) for v in p0], [0.0] * 16
    rho, c, dx, area = 1.204, 343.0, t.dx, 1e-4
    rc2, imp = rho * c * c, rho * c
    dt = 0.9 * t.max_stable_dt()
    for s in range(20):
        vh = dt if s else 0.5 * dt
        for i in range(16):
            u[i] -= vh / (rho * dx) * (p[i + 1] - p[i])
        lw = -1.0 * p[0] / imp
        rw = 1.0 * p[16] / z
        for i in range(17):
            inflow = lw if i == 0 else u[i - 1]
            outflow = rw if i == 16 else u[i]
            share = 2.0 if i in (0, 16) else 1.0
            p[i] -= dt * rc2 / dx * share * (outflow - inflow)
        absorbed = abs(p[0] * lw) * area * dt + abs(p[16] * rw) * area * dt
        got = t.step(dt)
        assert np.array_equal(t.p, np.array(p))
        assert np.array_equal(t.u, np.array(u))
        assert got == absorbed
