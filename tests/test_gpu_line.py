"""Bar1D and Tube on the GPU against their oracles: bit-exact cells, pressures and velocities,
the books each step, and the reference's own bar and tube tests run on the device
(crates/pantometry-thermal/src/lib.rs:1175-1290; crates/pantometry-acoustic/src/lib.rs:566-990)."""
import math

import numpy as np
import pytest

from tests.conftest import gpu_case

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).reshape(-1).view(np.uint64)


def _bar_pair(cells, dx=1e-3, area=1e-4, substance=None, initial=293.15):
    import pantometry_amd as pm
    from oracle.oracle import OracleBar1D
    s = substance or pm.Substance.aluminium_6061()
    return (pm.Bar1D("bar", s, cells, dx, area, initial),
            OracleBar1D(cells, dx, area, s.density, s.conductivity, s.specific_heat, initial))


@pytest.mark.parametrize("cells", [2, 13, 200, 2048, 3001])
def test_bar_cells_are_bit_exact(cells):
    gpu_case(f"bar parity {cells}")
    import pantometry_amd as pm
    g, o = _bar_pair(cells, dx=5e-4)
    rng = np.random.default_rng(cells)
    start = 293.15 + rng.uniform(-4.0, 4.0, size=o.n)
    g.set_cells(start)
    o.cells = start.copy()
    dt = 0.83 * o.max_stable_dt()
    assert g.max_stable_dt() == o.max_stable_dt()
    assert g.fourier_number(dt) == o.fourier_number(dt)
    for s in range(6):
        bus = pm.Exchange()
        gained = 0.003 * s
        if gained:
            bus.publish(pm.HEAT, gained)
            bus.covering(dt)
        joules = rng.uniform(0.0, 0.01, size=o.n) if s % 2 else None
        g.step(0.0, dt, bus, joules)
        o.step(dt, gained, joules)
        assert np.array_equal(bits(g.cells()), bits(o.cells))
    assert g.stored_heat() == o.stored()
    assert g.absorbed_energy() == o.absorbed
    totals, _ = g.advance(dt, 300)
    want = []
    for _ in range(300):
        o.step(dt)
        want.append(o.stored())
    assert np.array_equal(bits(g.cells()), bits(o.cells))
    assert np.allclose(totals, want, rtol=1e-12, atol=1e-12 * max(abs(w) for w in want))


def test_bar_refuses_past_the_fourier_limit_and_leaves_everything():
    # lib.rs:1178-1204: about 7.2 ms for aluminium on a millimetre grid.
    gpu_case("bar refusal")
    import pantometry_amd as pm
    from pantometry_amd.core import Violation
    g, _ = _bar_pair(20)
    limit = g.max_stable_dt()
    assert abs(limit * 1e3 - 7.2) < 0.2
    assert abs(g.fourier_number(limit) - 0.5) < 1e-12
    bus = pm.Exchange()
    g.step(0.0, limit, bus)
    before = g.cells()
    bus.publish(pm.HEAT, 1.0)
    bus.covering(limit * 1.5)
    with pytest.raises(Violation) as err:
        g.step(0.0, limit * 1.5, bus)
    assert err.value.quantity == "Fourier number" and err.value.after > 0.5
    assert np.array_equal(bits(g.cells()), bits(before))
    assert bus.peek(pm.HEAT) == 1.0
    glass = pm.Bar1D("g", pm.Substance.borosilicate_crown(), 10, 1e-3, 1e-4, 293.15)
    metal = pm.Bar1D("m", pm.Substance.aluminium_6061(), 10, 1e-3, 1e-4, 293.15)
    assert abs(glass.max_stable_dt() - 0.967) < 0.02
    assert glass.max_stable_dt() / metal.max_stable_dt() > 100.0


def test_bar_conserves_heat_flattens_and_decays_a_mode_exactly():
    # lib.rs:1232-1270 on the device, and the discrete cosine mode's closed-form decay.
    gpu_case("bar conservation and mode")
    g, _ = _bar_pair(21)
    c = np.full(21, 293.15)
    c[10] = 333.15
    g.set_cells(c)
    before = float(np.sum(c))
    dt = g.max_stable_dt() * 0.9
    g.advance(dt, 500)
    after = g.cells()
    assert abs(float(np.sum(after)) / before - 1.0) < 1e-12
    assert float(np.max(after) - np.min(after)) < 40.0 / 10.0
    assert g.readings()[0][2] > 21.0
    n, m, steps = 24, 3, 200
    bar, _ = _bar_pair(n, dx=2e-3)
    shape = np.array([math.cos(math.pi * m * (i + 0.5) / n) for i in range(n)])
    bar.set_cells(293.15 + 5.0 * shape)
    dt = 0.37 * bar.max_stable_dt()
    f = bar.fourier_number(dt)
    gain = 1.0 - 4.0 * f * math.sin(math.pi * m / (2.0 * n)) ** 2
    bar.advance(dt, steps)
    assert np.allclose((bar.cells() - 293.15) / 5.0, shape * gain ** steps, rtol=0.0, atol=1e-11)


def test_bar_sampling_readings_and_restore():
    gpu_case("bar sampling and restore")
    import pantometry_amd as pm
    g, o = _bar_pair(11)
    c = np.arange(11, dtype=np.float64) * 2.0 + 300.0
    g.set_cells(c)
    o.cells = c.copy()
    xs = np.concatenate([np.linspace(-2e-3, 13e-3, 97), [math.nan, 5.5e-3, 1.0]])
    got = g.sample(xs)
    want = np.array([o.sample(x) for x in xs])
    assert np.array_equal(bits(got), bits(want))
    mean, peak, absorbed = [r[2] for r in g.readings()]
    assert mean == sum(c.tolist()) / 11.0 - 273.15
    assert peak == c[-1] - 273.15 and absorbed == 0.0
    # The saved cells start as the initial ones (lib.rs:455-461).
    g.restore()
    assert np.array_equal(g.cells(), np.full(11, 293.15))
    g.set_cells(c)
    g.checkpoint()
    g.step(0.0, 0.5 * g.max_stable_dt(), pm.Exchange())
    g.restore()
    assert np.array_equal(bits(g.cells()), bits(c))


def test_bar_books_balance_under_the_multirate_scheduler():
    # The shape of lib.rs:1593-1638: a heater publishes 10 mW of heat, the bar takes it share
    # by substep, and the simulation's books hold at 1e-9 for 40 windows.
    gpu_case("bar under the scheduler")
    import pantometry_amd as pm

    class Heater:
        def __init__(self):
            self.paid = 0.0

        def name(self):
            return "heater"

        def kind(self):
            return pm.Kind.QUASI_STATIC

        def max_stable_dt(self, now=0.0):
            return math.inf

        def step(self, t, dt, bus):
            bus.publish(pm.HEAT, 0.01 * dt)
            self.paid += 0.01 * dt

        def ledger(self):
            return pm.Ledger().with_(pm.ENERGY, -self.paid)

        def books_balance(self):
            return True

    bar, _ = _bar_pair(21)
    sim = pm.Simulation(pm.Schedule.MULTIRATE).conservation_tolerance(1e-9)
    sim.with_(Heater()).with_(bar)
    for _ in range(40):
        sim.advance(0.05)
    assert abs(sim.ledger().get(pm.ENERGY)) < 1e-9
    assert bar.temperature_at(0) > bar.temperature_at(20)


def _tube_pair(cells, ends=(("c", 0.0), ("c", 0.0)), length=1.0):
    import pantometry_amd as pm
    from oracle.oracle import OracleTube
    kinds = {"c": 0, "o": 1, "m": 2}
    make = {"c": lambda z: pm.End.closed(), "o": lambda z: pm.End.open(),
            "m": lambda z: pm.End.matched(z)}
    g = pm.Tube.of_air("tube", length, cells, 1e-4).with_ends(make[ends[0][0]](ends[0][1]),
                                                              make[ends[1][0]](ends[1][1]))
    o = OracleTube(length, cells, 1e-4, 1.204, 343.0, (kinds[ends[0][0]], ends[0][1]),
                   (kinds[ends[1][0]], ends[1][1]))
    return g, o


def _bump(n, dx, centre, width, amplitude):
    return np.array([amplitude * math.exp(-((i * dx - centre) / width) ** 2) for i in range(n)])


@pytest.mark.parametrize("cells,ends", [
    (17, (("o", 0.0), ("m", 1.7 * 1.204 * 343.0))),
    (201, (("c", 0.0), ("c", 0.0))),
    (201, (("m", -1.0), ("o", 0.0))),
    (2048, (("c", 0.0), ("m", 413.0))),
    (3001, (("o", 0.0), ("c", 0.0))),
], ids=str)
def test_tube_fields_are_bit_exact(cells, ends):
    gpu_case(f"tube parity {cells} {ends}")
    import pantometry_amd as pm
    g, o = _tube_pair(cells, ends)
    p0 = _bump(o.n, o.dx, 0.4, 0.1, 80.0) + 3.0 * np.cos(np.arange(o.n) * o.dx * math.pi)
    g.released_from(p0)
    o.release(p0)
    assert g.max_stable_dt() == o.max_stable_dt()
    assert abs(g.energy() - o.energy()) <= 1e-12 * abs(o.energy())
    dt = 0.9 * o.max_stable_dt()
    for _ in range(12):
        bus = pm.Exchange()
        g.step(0.0, dt, bus)
        want = o.step(dt)
        assert bus.peek(pm.ENERGY) == want
    p, pp, u = g.fields()
    assert np.array_equal(bits(p), bits(o.p))
    assert np.array_equal(bits(pp), bits(o.pp))
    assert np.array_equal(bits(u), bits(o.u))
    assert g.radiated_energy() == o.radiated
    assert abs(g.startup_adjustment() - o.offset) <= 1e-12 * abs(o.energy())
    # An open or matched end's books carry the half-step stagger of p u A (lib.rs:925-927),
    # percents a step on a coarse tube, so the audit here is loose; parity is the point.
    totals, _ = g.advance(dt, 150, tolerance=1.0)
    want = []
    for _ in range(150):
        o.step(dt)
        want.append(o.energy() + o.radiated)
    p, pp, u = g.fields()
    assert np.array_equal(bits(p), bits(o.p))
    assert np.array_equal(bits(pp), bits(o.pp))
    assert np.array_equal(bits(u), bits(o.u))
    assert g.radiated_energy() == o.radiated
    assert np.allclose(totals, want, rtol=1e-12, atol=0.0)


def test_a_fresh_tube_advance_starts_the_leapfrog():
    gpu_case("tube advance from release")
    g, o = _tube_pair(101)
    p0 = np.array([math.cos(math.pi * i * o.dx) for i in range(o.n)])
    g.released_from(p0)
    o.release(p0)
    dt = o.max_stable_dt() * 0.7
    totals, _ = g.advance(dt, 40)
    want = []
    for _ in range(40):
        o.step(dt)
        want.append(o.energy())
    assert np.array_equal(bits(g.fields()[0]), bits(o.p))
    assert np.allclose(totals, want, rtol=1e-12, atol=0.0)


def test_tube_reference_behaviour_on_the_device():
    # lib.rs:572-610 (one step from rest is the Laplacian), 697-716 (Courant), 812-852
    # (fundamental), 856-882 (closed keeps its energy), 963-990 (degenerate).
    gpu_case("tube reference tests")
    import pantometry_amd as pm
    from pantometry_amd.core import Violation
    t, _ = _tube_pair(33)
    dx = t.spacing()
    t.released_from(lambda x: math.cos(math.pi * x))
    h = t.max_stable_dt() * 0.7
    before = t.pressure()
    t.step(0.0, h, pm.Exchange())
    after = t.pressure()
    scale = float(np.max(np.abs(before))) / (dx * dx)
    for i in range(33):
        left = before[1] if i == 0 else before[i - 1]
        right = before[31] if i == 32 else before[i + 1]
        expected = (left - 2.0 * before[i] + right) / (dx * dx)
        implied = (after[i] - before[i]) / (0.5 * h * h * 343.0 * 343.0)
        assert abs(implied - expected) < scale * 1e-12

    c, _ = _tube_pair(101)
    limit = c.max_stable_dt()
    assert abs(limit * 1e6 - 29.15) < 0.05
    c.step(0.0, limit, pm.Exchange())
    with pytest.raises(Violation) as err:
        c.step(0.0, limit * 1.01, pm.Exchange())
    assert err.value.quantity == "Courant number"

    f, _ = _tube_pair(201)
    f.released_from(lambda x: 100.0 * math.cos(math.pi * x))
    period = 1.0 / f.mode_frequency(1)
    dt = f.max_stable_dt()
    steps = int(round(period / dt))
    start = f.pressure_at(0)
    f.advance(dt, steps // 2)
    assert f.pressure_at(0) < -0.9 * start
    f.advance(dt, steps - steps // 2)
    assert abs(f.pressure_at(0) / start - 1.0) < 0.05

    k, _ = _tube_pair(201)
    k.released_from(lambda x: 200.0 * math.exp(-((x - 0.5) / 0.08) ** 2))
    e0 = k.energy()
    k.advance(k.max_stable_dt(), 2000)
    assert k.radiated_energy() == 0.0
    assert abs(k.energy() / e0 - 1.0) < 1e-9

    tiny = pm.Tube.of_air("tiny", 1.0, 1, 1e-4)
    assert tiny.cells() == 3
    frozen = pm.Tube("frozen", 1.0, 11, 1e-4, 1.0, 0.0)
    assert not math.isfinite(frozen.max_stable_dt())
    frozen.step(0.0, 1.0, pm.Exchange())
    assert frozen.peak_pressure() == 0.0


def test_a_matched_end_absorbs_and_publishes_what_it_took():
    # lib.rs:887-930, stepped through the bus.
    gpu_case("tube matched end")
    import pantometry_amd as pm
    t, _ = _tube_pair(201, (("c", 0.0), ("m", 1.204 * 343.0)))
    t.released_from(lambda x: 150.0 * math.exp(-((x - 0.3) / 0.05) ** 2))
    start = t.energy()
    dt = t.max_stable_dt()
    bus = pm.Exchange()
    collected = 0.0
    for _ in range(3000):
        t.step(0.0, dt, bus)
        collected += bus.take(pm.ENERGY)
    assert collected > 0.0
    assert collected == t.radiated_energy()
    left = t.energy()
    assert abs(collected / (start - left) - 1.0) < 2e-3
    assert left < start * 0.01


def test_the_tube_runs_under_the_multirate_scheduler():
    # lib.rs:933-960.
    gpu_case("tube under the scheduler")
    import pantometry_amd as pm
    t, _ = _tube_pair(201)
    t.released_from(lambda x: 200.0 * math.exp(-((x - 0.5) / 0.08) ** 2))
    assert abs(t.max_stable_dt() * 1e6 - 14.58) < 0.05
    sim = pm.Simulation(pm.Schedule.MULTIRATE).conservation_tolerance(1e-9).with_(t)
    report = sim.advance(1e-3)
    assert report.substeps[0] == ("tube", 69)
    for _ in range(20):
        sim.advance(1e-3)


def test_tube_checkpoint_restore():
    gpu_case("tube checkpoint")
    import pantometry_amd as pm
    t, _ = _tube_pair(51, (("o", 0.0), ("c", 0.0)))
    t.released_from(lambda x: 10.0 * math.cos(2.0 * math.pi * x))
    dt = 0.6 * t.max_stable_dt()
    for _ in range(4):
        t.step(0.0, dt, pm.Exchange())
    t.checkpoint()
    p, _, u = t.fields()
    for _ in range(9):
        t.step(0.0, dt, pm.Exchange())
    t.restore()
    p2, pp2, u2 = t.fields()
    assert np.array_equal(bits(p2), bits(p)) and np.array_equal(bits(pp2), bits(p))
    assert np.array_equal(bits(u2), bits(u))
