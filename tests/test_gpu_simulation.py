"""Simulation (sim.rs:947-1228) driving GPU domains, against the same run driving the oracle.

Each scenario is run twice, once with the GPU Solid3D / Hall and once with the oracle's behind the
same Domain protocol (tests/sim_domains.py), and must end the same way: the same report or the same
Violation (quantity, site and tolerance equal; before/after/scale equal, or within 1e-12 where they
are ledger totals, which the device sums in a fixed order and the oracle as a left fold), the same
fields bit for bit, and the clock unchanged after a refusal.
"""
import numpy as np
import pytest

from tests.conftest import gpu_case
from tests.sim_domains import Coupler, Lossy, OracleHallDomain, OracleSolidDomain

pytestmark = pytest.mark.gpu

COUNTS = (20, 17, 9)


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).reshape(-1).view(np.uint64)


def _blocks():
    import pantometry_amd as pm
    from oracle import oracle as o
    k, t = o.heterogeneous_fields(COUNTS, spots=2, k_seed=77, spot_seed=78)
    g = pm.Solid3D("block", pm.Substance("s", 4000.0, 1.0, 1000.0), COUNTS, 1e-3, 300.0)
    g.with_conductivity(k)
    g.set_cells(t)
    c = o.OracleSolid3D("block", 4000.0, 1.0, 1000.0, COUNTS, 1e-3, 300.0).with_conductivity(k)
    c.set_cells(t)
    return g, OracleSolidDomain(c)


def _halls(across=13, mode=(1, 1, 1)):
    import pantometry_amd as pm
    from oracle.oracle import OracleHall
    g = pm.Hall("hall", 4.4, 3.1, 2.4, across, 343.0, 1.204)
    o = OracleHall(4.4, 3.1, 2.4, across, 343.0, 1.204)
    o.release(o.mode(mode, 1.0))
    g.released_from(o.p.copy())
    return g, OracleHallDomain(o)


def _same_violation(a, b, rel=0.0):
    """Same quantity, site and tolerance; before, after and scale equal, or within `rel` of the
    violation's scale (ledger totals summed in different orders)."""
    assert (a.quantity, a.site, a.tolerance) == (b.quantity, b.site, b.tolerance)
    scale = max(abs(a.scale), abs(b.scale))
    for x, y in ((a.before, b.before), (a.after, b.after), (a.scale, b.scale)):
        assert abs(x - y) <= rel * max(scale, abs(x), abs(y)), (a, b)


def _sim(schedule, domains, **kw):
    from pantometry_amd.core import Simulation
    sim = Simulation(schedule, **kw)
    for d in domains:
        sim.with_(d)
    return sim


def test_iterative_schedule_restores_between_iterations():
    """Schedule::Iterative (sim.rs:1114-1153): checkpoint once, restore before every iteration
    after the first, re-sweep with multirate substeps until the residual settles (third
    iteration here); the block and what it absorbed match the oracle-driven run over two
    advances."""
    from pantometry_amd.core import Schedule
    from pantometry_amd.scene import Heater
    gpu_case("simulation iterative, block + heater")
    runs = []
    for side in ("gpu", "oracle"):
        gb, ob = _blocks()
        block = gb if side == "gpu" else ob
        coupler = Coupler([1e-1, 1e-2, 1e-4])
        sim = _sim(Schedule.ITERATIVE, [Heater("coil", 2.0, 100.0), block, coupler],
                   max_iter=6, tol=1e-3)
        dt = 2.5 * ob.max_stable_dt()
        reports = [sim.advance(dt), sim.advance(dt)]
        assert [r.iterations for r in reports] == [3, 3]
        assert coupler.checkpoints == 2 and coupler.restores == 4
        cells = gb.cells() if side == "gpu" else ob.b.cells
        absorbed = gb.absorbed_energy() if side == "gpu" else ob.absorbed
        runs.append((reports, cells.copy(), absorbed, sim.t))
    (rg, cg, ag, tg), (ro, co, ao, to) = runs
    assert [r.substeps for r in rg] == [r.substeps for r in ro]
    assert dict(rg[0].substeps)["block"] == 3
    assert np.array_equal(bits(cg), bits(co))
    # `absorbed` is a reading that restore does not rewind (solid.rs:1935-1950): three sweeps of
    # the heater's 2 W per advance.
    assert ag == ao and abs(ag - 3 * 2.0 * tg) <= 1e-12 * ag
    assert tg == to


def test_a_hall_under_the_iterative_schedule_breaks_its_books_as_the_reference_does():
    """Hall::restore (hall.rs:546-555) rewinds the fields and the staggered flag but not the
    startup offset, and the restarted first step sets a new one; so on the second iteration the
    hall's own books move by the first offset (3.4% of its energy at 13 nodes) and the per-domain
    check refuses it.  The GPU hall does what the reference does: the same Violation, and the
    block and hall fields as the second iteration left them."""
    from pantometry_amd.core import Schedule, Violation
    from pantometry_amd.scene import Heater
    gpu_case("simulation iterative, block + hall + heater")
    out = []
    for side in ("gpu", "oracle"):
        gb, ob = _blocks()
        gh, oh = _halls()
        block, hall = (gb, gh) if side == "gpu" else (ob, oh)
        coupler = Coupler([1e-1, 1e-2, 1e-4])
        sim = _sim(Schedule.ITERATIVE, [Heater("coil", 2.0, 100.0), block, hall, coupler],
                   max_iter=6, tol=1e-3)
        with pytest.raises(Violation) as err:
            sim.advance(2.5 * ob.max_stable_dt())
        assert sim.t == 0.0 and coupler.restores == 1
        cells = gb.cells() if side == "gpu" else ob.b.cells
        p = gh.fields()[0] if side == "gpu" else oh.h.p
        out.append((err.value, cells.copy(), p.copy()))
    (vg, cg, pg), (vo, co, po) = out
    assert vg.site == "hall (its own books, against what it moved on the bus)"
    _same_violation(vg, vo, rel=1e-12)
    assert abs(vg.after - vg.before) > 0.03 * vg.scale
    assert np.array_equal(bits(cg), bits(co))
    assert np.array_equal(bits(pg), bits(po))


def test_an_iteration_that_never_settles_is_refused():
    """After max_iter sweeps the residual is still above tol: Violation 'coupling residual' at
    'simulation (after N iterations)', the clock unchanged, the state the last sweep left."""
    from pantometry_amd.core import Schedule, Violation
    gpu_case("simulation iterative, no convergence")
    out = []
    for side in ("gpu", "oracle"):
        gb, ob = _blocks()
        block = gb if side == "gpu" else ob
        sim = _sim(Schedule.ITERATIVE, [block, Coupler([0.5])], max_iter=4, tol=1e-3)
        dt = 2.0 * ob.max_stable_dt()
        with pytest.raises(Violation) as err:
            sim.advance(dt)
        assert sim.t == 0.0
        out.append((err.value, (gb.cells() if side == "gpu" else ob.b.cells).copy()))
    (vg, cg), (vo, co) = out
    _same_violation(vg, vo)
    assert vg.quantity == "coupling residual" and vg.site == "simulation (after 4 iterations)"
    assert vg.after == 0.5 and vg.tolerance == 1e-3
    assert np.array_equal(bits(cg), bits(co))


def test_a_consumer_that_loses_what_it_took_breaks_its_own_books():
    """The per-domain books check (sim.rs:1169-1228): a block that took 2 J from the bus and kept
    none is refused at '<name> (its own books, against what it moved on the bus)'."""
    from pantometry_amd.core import Schedule, Violation
    from pantometry_amd.scene import Heater
    gpu_case("simulation lossy consumer")
    out = []
    for side in ("gpu", "oracle"):
        gb, ob = _blocks()
        block = gb if side == "gpu" else ob
        sim = _sim(Schedule.STAGGERED, [Heater("coil", 2.0, 100.0), Lossy(block)])
        with pytest.raises(Violation) as err:
            sim.advance(0.5 * ob.max_stable_dt())
        assert sim.t == 0.0
        out.append((err.value, (gb.cells() if side == "gpu" else ob.b.cells).copy()))
    (vg, cg), (vo, co) = out
    assert vg.site == "block (its own books, against what it moved on the bus)"
    _same_violation(vg, vo, rel=1e-12)
    assert np.array_equal(bits(cg), bits(co))


def test_a_hall_started_in_its_sharpest_mode_is_refused():
    """Hall's startup check (hall.rs:513-528): the checkerboard mode released and stepped at the
    Courant limit moves more than a quarter of the energy on the first half step."""
    from pantometry_amd.core import Schedule, Violation
    gpu_case("simulation hall startup refusal")
    out = []
    for side in ("gpu", "oracle"):
        gh, oh = _halls(7, (6, 4, 3))
        nodes = oh.h.nodes()
        assert nodes == (7, 5, 4)
        sim = _sim(Schedule.STAGGERED, [gh if side == "gpu" else oh])
        with pytest.raises(Violation) as err:
            sim.advance(oh.max_stable_dt())
        assert sim.t == 0.0
        out.append((err.value, (gh.fields()[0] if side == "gpu" else oh.h.p).copy()))
    (vg, pg), (vo, po) = out
    assert vg.quantity == "energy" and vg.site == "hall (starting the leapfrog)"
    _same_violation(vg, vo, rel=1e-12)
    assert np.array_equal(bits(pg), bits(po))


def test_refusals_through_advance_keep_the_state_and_the_clock():
    """A block stepped 1.5x past its limit (Fourier number) and a hall 1.5x past its Courant
    limit: the reference's Violation fields exactly, nothing written, the clock unchanged."""
    from pantometry_amd.core import Schedule, Violation
    gpu_case("simulation refusals")
    out = []
    for side in ("gpu", "oracle"):
        gb, ob = _blocks()
        gh, oh = _halls()
        block, hall = (gb, gh) if side == "gpu" else (ob, oh)
        vs = []
        for d, dt in ((block, 1.5 * ob.max_stable_dt()), (hall, 1.5 * oh.max_stable_dt())):
            sim = _sim(Schedule.STAGGERED, [d])
            with pytest.raises(Violation) as err:
                sim.advance(dt)
            assert sim.t == 0.0
            vs.append(err.value)
        cells = gb.cells() if side == "gpu" else ob.b.cells
        p = gh.fields()[0] if side == "gpu" else oh.h.p
        out.append((vs, cells.copy(), p.copy()))
    (vg, cg, pg), (vo, co, po) = out
    for a, b in zip(vg, vo):
        _same_violation(a, b)
    assert vg[0].quantity == "Fourier number" and vg[1].quantity == "Courant number"
    assert np.array_equal(bits(cg), bits(co)) and np.array_equal(bits(pg), bits(po))
