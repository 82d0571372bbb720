"""The device-resident advance of the extended heat step (films, sources, void and radiating gaps,
latent heat): every step's limit, sweep, books and audit run on the device with no host
synchronisation, so they must reproduce stepping one call at a time (each step checked, swept and
booked on the host) bit for bit: cells, the books (stored, lost, supplied) and every ledger total.
The oracle checks the cells bit for bit and the totals to 1e-12.  A refusal inside the run (a
radiating film whose limit tightens as a source heats the block) stops at the step the host path
refuses, with the state before it; a breach keeps the breaking step."""

import numpy as np
import pytest

from tests.conftest import gpu_case

pytestmark = pytest.mark.gpu

C0 = 273.15
COPPER = (8.96e3, 401.0, 385.0, 0.04)
PLA = (1.24e3, 0.13, 1800.0, 0.90)
ALUMINA = (3720.0, 24.0, 880.0, 0.8)


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).reshape(-1).view(np.uint64)


def box(counts, lo, hi):
    nx, ny, nz = counts
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    return ((i >= lo[0]) & (i < hi[0]) & (j >= lo[1]) & (j < hi[1]) & (k >= lo[2]) &
            (k < hi[2])).reshape(-1)


def film_module(counts=(9, 7, 6), dx=1.5e-3):
    """Copper with an alumina layer, two sources, every face exposed: (GPU maker, oracle)."""
    import pantometry_amd as pm
    from oracle.oracle import OracleSolid3D
    nx, ny, nz = counts
    side = lambda a, b: a * b * dx * dx
    faces = [(0, C0 + 25.0, 40.0, side(ny, nz)), (1, C0 + 30.0, 7.0, side(ny, nz)),
             (2, C0 + 20.0, 150.0, 2.0 * side(nx, nz)), (3, C0 + 22.0, 0.0, side(nx, nz)),
             (4, C0 + 40.0, 3000.0, side(nx, ny)), (5, C0 + 35.0, 12.0, side(nx, ny))]
    layer = box(counts, (0, 0, 1), (nx, ny, 2))
    sources = [(45.0, box(counts, (0, 0, nz - 1), (nx, ny, nz))),
               (3.0, box(counts, (2, 2, 3), (4, 5, 4)))]

    def make():
        g = pm.Solid3D("block", pm.Substance.of("cu", *COPPER), counts, dx, C0 + 60.0)
        g.fill(pm.Substance.of("alumina", *ALUMINA), layer)
        for f, amb, h, area in faces:
            g.losing_from(pm.Face(f), pm.Environment(amb, h, area))
        for w, m in sources:
            g.dissipating(w, m)
        return g

    o = OracleSolid3D("block", *COPPER[:3], counts, dx, C0 + 60.0, emissivity=COPPER[3])
    o.fill(*ALUMINA[:3], layer, emissivity=ALUMINA[3])
    for f, amb, h, area in faces:
        o.losing_from(f, amb, h, area)
    for w, m in sources:
        o.dissipating(w, m)
    return make, o


def void_gaps(counts=(13, 9, 8), dx=2e-3):
    """Copper and PLA with void cells (radiating gaps), films and a source, seeded start."""
    import pantometry_amd as pm
    from oracle.oracle import OracleSolid3D
    n = counts[0] * counts[1] * counts[2]
    rng = np.random.default_rng(11)
    nx, ny, nz = counts
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    second = ((i + 2 * j + k) % 3 == 0).reshape(-1).astype(np.uint8)
    voids = ((k == 3) | ((i == 4) & (j == 2))).reshape(-1).astype(np.uint8)
    voids |= (rng.random(n) < 0.06).astype(np.uint8)
    start = C0 + rng.uniform(20.0, 600.0, size=n)
    heat = (i == nx - 1).reshape(-1).astype(np.uint8)

    def make():
        g = pm.Solid3D("b", pm.Substance.of("Cu", *COPPER), counts, dx, 300.0)
        g.fill(pm.Substance.of("PLA", *PLA), second)
        g.set_cells(start)
        g.empty(voids)
        g.losing_from(pm.Face.XMIN, pm.Environment(290.0, 12.0, 4e-4))
        g.losing_from(pm.Face.ZMAX, pm.Environment(310.0, 3.0, 1e-3))
        g.dissipating(2.5, heat)
        return g

    o = OracleSolid3D("b", *COPPER[:3], counts, dx, 300.0, COPPER[3])
    o.fill(*PLA[:3], second, PLA[3])
    o.set_cells(start)
    o.empty(voids)
    o.losing_from(0, 290.0, 12.0, 4e-4)
    o.losing_from(5, 310.0, 3.0, 1e-3)
    o.dissipating(2.5, heat)
    assert len(o.gap_a) > 5
    return make, o


CASES = {"film and sources": film_module, "void, gaps, films and a source": void_gaps}


def ledger_total(books):
    stored, lost, supplied = books
    t = 0.0
    t += stored
    t += lost
    t += -supplied
    return t


@pytest.mark.parametrize("name", list(CASES))
def test_the_device_resident_advance_is_stepping_one_call_at_a_time(name):
    gpu_case(f"extended advance against stepping: {name}")
    from pantometry_amd.core import Exchange
    make, o = CASES[name]()
    run, stepped = make(), make()
    dt = 0.7 * o.max_stable_dt()
    assert run.max_stable_dt() == o.max_stable_dt()
    steps = 60
    totals, _ = run.advance(dt, steps, tolerance=1e-9)
    bus = Exchange()
    for n in range(steps):
        stepped.step(0.0, dt, bus)
        o.step(dt)
        assert totals[n] == ledger_total(stepped.books()), n
        ref, scale = o.ledger()
        assert abs(totals[n] - ref) <= 1e-12 * max(scale, abs(ref)), n
    assert np.array_equal(bits(run.cells()), bits(stepped.cells()))
    assert np.array_equal(bits(run.cells()), bits(o.cells))
    assert run.books() == stepped.books()
    assert run.books()[2] == o.supplied
    # the run continues from where it ended: a second advance, and a step after it
    more, _ = run.advance(dt, 7, tolerance=1e-9)
    for n in range(7):
        stepped.step(0.0, dt, bus)
        assert more[n] == ledger_total(stepped.books())
    assert np.array_equal(bits(run.cells()), bits(stepped.cells()))


def test_a_tightening_film_refuses_inside_the_run_where_stepping_does():
    """A coarse PLA block (1 cm cells) at 400 K under two radiating films, heated by 10 W: the
    films' conductance grows with T^3, so a dt just under the starting limit is refused some steps
    later (the oracle refuses at step 12)."""
    gpu_case("extended advance: refusal inside the run")
    import pantometry_amd as pm
    from pantometry_amd.core import Exchange, Violation
    counts, dx = (4, 4, 4), 1e-2
    side = counts[0] * dx

    def make():
        g = pm.Solid3D("part", pm.Substance.of("pla", *PLA), counts, dx, 400.0)
        g.losing_from(pm.Face.ZMAX, pm.Environment(C0 + 20.0, 0.0, side * side))
        g.losing_from(pm.Face.XMIN, pm.Environment(C0 + 20.0, 0.0, side * side))
        g.dissipating(10.0, np.ones(64, dtype=np.uint8))
        return g

    run, stepped = make(), make()
    dt = 0.9995 * run.max_stable_dt()
    bus = Exchange()
    refused = None
    for n in range(400):
        try:
            stepped.step(0.0, dt, bus)
        except Violation as e:
            refused = (n, e)
            break
    assert refused is not None and 2 <= refused[0] < 400, "the limit should tighten within the run"
    m, want = refused
    with pytest.raises(Violation) as err:
        run.advance(dt, 400, tolerance=1e-9)
    got = err.value
    assert got.step == m
    assert (got.quantity, got.site, got.before, got.after, got.scale, got.tolerance) == \
        (want.quantity, want.site, want.before, want.after, want.scale, want.tolerance)
    assert np.array_equal(bits(run.cells()), bits(stepped.cells()))
    assert run.books() == stepped.books()
    assert run.max_stable_dt() == stepped.max_stable_dt() < dt


def test_a_breach_keeps_the_breaking_step():
    gpu_case("extended advance: breach")
    from pantometry_amd.core import Exchange, Violation
    make, _ = film_module()
    run, stepped = make(), make()
    dt = 0.6 * run.max_stable_dt()
    bus = Exchange()
    for _ in range(3):
        stepped.step(0.0, dt, bus)
    run.advance(dt, 2, tolerance=1e-9)
    before = ledger_total(run.books())
    with pytest.raises(Violation) as err:
        run.advance(dt, 30, tolerance=-1.0)
    assert err.value.step == 0 and err.value.quantity == "energy"
    assert err.value.before == before
    assert err.value.after == ledger_total(stepped.books())
    assert np.array_equal(bits(run.cells()), bits(stepped.cells()))
    assert run.books() == stepped.books()
