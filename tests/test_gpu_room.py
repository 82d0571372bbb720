"""Room on the GPU (the hall's kernels one node deep, under the room's rules) against the room
oracle: bit-exact fields, the room's invariant to 1e-12, the two-dimensional Courant limit, and
the reference's own room tests (crates/pantometry-acoustic/src/room.rs:755-1335) run on the
device."""
import math

import numpy as np
import pytest

from tests.conftest import gpu_case

pytestmark = pytest.mark.gpu

SPEED, DENSITY = 343.0, 1.204


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).reshape(-1).view(np.uint64)


def _pair(width, height, cells, depth=1.0):
    import pantometry_amd as pm
    from oracle.oracle import OracleRoom
    return (pm.Room("room", width, height, cells, depth, DENSITY, SPEED),
            OracleRoom(width, height, cells, depth, SPEED, DENSITY))


def _gauss(room, cx, cy, width2, amplitude=200.0):
    nx, ny = room.nx, room.ny
    x = np.arange(nx) * room.dx
    y = np.arange(ny) * room.dx
    r2 = (x[None, :] - cx) ** 2 + (y[:, None] - cy) ** 2
    return np.array([[amplitude * math.exp(-v / width2) for v in row] for row in r2]).reshape(-1)


def _same_fields(g, o):
    p, pp, vx, vy, _ = g.fields()
    assert np.array_equal(bits(p), bits(o.p))
    assert np.array_equal(bits(pp), bits(o.pp))
    assert np.array_equal(bits(vx), bits(o.vx))
    assert np.array_equal(bits(vy), bits(o.vy))


def test_room_new_and_the_two_dimensional_limit():
    gpu_case("room counts and limit")
    g, o = _pair(5.0, 3.0, 51)
    assert g.nodes() == (o.nx, o.ny, 1) == (51, 31, 1)
    assert g.spacing() == o.dx
    assert g.max_stable_dt() == o.max_stable_dt()
    sq, _ = _pair(4.0, 4.0, 41)
    limit = sq.max_stable_dt()
    assert abs(limit * 1e6 - 206.2) < 0.5
    assert abs((0.1 / 343.0) / limit - math.sqrt(2.0)) < 1e-9
    assert abs(sq.courant(limit) - 1.0) < 1e-12
    import pantometry_amd as pm
    from pantometry_amd.core import Violation
    sq.released_from(lambda x, y: 0.0)
    sq.step(0.0, limit, pm.Exchange())
    with pytest.raises(Violation) as err:
        sq.step(0.0, limit * 1.01, pm.Exchange())
    assert err.value.quantity == "Courant number"
    assert "two-dimensional" in err.value.site
    tiny, _ = _pair(0.1, 0.1, 1)
    assert tiny.cells()[0] >= 3 and tiny.cells()[1] >= 3
    with pytest.raises(TypeError):
        tiny.set_source(0, np.ones(3))


@pytest.mark.parametrize("shape", [(4.0, 4.0, 41, 1.0), (4.4, 3.1, 45, 0.35), (1.3, 0.9, 9, 2.0)],
                         ids=str)
def test_fields_and_energy_match_the_oracle(shape):
    gpu_case(f"room parity {shape}")
    import pantometry_amd as pm
    g, o = _pair(*shape)
    p0 = _gauss(o, 0.45 * (o.nx - 1) * o.dx, 0.6 * (o.ny - 1) * o.dx, 0.09) + o.mode2(2, 1, 3.0)
    g.released_from(p0.reshape(o.ny, o.nx))
    o.release(p0)
    assert abs(g.energy() - o.energy()) <= 1e-12 * abs(o.energy())
    dt = 0.8 * o.max_stable_dt()
    for _ in range(30):
        g.step(0.0, dt, pm.Exchange())
        o.step(dt)
    _same_fields(g, o)
    e = o.energy()
    assert abs(g.energy() - e) <= 1e-12 * abs(e)
    assert abs(g.startup_adjustment() - o.offset) <= 1e-12 * abs(e)
    # The device loop, audited on the host each step, against the oracle's ledger per step.
    totals, _ = g.advance(dt, 70, 1e-9)
    want = []
    for _ in range(70):
        o.step(dt)
        want.append(o.ledger()[0])
    _same_fields(g, o)
    assert np.allclose(totals, want, rtol=1e-12, atol=0.0)


def test_a_fresh_room_advance_starts_the_leapfrog():
    gpu_case("room advance from release")
    g, o = _pair(4.4, 3.1, 31)
    p0 = o.mode2(1, 1, 1.0)
    g.released_in_mode((1, 1), 1.0)
    o.release(p0)
    assert np.array_equal(bits(g.fields()[0]), bits(p0))
    dt = o.max_stable_dt()
    totals, _ = g.advance(dt, 12, 1e-9)
    want = []
    for _ in range(12):
        o.step(dt)
        want.append(o.ledger()[0])
    _same_fields(g, o)
    assert np.allclose(totals, want, rtol=1e-12, atol=0.0)
    # The startup adjustment is real and small (room.rs:1271-1296): 0.42 % at 31 cells.
    assert 1e-4 < abs(g.startup_adjustment()) / abs(totals[-1]) < 1e-2


@pytest.mark.parametrize("mode", [(1, 0), (1, 1), (2, 1)])
def test_an_excited_mode_oscillates_at_its_predicted_frequency(mode):
    # room.rs:1104-1141, on the device.
    gpu_case(f"room mode {mode}")
    import pantometry_amd as pm
    g, _ = _pair(4.0, 4.0, 81)
    g.released_in_mode(mode, 100.0)
    period = 1.0 / g.mode_frequency(mode)
    dt = g.max_stable_dt() * 0.5
    steps = int(round(period / dt))
    start = g.pressure_at(0, 0)
    assert abs(start) > 50.0
    g.advance(dt, steps // 2, 1e-9)
    assert g.pressure_at(0, 0) < -0.85 * start
    for _ in range(steps - steps // 2):
        g.step(0.0, dt, pm.Exchange())
    assert abs(g.pressure_at(0, 0) / start - 1.0) < 0.1


def test_the_domain_runs_under_the_multirate_scheduler():
    # room.rs:1203-1222: five substeps for a millisecond, twenty more held to 1e-9.
    gpu_case("room under the scheduler")
    import pantometry_amd as pm
    g, _ = _pair(4.0, 4.0, 41)
    g.released_from(lambda x, y: 200.0 * math.exp(-((x - 2.0) ** 2 + (y - 2.0) ** 2) / 0.5))
    sim = pm.Simulation(pm.Schedule.MULTIRATE).conservation_tolerance(1e-9).with_(g)
    report = sim.advance(1e-3)
    assert report.substeps[0] == ("room", 5)
    for _ in range(20):
        sim.advance(1e-3)


def test_sampling_is_bilinear_and_clamped():
    # room.rs:587-600 and its test at room.rs:764-797.
    gpu_case("room sampling")
    g, o = _pair(4.0, 4.0, 33)
    p0 = o.mode2(1, 0, 1.0)
    g.released_in_mode((1, 0), 1.0)
    o.release(p0)
    lx = (o.nx - 1) * o.dx
    rng = np.random.default_rng(3)
    pts = np.concatenate([rng.uniform(-1.0, lx + 1.0, size=(200, 2)),
                          [[0.0, 2.0], [lx, 2.0], [lx / 2, 2.0], [-9.0, 2.0], [lx + 9.0, 2.0],
                           [math.nan, math.nan]]])
    got = g.sample(pts)
    want = np.array([o.sample(x, y, 0.0) for x, y in pts])
    assert np.array_equal(bits(got), bits(want))
    assert abs(got[-6] - 1.0) < 1e-12 and abs(got[-5] + 1.0) < 1e-12 and abs(got[-4]) < 1e-12
    deep = g.sample([[1.0, 2.0, 0.0], [1.0, 2.0, 55.0]])
    assert deep[0] == deep[1]
    assert g.readings() == [("room", "peak", g.peak_pressure(), "Pa")]
    other, _ = _pair(4.4, 3.1, 41)
    other.released_in_mode((1, 1), 1.0)
    assert abs(other.readings()[0][2] - 1.0) < 1e-9


def test_checkpoint_restore_and_degenerate_rooms():
    gpu_case("room checkpoint and degenerate")
    import pantometry_amd as pm
    g, o = _pair(4.0, 4.0, 21)
    g.released_from(lambda x, y: 7.0)
    before = g.pressure_at(5, 5)
    g.step(0.0, 0.0, pm.Exchange())
    assert g.pressure_at(5, 5) == before
    g.released_in_mode((1, 1), 2.0)
    dt = 0.7 * g.max_stable_dt()
    for _ in range(5):
        g.step(0.0, dt, pm.Exchange())
    g.checkpoint()
    saved = g.fields()
    for _ in range(7):
        g.step(0.0, dt, pm.Exchange())
    g.restore()
    p, pp, vx, vy, _ = g.fields()
    # Domain::restore, room.rs:493-501: pressure_prev comes back as the saved pressure.
    assert np.array_equal(bits(p), bits(saved[0]))
    assert np.array_equal(bits(pp), bits(saved[0]))
    assert np.array_equal(bits(vx), bits(saved[2]))
    assert np.array_equal(bits(vy), bits(saved[3]))
    silent = pm.Room("silent", 1.0, 1.0, 11, 1.0, 1.0, 0.0)
    assert not math.isfinite(silent.max_stable_dt())
    silent.step(0.0, 1.0, pm.Exchange())
    assert silent.peak_pressure() == 0.0


def test_a_breach_leaves_the_state_of_the_breaking_step():
    # The room's audit runs on the device; once it halts, later sweeps do nothing and the handle
    # points at the buffers the breaking step wrote.  A tolerance of 1e-300 breaks on the
    # first step whose energy moves by rounding at all.
    gpu_case("room breach")
    from pantometry_amd.core import Violation
    g, o = _pair(4.4, 3.1, 37)
    p0 = o.mode2(2, 1, 5.0) + _gauss(o, 1.7, 1.2, 0.2)
    g.released_from(p0.reshape(o.ny, o.nx))
    o.release(p0)
    dt = 0.8 * o.max_stable_dt()
    g.advance(dt, 3, 1e-9)
    for _ in range(3):
        o.step(dt)
    try:
        g.advance(dt, 40, 1e-300)
        taken = 40
    except Violation as err:
        assert err.quantity == "energy"
        taken = err.step + 1
    for _ in range(taken):
        o.step(dt)
    _same_fields(g, o)
