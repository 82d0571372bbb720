"""The Room oracle pinned against the reference's own room tests
(crates/pantometry-acoustic/src/room.rs:755-1335) and against the mathematics: the Courant
limit's closed form, the mode frequency a stepped mode must oscillate at, exact conservation of
the time-centred invariant, and second-order convergence of the startup."""
import math

import numpy as np
import pytest

from oracle.oracle import OracleRoom, audit_breaks

AIR = dict(density=1.204, speed=343.0)


def of_air(width, height, cells):
    """Room::of_air, room.rs:122-137: one metre deep."""
    return OracleRoom(width, height, cells, 1.0, AIR["speed"], AIR["density"])


def square(cells):
    """room.rs:755-757."""
    return of_air(4.0, 4.0, cells)


def released(room, profile):
    nx, ny = room.nx, room.ny
    p = np.empty(nx * ny)
    for j in range(ny):
        for i in range(nx):
            p[i + nx * j] = profile(float(i) * room.dx, float(j) * room.dx)
    room.release(p)
    return room


def in_mode(room, mx, my, amplitude):
    room.release(room.mode2(mx, my, amplitude))
    return room


def mode_frequency(room, mx, my):
    """room.rs:291-300."""
    lx, ly = (room.nx - 1) * room.dx, (room.ny - 1) * room.dx
    if lx <= 0.0 or ly <= 0.0:
        return 0.0
    a, b = mx / lx, my / ly
    return room.speed / 2.0 * math.sqrt(a * a + b * b)


def test_node_counts_follow_room_new():
    # room.rs:1008-1020: the width is exact, the height quantised to whole cells.
    room = of_air(5.0, 3.0, 51)
    assert abs((room.nx - 1) * room.dx - 5.0) < 1e-12
    assert abs((room.ny - 1) * room.dx - 3.0) < 0.11
    assert (room.nx, room.ny, room.nz) == (51, 31, 1)
    # room.rs:1304-1309: at least three nodes each way.
    tiny = of_air(0.1, 0.1, 1)
    assert tiny.nx >= 3 and tiny.ny >= 3


def test_two_dimensions_tighten_the_courant_limit_by_root_two():
    # room.rs:1074-1102.
    room = square(41)
    limit = room.max_stable_dt()
    assert abs(limit * 1e6 - 206.2) < 0.5
    assert abs((0.1 / 343.0) / limit - math.sqrt(2.0)) < 1e-9
    assert abs(room.courant(limit) - 1.0) < 1e-12
    room.step(limit)
    with pytest.raises(ValueError, match="Courant"):
        room.step(limit * 1.01)


@pytest.mark.parametrize("mode", [(1, 0), (1, 1), (2, 1)])
def test_an_excited_mode_oscillates_at_its_predicted_frequency(mode):
    # room.rs:1104-1141: half a period inverts the corner, a whole one restores it.
    room = in_mode(square(81), mode[0], mode[1], 100.0)
    period = 1.0 / mode_frequency(room, *mode)
    dt = room.max_stable_dt() * 0.5
    steps = int(round(period / dt))
    start = room.p[0]
    assert abs(start) > 50.0
    for _ in range(steps // 2):
        room.step(dt)
    assert room.p[0] < -0.85 * start
    for _ in range(steps - steps // 2):
        room.step(dt)
    assert abs(room.p[0] / start - 1.0) < 0.1


def test_a_closed_room_keeps_its_energy():
    # room.rs:1143-1167 (2 % over 3000 steps), and, tighter, what makes the scheme's audit
    # possible: the time-centred invariant is conserved step to step to rounding.
    def pulse(x, y):
        r2 = (x - 2.0) ** 2 + (y - 2.0) ** 2
        return 200.0 * math.exp(-r2 / 0.09)

    room = released(square(61), pulse)
    start = room.energy()
    assert start > 0.0
    dt = room.max_stable_dt() * 0.5
    room.step(dt)
    prev = room.energy()
    worst = 0.0
    for _ in range(2999):
        room.step(dt)
        now = room.energy()
        worst = max(worst, abs(now - prev) / abs(now))
        prev = now
    assert abs(prev / start - 1.0) < 0.02
    assert worst < 1e-12


def test_the_startup_adjustment_is_second_order_and_declared():
    # room.rs:1271-1296.
    def relative(cells):
        room = in_mode(square(cells), 1, 1, 1.0)
        total = room.energy()
        room.step(room.max_stable_dt())
        return abs(room.offset) / total

    a, b = relative(31), relative(121)
    assert a > 1e-4
    assert 10.0 <= a / b < 24.0


def test_starting_the_leapfrog_correctly_makes_the_scheme_second_order():
    # room.rs:1225-1266: a 4.4 x 3.1 room's (1,1) mode against |cos(2 pi f t)| over 20 ms;
    # the worst departure falls by more than 30x from 31 to 241 cells (64x for second order).
    def worst(cells):
        room = in_mode(of_air(4.4, 3.1, cells), 1, 1, 1.0)
        f = mode_frequency(room, 1, 1)
        steps = int(math.ceil(0.02 / room.max_stable_dt()))
        h = 0.02 / steps
        t, bad = 0.0, 0.0
        for _ in range(steps):
            room.step(h)
            t += h
            want = abs(math.cos(2.0 * math.pi * f * t))
            bad = max(bad, abs(float(np.max(np.abs(room.p))) - want))
        return bad

    assert worst(31) / worst(241) > 30.0


def test_the_domain_holds_its_books_under_the_multirate_schedule():
    # room.rs:1203-1222: 1 ms at a 206 us limit is 5 substeps, and 21 ms of them hold the
    # ledger to 1e-9 step after step.
    def pulse(x, y):
        return 200.0 * math.exp(-((x - 2.0) ** 2 + (y - 2.0) ** 2) / 0.5)

    room = released(square(41), pulse)
    substeps = int(math.ceil(1e-3 / room.max_stable_dt() - 1e-12))
    assert substeps == 5
    before = room.ledger()
    for _ in range(21 * substeps):
        room.step(1e-3 / substeps)
        after = room.ledger()
        assert not audit_breaks(before, after, 1e-9)
        before = after


def test_the_field_samples_the_room_and_stops_at_its_walls():
    # room.rs:764-797.
    room = in_mode(square(33), 1, 0, 1.0)
    lx = (room.nx - 1) * room.dx
    assert abs(room.sample(0.0, 2.0, 0.0) - 1.0) < 1e-12
    assert abs(room.sample(lx, 2.0, 0.0) + 1.0) < 1e-12
    assert abs(room.sample(lx / 2.0, 2.0, 0.0)) < 1e-12
    dx = lx / 32.0
    ends = (room.sample(0.0, 2.0, 0.0) + room.sample(dx, 2.0, 0.0)) / 2.0
    assert abs(room.sample(dx * 0.5, 2.0, 0.0) - ends) < 1e-12
    assert abs(room.sample(-9.0, 2.0, 0.0) - 1.0) < 1e-12
    assert abs(room.sample(lx + 9.0, 2.0, 0.0) + 1.0) < 1e-12
    assert math.isfinite(room.sample(math.nan, math.nan, 0.0))
    assert room.sample(1.0, 2.0, 0.0) == room.sample(1.0, 2.0, 55.0)


def test_the_field_is_signed_about_zero_and_peaks_at_the_amplitude():
    # room.rs:986-1003 and crates/pantometry/tests/domain_readings.rs:68-82.
    room = in_mode(square(33), 1, 1, 2.0)
    lx, ly = (room.nx - 1) * room.dx, (room.ny - 1) * room.dx
    corners = [room.sample(x, y, 0.0) for x, y in ((0, 0), (lx, 0), (0, ly), (lx, ly))]
    assert any(v > 1.9 for v in corners) and any(v < -1.9 for v in corners)
    assert abs(room.p.sum() / room.p.size) < 1e-12
    other = in_mode(of_air(4.4, 3.1, 41), 1, 1, 1.0)
    assert abs(float(np.max(np.abs(other.p))) - 1.0) < 1e-9


def test_degenerate_rooms_are_handled():
    # room.rs:1303-1334.
    silent = OracleRoom(1.0, 1.0, 11, 1.0, 0.0, 1.0)
    assert not math.isfinite(silent.max_stable_dt())
    silent.step(1.0)
    assert float(np.max(np.abs(silent.p))) == 0.0
    room = released(square(21), lambda x, y: 7.0)
    before = room.p[5 + room.nx * 5]
    room.step(0.0)
    assert room.p[5 + room.nx * 5] == before


def _room_step_by_hand(nx, ny, dx, speed, density, h, vh, p, vx, vy):
    """Room::step's two sweeps (room.rs:384-440), restated loop by loop in Python floats."""
    rc2 = density * speed * speed

    def ww(k, n):
        return 2.0 if k == 0 or k == n - 1 else 1.0

    for j in range(ny):
        for i in range(nx - 1):
            vx[j * (nx - 1) + i] -= vh / (density * dx) * (p[j * nx + i + 1] - p[j * nx + i])
    for j in range(ny - 1):
        for i in range(nx):
            vy[j * nx + i] -= vh / (density * dx) * (p[(j + 1) * nx + i] - p[j * nx + i])
    prev = list(p)
    for j in range(ny):
        for i in range(nx):
            left = 0.0 if i == 0 else vx[j * (nx - 1) + i - 1]
            right = 0.0 if i == nx - 1 else vx[j * (nx - 1) + i]
            below = 0.0 if j == 0 else vy[(j - 1) * nx + i]
            above = 0.0 if j == ny - 1 else vy[j * nx + i]
            p[j * nx + i] -= h * rc2 / dx * (ww(i, nx) * (right - left) + ww(j, ny) * (above - below))
    return prev


def test_the_room_sweep_is_room_step_to_the_bit():
    # The oracle steps a room with the hall's compiled sweep at one node deep; a restatement
    # of Room::step in Python on a small ragged room, started with the half step, agrees to
    # the bit after every step, and the room's invariant follows room.rs:243-281.
    rng = np.random.default_rng(7)
    room = OracleRoom(1.3, 0.9, 9, 0.7, 343.0, 1.204)
    nx, ny, dx = room.nx, room.ny, room.dx
    assert (nx, ny) == (9, 7)
    a = rng.uniform(0.5, 2.0, size=3)
    p0 = a[0] * room.mode2(1, 0, 1.0) + a[1] * room.mode2(0, 1, 1.0) + a[2] * room.mode2(1, 1, 1.0)
    room.release(p0)
    p, vx, vy = list(p0), [0.0] * ((nx - 1) * ny), [0.0] * (nx * (ny - 1))
    h = room.max_stable_dt() * 0.93
    for n in range(25):
        prev = _room_step_by_hand(nx, ny, dx, 343.0, 1.204, h, h if n else 0.5 * h, p, vx, vy)
        room.step(h)
        assert np.array_equal(room.p, np.array(p))
        assert np.array_equal(room.pp, np.array(prev))
        assert np.array_equal(room.vx, np.array(vx))
        assert np.array_equal(room.vy, np.array(vy))
    volume, rc2 = dx * dx * 0.7, 1.204 * 343.0 ** 2

    def ww(k, n):
        return 2.0 if k == 0 or k == n - 1 else 1.0

    pot = sum(1.0 / (ww(k % nx, nx) * ww(k // nx, ny)) * p[k] * prev[k] / (2.0 * rc2) * volume
              for k in range(nx * ny))
    kx = sum(1.204 * u * u / 2.0 * volume / ww(k // (nx - 1), ny) for k, u in enumerate(vx))
    ky = sum(1.204 * u * u / 2.0 * volume / ww(k % nx, nx) for k, u in enumerate(vy))
    assert room.invariant() == pot + kx + ky
