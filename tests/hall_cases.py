"""Hall's closed-form checks (crates/pantometry-acoustic/tests/hall_closed_forms.rs), written once
against an adapter so the same bodies pin the CPU oracle and check the GPU domain."""
from __future__ import annotations

import math

import numpy as np


class Refused(Exception):
    def __init__(self, quantity, after=None):
        super().__init__(quantity)
        self.quantity = quantity
        self.after = after


class OracleRoom:
    def __init__(self, width, height, depth, across, speed=343.0, density=1.204):
        from oracle.oracle import OracleHall
        self.h = OracleHall(width, height, depth, across, speed, density)
        self.speed = speed

    def nodes(self):
        return self.h.nodes()

    def spacing(self):
        return self.h.dx

    def release_mode(self, mode, amp):
        self.h.release(self.h.mode(mode, amp))

    def max_stable_dt(self):
        return self.h.max_stable_dt()

    def courant(self, dt):
        return self.h.courant(dt)

    def step(self, dt):
        if self.h.courant(dt) > 1.0 + 1e-12:
            raise Refused("Courant number", self.h.courant(dt))
        self.h.step(dt)

    def peak(self):
        return float(np.abs(self.h.p).max())

    def energy(self):
        return self.h.energy()

    def startup(self):
        return self.h.offset

    def pressure(self):
        return self.h.p.copy()

    def mode_frequency(self, mode):
        return mode_frequency(self, mode)

    def modes_below(self, limit):
        """Hall::modes_below, hall.rs:286-316, over the oracle's node counts."""
        nx, ny, nz = self.nodes()
        dx = self.spacing()
        bounds = [int(math.floor(2.0 * limit * ((n - 1) * dx) / self.speed)) if n > 1 else 0
                  for n in (nx, ny, nz)]
        found = []
        for a in range(bounds[0] + 1):
            for b in range(bounds[1] + 1):
                for c in range(bounds[2] + 1):
                    if a == b == c == 0:
                        continue
                    f = mode_frequency(self, (a, b, c))
                    if f < limit:
                        found.append(((a, b, c), f))
        found.sort(key=lambda m: m[1])
        return found


class GpuRoom:
    def __init__(self, width, height, depth, across, speed=343.0, density=1.204):
        import pantometry_amd as pm
        self.pm = pm
        self.h = pm.Hall("hall", width, height, depth, across, speed, density)
        self.speed = speed

    def nodes(self):
        return self.h.nodes()

    def spacing(self):
        return self.h.spacing()

    def release_mode(self, mode, amp):
        self.h.released_in_mode(mode, amp)

    def max_stable_dt(self):
        return self.h.max_stable_dt()

    def courant(self, dt):
        return self.h.courant(dt)

    def step(self, dt):
        try:
            self.h.step(0.0, dt, self.pm.Exchange())
        except self.pm.Violation as v:
            raise Refused(v.quantity, v.after)

    def peak(self):
        return self.h.peak_pressure()

    def energy(self):
        return self.h.energy()

    def startup(self):
        return self.h.startup_adjustment()

    def pressure(self):
        return self.h.fields()[0]

    def mode_frequency(self, mode):
        return self.h.mode_frequency(mode)

    def modes_below(self, limit):
        return self.h.modes_below(limit)


def hall(make, across):
    return make(4.4, 3.1, 2.4, across)


def mode_frequency(room, mode):
    nx, ny, nz = room.nodes()
    dx = room.spacing()
    total = 0.0
    for n, count in zip(mode, (nx, ny, nz)):
        length = (count - 1) * dx
        total += (n / length) ** 2 if length > 0.0 else 0.0
    return room.speed / 2.0 * math.sqrt(total)


def a_released_mode_tracks_the_cosine_at_second_order(make):
    """hall_closed_forms.rs:99-123."""
    def worst(across):
        h = hall(make, across)
        h.release_mode((1, 1, 1), 1.0)
        f = mode_frequency(h, (1, 1, 1))
        dt = h.max_stable_dt() * 0.9
        steps = int(round(0.006 / dt))
        t = 0.0
        w = 0.0
        for _ in range(steps):
            h.step(dt)
            t += dt
            w = max(w, abs(h.peak() - abs(math.cos(2.0 * math.pi * f * t))))
        return w

    coarse, fine = worst(13), worst(97)
    assert coarse / fine > 30.0, (coarse, fine)
    assert fine < 1e-3


def a_rigid_hall_conserves_its_invariant_and_reports_its_startup(make):
    """hall_closed_forms.rs:132-178."""
    def adjustment(across):
        h = hall(make, across)
        h.release_mode((1, 1, 1), 2.0)
        dt = h.max_stable_dt() * 0.9
        for _ in range(200):
            h.step(dt)
        energies = []
        for _ in range(200):
            h.step(dt)
            energies.append(h.energy())
        first = energies[0]
        drift = max(abs(e - first) for e in energies)
        return drift / abs(first), abs(h.startup()) / abs(first)

    drift, coarse = adjustment(13)
    assert drift < 1e-12, drift
    _, fine = adjustment(25)
    assert coarse > 1e-6
    assert 2.5 < coarse / fine < 6.0, (coarse, fine)


def the_courant_limit_is_the_diagonal_of_a_cube(make):
    """hall_closed_forms.rs:186-230."""
    h = hall(make, 45)
    dx = h.spacing()
    limit = h.max_stable_dt()
    assert abs(limit / (dx / (343.0 * math.sqrt(3.0))) - 1.0) < 1e-12
    assert abs(h.courant(limit) - 1.0) < 1e-12
    h = hall(make, 45)
    h.release_mode((1, 0, 0), 1.0)
    try:
        h.step(limit * 1.05)
        raise AssertionError("5% past the limit must be refused")
    except Refused as r:
        assert r.quantity == "Courant number"
        assert abs(r.after - 1.05) < 1e-9
    h = hall(make, 45)
    h.release_mode((1, 0, 0), 1.0)
    h.step(limit)


def collapsing_an_axis_reduces_to_the_lower_dimension(make):
    """hall_closed_forms.rs:239-264."""
    flat = make(4.4, 3.1, 0.0, 45)
    assert flat.nodes()[2] == 1
    flat.release_mode((1, 1, 0), 1.0)
    dt = flat.max_stable_dt() * 0.9
    peaks = []
    for _ in range(400):
        flat.step(dt)
        peaks.append(flat.peak())
    assert min(peaks) < 0.2
    assert max(peaks) <= 1.0 + 1e-9


def the_mode_frequencies_are_the_closed_form(make):
    """hall_closed_forms.rs:41-71: (c/2) sqrt((a/Lx)^2 + (b/Ly)^2 + (c/Lz)^2) including the
    vertical mode c/2Lz; and each mode, released and stepped, reaches its first zero at a quarter
    of that period (the projection onto the mode shape, interpolated between steps, within 1%)."""
    h = hall(make, 45)
    lx, ly, lz = 4.4, 3.1, 2.4

    def want(a, b, c):
        return 343.0 / 2.0 * math.sqrt((a / lx) ** 2 + (b / ly) ** 2 + (c / lz) ** 2)

    for mode in [(1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 1), (3, 2, 1)]:
        assert abs(h.mode_frequency(mode) - want(*mode)) < 1e-9, mode
    vertical = h.mode_frequency((0, 0, 1))
    assert abs(vertical - 343.0 / (2.0 * lz)) < 1e-9 and 60.0 < vertical < 85.0
    nx, ny, nz = h.nodes()
    for mode in [(1, 0, 0), (0, 0, 1), (3, 2, 1)]:
        g = hall(make, 45)
        g.release_mode(mode, 1.0)
        shape = np.einsum("k,j,i->kji", *[np.cos(m * math.pi * np.arange(n) / (n - 1)) if n > 1
                                          else np.ones(1) for m, n in zip(mode[::-1], (nz, ny, nx))])
        shape = shape.reshape(-1)
        norm = float(shape @ shape)
        dt = g.max_stable_dt() * 0.5
        quarter = 1.0 / (4.0 * want(*mode))
        t, prev = 0.0, 1.0
        while True:
            g.step(dt)
            t += dt
            amp = float(g.pressure().reshape(-1) @ shape) / norm
            if amp <= 0.0:
                crossing = t - dt * amp / (amp - prev)
                break
            prev = amp
            assert t < 2.0 * quarter, mode
        assert abs(crossing / quarter - 1.0) < 0.01, (mode, crossing, quarter)


def the_mode_count_grows_as_the_cube_of_frequency(make):
    """hall_closed_forms.rs:284-325: modes below 150 and 300 Hz, their ratio, the three-term Weyl
    estimate within 10%, the leading term alone 20% short, ascending order, (1,0,0) first."""
    h = hall(make, 45)
    low, high = len(h.modes_below(150.0)), len(h.modes_below(300.0))
    assert low > 10
    assert 5.5 <= high / low < 10.0
    lx, ly, lz = 4.4, 3.1, 2.4
    v = lx * ly * lz
    surface = 2.0 * (lx * ly + ly * lz + lz * lx)
    edges = 4.0 * (lx + ly + lz)
    k = 300.0 / 343.0
    weyl = 4.0 * math.pi / 3.0 * v * k ** 3 + math.pi / 4.0 * surface * k ** 2 + edges * k / 8.0
    assert 0.9 <= high / weyl < 1.1, (high, weyl)
    assert high / (4.0 * math.pi / 3.0 * v * k ** 3) > 1.2
    modes = h.modes_below(120.0)
    assert all(a[1] <= b[1] for a, b in zip(modes, modes[1:]))
    assert tuple(modes[0][0]) == (1, 0, 0)


HALL_CASES = [
    a_released_mode_tracks_the_cosine_at_second_order,
    a_rigid_hall_conserves_its_invariant_and_reports_its_startup,
    the_courant_limit_is_the_diagonal_of_a_cube,
    collapsing_an_axis_reduces_to_the_lower_dimension,
    the_mode_frequencies_are_the_closed_form,
    the_mode_count_grows_as_the_cube_of_frequency,
]
