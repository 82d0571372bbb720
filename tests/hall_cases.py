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


HALL_CASES = [
    a_released_mode_tracks_the_cosine_at_second_order,
    a_rigid_hall_conserves_its_invariant_and_reports_its_startup,
    the_courant_limit_is_the_diagonal_of_a_cube,
    collapsing_an_axis_reduces_to_the_lower_dimension,
]
