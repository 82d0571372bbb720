"""Solid3D's closed-form checks (crates/pantometry-thermal/tests/solid_closed_forms.rs,
a_layered_wall.rs, a_composite.rs), written once against a small adapter so the same bodies pin
the CPU oracle and check the GPU domain.

Every check is against a closed form or an invariant, never against another implementation.
"""
from __future__ import annotations

import math

import numpy as np

ALU = (2.70 * 1e3, 167.0, 896.0)          # aluminium_6061: density, conductivity, c_p
GLASS = (2.51 * 1e3, 1.114, 858.0)        # borosilicate_crown


def celsius(v):
    return v + 273.15


class Refused(Exception):
    def __init__(self, quantity, after=None):
        super().__init__(quantity)
        self.quantity = quantity
        self.after = after


class OracleBlock:
    def __init__(self, counts, dx_mm, sub=ALU, initial=celsius(20.0)):
        from oracle.oracle import OracleSolid3D
        self.b = OracleSolid3D("block", sub[0], sub[1], sub[2], counts, dx_mm * 1e-3, initial)
        self.sub = sub

    def release_mode(self, mode, mean, amp):
        self.b.release_mode(mode, mean, amp)

    def step(self, dt, gained=0.0):
        why = self.b.refuses(dt)
        if why:
            raise Refused(why, dt * self.b.worst_rate * (1.0 / 6.0))
        self.b.step(dt, gained)

    def max_stable_dt(self):
        return self.b.max_stable_dt()

    def mode_amplitude(self, mode):
        return self.b.mode_amplitude(mode)

    def mode_amplification(self, mode, dt):
        return self.b.mode_amplification(mode, dt)

    def grid(self):
        return self.b.grid()

    def deposit(self, i, j, k, joules):
        self.b.deposit(i, j, k, joules)

    def fill(self, sub, mask):
        self.b.fill(sub[0], sub[1], sub[2], mask)

    def face_z(self, k):
        """Conductance of the z face between cells k-1 and k of column (0, 0): k_face * dx."""
        nx, ny, nz = self.b.counts_
        return self.b.kz[nx * ny * k] * self.b.dx

    def stored(self):
        return self.b.stored_heat()

    def set_temperature(self, i, j, k, t):
        nx, ny, _ = self.b.counts_
        self.b.cells[i + nx * (j + ny * k)] = t


class GpuBlock:
    def __init__(self, counts, dx_mm, sub=ALU, initial=celsius(20.0)):
        import pantometry_amd as pm
        self.pm = pm
        s = pm.Substance("s", *sub)
        self.b = pm.Solid3D("block", s, counts, dx_mm * 1e-3, initial)
        self.sub = sub

    def release_mode(self, mode, mean, amp):
        self.b.release_mode(mode, mean, amp)

    def step(self, dt, gained=0.0):
        bus = self.pm.Exchange()
        if gained:
            bus.publish(self.pm.HEAT, gained)
        try:
            self.b.step(0.0, dt, bus)
        except self.pm.Violation as v:
            raise Refused(v.quantity, v.after)

    def max_stable_dt(self):
        return self.b.max_stable_dt()

    def mode_amplitude(self, mode):
        return self.b.mode_amplitude(mode)

    def mode_amplification(self, mode, dt):
        return self.b.mode_amplification(mode, dt)

    def grid(self):
        return self.b.cells()

    def deposit(self, i, j, k, joules):
        self.b.deposit(i, j, k, joules)

    def fill(self, sub, mask):
        self.b.fill(self.pm.Substance("t", *sub), mask)

    def face_z(self, k):
        return None

    def stored(self):
        return self.b.books()[0]

    def set_temperature(self, i, j, k, t):
        self.b.set_temperature(i, j, k, t)


def run(block, dt, steps):
    for _ in range(steps):
        block.step(dt)


# ---- solid_closed_forms.rs ---------------------------------------------------------------

def a_separable_mode_decays_at_its_exact_discrete_rate(make):
    """solid_closed_forms.rs:46-72."""
    for mode in [(1, 0, 0), (0, 1, 0), (0, 0, 1), (2, 1, 1)]:
        b = make((7, 5, 3), 2.0)
        b.release_mode(mode, celsius(20.0), 4.0)
        dt = b.max_stable_dt() * 0.8
        per_step = b.mode_amplification(mode, dt)
        assert abs(per_step) <= 1.0
        run(b, dt, 25)
        want = 4.0 * per_step ** 25
        got = b.mode_amplitude(mode)
        assert abs(got - want) < 1e-12 * 4.0, (mode, got, want)
        assert abs(got) < 0.9 * 4.0


def the_axes_are_the_same_physics_in_a_different_order(make):
    """solid_closed_forms.rs:84-109."""
    def after(counts, mode):
        b = make(counts, 2.0)
        b.release_mode(mode, celsius(20.0), 4.0)
        run(b, 5e-3, 200)
        return b.mode_amplitude(mode)

    x = after((7, 5, 3), (1, 0, 0))
    z = after((5, 3, 7), (0, 0, 1))
    y = after((3, 7, 5), (0, 1, 0))
    assert abs(x - z) < 1e-12 and abs(x - y) < 1e-12
    assert abs(x) < 0.2 * 4.0


def an_insulated_block_conserves_exactly(make):
    """solid_closed_forms.rs:156-192."""
    b = make((9, 7, 5), 1.5)
    b.release_mode((2, 1, 1), celsius(20.0), 30.0)
    g = b.grid()
    volume = 9 * 7 * 5 * (1.5e-3 ** 3)
    capacity = (ALU[0] * volume) * ALU[2]
    before = float(np.cumsum(g.reshape(-1))[-1]) / g.size
    spread = g.max() - g.min()
    assert spread > 50.0
    run(b, 1e-4, 600)
    g = b.grid()
    after = float(np.cumsum(g.reshape(-1))[-1]) / g.size
    assert capacity * abs(after - before) < 1e-9 * capacity * spread
    assert g.max() - g.min() < 0.2 * spread


def the_limit_is_one_sixth_and_it_is_refused(make):
    """solid_closed_forms.rs:199-242."""
    b = make((5, 5, 5), 2.0)
    dx = 2e-3
    alpha = ALU[1] / (ALU[0] * ALU[2])
    limit = b.max_stable_dt()
    exact = dx * dx / (6.0 * alpha)
    assert abs(limit / exact - 1.0) < 1e-12
    assert abs(dx * dx / (2.0 * alpha) / limit - 3.0) < 1e-9
    b = make((5, 5, 5), 2.0)
    g0 = b.grid().copy()
    try:
        b.step(limit * 1.05)
        raise AssertionError("5% past the limit must be refused")
    except Refused as r:
        assert r.quantity == "Fourier number"
        assert abs(r.after / (1.0 / 6.0) - 1.05) < 1e-9
    assert np.array_equal(b.grid(), g0), "a refused step mutates nothing"
    b = make((5, 5, 5), 2.0)
    b.step(limit)


def placeless_heat_leaves_the_block_uniform(make):
    """solid_closed_forms.rs:254-287."""
    b = make((4, 3, 2), 3.0)
    joules = 12.0
    b.step(1e-5, gained=joules)
    g = b.grid()
    first = g[0, 0, 0]
    assert np.all(np.abs(g - first) < 1e-12)
    volume = 4 * 3 * 2 * (3e-3 ** 3)
    capacity = (ALU[0] * volume) * ALU[2]
    assert abs((first - celsius(20.0)) - joules / capacity) < 1e-12


def a_hot_spot_spreads_the_same_way_along_every_axis(make):
    """solid_closed_forms.rs:296-329."""
    b = make((9, 9, 9), 1.0)
    b.deposit(4, 4, 4, 2.0)
    run(b, 2e-6, 5)
    g = b.grid()
    arms = [g[4, 4, 3], g[4, 4, 5], g[4, 3, 4], g[4, 5, 4], g[3, 4, 4], g[5, 4, 4]]
    ambient = celsius(20.0)
    for a in arms:
        assert a > ambient + 1e-6
        assert abs(a - arms[0]) < 1e-12 * (arms[0] - ambient)
    corner = g[4, 3, 3]
    assert ambient < corner < arms[0]


def one_cell_thick_reduces_to_a_bar(make):
    """solid_closed_forms.rs:341-363."""
    n = 12
    b = make((n, 1, 1), 2.0)
    b.release_mode((1, 0, 0), celsius(20.0), 5.0)
    dt = 2e-5
    alpha = ALU[1] / (ALU[0] * ALU[2])
    f = alpha * dt / (2e-3 * 2e-3)
    lam = -4.0 * math.sin(math.pi / (2.0 * n)) ** 2
    want_per_step = 1.0 + f * lam
    assert abs(b.mode_amplification((1, 0, 0), dt) - want_per_step) < 1e-15
    run(b, dt, 30)
    assert abs(b.mode_amplitude((1, 0, 0)) - 5.0 * want_per_step ** 30) < 1e-12 * 5.0


def the_limit_is_where_the_sharpest_mode_stops_growing(make):
    """solid_closed_forms.rs:503-542."""
    n = 5
    sharp = (n - 1, n - 1, n - 1)
    b = make((n, n, n), 2.0)
    b.release_mode(sharp, celsius(20.0), 1.0)
    limit = b.max_stable_dt()
    run(b, limit, 60)
    assert abs(b.mode_amplitude(sharp)) <= 1.0 + 1e-9
    per_step = make((n, n, n), 2.0).mode_amplification(sharp, limit)
    assert -1.0 <= per_step < -0.5
    b = make((n, n, n), 2.0)
    try:
        b.step(limit * 3.0)
        raise AssertionError("three times the limit must be refused")
    except Refused:
        pass


# ---- a_layered_wall.rs / a_composite.rs --------------------------------------------------

def filling_a_block_with_its_own_substance_changes_nothing(make):
    """a_layered_wall.rs: a uniform fill is a no-op, to the bit."""
    def build():
        b = make((5, 4, 3), 2.0)
        b.deposit(2, 2, 1, 4.0)
        return b
    plain, filled = build(), build()
    filled.fill(ALU, np.ones((3, 4, 5), dtype=bool))
    dt = plain.max_stable_dt() * 0.5
    run(plain, dt, 40)
    run(filled, dt, 40)
    assert np.array_equal(plain.grid(), filled.grid())


def a_layered_wall_conserves_over_many_steps(make):
    """a_layered_wall.rs:363-416: two materials, 4000 steps, books within 4000 eps."""
    cells = 24
    b = make((1, 1, cells), 1.0)
    mask = np.zeros((cells, 1, 1), dtype=bool)
    mask[cells // 2:, 0, 0] = True
    b.fill(GLASS, mask)
    b.deposit(0, 0, 3, 5.0)
    before = b.stored()
    dt = b.max_stable_dt() * 0.9
    run(b, dt, 4000)
    after = b.stored()
    assert abs(after - before) <= 4000 * 2.2e-16 * max(abs(before), 5.0)


def a_layered_wall_has_the_series_resistance_of_its_layers(make):
    """a_layered_wall.rs:136: sum of 1/G over the z faces equals sum of L/(kA), exactly."""
    dx = 1e-3
    area = dx * dx
    for cells in [12, 24, 48, 96]:
        b = make((1, 1, cells), 1.0)
        mask = np.zeros((cells, 1, 1), dtype=bool)
        mask[cells // 2:, 0, 0] = True
        b.fill(GLASS, mask)
        if b.face_z(1) is None:
            return
        half = cells // 2
        closed = (half - 0.5) * dx / (ALU[1] * area) + (cells - half - 0.5) * dx / (GLASS[1] * area)
        got = sum(1.0 / b.face_z(k) for k in range(1, cells))
        assert abs(got / closed - 1.0) < 1e-14


SOLID_CASES = [
    a_separable_mode_decays_at_its_exact_discrete_rate,
    the_axes_are_the_same_physics_in_a_different_order,
    an_insulated_block_conserves_exactly,
    the_limit_is_one_sixth_and_it_is_refused,
    placeless_heat_leaves_the_block_uniform,
    a_hot_spot_spreads_the_same_way_along_every_axis,
    one_cell_thick_reduces_to_a_bar,
    the_limit_is_where_the_sharpest_mode_stops_growing,
    filling_a_block_with_its_own_substance_changes_nothing,
    a_layered_wall_conserves_over_many_steps,
    a_layered_wall_has_the_series_resistance_of_its_layers,
]
