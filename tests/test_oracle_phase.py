"""The Solid3D oracle's enthalpy phase change, pinned against the reference's own tests
(crates/pantometry-thermal/tests/a_freezing_front.rs, two_phase_stefan.rs): Neumann's solution for
the front position and the profile behind it, the latent plateau, the inverse map across the whole
latent heat, the ledger, checkpoints, and the two-phase Neumann solution."""
import math

import numpy as np
import pytest

from oracle.oracle import OracleSolid3D

K, RHO, CP, LATENT = 2.22, 917.0, 2050.0, 333_550.0
ALPHA = K / (RHO * CP)
K_L, C_L = 0.598, 4182.0
ALPHA_L = K_L / (RHO * C_L)
C0 = 273.15
# Substance::ice (crates/pantometry-core/src/substance.rs:613-650): 0.917 g/cm3, one-phase.
ICE = (0.917 * 1000.0, 2.22, 2050.0, 0.97)
ICE_FUSION = (C0, 333.55 * 1000.0, None)
TWO_PHASE_FUSION = (C0, 333.55 * 1000.0, (K_L, C_L))
DEPTH, DX = 40e-3, 1e-3


def ice_column(dx, fusion=ICE_FUSION, initial_c=0.0):
    """a_freezing_front.rs:123-134: liquid at exactly its melting point."""
    cells = int(round(DEPTH / dx))
    c = OracleSolid3D("ice", ICE[0], ICE[1], ICE[2], (1, 1, cells), dx, initial_c + C0, ICE[3],
                      fusion)
    for k in range(cells):
        c.set_melted_fraction(0, 0, k, 1.0)
    return c


def frozen_depth(c, dx):
    """a_freezing_front.rs:144-149."""
    return DEPTH - c.melted_volume() / (dx * dx) - 0.5 * dx


def neumann_lambda(stefan):
    target = stefan / math.sqrt(math.pi)
    lo, hi = 0.0, 1.0
    for _ in range(60):
        mid = 0.5 * (lo + hi)
        if mid * math.exp(mid * mid) * math.erf(mid) - target < 0.0:
            lo = mid
        else:
            hi = mid
    return 0.5 * (lo + hi)


def march(dx, undercooling, until):
    cold = C0 - undercooling
    c = ice_column(dx)
    dt = c.max_stable_dt() * 0.9
    out, t = [], 0.0
    for want_at in until:
        while t < want_at:
            c.set_temperature(0, 0, 0, cold)
            c.step(dt)
            t += dt
        c.set_temperature(0, 0, 0, cold)
        out.append((t, frozen_depth(c, dx)))
    return out, c


def test_the_front_advances_as_neumanns_solution_says():
    # a_freezing_front.rs:181-262.
    lam = neumann_lambda(CP * 20.0 / LATENT)
    coefficient = 2.0 * lam * math.sqrt(ALPHA)
    times = [100.0, 400.0, 900.0]
    by_dx = []
    for dx in (DX, DX / 2.0):
        errors = []
        for t, got in march(dx, 20.0, times)[0]:
            closed = coefficient * math.sqrt(t)
            off = abs(got / closed - 1.0)
            assert off < (dx / closed) ** 2
            errors.append(off)
        by_dx.append(errors)
    for coarse, fine in zip(*by_dx):
        assert fine < coarse
    fine = march(DX / 2.0, 20.0, times)[0]
    assert abs(fine[2][1] / fine[0][1] / 3.0 - 1.0) < 1e-3


def test_behind_the_front_the_profile_is_an_error_function():
    # a_freezing_front.rs:267-385: fixed probes behind the front, the liquid ahead untouched to
    # the bit, first-to-second order over a fourfold refinement.
    lam = neumann_lambda(CP * 20.0 / LATENT)
    means = []
    for dx in (DX, DX / 2.0, DX / 4.0):
        cells = int(round(DEPTH / dx))
        c = ice_column(dx)
        dt = c.max_stable_dt() * 0.9
        t, errors = 0.0, []
        ahead = 0
        for want_at in (300.0, 450.0, 600.0, 750.0, 900.0):
            while t < want_at:
                c.set_temperature(0, 0, 0, C0 - 20.0)
                c.step(dt)
                t += dt
            c.set_temperature(0, 0, 0, C0 - 20.0)
            front = frozen_depth(c, dx)
            scale = 2.0 * math.sqrt(ALPHA * t)
            worst = 0.0
            for x in (2e-3, 4e-3, 6e-3, 8e-3):
                k = int(round(x / dx))
                want = C0 - 20.0 + 20.0 * math.erf(x / scale) / math.erf(lam)
                worst = max(worst, abs(c.temperature_at(0, 0, k) - want))
            errors.append(worst)
            ahead = 0
            for k in range(1, cells):
                if k * dx > front + dx:
                    assert c.temperature_at(0, 0, k) == C0
                    assert c.melted_fraction_at(0, 0, k) == 1.0
                    ahead += 1
        assert ahead > 10
        means.append(sum(errors) / len(errors))
    assert means[1] < means[0] and means[2] < means[1]
    overall = math.log2(means[0] / means[2]) / 2.0
    assert 1.0 < overall < 2.0
    assert means[0] < 0.01 * 20.0


def _cube():
    return OracleSolid3D("cube", ICE[0], ICE[1], ICE[2], (1, 1, 1), DX, C0, ICE[3], ICE_FUSION)


def test_a_cell_at_its_melting_point_holds_there_for_exactly_the_latent_heat():
    # a_freezing_front.rs:395-455.
    c = _cube()
    latent, capacity = c.latent[0], c.c_solid[0]
    assert abs(latent / capacity - LATENT / CP) < 1e-9
    power, dt = 1e-3, 0.1
    want = latent / power
    t, left_at = 0.0, None
    while t < 2.0 * want:
        c.deposit(0, 0, 0, power * dt)
        t += dt
        if left_at is None and c.temperature_at(0, 0, 0) > C0:
            left_at = t
    assert abs(left_at - want) <= dt * 1.001
    assert c.temperature_at(0, 0, 0) > C0


@pytest.mark.parametrize("multiple", [0.5, 1.0, 1.0000001, 3.0, 10.0])
def test_a_step_that_crosses_the_whole_latent_heat_deposits_the_remainder(multiple):
    # a_freezing_front.rs:465-527.
    c = _cube()
    latent, capacity = c.latent[0], c.c_solid[0]
    joules = multiple * latent
    opening = c.ledger()[0]
    c.deposit(0, 0, 0, joules)
    want_phi, want_t = (1.0, C0 + (joules - latent) / capacity) if multiple >= 1.0 else (multiple, C0)
    assert abs(c.melted_fraction_at(0, 0, 0) - want_phi) < 1e-15
    assert abs(c.temperature_at(0, 0, 0) - want_t) < 1e-12 * (1.0 + abs(want_t - C0))
    assert abs((c.ledger()[0] - opening) / joules - 1.0) < 1e-14


def test_the_ledger_counts_the_latent_heat_and_not_only_the_cooling():
    # a_freezing_front.rs:537-589.
    c = ice_column(DX)
    opening = c.ledger()[0]
    dt = c.max_stable_dt() * 0.9
    t = 0.0
    while t < 600.0:
        c.set_temperature(0, 0, 0, C0 - 20.0)
        c.step(dt)
        t += dt
    c.set_temperature(0, 0, 0, C0 - 20.0)
    held = c.ledger()[0] - opening
    cap, lat = RHO * CP * DX ** 3, RHO * LATENT * DX ** 3
    sensible = fusion = 0.0
    for k in range(40):
        sensible += cap * (c.temperature_at(0, 0, k) - C0)
        fusion += lat * (c.melted_fraction_at(0, 0, k) - 1.0)
    assert abs(held - (sensible + fusion)) < 1e-12 * abs(fusion)
    assert abs(fusion) > 2.0 * abs(sensible)


def test_a_checkpoint_carries_the_phase_and_not_only_the_temperature():
    # a_freezing_front.rs:595-650.
    c = ice_column(DX)
    dt = c.max_stable_dt() * 0.9
    t = 0.0
    while t < 120.0:
        c.set_temperature(0, 0, 0, C0 - 20.0)
        c.step(dt)
        t += dt
    c.checkpoint()
    depth, held = frozen_depth(c, DX), c.ledger()[0]
    assert depth > 3e-3
    for _ in range(200):
        c.set_temperature(0, 0, 0, C0 - 20.0)
        c.step(dt)
    assert frozen_depth(c, DX) > depth
    c.restore()
    assert frozen_depth(c, DX) == depth
    assert c.ledger()[0] == held


def test_a_block_of_something_that_does_not_melt_is_unchanged():
    # a_freezing_front.rs:658-697: aluminium never melts, and its arithmetic is the plain sweep's.
    def build():
        b = OracleSolid3D("b", 2.70 * 1000.0, 167.0, 896.0, (4, 3, 5), 2e-3, 20.0 + C0, 0.09)
        b.deposit(2, 1, 2, 3.0)
        return b

    plain, reference = build(), build()
    assert plain.melted_fraction_at(2, 1, 2) == 0.0 and plain.melted_volume() == 0.0
    plain.set_melted_fraction(2, 1, 2, 1.0)
    assert plain.melted_fraction_at(2, 1, 2) == 0.0
    dt = reference.max_stable_dt() * 0.5
    for _ in range(60):
        plain.step(dt)
        reference.step(dt)
    assert np.array_equal(plain.cells, reference.cells)
    assert plain.melted_volume() == 0.0


def _two_phase_lambda(superheat, drop=20.0):
    """two_phase_stefan.rs:103-127: the two-phase Neumann condition by bisection."""
    nu = math.sqrt(ALPHA / ALPHA_L)
    rhs = math.sqrt(math.pi) / (CP * drop / LATENT)

    def f(lam):
        first = math.exp(-lam * lam) / math.erf(lam)
        second = (K_L / K) * nu * (superheat / drop) * math.exp(-nu * nu * lam * lam) / math.erfc(nu * lam)
        return first - second - rhs * lam

    lo, hi = 1e-9, 0.45
    for _ in range(80):
        mid = 0.5 * (lo + hi)
        if f(mid) > 0.0:
            lo = mid
        else:
            hi = mid
    return 0.5 * (lo + hi)


def test_a_superheated_liquid_slows_the_front_by_what_the_two_phase_form_says():
    # two_phase_stefan.rs:185-275.
    one = neumann_lambda(CP * 20.0 / LATENT)
    probe = OracleSolid3D("probe", ICE[0], ICE[1], ICE[2], (1, 1, 8), 1e-3, C0 - 1.0, ICE[3],
                          ICE_FUSION)
    dt = probe.max_stable_dt() * 0.9
    worst_two, worst_one = 0.0, 0.0
    for superheat in (0.0, 2.0, 5.0, 10.0, 20.0):
        want = 2.0 * _two_phase_lambda(max(superheat, 1e-12)) * math.sqrt(ALPHA * 900.0)
        one_says = 2.0 * one * math.sqrt(ALPHA * 900.0)
        c = ice_column(1e-3, TWO_PHASE_FUSION, initial_c=superheat)
        assert c.two_phase
        t = 0.0
        while t < 900.0:
            c.set_temperature(0, 0, 0, C0 - 20.0)
            c.step(dt)
            t += dt
        c.set_temperature(0, 0, 0, C0 - 20.0)
        got = DEPTH - c.melted_volume() / 1e-6 - 0.5e-3
        worst_two = max(worst_two, abs(got / want - 1.0))
        if superheat >= 20.0:
            worst_one = abs(got / one_says - 1.0)
    allowed = 1e-3 / (2.0 * one * math.sqrt(ALPHA * 900.0))
    assert worst_two < allowed
    assert worst_one > 3.0 * allowed


def test_the_two_models_differ_by_one_cell_and_refining_closes_it():
    # two_phase_stefan.rs:281-332.
    def run(fusion, cells, dx):
        c = OracleSolid3D("column", ICE[0], ICE[1], ICE[2], (1, 1, cells), dx, C0, ICE[3], fusion)
        for k in range(cells):
            c.set_melted_fraction(0, 0, k, 1.0)
        dt = 0.4 * dx * dx / (2.0 * ALPHA) * 0.9
        t = 0.0
        while t < 120.0:
            c.set_temperature(0, 0, 0, C0 - 20.0)
            c.step(dt)
            t += dt
        return c.melted_volume() / (dx * dx)

    gaps = []
    for cells, dx in ((20, 1e-3), (40, 5e-4), (80, 2.5e-4)):
        one, two = run(ICE_FUSION, cells, dx), run(TWO_PHASE_FUSION, cells, dx)
        gaps.append(abs(two / one - 1.0))
    assert gaps[1] < gaps[0] and gaps[2] < gaps[1]
    assert gaps[0] / gaps[2] > 2.5


def test_the_enthalpy_step_to_the_bit():
    # A column of six cells, five of ice in every state (solid, mushy, liquid, either side of
    # the melting point) and one of aluminium, stepped once by the oracle and by a restatement of
    # the sweep (solid.rs:1776-1826) and add_kelvin (solid.rs:1620-1648) in Python floats.
    n, dx = 6, 1e-3
    b = OracleSolid3D("b", ICE[0], ICE[1], ICE[2], (1, 1, n), dx, C0 - 3.0, ICE[3], ICE_FUSION)
    al = np.zeros(n, dtype=bool)
    al[4] = True
    b.fill(2.70 * 1000.0, 167.0, 896.0, al, 0.09)
    for k, t in enumerate((C0 - 5.0, C0, C0, C0 + 4.0, C0 + 9.0, C0 - 1.0)):
        b.set_temperature(0, 0, k, t)
    b.set_melted_fraction(0, 0, 1, 0.37)
    b.set_melted_fraction(0, 0, 2, 1.0)
    old, phi = [float(v) for v in b.cells], [float(v) for v in b.melted]
    vol = dx * dx * dx
    cs = [(ICE[0] * vol) * ICE[2]] * n
    cs[4] = (2.70 * 1000.0 * vol) * 896.0
    ks = [ICE[1]] * n
    ks[4] = 167.0
    latent = [(ICE[0] * vol) * 333.55 * 1000.0] * n
    latent[4] = 0.0

    def series(x, y):
        total = x + y
        return 2.0 * x * y / total if total > 0.0 else 0.0

    dt = 0.8 * b.max_stable_dt()
    b.step(dt)
    cells = list(old)
    for c in range(n):
        t = old[c]
        flux = 0.0
        if c > 0:
            flux += series(ks[c - 1], ks[c]) * (old[c - 1] - t)
        if c + 1 < n:
            flux += series(ks[c], ks[c + 1]) * (old[c + 1] - t)
        rise = dt * (dx / cs[c]) * (flux - 0.0)
        if latent[c] > 0.0:
            if phi[c] >= 1.0:
                e = latent[c] + cs[c] * (t - C0)
            elif phi[c] <= 0.0:
                e = cs[c] * (t - C0)
            else:
                e = phi[c] * latent[c]
            e = e + rise * cs[c]
            if e <= 0.0:
                cells[c], phi[c] = C0 + e / cs[c], 0.0
            elif e >= latent[c]:
                cells[c], phi[c] = C0 + (e - latent[c]) / cs[c], 1.0
            else:
                cells[c], phi[c] = C0, e / latent[c]
        else:
            cells[c] = t + rise
    assert np.array_equal(b.cells, np.array(cells))
    assert np.array_equal(b.melted, np.array(phi))
