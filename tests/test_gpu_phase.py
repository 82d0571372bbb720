"""Phase change on the GPU against the Solid3D oracle: bit-exact temperatures and melted fractions
through the enthalpy sweep, sources, gaps and bus heat, one- and two-phase; and the reference's
freezing-front and two-phase Stefan tests on the device
(crates/pantometry-thermal/tests/a_freezing_front.rs, two_phase_stefan.rs)."""
import math

import numpy as np
import pytest

from tests.conftest import gpu_case

pytestmark = pytest.mark.gpu

C0 = 273.15
K, RHO, CP, LATENT = 2.22, 917.0, 2050.0, 333_550.0
ALPHA = K / (RHO * CP)
K_L, C_L = 0.598, 4182.0
ICE = (0.917 * 1000.0, 2.22, 2050.0, 0.97)
AL = (2.70 * 1000.0, 167.0, 896.0, 0.09)
ICE_FUSION = (C0, 333.55 * 1000.0, None)
TWO_PHASE = (C0, 333.55 * 1000.0, (K_L, C_L))
DEPTH = 40e-3


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).reshape(-1).view(np.uint64)


def substance(t, fusion=None):
    import pantometry_amd as pm
    f = None if fusion is None else pm.Fusion(fusion[0], fusion[1], fusion[2])
    return pm.Substance.of("s", t[0], t[1], t[2], t[3], f)


def pair(counts, dx, initial, first=ICE, fusion=ICE_FUSION):
    import pantometry_amd as pm
    from oracle.oracle import OracleSolid3D
    g = pm.Solid3D("b", substance(first, fusion), counts, dx, initial)
    o = OracleSolid3D("b", first[0], first[1], first[2], counts, dx, initial, first[3], fusion)
    return g, o


def same(g, o):
    assert np.array_equal(bits(g.cells()), bits(o.cells))
    assert np.array_equal(bits(g.melted_fractions()), bits(o.melted))


def mask_of(counts, pred):
    nx, ny, nz = counts
    m = np.zeros(nx * ny * nz, dtype=np.uint8)
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                m[i + nx * (j + ny * k)] = 1 if pred(i, j, k) else 0
    return m


@pytest.mark.parametrize("two_phase", [False, True], ids=["one-phase", "two-phase"])
def test_a_melting_block_whose_sweep_marches_columns_is_bit_exact(two_phase):
    """70 x 20 x 70 cells (2 x 2 column tiles x 70 planes, at least the 256 work items of the
    sweep that marches z columns): ice with an aluminium slab, a source and a start around the
    melting point, temperatures and melted fractions bit for bit over a few steps."""
    gpu_case(f"phase, marching sweep, two_phase={two_phase}")
    from pantometry_amd.core import Exchange
    counts = (70, 20, 70)
    nx, ny, nz = counts
    n = nx * ny * nz
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    rng = np.random.default_rng(5 + two_phase)
    g, o = pair(counts, 1e-3, C0 - 2.0)
    al = ((i < 3) | ((j == 9) & (k > 30))).reshape(-1).astype(np.uint8)
    g.fill(substance(AL), al)
    o.fill(AL[0], AL[1], AL[2], al, AL[3])
    if two_phase:
        wet = (k >= 50).reshape(-1).astype(np.uint8)
        g.fill(substance(ICE, TWO_PHASE), wet)
        o.fill(ICE[0], ICE[1], ICE[2], wet, ICE[3], TWO_PHASE)
    start = C0 + rng.uniform(-3.0, 3.0, size=n)
    g.set_cells(start)
    o.set_cells(start)
    src = ((i > 60) & (k < 4)).reshape(-1).astype(np.uint8)
    g.dissipating(40.0, src)
    o.dissipating(40.0, src)
    dt = 0.5 * o.max_stable_dt()
    assert g.max_stable_dt() == o.max_stable_dt()
    bus = Exchange()
    for _ in range(4):
        g.step(0.0, dt, bus)
        o.step(dt)
    same(g, o)


@pytest.mark.parametrize("two_phase", [False, True], ids=["one-phase", "two-phase"])
def test_a_melting_block_with_everything_is_bit_exact(two_phase):
    gpu_case(f"phase parity two_phase={two_phase}")
    import pantometry_amd as pm
    counts = (9, 7, 6)
    n = 9 * 7 * 6
    rng = np.random.default_rng(31 + two_phase)
    g, o = pair(counts, 1e-3, C0 - 2.0)
    al = mask_of(counts, lambda i, j, k: i == 0 or (j == 3 and k == 1))
    g.fill(substance(AL), al)
    o.fill(AL[0], AL[1], AL[2], al, AL[3])
    if two_phase:
        wet = mask_of(counts, lambda i, j, k: k >= 4)
        g.fill(substance(ICE, TWO_PHASE), wet)
        o.fill(ICE[0], ICE[1], ICE[2], wet, ICE[3], TWO_PHASE)
    start = C0 + rng.uniform(-8.0, 8.0, size=n)
    g.set_cells(start)
    o.set_cells(start)
    for c in range(0, n, 7):
        i, j, k = c % 9, (c // 9) % 7, c // 63
        g.set_melted_fraction(i, j, k, 0.25 + 0.5 * (c % 3) / 2.0)
        o.set_melted_fraction(i, j, k, 0.25 + 0.5 * (c % 3) / 2.0)
    voids = mask_of(counts, lambda i, j, k: k == 2 and 2 <= i <= 6)
    g.empty(voids)
    o.empty(voids)
    g.losing_from(pm.Face.YMAX, pm.Environment(C0 + 25.0, 10.0, 6e-5))
    o.losing_from(3, C0 + 25.0, 10.0, 6e-5)
    src = mask_of(counts, lambda i, j, k: i == 8 and k == 0)
    g.dissipating(0.4, src)
    o.dissipating(0.4, src)
    same(g, o)
    assert g.max_stable_dt() == o.max_stable_dt()
    dt = 0.8 * o.max_stable_dt()
    for s in range(6):
        bus = pm.Exchange()
        gained = 0.01 * s
        if gained:
            bus.publish(pm.HEAT, gained)
            bus.covering(dt)
        g.step(0.0, dt, bus)
        o.step(dt, gained)
        same(g, o)
    assert abs(g.books()[0] - o.stored_heat()) <= 1e-12 * max(1.0, abs(o.stored_heat()))
    g.checkpoint()
    o.checkpoint()
    g.advance(dt, 25, tolerance=1e-6)
    for _ in range(25):
        o.step(dt)
    same(g, o)
    assert g.max_stable_dt() == o.max_stable_dt()
    assert g.melted_volume() == o.melted_volume()
    g.restore()
    o.restore()
    same(g, o)
    g.deposit(4, 3, 5, 0.05)
    o.deposit(4, 3, 5, 0.05)
    same(g, o)
    readings = dict((r[1], r[2]) for r in g.readings())
    assert readings["melted"] == o.melted_volume() * 1e9


def _column(dx, fusion=ICE_FUSION, initial_c=0.0):
    import pantometry_amd as pm
    from oracle.oracle import OracleSolid3D
    cells = int(round(DEPTH / dx))
    g = pm.Solid3D("ice", substance(ICE, fusion), (1, 1, cells), dx, initial_c + C0)
    o = OracleSolid3D("ice", ICE[0], ICE[1], ICE[2], (1, 1, cells), dx, initial_c + C0, ICE[3],
                      fusion)
    for k in range(cells):
        g.set_melted_fraction(0, 0, k, 1.0)
        o.set_melted_fraction(0, 0, k, 1.0)
    return g, o


def _lambda(stefan):
    target = stefan / math.sqrt(math.pi)
    lo, hi = 0.0, 1.0
    for _ in range(60):
        mid = 0.5 * (lo + hi)
        if mid * math.exp(mid * mid) * math.erf(mid) - target < 0.0:
            lo = mid
        else:
            hi = mid
    return 0.5 * (lo + hi)


def test_the_freezing_front_on_the_device_is_the_oracles_and_neumanns():
    # a_freezing_front.rs:181-262 at 1 mm: the GPU column and the oracle's, stepped side by side
    # with the cold face clamped, agree to the bit, and the front is within (dx/X)^2 of Neumann.
    gpu_case("freezing front")
    import pantometry_amd as pm
    coefficient = 2.0 * _lambda(CP * 20.0 / LATENT) * math.sqrt(ALPHA)
    g, o = _column(1e-3)
    dt = g.max_stable_dt() * 0.9
    assert dt == o.max_stable_dt() * 0.9
    t = 0.0
    for want_at in (100.0, 400.0, 900.0):
        while t < want_at:
            g.set_temperature(0, 0, 0, C0 - 20.0)
            o.set_temperature(0, 0, 0, C0 - 20.0)
            g.step(t, dt, pm.Exchange())
            o.step(dt)
            t += dt
        g.set_temperature(0, 0, 0, C0 - 20.0)
        o.set_temperature(0, 0, 0, C0 - 20.0)
        same(g, o)
        got = DEPTH - g.melted_volume() / 1e-6 - 0.5e-3
        closed = coefficient * math.sqrt(t)
        assert abs(got / closed - 1.0) < (1e-3 / closed) ** 2


def test_the_latent_heat_tests_on_the_device():
    # a_freezing_front.rs:395-650.
    gpu_case("latent heat reference tests")
    import pantometry_amd as pm
    cube = pm.Solid3D("cube", substance(ICE, ICE_FUSION), (1, 1, 1), 1e-3, C0)
    vol = 1e-9
    latent, capacity = (ICE[0] * vol) * 333.55e3, (ICE[0] * vol) * ICE[2]
    t, left_at = 0.0, None
    want = latent / 1e-3
    while t < 2.0 * want:
        cube.deposit(0, 0, 0, 1e-3 * 0.1)
        t += 0.1
        if left_at is None and cube.temperature_at(0, 0, 0) > C0:
            left_at = t
    assert abs(left_at - want) <= 0.1 * 1.001

    for multiple in (0.5, 1.0, 1.0000001, 3.0, 10.0):
        c = pm.Solid3D("cube", substance(ICE, ICE_FUSION), (1, 1, 1), 1e-3, C0)
        joules = multiple * latent
        opening = c.ledger().get(pm.ENERGY) or 0.0
        c.deposit(0, 0, 0, joules)
        want_phi, want_t = (1.0, C0 + (joules - latent) / capacity) if multiple >= 1.0 else (multiple, C0)
        assert abs(c.melted_fraction_at(0, 0, 0) - want_phi) < 1e-15
        assert abs(c.temperature_at(0, 0, 0) - want_t) < 1e-12 * (1.0 + abs(want_t - C0))
        assert abs(((c.ledger().get(pm.ENERGY) or 0.0) - opening) / joules - 1.0) < 1e-14

    g, _ = _column(1e-3)
    opening = g.ledger().get(pm.ENERGY)
    dt = g.max_stable_dt() * 0.9
    t = 0.0
    while t < 120.0:
        g.set_temperature(0, 0, 0, C0 - 20.0)
        g.step(t, dt, pm.Exchange())
        t += dt
    g.checkpoint()
    depth, held = g.melted_volume(), g.ledger().get(pm.ENERGY)
    for _ in range(200):
        g.set_temperature(0, 0, 0, C0 - 20.0)
        g.step(t, dt, pm.Exchange())
    assert g.melted_volume() < depth
    g.restore()
    assert g.melted_volume() == depth and g.ledger().get(pm.ENERGY) == held
    assert held < opening

    plain = pm.Solid3D("b", pm.Substance.aluminium_6061(), (4, 3, 5), 2e-3, 20.0 + C0)
    plain.deposit(2, 1, 2, 3.0)
    plain.set_melted_fraction(2, 1, 2, 1.0)
    assert plain.melted_fraction_at(2, 1, 2) == 0.0 and plain.melted_volume() == 0.0
    assert all(r[1] != "melted" for r in plain.readings())


def test_a_superheated_liquid_follows_the_two_phase_solution():
    # two_phase_stefan.rs:185-275 at 0 and 20 K of superheat, the GPU against the oracle to the
    # bit and the front within dx/X of the two-phase Neumann solution.
    gpu_case("two-phase Stefan")
    import pantometry_amd as pm
    from tests.test_oracle_phase import _two_phase_lambda
    one = _lambda(CP * 20.0 / LATENT)
    probe = pm.Solid3D("probe", substance(ICE, ICE_FUSION), (1, 1, 8), 1e-3, C0 - 1.0)
    dt = probe.max_stable_dt() * 0.9
    allowed = 1e-3 / (2.0 * one * math.sqrt(ALPHA * 900.0))
    for superheat in (0.0, 20.0):
        g, o = _column(1e-3, TWO_PHASE, superheat)
        t = 0.0
        while t < 900.0:
            g.set_temperature(0, 0, 0, C0 - 20.0)
            o.set_temperature(0, 0, 0, C0 - 20.0)
            g.step(t, dt, pm.Exchange())
            o.step(dt)
            t += dt
        same(g, o)
        want = 2.0 * _two_phase_lambda(max(superheat, 1e-12)) * math.sqrt(ALPHA * 900.0)
        got = DEPTH - g.melted_volume() / 1e-6 - 0.5e-3
        assert abs(got / want - 1.0) < allowed
