"""Hall on the GPU against the CPU oracle: bit-exact fields, energy to 1e-12, through the C ABI."""
import math

import numpy as np
import pytest

from tests import hall_cases
from tests.conftest import gpu_case

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).reshape(-1).view(np.uint64)


@pytest.mark.parametrize("case", hall_cases.HALL_CASES, ids=lambda c: c.__name__)
def test_closed_forms_on_the_gpu(case):
    gpu_case(f"hall closed form {case.__name__}")
    case(hall_cases.GpuRoom)


def _pair(width, height, depth, across, speed=343.0, density=1.204):
    import pantometry_amd as pm
    from oracle.oracle import OracleHall
    return (pm.Hall("hall", width, height, depth, across, speed, density),
            OracleHall(width, height, depth, across, speed, density))


def _same_fields(g, o):
    p, pp, vx, vy, vz = g.fields()
    assert np.array_equal(bits(p), bits(o.p))
    assert np.array_equal(bits(pp), bits(o.pp))
    n = len(vx)
    assert np.array_equal(bits(vx), bits(o.vx[:n]))
    assert np.array_equal(bits(vy), bits(o.vy[:len(vy)]))
    assert np.array_equal(bits(vz), bits(o.vz[:len(vz)]))


@pytest.mark.parametrize("dims,across,mode", [
    ((4.4, 3.1, 2.4), 13, (1, 1, 1)),
    ((4.4, 3.1, 2.4), 45, (3, 2, 1)),
    ((4.4, 3.1, 0.0), 45, (1, 1, 0)),
    ((0.5, 0.07, 0.3), 130, (2, 1, 3)),
], ids=str)
def test_fields_are_bit_exact(dims, across, mode):
    gpu_case(f"hall parity {dims} {across}")
    g, o = _pair(*dims, across)
    assert g.nodes() == o.nodes()
    g.released_in_mode(mode, 2.0)
    o.release(o.mode(mode, 2.0))
    assert np.array_equal(bits(g.fields()[0]), bits(o.p))
    dt = 0.9 * o.max_stable_dt()
    assert g.max_stable_dt() == o.max_stable_dt()
    import pantometry_amd as pm
    for n in range(150):
        g.step(0.0, dt, pm.Exchange())
        o.step(dt, threads_=16)
    _same_fields(g, o)
    e_ref = o.energy()
    assert abs(g.energy() - e_ref) <= 1e-12 * abs(e_ref)
    assert abs(g.startup_adjustment() - o.offset) <= 1e-12 * abs(e_ref)


@pytest.mark.parametrize("tile_y,z_chunk", [(8, 1), (8, 3), (16, 0), (16, 5), (8, 64)])
def test_every_tile_shape_gives_the_same_bits(tile_y, z_chunk):
    gpu_case(f"hall tile {tile_y} x {z_chunk}")
    import pantometry_amd as pm
    g, o = _pair(0.5, 0.07, 0.3, 130)
    g.set_tile(tile_y, z_chunk)
    g.released_in_mode((2, 1, 3), 2.0)
    o.release(o.mode((2, 1, 3), 2.0))
    dt = 0.9 * o.max_stable_dt()
    for n in range(40):
        g.step(0.0, dt, pm.Exchange())
        o.step(dt, threads_=16)
    _same_fields(g, o)


@pytest.mark.parametrize("dims,across,mode", [
    ((4.4, 3.1, 2.4), 13, (1, 1, 1)),
    ((4.4, 3.1, 0.0), 45, (1, 1, 0)),
    ((0.5, 0.07, 0.3), 130, (2, 1, 3)),
], ids=str)
@pytest.mark.parametrize("per_pass", [1, 2])
def test_advance_is_bit_exact(dims, across, mode, per_pass):
    """advance runs pairs of leapfrog steps per pass by default: an odd number of audited
    steps (the half step, pairs, a trailing single step) against the oracle."""
    gpu_case(f"hall advance {dims} {across} per pass {per_pass}")
    g, o = _pair(*dims, across)
    g.set_steps_per_pass(per_pass)
    g.released_in_mode(mode, 2.0)
    o.release(o.mode(mode, 2.0))
    dt = 0.9 * o.max_stable_dt()
    totals, _ = g.advance(dt, 37, tolerance=1e-12)
    for n in range(37):
        o.step(dt, threads_=16)
        t_ref, s_ref = o.ledger()
        assert abs(totals[n] - t_ref) <= 1e-12 * s_ref, n
    _same_fields(g, o)


@pytest.mark.parametrize("across", [113, 97, 65])
@pytest.mark.parametrize("tile_x", [64, 56, 48, 0])
@pytest.mark.parametrize("tile_y", [8, 16])
def test_every_paired_tile_width_gives_the_same_bits(across, tile_x, tile_y):
    """The paired step on 64-, 56- and 48-wide column tiles and the default width: 113 = 2 x 56
    + 1, 97 = 2 x 48 + 1 and 65 = 64 + 1 nodes leave one column in the last tile of one width
    each."""
    gpu_case(f"hall paired tile {tile_x} x {tile_y}, {across} nodes across")
    g, o = _pair(0.01 * (across - 1), 0.2, 0.15, across)
    assert o.nodes() == (across, 21, 16)
    g.set_tile(tile_y, 0)
    g.set_tile_width(tile_x)
    g.released_in_mode((2, 1, 1), 2.0)
    o.release(o.mode((2, 1, 1), 2.0))
    dt = 0.9 * o.max_stable_dt()
    totals, _ = g.advance(dt, 37, tolerance=1e-12)
    for n in range(37):
        o.step(dt, threads_=16)
        t_ref, s_ref = o.ledger()
        assert abs(totals[n] - t_ref) <= 1e-12 * s_ref, n
    _same_fields(g, o)


def test_tile_widths_outside_the_set_are_refused():
    gpu_case("hall tile width refusal")
    from pantometry_amd._capi import DeviceError
    g, _ = _pair(1.12, 0.2, 0.15, 113)
    for w in (32, 63, 65, -1):
        with pytest.raises(DeviceError):
            g.set_tile_width(w)


def test_a_breach_inside_a_pair_leaves_the_state_of_the_breaking_step():
    gpu_case("hall breach inside a pair")
    import pantometry_amd as pm
    g, o = _pair(4.4, 3.1, 2.4, 31)
    g.released_in_mode((2, 1, 1), 1.0)
    p0 = o.mode((2, 1, 1), 1.0)
    dt = 0.9 * g.max_stable_dt()
    # a negative tolerance breaks on the first audit of the device loop: step 1, the first of a
    # pair (step 0, the half step, runs alone)
    g.checkpoint()
    with pytest.raises(pm.Violation) as err:
        g.advance(dt, 20, tolerance=-1.0)
    assert err.value.step == 1
    o.release(p0)
    for _ in range(2):
        o.step(dt, threads_=16)
    _same_fields(g, o)

    # A tolerance met by every earlier step and broken by a second step of a pair (step 2, 4,
    # ...).  A hall released in a mode with a little seeded noise, whose ledger (its energy alone)
    # moves by rounding every step, gives such a step for some seed: the first seed whose run has
    # an even step drifting more than every step before it.
    for seed in range(7, 47):
        g, o = _pair(4.4, 3.1, 2.4, 31)
        nx, ny, nz = g.nodes()
        p0 = o.mode((2, 1, 1), 1.0) + 1e-3 * np.random.default_rng(seed).standard_normal(nx * ny * nz)
        g.released_from(p0.reshape(nz, ny, nx))
        dt = 0.9 * g.max_stable_dt()
        g.checkpoint()
        totals, _ = g.advance(dt, 80, tolerance=1.0)
        prev = np.concatenate([[totals[0]], totals[:-1]])
        scale = np.maximum(np.abs(totals), np.abs(prev))
        drift = np.where(scale > 0, np.abs(totals - prev) / np.where(scale > 0, scale, 1.0), 0.0)
        js = [j for j in range(2, 80, 2) if drift[j] > drift[1:j].max()]
        if js:
            break
    assert js, drift
    j = js[0]
    tol = 0.5 * (drift[j] + drift[1:j].max())
    g.restore()
    with pytest.raises(pm.Violation) as err:
        g.advance(dt, 80, tolerance=tol)
    assert err.value.step == j
    o.release(p0)
    for _ in range(j + 1):
        o.step(dt, threads_=16)
    _same_fields(g, o)


def test_a_sourced_hall_is_bit_exact_and_books_its_source():
    gpu_case("hall with source 41^3")
    import pantometry_amd as pm
    from pantometry_amd.workloads import ricker
    g, o = _pair(0.04, 0.04, 0.04, 41, 343.0, 1.2)
    g.released_from(np.zeros(41 ** 3))
    o.release(np.zeros(41 ** 3))
    dt = 0.5 * o.max_stable_dt()
    f = 343.0 / (20.0 * o.dx)
    inc = np.array([ricker((n + 1) * dt, f, 1.0 / f) for n in range(200)])
    node = 17 + 41 * (23 + 41 * 11)
    g.set_source(node, inc)
    o.set_source(node, inc)
    totals, _ = g.advance(dt, 260, tolerance=1e-12)
    for _ in range(260):
        o.step(dt)
    _same_fields(g, o)
    ref_total, ref_scale = o.ledger()
    assert abs(totals[-1] - ref_total) <= 1e-12 * ref_scale
    assert np.all(np.abs(totals - totals[0]) <= 1e-12 * ref_scale)
    assert abs(g.supplied() - o.supplied) <= 1e-12 * abs(o.supplied)


@pytest.mark.parametrize("at", [(64, 7, 5), (64, 8, 0), (64, 18, 11)], ids=str)
@pytest.mark.parametrize("tile_y", [8, 16])
def test_a_source_in_the_last_column_of_a_wide_tile(at, tile_y):
    """65 = 64 + 1 nodes across: the paired step's last 64-wide tile also owns column 64, whose
    second step, stores and audit its ring threads take.  A source there (beside a tile row's
    edge, on one, at the far corner) is booked by those threads in both steps of a pair."""
    gpu_case(f"hall wide tile with a source at {at}, tile_y {tile_y}")
    from pantometry_amd.workloads import ricker
    g, o = _pair(0.64, 0.18, 0.11, 65, 343.0, 1.2)
    assert o.nodes() == (65, 19, 12)
    g.set_tile(tile_y, 0)
    n = 65 * 19 * 12
    g.released_from(np.zeros(n))
    o.release(np.zeros(n))
    dt = 0.5 * o.max_stable_dt()
    f = 343.0 / (10.0 * o.dx)
    inc = np.array([ricker((k + 1) * dt, f, 1.0 / f) for k in range(60)])
    node = at[0] + 65 * (at[1] + 19 * at[2])
    g.set_source(node, inc)
    o.set_source(node, inc)
    totals, _ = g.advance(dt, 75, tolerance=1e-12)
    for k in range(75):
        o.step(dt)
        t_ref, s_ref = o.ledger()
        assert abs(totals[k] - t_ref) <= 1e-12 * s_ref, k
    _same_fields(g, o)
    assert o.supplied > 0.0
    assert abs(g.supplied() - o.supplied) <= 1e-12 * abs(o.supplied)


def test_checkpoint_and_restore():
    gpu_case("hall checkpoint")
    import pantometry_amd as pm
    g, _ = _pair(4.4, 3.1, 2.4, 25)
    g.released_in_mode((1, 1, 1), 1.0)
    dt = 0.9 * g.max_stable_dt()
    for _ in range(5):
        g.step(0.0, dt, pm.Exchange())
    g.checkpoint()
    saved = g.fields()
    for _ in range(9):
        g.step(0.0, dt, pm.Exchange())
    after = g.fields()
    g.restore()
    p, pp, vx, vy, vz = g.fields()
    assert np.array_equal(bits(p), bits(saved[0])) and np.array_equal(bits(pp), bits(saved[0]))
    assert np.array_equal(bits(vx), bits(saved[2]))


@pytest.mark.timeout(900)
def test_the_acoustic_hall_at_384():
    """385^3 nodes: 20 steps bit-exact against the oracle, then 4000 audited steps."""
    gpu_case("hall 385^3")
    from oracle.oracle import OracleHall
    from pantometry_amd import workloads as w
    case = w.acoustic_hall(384, 4000)
    g = case.hall
    o = OracleHall(384 * w.DX, 384 * w.DX, 384 * w.DX, 385, 343.0, 1.2)
    assert o.nodes() == g.nodes()
    o.release(np.zeros(385 ** 3))
    o.set_source(case.source_node, case.increments)
    g.checkpoint()
    totals, ms = g.advance(case.dt, 20, tolerance=1e-12)
    for _ in range(20):
        o.step(case.dt, threads_=16)
    _same_fields(g, o)
    t_ref, s_ref = o.ledger()
    assert abs(totals[-1] - t_ref) <= 1e-12 * s_ref
    del o
    totals, ms = g.advance(case.dt, 3980, tolerance=1e-12)
    scale = max(abs(g.energy()), abs(g.supplied()))
    assert g.supplied() > 0.0
    assert np.all(np.abs(totals - t_ref) <= 1e-12 * scale)
