"""Exposed faces and volumetric sources on the GPU (Solid3D::losing_from / dissipating): every
field bit-identical to the CPU oracle step for step, the state-dependent limit identical to the
oracle's, the reference's limit behaviour (a stiff film tightens and is enforced, a gentle one
leaves the limit alone), and the books: `supplied` identical, `lost` and the stored heat to
1e-12 (the device sums them in a fixed order rather than the reference's left fold)."""
import math

import numpy as np
import pytest

from tests.conftest import gpu_case

pytestmark = pytest.mark.gpu

COPPER = (8.96 * 1e3, 401.0, 385.0, 0.04)
PLA = (1.24 * 1e3, 0.13, 1800.0, 0.90)
ALUMINA = (3720.0, 24.0, 880.0, 0.8)


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).reshape(-1).view(np.uint64)


def celsius(c):
    return c + 273.15


def pair(oracle, counts, bulk, dx, start, regions=(), faces=(), sources=()):
    """The same block on the GPU and in the oracle: `regions` (substance, mask), `faces`
    (face, ambient, h, area), `sources` (watts, mask)."""
    from pantometry_amd.thermal import Environment, Solid3D, Substance
    sub = Substance.of("bulk", bulk[0], bulk[1], bulk[2], bulk[3])
    g = Solid3D("block", sub, counts, dx, start)
    o = oracle.OracleSolid3D("block", bulk[0], bulk[1], bulk[2], counts, dx, start,
                             emissivity=bulk[3])
    for (d, k, cp, eps), mask in regions:
        g.fill(Substance.of("region", d, k, cp, eps), mask)
        o.fill(d, k, cp, mask, emissivity=eps)
    for face, ambient, h, area in faces:
        g.losing_from(face, Environment(ambient, h, area))
        o.losing_from(face, ambient, h, area)
    for watts, mask in sources:
        g.dissipating(watts, mask)
        o.dissipating(watts, mask)
    return g, o


def box(counts, lo, hi):
    nx, ny, nz = counts
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    return ((i >= lo[0]) & (i < hi[0]) & (j >= lo[1]) & (j < hi[1]) & (k >= lo[2]) &
            (k < hi[2])).reshape(-1)


def module(oracle, counts=(9, 7, 6), dx=1.5e-3):
    """A stack in the spirit of the power-module scene: copper with an alumina layer, a source in
    the top planes, every face exposed to a different environment."""
    nx, ny, nz = counts
    side = lambda a, b: a * b * dx * dx
    faces = [(0, celsius(25.0), 40.0, side(ny, nz)), (1, celsius(30.0), 7.0, side(ny, nz)),
             (2, celsius(20.0), 150.0, 2.0 * side(nx, nz)), (3, celsius(22.0), 0.0, side(nx, nz)),
             (4, celsius(40.0), 3000.0, side(nx, ny)), (5, celsius(35.0), 12.0, side(nx, ny))]
    regions = [(ALUMINA, box(counts, (0, 0, 1), (nx, ny, 2)))]
    sources = [(45.0, box(counts, (0, 0, nz - 1), (nx, ny, nz))),
               (3.0, box(counts, (2, 2, 3), (4, 5, 4)))]
    g, o = pair(oracle, counts, COPPER, dx, celsius(60.0), regions, faces, sources)
    return g, o


def test_limit_matches_the_oracle_bit_for_bit(oracle):
    gpu_case("film limit")
    g, o = module(oracle)
    assert g.max_stable_dt() == o.max_stable_dt()
    # A polymer under a stiff film, hot and cool: there the film sets the limit.
    side = 4 * 2.0e-3
    for start in (celsius(30.0), celsius(900.0)):
        g, o = pair(oracle, (4, 4, 4), PLA, 2e-3, start,
                    faces=[(5, celsius(20.0), 2000.0, side * side),
                           (0, celsius(20.0), 900.0, side * side)])
        assert o.max_stable_dt() < 0.9 / o.worst_rate
        assert g.max_stable_dt() == o.max_stable_dt()


@pytest.mark.parametrize("steps", [1, 2, 25])
def test_steps_match_the_oracle(oracle, steps):
    gpu_case(f"film + sources, {steps} steps")
    from pantometry_amd.core import Exchange
    g, o = module(oracle)
    dt = 0.6 * o.max_stable_dt()
    bus = Exchange()
    for _ in range(steps):
        g.step(0.0, dt, bus)
        o.step(dt)
    assert np.array_equal(bits(g.cells()), bits(o.cells))
    stored, lost, supplied = g.books()
    assert supplied == o.supplied
    assert lost > 0.0 and abs(lost - o.lost) <= 1e-12 * abs(o.lost)
    assert abs(stored - o.stored_heat()) <= 1e-12 * max(abs(o.stored_heat()), supplied)
    assert g.generated_power() == pytest.approx(48.0, rel=1e-15)
    assert g.max_stable_dt() == o.max_stable_dt()


def test_uniform_block_with_film_and_source(oracle):
    """The uniform form (one substance: kernel-scalar faces and capacity)."""
    gpu_case("film uniform block")
    from pantometry_amd.core import Exchange
    counts = (5, 6, 7)
    dx = 2e-3
    faces = [(f, celsius(20.0), 25.0 * (f + 1), 1e-4 * (f + 1)) for f in range(6)]
    g, o = pair(oracle, counts, PLA, dx, celsius(150.0), faces=faces,
                sources=[(2.0, box(counts, (1, 1, 1), (3, 3, 3)))])
    dt = 0.5 * o.max_stable_dt()
    bus = Exchange()
    for _ in range(30):
        g.step(0.0, dt, bus)
        o.step(dt)
    assert np.array_equal(bits(g.cells()), bits(o.cells))
    assert g.books()[2] == o.supplied


@pytest.mark.parametrize("counts,uniform", [((70, 20, 130), False), ((67, 17, 70), True)], ids=str)
def test_blocks_whose_sweep_marches_columns_match_the_oracle(oracle, counts, uniform):
    """Blocks of at least 256 column tiles (64 x 16) x planes take the sweep that marches z
    columns: 2 x 2 tiles x 130 planes with the module's per-cell conductivity, and 2 x 2 x 70
    planes of one substance (the uniform face).  Steps one call at a time, then a
    device-resident advance, every cell bit for bit."""
    gpu_case(f"film sweep marching columns, {counts}, uniform {uniform}")
    from pantometry_amd.core import Exchange
    if uniform:
        dx = 2e-3
        faces = [(f, celsius(20.0), 25.0 * (f + 1), 1e-4 * (f + 1)) for f in range(6)]
        g, o = pair(oracle, counts, PLA, dx, celsius(150.0), faces=faces,
                    sources=[(2.0, box(counts, (1, 1, 1), (30, 9, 33)))])
    else:
        g, o = module(oracle, counts, dx=1e-3)
    dt = 0.6 * o.max_stable_dt()
    assert g.max_stable_dt() == o.max_stable_dt()
    bus = Exchange()
    for _ in range(3):
        g.step(0.0, dt, bus)
        o.step(dt)
    assert np.array_equal(bits(g.cells()), bits(o.cells))
    totals, _ = g.advance(dt, 8, tolerance=1e-9)
    for _ in range(8):
        o.step(dt)
    assert np.array_equal(bits(g.cells()), bits(o.cells))
    assert g.books()[2] == o.supplied
    assert abs(g.books()[1] - o.lost) <= 1e-12 * abs(o.lost)


def test_advance_matches_and_audits(oracle):
    gpu_case("film advance")
    g, o = module(oracle)
    dt = 0.6 * o.max_stable_dt()
    totals, _ = g.advance(dt, 40, tolerance=1e-9)
    for _ in range(40):
        o.step(dt)
    assert np.array_equal(bits(g.cells()), bits(o.cells))
    assert len(totals) == 40
    assert abs(totals[-1] - o.ledger()[0]) <= 1e-12 * max(abs(o.ledger()[1]), 1.0)


def test_a_stiff_film_tightens_the_limit_and_is_enforced(oracle):
    gpu_case("film stiff limit")
    from pantometry_amd.core import Exchange, Violation
    from pantometry_amd.thermal import Environment, Face, Solid3D, Substance
    pla = Substance.of("pla", *PLA)
    side = 4 * 2.0e-3
    bare = Solid3D("part", pla, (4, 4, 4), 2e-3, celsius(300.0)).max_stable_dt()
    cooled = Solid3D("part", pla, (4, 4, 4), 2e-3, celsius(300.0))
    cooled.losing_from(Face.ZMAX, Environment(celsius(20.0), 2000.0, side * side))
    limit = cooled.max_stable_dt()
    assert limit < 0.9 * bare
    before = cooled.cells().copy()
    with pytest.raises(Violation):
        cooled.step(0.0, 1.5 * limit, Exchange())
    assert np.array_equal(bits(cooled.cells()), bits(before))


def test_a_gentle_film_leaves_the_limit_alone():
    gpu_case("film gentle limit")
    from pantometry_amd.thermal import Environment, Face, Solid3D, Substance
    cu = Substance.of("copper", *COPPER)
    side = 4 * 5.0e-3
    bare = Solid3D("block", cu, (4, 4, 4), 5e-3, celsius(200.0)).max_stable_dt()
    cooled = Solid3D("block", cu, (4, 4, 4), 5e-3, celsius(200.0))
    cooled.losing_from(Face.ZMAX, Environment.still_air(celsius(20.0), side * side))
    assert cooled.max_stable_dt() == bare


def test_checkpoint_restores_the_books(oracle):
    gpu_case("film checkpoint")
    from pantometry_amd.core import Exchange
    g, o = module(oracle)
    dt = 0.6 * o.max_stable_dt()
    bus = Exchange()
    g.step(0.0, dt, bus)
    g.checkpoint()
    saved = (g.cells().copy(), g.books())
    for _ in range(3):
        g.step(0.0, dt, bus)
    g.restore()
    assert np.array_equal(bits(g.cells()), bits(saved[0]))
    assert g.books()[1:] == saved[1][1:]
