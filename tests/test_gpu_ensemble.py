"""The batched heat ensemble on the GPU: every member bit-exact against a standalone oracle block,
under both schedules; the full 512-member run through per-member conservation and determinism."""
import numpy as np
import pytest

from tests.conftest import gpu_case

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).reshape(-1).view(np.uint64)


def _oracle_member(oracle, edge, alpha, seed, steps, dt):
    from pantometry_amd import workloads as w
    o = oracle.OracleSolid3D("m", w.RHO, alpha * w.RHO * w.CP, w.CP, (edge, edge, edge), w.DX, 300.0)
    o.set_cells(oracle.hot_spot_cells((edge, edge, edge), seed=seed))
    assert dt == 0.9 * o.max_stable_dt()
    for _ in range(steps):
        o.step(dt, threads_=16)
    return o


@pytest.mark.parametrize("resident_edge,z_chunk,per_pass,tiles",
                         [(0, 8, 2, 0), (0, 32, 2, 0), (24, 16, 2, 0), (40, 5, 2, 0), (0, 16, 1, 0),
                          (20, 3, 1, 0), (0, 16, 2, 1), (0, 16, 2, 2), (0, 16, 2, 3), (0, 5, 1, 3)])
def test_members_are_bit_exact_under_every_schedule(resident_edge, z_chunk, per_pass, tiles, oracle):
    gpu_case(f"ensemble schedule {resident_edge}/{z_chunk}/{per_pass} tiles {tiles}")
    from pantometry_amd import workloads as w
    ens, shapes = w.heat_ensemble(24, lo_edge=14.0, hi_edge=40.0)
    ens.set_tiles(tiles)
    ens.set_schedule(resident_edge, z_chunk)
    ens.set_steps_per_pass(per_pass)
    before = ens.ledgers()
    ens.run(37)
    after = ens.ledgers()
    for m, (edge, alpha) in enumerate(shapes):
        o = _oracle_member(oracle, edge, alpha, w.ENSEMBLE_SEED + 1 + m, 37, ens.members[m].dt)
        assert np.array_equal(bits(ens.cells(m)), bits(o.cells)), m
        assert abs(after[m] - o.stored_heat()) <= 1e-12 * abs(o.stored_heat())
        assert abs(after[m] - before[m]) <= 1e-12 * abs(before[m])


def test_runs_continue_across_calls(oracle):
    gpu_case("ensemble across calls")
    from pantometry_amd import workloads as w
    ens, shapes = w.heat_ensemble(6, lo_edge=16.0, hi_edge=30.0)
    ens.set_schedule(20, 8)
    ens.run(5)
    ens.run(6)
    ens.run(6)
    for m, (edge, alpha) in enumerate(shapes):
        o = _oracle_member(oracle, edge, alpha, w.ENSEMBLE_SEED + 1 + m, 17, ens.members[m].dt)
        assert np.array_equal(bits(ens.cells(m)), bits(o.cells)), m


# Uneven shapes: rows over 128 cells (64 x 8 tiles, partial in x and y), rows of 1-3 cells
# (tiles of hundreds of rows), single-plane and single-column members, z-chunks of one plane.
ODD_SHAPES = [(130, 9, 7), (200, 17, 3), (3, 50, 20), (1, 1, 40), (97, 6, 33), (128, 5, 9),
              (40, 1, 1), (2, 2, 2), (65, 70, 2)]


@pytest.mark.parametrize("tiles", [0, 1, 2, 3])
@pytest.mark.parametrize("steps", [1, 2, 3, 4, 7])
@pytest.mark.parametrize("z_chunk", [1, 16])
def test_uneven_members_are_bit_exact(steps, z_chunk, tiles, oracle):
    gpu_case(f"ensemble uneven members steps={steps} z_chunk={z_chunk} tiles={tiles}")
    from pantometry_amd import workloads as w
    from pantometry_amd.ensemble import Ensemble, Member
    from pantometry_amd.thermal import Substance
    members, blocks = [], []
    for n, counts in enumerate(ODD_SHAPES):
        alpha = 2e-5 * (1.0 + 0.37 * n)
        o = oracle.OracleSolid3D("m", w.RHO, alpha * w.RHO * w.CP, w.CP, counts, w.DX, 300.0)
        o.set_cells(oracle.hot_spot_cells(counts, seed=0x0DD0 + n))
        dt = 0.9 * o.max_stable_dt()
        members.append(Member(counts, w.DX, Substance.of("m", w.RHO, alpha * w.RHO * w.CP, w.CP),
                              300.0, dt))
        blocks.append((o, dt))
    ens = Ensemble(members)
    for n, (o, _) in enumerate(blocks):
        ens.set_cells(n, o.cells)
    ens.set_tiles(tiles)
    ens.set_schedule(0, z_chunk)
    ens.run(steps)
    for n, (o, dt) in enumerate(blocks):
        for _ in range(steps):
            o.step(dt, threads_=16)
        assert np.array_equal(bits(ens.cells(n)), bits(o.cells)), ODD_SHAPES[n]


def test_one_and_two_steps_per_pass_agree():
    gpu_case("ensemble per-pass 1 vs 2")
    from pantometry_amd import workloads as w
    runs = []
    for per_pass in (1, 2):
        ens, shapes = w.heat_ensemble(40, lo_edge=16.0, hi_edge=90.0)
        ens.set_schedule(0, 24)
        ens.set_steps_per_pass(per_pass)
        ens.run(9)
        ens.run(12)
        runs.append([ens.cells(m).copy() for m in range(len(shapes))])
    for a, b in zip(*runs):
        assert np.array_equal(bits(a), bits(b))


@pytest.mark.timeout(900)
def test_the_full_ensemble(oracle):
    """512 members, edges 16-96, 500 steps in one graph launch; the tuned schedule."""
    gpu_case("ensemble 512 x 500")
    from pantometry_amd import workloads as w
    ens, shapes = w.heat_ensemble(512)
    report = ens.tune()
    print(report)
    before = ens.ledgers()
    ens.run(500)
    after = ens.ledgers()
    assert np.all(np.abs(after - before) <= 1e-12 * np.abs(before))
    # A few members against the oracle: the smallest, the largest and one in between.
    order = sorted(range(len(shapes)), key=lambda m: shapes[m][0])
    for m in [order[0], order[len(order) // 2], order[-1]]:
        edge, alpha = shapes[m]
        o = _oracle_member(oracle, edge, alpha, w.ENSEMBLE_SEED + 1 + m, 500, ens.members[m].dt)
        assert np.array_equal(bits(ens.cells(m)), bits(o.cells)), (m, edge)
