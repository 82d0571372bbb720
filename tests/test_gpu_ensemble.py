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


@pytest.mark.parametrize("resident_edge,z_chunk", [(0, 8), (0, 32), (24, 16), (40, 5)])
def test_members_are_bit_exact_under_every_schedule(resident_edge, z_chunk, oracle):
    gpu_case(f"ensemble schedule {resident_edge}/{z_chunk}")
    from pantometry_amd import workloads as w
    ens, shapes = w.heat_ensemble(24, lo_edge=14.0, hi_edge=40.0)
    ens.set_schedule(resident_edge, z_chunk)
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
