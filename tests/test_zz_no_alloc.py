"""pm_solid3d_advance_static, the advance that allocates nothing, against pm_solid3d_advance.

Both entry points run the same kernels; the static one keeps its audit state, partials and staged
totals in the library's static scratch.  Two calls in a row (the second on a smaller block) must
give fields bit-identical to the allocating advance and identical per-step totals; a breach must
stop at the same step with the same state; totals staged over several chunks must match; a
block whose partials exceed the scratch, and a block on the extended path, are refused.
"""
import ctypes as C

import numpy as np
import pytest

from tests.conftest import gpu_case

pytestmark = pytest.mark.gpu

PM_OK, PM_VIOLATION, PM_ERR_ARGUMENT = 0, 1, 3


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).reshape(-1).view(np.uint64)


def _static(block, dt, steps, tolerance):
    import pantometry_amd as pm
    from pantometry_amd import _capi
    fn = pm.lib().pm_solid3d_advance_static
    totals = np.full(max(steps, 1), -1.0)
    ms = C.c_double(0.0)
    v = _capi.pm_violation()
    rc = fn(block._h, float(dt), int(steps), float(tolerance),
            totals.ctypes.data_as(C.POINTER(C.c_double)), C.byref(ms), C.byref(v))
    return rc, totals[:steps], v


def _allocating(block, dt, steps, tolerance):
    import pantometry_amd as pm
    from pantometry_amd import _capi
    totals = np.full(max(steps, 1), -1.0)
    ms = C.c_double(0.0)
    v = _capi.pm_violation()
    rc = pm.lib().pm_solid3d_advance(block._h, float(dt), int(steps), float(tolerance),
                                     totals.ctypes.data_as(C.POINTER(C.c_double)), C.byref(ms),
                                     C.byref(v))
    return rc, totals[:steps], v


def _twins(counts, seed, spots=3):
    """Two GPU blocks with the same per-cell conductivity and cells."""
    import pantometry_amd as pm
    from oracle import oracle as o
    k, t = o.heterogeneous_fields(counts, spots=spots, k_seed=seed, spot_seed=seed + 1000)
    blocks = []
    for _ in range(2):
        g = pm.Solid3D("block", pm.Substance("s", 4000.0, 1.0, 1000.0), counts, 1e-3, 300.0)
        g.with_conductivity(k)
        g.set_cells(t)
        blocks.append(g)
    return blocks, k, t


def test_two_calls_match_the_allocating_advance():
    """A 96x80x70 block for 13 steps, then a 40x33x21 block for 7: fields and totals identical
    to pm_solid3d_advance, and the fields equal to the oracle's."""
    from oracle import oracle as o
    for counts, steps, seed in (((96, 80, 70), 13, 11), ((40, 33, 21), 7, 12)):
        gpu_case(f"static advance {counts} x {steps}")
        (a, b), k, t = _twins(counts, seed)
        dt = 0.9 * a.max_stable_dt()
        rc_s, tot_s, _ = _static(a, dt, steps, 1e-12)
        rc_a, tot_a, _ = _allocating(b, dt, steps, 1e-12)
        assert rc_s == PM_OK and rc_a == PM_OK
        assert np.array_equal(bits(a.cells()), bits(b.cells()))
        assert np.array_equal(bits(tot_s), bits(tot_a))
        ref = o.OracleSolid3D("block", 4000.0, 1.0, 1000.0, counts, 1e-3, 300.0).with_conductivity(k)
        ref.set_cells(t)
        for _ in range(steps):
            ref.step(dt)
        assert np.array_equal(bits(a.cells()), bits(ref.cells))
        stored = ref.stored_heat()
        assert abs(tot_s[-1] - stored) <= 1e-12 * abs(stored)


def test_breach_stops_where_the_allocating_advance_does():
    gpu_case("static advance breach at tolerance 1e-300")
    (a, b), _, _ = _twins((70, 37, 23), 21)
    dt = 0.9 * a.max_stable_dt()
    rc_s, tot_s, v_s = _static(a, dt, 9, 1e-300)
    rc_a, tot_a, v_a = _allocating(b, dt, 9, 1e-300)
    assert rc_s == PM_VIOLATION and rc_a == PM_VIOLATION
    assert v_s.step == v_a.step
    assert (v_s.before, v_s.after, v_s.scale) == (v_a.before, v_a.after, v_a.scale)
    assert np.array_equal(bits(a.cells()), bits(b.cells()))
    assert np.array_equal(bits(tot_s), bits(tot_a))


def test_totals_staged_over_several_chunks():
    """More steps than one chunk of staged totals (8192) holds."""
    gpu_case("static advance 16^3 x 9001 steps")
    (a, b), _, _ = _twins((16, 16, 16), 31, spots=1)
    dt = 0.9 * a.max_stable_dt()
    rc_s, tot_s, _ = _static(a, dt, 9001, 1e-12)
    rc_a, tot_a, _ = _allocating(b, dt, 9001, 1e-12)
    assert rc_s == PM_OK and rc_a == PM_OK
    assert np.array_equal(bits(a.cells()), bits(b.cells()))
    assert np.array_equal(bits(tot_s), bits(tot_a))


def test_refusals():
    import pantometry_amd as pm
    gpu_case("static advance refusals")
    # 64 x 8 x 32769 with one plane per z-chunk: 2 x 32769 partials per paired launch.
    big = pm.Solid3D("big", pm.Substance("s", 4000.0, 1.0, 1000.0), (64, 8, 32769), 1e-3, 300.0)
    big.with_conductivity(np.ones(64 * 8 * 32769))
    big.set_tile(8, 1)
    rc, _, _ = _static(big, 0.5 * big.max_stable_dt(), 2, 1e-12)
    assert rc == PM_ERR_ARGUMENT
    assert "static scratch" in pm.lib().pm_last_error().decode()
    del big
    film = pm.Solid3D("film", pm.Substance("s", 4000.0, 1.0, 1000.0), (8, 8, 8), 1e-3, 300.0)
    film.losing_from(pm.Face.XMIN, pm.Environment.still_air(290.0, 64e-6))
    rc, _, _ = _static(film, 0.5 * film.max_stable_dt(), 2, 1e-12)
    assert rc == PM_ERR_ARGUMENT
