"""Pins the CPU oracle: the reference's closed-form tests run against it, plus the properties the
mathematics fixes where those tests leave gaps (a brute-force restatement on tiny grids, the
harmonic-mean face, symmetry of the face mean, the exact-sum diagnostic)."""
import math

import numpy as np
import pytest

from tests import hall_cases, solid_cases


@pytest.mark.parametrize("case", solid_cases.SOLID_CASES, ids=lambda c: c.__name__)
def test_solid_closed_forms_on_the_oracle(case, oracle):
    case(solid_cases.OracleBlock)


@pytest.mark.parametrize("case", hall_cases.HALL_CASES, ids=lambda c: c.__name__)
def test_hall_closed_forms_on_the_oracle(case, oracle):
    case(hall_cases.OracleRoom)


def _brute_step(T, k, cap, dx, dt):
    """Solid3D's plain sweep by brute force over a dict of faces: every face computed once from
    both sides' conductivities, every cell's flux summed x-, x+, y-, y+, z-, z+."""
    nz, ny, nx = T.shape
    out = T.copy()

    def series(a, b):
        s = a + b
        return 2.0 * a * b / s if s > 0.0 else 0.0

    for kk in range(nz):
        for j in range(ny):
            for i in range(nx):
                t = T[kk, j, i]
                flux = 0.0
                for (di, dj, dk) in [(-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1)]:
                    a, b, c = i + di, j + dj, kk + dk
                    if 0 <= a < nx and 0 <= b < ny and 0 <= c < nz:
                        lo, hi = ((a, b, c), (i, j, kk)) if (di + dj + dk) < 0 else ((i, j, kk), (a, b, c))
                        face = series(k[lo[2], lo[1], lo[0]], k[hi[2], hi[1], hi[0]])
                        flux = flux + face * (T[c, b, a] - t)
                mob = dx / cap[kk, j, i]
                out[kk, j, i] = t + dt * mob * (flux - 0.0)
    return out


def test_the_sweep_matches_a_brute_force_restatement(oracle):
    rng = np.random.default_rng(5)
    for counts in [(1, 1, 1), (2, 1, 1), (3, 4, 2), (5, 3, 4)]:
        nx, ny, nz = counts
        k = rng.lognormal(0.0, 0.5, size=(nz, ny, nx))
        cap = rng.uniform(1e-3, 3e-3, size=(nz, ny, nx))
        T = 300.0 + rng.uniform(0, 50, size=(nz, ny, nx))
        b = oracle.OracleSolid3D("b", 1.0, 1.0, 1.0, counts, 1e-3, 300.0)
        b.kcell = k.reshape(-1).copy()
        b.cap = cap.reshape(-1).copy()
        b.resolve()
        b.set_cells(T)
        dt = min(0.5 * b.max_stable_dt(), 1.0)
        b.step(dt)
        want = _brute_step(T, k, cap, 1e-3, dt)
        assert np.array_equal(b.grid(), want)


def test_a_dropped_term_is_caught_by_the_brute_force_check(oracle):
    # The brute-force restatement is sensitive: dropping the z+ face changes the result.
    rng = np.random.default_rng(6)
    T = 300.0 + rng.uniform(0, 50, size=(3, 3, 3))
    k = np.ones((3, 3, 3))
    cap = np.full((3, 3, 3), 2e-3)
    good = _brute_step(T, k, cap, 1e-3, 1e-4)
    T2 = T.copy()
    T2[2] = T2[1]  # as if the top plane were a mirror of the middle one
    bad = _brute_step(T2, k, cap, 1e-3, 1e-4)
    assert not np.array_equal(good[1], bad[1])


def test_the_face_mean_is_harmonic_and_symmetric(oracle):
    rng = np.random.default_rng(7)
    for _ in range(2000):
        a, b = rng.lognormal(0, 2, size=2)
        s = oracle.lib().orc_series(a, b)
        assert s == oracle.lib().orc_series(b, a)
        assert abs(s - 2.0 / (1.0 / a + 1.0 / b)) <= 4e-16 * s
    assert oracle.lib().orc_series(0.0, 0.0) == 0.0
    assert oracle.lib().orc_series(0.0, 5.0) == 0.0
    assert oracle.lib().orc_series(3.0, 3.0) == 3.0


def test_stored_heat_is_a_left_fold_and_close_to_the_exact_sum(oracle):
    rng = np.random.default_rng(8)
    n = 200_000
    cap = np.full(n, 4.0)
    cells = 300.0 + rng.exponential(5.0, size=n)
    seq = oracle.lib().orc_stored_heat(n, oracle.p(cap), oracle.p(cells), 300.0)
    comp = oracle.lib().orc_stored_heat_compensated(n, oracle.p(cap), oracle.p(cells), 300.0)
    exact = math.fsum(4.0 * (c - 300.0) for c in cells)
    assert comp == exact or abs(comp - exact) <= 1e-15 * abs(exact)
    assert abs(seq - exact) <= 1e-12 * abs(exact)
    left = 0.0
    for c in cells[:1000]:
        left += 4.0 * (c - 300.0) - 0.0
    assert oracle.lib().orc_stored_heat(1000, oracle.p(cap), oracle.p(cells), 300.0) == left


def test_the_parallel_sweep_is_bit_identical(oracle):
    k, t = oracle.heterogeneous_fields((40, 33, 21))
    a = oracle.OracleSolid3D("a", 4000.0, 1.0, 1000.0, (40, 33, 21), 1e-3, 300.0).with_conductivity(k)
    b = oracle.OracleSolid3D("b", 4000.0, 1.0, 1000.0, (40, 33, 21), 1e-3, 300.0).with_conductivity(k)
    a.set_cells(t)
    b.set_cells(t)
    dt = 0.9 * a.max_stable_dt()
    for _ in range(5):
        a.step(dt)
        b.step(dt, threads_=4)
    assert np.array_equal(a.cells, b.cells)


def test_the_hall_source_is_booked(oracle):
    # With a source, energy - supplied is conserved to rounding over many steps.
    h = oracle.OracleHall(0.032, 0.032, 0.032, 33, 343.0, 1.2)
    h.release(np.zeros(33 ** 3))
    dt = 0.5 * h.max_stable_dt()
    inc = np.array([math.sin(0.3 * n) for n in range(300)])
    h.set_source(16 + 33 * (16 + 33 * 16), inc)
    totals = []
    for _ in range(300):
        h.step(dt)
        totals.append(h.ledger()[0])
    scale = max(abs(h.energy()), abs(h.supplied))
    assert h.supplied != 0.0
    assert max(abs(x - totals[0]) for x in totals) <= 1e-12 * scale


def test_the_compensated_hall_invariant_is_the_exact_sum_of_its_terms(oracle):
    """orc_hall_invariant_compensated_omp against math.fsum (correctly rounded) of the same
    terms, evaluated term by term as the fold evaluates them, on a hall with a source running:
    within 2 ulps, for 1 and 5 threads; the left fold stays within its own error bound."""
    h = oracle.OracleHall(0.024, 0.018, 0.013, 25, 343.0, 1.2)
    n = h.nx * h.ny * h.nz
    h.release(np.zeros(n))
    dt = 0.5 * h.max_stable_dt()
    h.set_source(12 + h.nx * (9 + h.ny * 6), np.array([math.sin(0.2 * s) for s in range(60)]))
    for _ in range(60):
        h.step(dt)
    nx, ny, nz = h.nx, h.ny, h.nz
    wall = lambda i, m: 2.0 if (i == 0 or i + 1 == m) else 1.0  # noqa: E731
    vol, rc2, hr = h.dx ** 3, h.density * h.speed ** 2, 0.5 * h.density
    terms = []
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                q = i + nx * (j + ny * k)
                terms.append(1.0 / (wall(i, nx) * wall(j, ny) * wall(k, nz)) * h.p[q] * h.pp[q]
                             / (2.0 * rc2) * vol)
    groups = ((h.vx, nx - 1, ny, nz, lambda i, j, k: wall(j, ny) * wall(k, nz)),
              (h.vy, nx, ny - 1, nz, lambda i, j, k: wall(i, nx) * wall(k, nz)),
              (h.vz, nx, ny, nz - 1, lambda i, j, k: wall(i, nx) * wall(j, ny)))
    for v, ex, ey, ez, w in groups:
        for k in range(ez):
            for j in range(ey):
                for i in range(ex):
                    x = v[i + ex * (j + ey * k)]
                    terms.append(1.0 / w(i, j, k) * hr * x * x * vol)
    exact = math.fsum(terms)
    for threads in (1, 5):
        got = h.invariant_compensated(threads)
        assert abs(got - exact) <= 2 * np.spacing(abs(exact)), (threads, got, exact)
    fold = h.invariant()
    bound = len(terms) * 2.0 ** -53 * math.fsum(abs(t) for t in terms)
    assert abs(fold - exact) <= bound


def test_the_hall_mode_frequency_closed_form():
    # hall_closed_forms.rs:41-71, on the geometry the oracle builds.
    from oracle.oracle import OracleHall
    h = OracleHall(4.4, 3.1, 2.4, 45, 343.0, 1.204)
    lx, ly, lz = (h.nx - 1) * h.dx, (h.ny - 1) * h.dx, (h.nz - 1) * h.dx
    assert 60.0 < 343.0 / (2.0 * lz) < 85.0
    assert h.nx == 45 and h.ny == int(round(3.1 / h.dx)) + 1
