"""The oracle's ScalarField::at for a block (solid.rs:2045-2102), pinned by what the sampling must
be: a sample at a cell centre is that cell exactly; a field linear in each axis is reproduced
between centres; outside the centres the outer half cell is flat; a NaN point answers NaN."""
import math

import numpy as np


def _block(oracle, counts, dx, cells):
    b = oracle.OracleSolid3D("b", 1000.0, 1.0, 1000.0, counts, dx, 300.0)
    b.set_cells(cells)
    return b


def test_cell_centres_are_exact(oracle):
    counts, dx = (5, 4, 3), 2e-3
    rng = np.random.default_rng(1)
    b = _block(oracle, counts, dx, 300.0 + rng.standard_normal(60))
    nx, ny, nz = counts
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                got = b.sample((i + 0.5) * dx, (j + 0.5) * dx, (k + 0.5) * dx)
                assert got == b.cells[i + nx * (j + ny * k)]


def test_a_linear_field_is_reproduced(oracle):
    counts, dx = (6, 5, 4), 1e-3
    nx, ny, nz = counts
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    b = _block(oracle, counts, dx, (3.0 + 2.0 * i - 1.5 * j + 0.25 * k).reshape(-1))
    rng = np.random.default_rng(2)
    for _ in range(200):
        q = rng.uniform(0.0, 1.0, 3) * (np.array(counts) - 1)   # between the centres
        x, y, z = (q + 0.5) * dx
        expect = 3.0 + 2.0 * q[0] - 1.5 * q[1] + 0.25 * q[2]
        assert abs(b.sample(x, y, z) - expect) < 1e-12


def test_outside_the_centres_is_flat_and_nan_is_nan(oracle):
    counts, dx = (4, 4, 4), 1e-3
    b = _block(oracle, counts, dx, np.arange(64, dtype=np.float64))
    assert b.sample(0.0, 0.0, 0.0) == b.cells[0]
    assert b.sample(-1.0, -1.0, -1.0) == b.cells[0]
    assert b.sample(1.0, 1.0, 1.0) == b.cells[63]
    assert math.isnan(b.sample(math.nan, 0.0, 0.0))


def _hall(oracle, nodes_across, p):
    h = oracle.OracleHall(0.4, 0.3, 0.2, nodes_across, 343.0, 1.204)
    h.p = np.ascontiguousarray(p(h), dtype=np.float64)
    return h


def test_hall_nodes_are_exact_and_a_linear_field_is_reproduced(oracle):
    """Hall's ScalarField::at (hall.rs:590-617): nodes exact, trilinear reproduces a field linear
    along each axis, clamped beyond the walls, p[0] for a NaN point."""
    def linear(h):
        k, j, i = np.meshgrid(np.arange(h.nz), np.arange(h.ny), np.arange(h.nx), indexing="ij")
        return (1.0 + 0.5 * i - 0.25 * j + 2.0 * k).reshape(-1)

    h = _hall(oracle, 9, linear)
    dx = h.dx
    for (i, j, k) in [(0, 0, 0), (3, 2, 1), (h.nx - 1, h.ny - 1, h.nz - 1)]:
        assert h.sample(i * dx, j * dx, k * dx) == h.p[i + h.nx * (j + h.ny * k)]
    rng = np.random.default_rng(3)
    for _ in range(100):
        q = rng.uniform(0.0, 1.0, 3) * (np.array([h.nx, h.ny, h.nz]) - 1)
        assert abs(h.sample(*(q * dx)) - (1.0 + 0.5 * q[0] - 0.25 * q[1] + 2.0 * q[2])) < 1e-12
    assert h.sample(-1.0, -1.0, -1.0) == h.p[0]
    assert h.sample(math.nan, 0.0, 0.0) == h.p[0]
    assert h.sample(9.0, 9.0, 9.0) == h.p[-1]
