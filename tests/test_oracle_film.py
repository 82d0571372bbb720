"""The oracle's exposed faces and volumetric sources, pinned by the closed forms and identities the
reference's own tests assert (crates/pantometry-thermal/tests/a_block_that_cools.rs and
a_block_that_generates.rs): nothing lost without exposure, Newton's law for a lumped block, the
ledger identity stored + lost (- supplied) to 1e-12 of what moved, the stability limit tightened by
a stiff film and enforced, left alone by a gentle one, shorter for a hot block, and unmoved by a
constant source."""
import math

import numpy as np
import pytest

# Catalogue substances, crates/pantometry-core/src/substance.rs (density g/cm^3 * 1e3, k, c_p, eps).
COPPER = (8.96 * 1e3, 401.0, 385.0, 0.04)
PLA = (1.24 * 1e3, 0.13, 1800.0, 0.90)
BOROSILICATE = (2.51 * 1e3, 1.114, 858.0, 0.90)


def celsius(c):
    return c + 273.15


def block(oracle, sub, n, cell_mm, start_c):
    d, k, cp, eps = sub
    return oracle.OracleSolid3D("block", d, k, cp, (n, n, n), cell_mm * 1e-3, celsius(start_c),
                                emissivity=eps)


def run(b, seconds):
    dt = b.max_stable_dt() * 0.4
    for _ in range(int(math.ceil(seconds / dt))):
        b.step(dt)
    return b


def stored(b):
    return b.stored_heat()


def test_an_unexposed_block_loses_nothing(oracle):
    b = block(oracle, COPPER, 4, 5.0, 200.0)
    opening = stored(b)
    run(b, 60.0)
    assert b.lost == 0.0
    assert abs(stored(b) - opening) < 1e-9


def test_a_cooling_block_follows_newtons_law(oracle):
    n, cell = 4, 5.0e-3
    side = n * cell
    b = block(oracle, COPPER, n, 5.0, 120.0)
    for f in range(6):
        b.losing_from(f, celsius(20.0), 7.0, side * side)   # Environment::still_air
    capacity = float(b.cap.sum())
    mid, amb = celsius(70.0), celsius(20.0)
    h_a = oracle.lib().orc_loss_from(mid, amb, 7.0, 6.0 * side * side, COPPER[3]) / (mid - amb)
    tau = capacity / h_a
    run(b, tau)
    expect = amb + (celsius(120.0) - amb) * math.exp(-1.0)
    assert abs(float(b.cells.mean()) - expect) < 3.0


def test_every_joule_that_leaves_is_counted(oracle):
    b = block(oracle, COPPER, 3, 4.0, 300.0)
    side = 3 * 4.0e-3
    b.losing_from(5, celsius(20.0), 7.0, side * side)
    opening = stored(b) + b.lost
    run(b, 2000.0)
    closing = stored(b) + b.lost
    assert b.lost > 0.0
    assert abs(closing - opening) / b.lost < 1e-12


def test_a_stiff_film_tightens_the_limit_and_the_limit_is_enforced(oracle):
    side = 4 * 2.0e-3
    bare = block(oracle, PLA, 4, 2.0, 300.0).max_stable_dt()
    cooled = block(oracle, PLA, 4, 2.0, 300.0).losing_from(5, celsius(20.0), 2000.0, side * side)
    limit = cooled.max_stable_dt()
    assert limit < bare * 0.9
    with pytest.raises(ValueError):
        cooled.step(limit * 1.5)


def test_a_gentle_film_leaves_the_limit_alone(oracle):
    side = 4 * 5.0e-3
    bare = block(oracle, COPPER, 4, 5.0, 200.0).max_stable_dt()
    cooled = block(oracle, COPPER, 4, 5.0, 200.0).losing_from(5, celsius(20.0), 7.0, side * side)
    assert cooled.max_stable_dt() == bare


def test_a_hot_block_gets_a_shorter_limit_than_a_cool_one(oracle):
    side = 4 * 3.0e-3

    def limit(c, exposed=True):
        b = block(oracle, BOROSILICATE, 4, 3.0, c)
        if exposed:
            b.losing_from(5, celsius(20.0), 600.0, side * side)
        return b.max_stable_dt()

    assert limit(1000.0) < limit(30.0)
    assert limit(30.0, False) == limit(1000.0, False)


def _die(oracle, watts):
    b = block(oracle, COPPER, 3, 5.0, 20.0)
    mask = np.zeros(27, dtype=bool)
    mask[1 + 3 * (1 + 3 * 1)] = True
    return b.dissipating(watts, mask)


def test_a_source_leaves_the_books_balanced(oracle):
    b = _die(oracle, 30.0)
    opening = oracle.ledger_total([stored(b), b.lost, -b.supplied])[0]
    hot_spot = 0.0
    for n in range(2000):
        b.step(b.max_stable_dt())
        if n == 20:
            hot_spot = b.cells[13] - b.cells[0]
    closing = oracle.ledger_total([stored(b), b.lost, -b.supplied])[0]
    assert b.supplied > 100.0
    assert abs(closing - opening) / b.supplied < 1e-12
    assert hot_spot > 1.0


def test_a_hard_source_is_stepped_like_a_gentle_one(oracle):
    assert abs(_die(oracle, 1.0).max_stable_dt() - _die(oracle, 1000.0).max_stable_dt()) < 1e-15


def test_watts_spread_evenly_and_replace(oracle):
    """dissipating: the total is the watts given, spread over the selected cells; a later call
    replaces what those cells generated (solid.rs:1058-1083)."""
    b = block(oracle, COPPER, 4, 5.0, 20.0)
    everything = np.ones(64, dtype=bool)
    b.dissipating(10.0, everything)
    assert abs(b.source.sum() - 10.0) < 1e-12
    half = np.arange(64) < 32
    b.dissipating(4.0, half)
    assert abs(b.source[:32].sum() - 4.0) < 1e-12 and abs(b.source[32:].sum() - 5.0) < 1e-12
