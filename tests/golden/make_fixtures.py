"""Writes tests/golden/stencil_fixtures.json: small known-answer cases of the stencil path, computed
by the CPU oracle (oracle/), with every field stored as the hex of its IEEE-754 binary64 bits.

    python tests/golden/make_fixtures.py

The reference holds no output arrays for this path, so these fixtures are the oracle's, pinned
by what the reference's own tests fix (the RNG digest of rng.rs:295-308, the eigen-decay of
solid_closed_forms.rs:46-72, conservation, Hall's invariant) and checked again by
tests/test_golden.py; the GPU must reproduce them bit for bit (tests/test_gpu_golden.py).

Cases:
  rng          Rng::new(0x5A17_7E3D): the 10,000-draw digest and the first eight outputs.
  mode_7x5x3   Al 6061 block 7x5x3, dx 1 mm, mode (2,1,1) of amplitude 1 K about 300 K, 30 steps at
               0.9 of the limit.
  scene15      crates/pantometry-world/scenes/15-a-hot-spot-in-a-block.json: Al 6061 9^3 at 20 C
               with cell (4,4,4) 60 K above, 50 steps at 0.9 of the limit.
  hetero_16    16^3 heterogeneous block (lognormal k, three hot spots), 40 steps at 0.9 of the
               limit.
  hall_13      13-node rigid hall (air: c 343, rho 1.2, dx 1 mm) released in mode (1,1,1) of
               amplitude 1 Pa, 40 steps at half the Courant limit.
"""
from __future__ import annotations

import json
import struct
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

OUT = Path(__file__).resolve().parent / "stencil_fixtures.json"

AL = (2.70e3, 167.0, 896.0)   # Al 6061: density, conductivity, specific heat (substance.rs:362-382)
DX = 1e-3


def hexs(a) -> str:
    """The array's binary64 bits, big-endian hex, 16 characters per value."""
    a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
    return a.astype(">f8").tobytes().hex()


def hex1(x: float) -> str:
    return struct.pack(">d", float(x)).hex()


def solid_case(o, counts, substance, initial, setup, steps, **meta):
    b = o.OracleSolid3D("block", *substance, counts, DX, initial)
    setup(b)
    start = b.cells.copy()
    k = b.kcell.copy()
    dt = 0.9 * b.max_stable_dt()
    before = b.ledger()
    for _ in range(steps):
        b.step(dt)
    out = dict(meta, counts=list(counts), dx=DX, density=substance[0],
               conductivity=substance[1], specific_heat=substance[2], initial=initial,
               dt=hex1(dt), steps=steps, start=hexs(start), final=hexs(b.cells),
               stored_before=hex1(before[0]), stored_after=hex1(b.ledger()[0]))
    if meta.get("per_cell_k"):
        out["k"] = hexs(k)
    return out


def main() -> int:
    from oracle import oracle as o
    o.lib()
    fx = {}

    r = o.OracleRng(0x5A177E3D)
    fx["rng"] = {"seed": 0x5A177E3D, "draws": 10000,
                 "digest": str(o.rng_digest(0x5A177E3D, 10000)),
                 "first": [format(r.next_u64(), "016x") for _ in range(8)]}

    fx["mode_7x5x3"] = solid_case(
        o, (7, 5, 3), AL, 300.0, lambda b: b.release_mode((2, 1, 1), 300.0, 1.0), 30,
        mode=[2, 1, 1], mean=300.0, amplitude=1.0)

    def spot(b):
        b.set_cells(np.full(9 ** 3, 20.0 + 273.15))
        c = 4 + 9 * (4 + 9 * 4)
        cells = b.cells.copy()
        cells[c] = (20.0 + 60.0) + 273.15
        b.set_cells(cells)
    fx["scene15"] = solid_case(o, (9, 9, 9), AL, 20.0 + 273.15, spot, 50, hot=[4, 4, 4],
                               above_k=60.0)

    k16, t16 = o.heterogeneous_fields((16, 16, 16), spots=3)

    def hetero(b):
        b.with_conductivity(k16)
        b.set_cells(t16)
    fx["hetero_16"] = solid_case(o, (16, 16, 16), (4000.0, 1.0, 1000.0), 300.0, hetero, 40,
                                 spots=3, k_seed=0x5EED0001, spot_seed=0x5EED0002,
                                 per_cell_k=True)

    h = o.OracleHall(12 * DX, 12 * DX, 12 * DX, 13, 343.0, 1.2)
    p0 = h.mode((1, 1, 1), 1.0)
    h.release(p0)
    dt = 0.5 * h.max_stable_dt()
    e0 = h.energy()
    for _ in range(40):
        h.step(dt)
    fx["hall_13"] = {"nodes": [h.nx, h.ny, h.nz], "width": 12 * DX, "speed": 343.0,
                     "density": 1.2, "mode": [1, 1, 1], "amplitude": 1.0, "dt": hex1(dt),
                     "steps": 40, "start": hexs(p0), "p": hexs(h.p), "p_prev": hexs(h.pp),
                     "vx": hexs(h.vx), "vy": hexs(h.vy), "vz": hexs(h.vz),
                     "energy_before": hex1(e0), "energy_after": hex1(h.energy())}

    OUT.write_text(json.dumps(fx, indent=1) + "\n")
    print(f"wrote {OUT} ({OUT.stat().st_size} bytes)")
    return 0


if __name__ == "__main__":
    sys.exit(main())
