"""Writes the full-length digests of the three stencil configurations, computed by the oracle on
this machine's cores: `full_length.json` (SHA-256 of the inputs and of the final fields' bits,
time steps, source node) and `full_length_totals.npz` (the oracle's ledger totals after every
step: the left fold the reference computes and, for the heat block, the compensated sum).

  heat      512^3 heterogeneous block (k lognormal(0, 0.5), 8 hot spots), 2000 steps at 0.9 of
            its limit, the ledger after every step (solid.rs:1773-1826, 1557-1561);
  hall      385^3 nodes, CFL 0.5, Ricker source at the seeded node, 4000 steps, the energy ledger
            after every step (hall.rs:408-530, 334-384): the reference's left fold, the same
            terms summed with compensation, and the ledger's scale;
  ensemble  512 uniform members (edges 16-96, alpha 1e-5..1e-3 log-uniform, hot spot + noise),
            500 steps each (ensemble.rs:102-147), each member's stored heat at the end.

Every input comes from the oracle's own generator (the reference's Rng and host libm), so the
GPU side checks its inputs' digests before its results'.  Run from the repository root:
`python tests/golden/make_digests.py [--resume DIR] [heat|hall|ensemble ...]` (all three by
default; about three hours on 8 cores and 12 GB of memory).  With --resume the heat and hall runs
save their state to DIR every few hundred steps and pick up from there when started again.
"""
from __future__ import annotations

import hashlib
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle import oracle as o  # noqa: E402

OUT_JSON = Path(__file__).with_name("full_length.json")
OUT_NPZ = Path(__file__).with_name("full_length_totals.npz")

DX = 1e-3
RHO, CP = 4000.0, 1000.0
HALL_SOURCE_SEED = 0x5EED0003
ENSEMBLE_SEED = 0x5EED0004
THREADS = len(os.sched_getaffinity(0))


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.float64).tobytes()).hexdigest()


def ricker(t: float, f: float, t0: float, amplitude: float = 1.0) -> float:
    tau = t - t0
    a = math.pi * math.pi * f * f * tau * tau
    return amplitude * (1.0 - 2.0 * a) * math.exp(-a)


RESUME = None       # directory of restart points, or None
SAVE_EVERY = 100


def _restart(name: str):
    """The saved restart point of run `name` (a dict of arrays), or None."""
    if RESUME is None or not (RESUME / f"{name}.npz").exists():
        return None
    with np.load(RESUME / f"{name}.npz") as z:
        return {k: z[k] for k in z.files}


def _save_restart(name: str, **state) -> None:
    if RESUME is None:
        return
    RESUME.mkdir(parents=True, exist_ok=True)
    tmp = RESUME / f"{name}.tmp.npz"
    np.savez(tmp, **state)
    os.replace(tmp, RESUME / f"{name}.npz")


def heat(rec: dict, arrays: dict, n: int = 512, steps: int = 2000) -> None:
    k, t = o.heterogeneous_fields((n, n, n))
    b = o.OracleSolid3D("block", RHO, 1.0, CP, (n, n, n), DX, 300.0).with_conductivity(k)
    b.set_cells(t)
    out = {"counts": [n, n, n], "steps": steps, "k_sha256": sha(k), "t0_sha256": sha(t)}
    del k, t
    dt = 0.9 * b.max_stable_dt()
    out["dt"] = dt.hex()
    out["stored0"] = b.stored_heat().hex()
    fold, exact = np.empty(steps), np.empty(steps)
    start = 0
    saved = _restart("heat")
    if saved is not None:
        start = int(saved["done"])
        b.cells[:] = saved["cells"]
        fold[:start], exact[:start] = saved["fold"][:start], saved["exact"][:start]
        print(f"heat: resuming after step {start}", flush=True)
    t0 = time.time()
    for s in range(start, steps):
        b.step(dt, threads_=THREADS)
        fold[s] = b.stored_heat()
        exact[s] = b.stored_heat_compensated(THREADS)
        if s % SAVE_EVERY == SAVE_EVERY - 1:
            print(f"heat step {s + 1} ({time.time() - t0:.0f} s)", flush=True)
            _save_restart("heat", done=s + 1, cells=b.cells, fold=fold, exact=exact)
    out["final_sha256"] = sha(b.cells)
    rec["heat"] = out
    arrays["heat_fold"], arrays["heat_exact"] = fold, exact


def hall(rec: dict, arrays: dict, cells: int = 384, steps: int = 4000) -> None:
    h = o.OracleHall(cells * DX, cells * DX, cells * DX, cells + 1, 343.0, 1.2)
    nx, ny, nz = h.nodes()
    dt = 0.5 * h.max_stable_dt()
    r = o.OracleRng.for_index(HALL_SOURCE_SEED, 0)
    i = 1 + int(r.unit() * (nx - 2))
    j = 1 + int(r.unit() * (ny - 2))
    k = 1 + int(r.unit() * (nz - 2))
    node = i + nx * (j + ny * k)
    f = 343.0 / (20.0 * h.dx)
    inc = np.array([ricker((s + 1) * dt, f, 1.0 / f) for s in range(steps)])
    h.release(np.zeros(nx * ny * nz))
    h.set_source(node, inc)
    out = {"nodes": [nx, ny, nz], "steps": steps, "dt": dt.hex(), "source_node": node,
           "increments_sha256": sha(inc)}
    totals, exact, scale = np.empty(steps), np.empty(steps), np.empty(steps)
    start = 0
    saved = _restart("hall")
    if saved is not None:
        start = int(saved["done"])
        for name in ("p", "pp", "vx", "vy", "vz"):
            getattr(h, name)[:] = saved[name]
        h.staggered = bool(saved["staggered"])
        h.offset, h.datum, h.supplied = (float(saved[x]) for x in ("offset", "datum", "supplied"))
        h.cursor = int(saved["cursor"])
        totals[:start] = saved["totals"][:start]
        exact[:start], scale[:start] = saved["exact"][:start], saved["scale"][:start]
        print(f"hall: resuming after step {start}", flush=True)
    t0 = time.time()
    for s in range(start, steps):
        h.step(dt, threads_=THREADS)
        totals[s], scale[s] = h.ledger()
        # Ledger::with over [energy, -supplied] with the energy's terms summed exactly
        exact[s] = (h.invariant_compensated(THREADS) + h.offset) + (-h.supplied)
        if s % SAVE_EVERY == SAVE_EVERY - 1:
            print(f"hall step {s + 1} ({time.time() - t0:.0f} s)", flush=True)
            _save_restart("hall", done=s + 1, p=h.p, pp=h.pp, vx=h.vx, vy=h.vy, vz=h.vz,
                          staggered=h.staggered, offset=h.offset, datum=h.datum,
                          supplied=h.supplied, cursor=h.cursor, totals=totals, exact=exact,
                          scale=scale)
    out["final_sha256"] = {name: sha(getattr(h, name)) for name in ("p", "pp", "vx", "vy", "vz")}
    out["supplied"] = float(h.supplied).hex()
    rec["hall"] = out
    arrays["hall_totals"], arrays["hall_exact"], arrays["hall_scale"] = totals, exact, scale


def _member(m: int, steps: int):
    edge, alpha = o.member_shape(ENSEMBLE_SEED, m)
    b = o.OracleSolid3D("m", RHO, alpha * RHO * CP, CP, (edge, edge, edge), DX, 300.0)
    b.set_cells(o.hot_spot_cells((edge, edge, edge), seed=ENSEMBLE_SEED + 1 + m))
    t0_sha = sha(b.cells)
    dt = 0.9 * b.max_stable_dt()
    for _ in range(steps):
        b.step(dt)
    return edge, alpha, dt, t0_sha, sha(b.cells), b.stored_heat()


def ensemble(rec: dict, arrays: dict, count: int = 512, steps: int = 500) -> None:
    from concurrent.futures import ThreadPoolExecutor
    t0 = time.time()
    # The oracle's single-thread step releases the GIL inside the C call, so members run side by
    # side; each member's arithmetic is that of ensemble.rs's one-sample-per-thread run.
    with ThreadPoolExecutor(max_workers=THREADS) as pool:
        rows = list(pool.map(lambda m: _member(m, steps), range(count)))
    print(f"ensemble ({time.time() - t0:.0f} s)", flush=True)
    rec["ensemble"] = {
        "count": count, "steps": steps,
        "members": [{"edge": e, "alpha": a.hex(), "dt": d.hex(), "t0_sha256": s0,
                     "final_sha256": s1} for e, a, d, s0, s1, _ in rows],
    }
    arrays["ensemble_stored"] = np.array([r[5] for r in rows])


def main(which) -> None:
    rec = json.loads(OUT_JSON.read_text()) if OUT_JSON.exists() else {}
    arrays = dict(np.load(OUT_NPZ)) if OUT_NPZ.exists() else {}
    for name in which:
        {"heat": heat, "hall": hall, "ensemble": ensemble}[name](rec, arrays)
        rec["generator"] = ("tests/golden/make_digests.py: the oracle (oracle/pm_oracle.c, gcc -O2 "
                            "-ffp-contract=off, OpenMP over planes), host libm")
        OUT_JSON.write_text(json.dumps(rec, indent=1, sort_keys=True) + "\n")
        np.savez_compressed(OUT_NPZ, **arrays)
        print(f"wrote {name}", flush=True)


if __name__ == "__main__":
    args = sys.argv[1:]
    if args[:1] == ["--resume"]:
        RESUME = Path(args[1])
        args = args[2:]
    main(args or ["ensemble", "hall", "heat"])
