"""Benchmark of the explicit 3D heat step on one MI355X.

Workload: the heterogeneous heat block, 512^3 cells of f64, per-cell conductivity lognormal(0,
0.5) W/m/K with harmonic-mean faces, rho c = 4e6 J/m^3/K, dx = 1 mm, eight seeded Gaussian hot
spots, stepped at 0.9 of its stability limit with the conservation audit after every step.  A
"step" is one sweep over the 512^3 cells plus that audit; inputs are resident in HBM before the
timed region starts.

Prints one JSON line: GCUPS (cell updates per second / 1e9) over the timed steps, the roofline of
the dominant kernel (the sweep) from HIP events around each of its launches, and the CPU oracle
timed on this host's cores over a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

N = 512
BYTES_PER_CELL = 24          # T read, k read, T written (f64 each)
HBM_PEAK_GBS = 8000.0        # MI355X HBM3E
METRIC = "GCUPS + %HBM roofline, FP64 heat 512³; surrogate vs stencil speedup @1e-3 L2"


def _traffic_from_profiles(steps_per_launch: int = 1):
    """HBM bytes per sweep launch, from the committed rocprofv3 counter summary of the kernel
    that ran (one or two steps per launch), if present."""
    name = "heat_sweep2_512_counters.json" if steps_per_launch == 2 else "heat_sweep_512_counters.json"
    p = ROOT / "profiles" / name
    if not p.exists():
        return None
    try:
        return json.loads(p.read_text()).get("hbm_bytes_per_launch")
    except (ValueError, OSError):
        return None


def _ensemble_traffic(steps_per_launch: int):
    """HBM bytes per ens_tile launch from the committed counter summary, if present."""
    p = ROOT / "profiles" / "ensemble_512_counters.json"
    if steps_per_launch != 2 or not p.exists():
        return None
    try:
        return json.loads(p.read_text()).get("hbm_bytes_per_launch")
    except (ValueError, OSError):
        return None


def host_cpu() -> dict:
    """The host this runs on: logical CPUs, the CPUs this process may run on, the OpenMP share
    the environment sets, and the CPU model from /proc/cpuinfo."""
    model = "unknown"
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)),
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS"), "model": model}


def _best_step(b, dt, threads: int, trials: int = 5) -> tuple:
    """where_the_time_goes.rs:22-52: one untimed warm-up, then the best of `trials` timed
    trials; at 512^3 one step (plus its ledger) is one trial of well over 0.15 s.  Returns the
    best (step + ledger) and the best step alone, in seconds."""
    b.step(dt, threads_=threads)
    b.stored_heat()
    best, best_sweep = float("inf"), float("inf")
    for _ in range(trials):
        t0 = time.perf_counter()
        b.step(dt, threads_=threads)
        t1 = time.perf_counter()
        b.stored_heat()
        t2 = time.perf_counter()
        best = min(best, t2 - t0)
        best_sweep = min(best_sweep, t1 - t0)
    return best, best_sweep


def cpu_baseline() -> dict:
    """The oracle (the reference's arithmetic and order, gcc -O2 -ffp-contract=off) on the
    heterogeneous 512^3 block, timed on this host: one step plus the ledger (the left-fold
    stored heat of solid.rs:1557-1561), as the workload audits every step.  Variant 1 is the
    reference's own single thread, with its per-step clone of the cells (solid.rs:1774);
    variant 2 spreads the sweep over the host cores (OpenMP over z-planes, bit-identical)."""
    from oracle import oracle as o
    from pantometry_amd import workloads as w
    host = host_cpu()
    share = int(host["omp_num_threads"] or 0) or host["affinity"]
    threads = max(1, min(host["affinity"], share))
    k, t = o.heterogeneous_fields((N, N, N))
    b = o.OracleSolid3D("block", w.RHO, 1.0, w.CP, (N, N, N), w.DX, 300.0)
    b.with_conductivity(k)
    b.set_cells(t)
    del k, t
    dt = 0.9 * b.max_stable_dt()
    t0 = time.perf_counter()
    multi, multi_sweep = _best_step(b, dt, threads)
    single, single_sweep = _best_step(b, dt, 1)
    el = time.perf_counter() - t0
    g = lambda sec: round(N ** 3 / sec / 1e9, 4)
    return {"value": g(multi), "unit": "GCUPS", "cores": threads, "kind": "port",
            "sample": f"512^3 heterogeneous block, best of 5 single steps each followed by the "
                      f"ledger, after one warm-up, per variant ({el:.1f} s in all); oracle "
                      f"sweep with OpenMP over z-planes on {threads} threads (the box's CPU "
                      f"share), stored heat as the reference's single-thread left fold",
            "sweep_only_gcups": g(multi_sweep),
            "single_thread": {"value": g(single), "unit": "GCUPS", "cores": 1,
                              "sweep_only_gcups": g(single_sweep),
                              "note": "the reference's own path: one thread, per-step clone of "
                                      "the cells, then the left-fold ledger"},
            "host": host}


def bench_hall(args) -> dict:
    """The rigid hall, 385^3 nodes (384^3 cells), Ricker source, energy audit every step."""
    from pantometry_amd import workloads as w
    steps = args.steps
    case = w.acoustic_hall(384, steps + args.warmup + 1)
    h = case.hall
    if args.tile_y or args.z_chunk:
        h.set_tile(args.tile_y, args.z_chunk)
    h.set_tile_width(args.tile_x)
    h.advance(case.dt, 1 + args.warmup, tolerance=1e-12)  # startup step + warm-up
    t0 = time.perf_counter()
    totals, ms = h.advance(case.dt, steps, tolerance=1e-12)
    wall = time.perf_counter() - t0
    nodes = 385 ** 3
    return {
        "metric": "G node-updates/s, FP64 hall 385^3 nodes, energy audit every step",
        "value": round(nodes * steps / wall / 1e9, 3), "unit": "Gnode/s", "n_gpus": 1,
        "steps": steps, "warmup": args.warmup, "ms_per_step": round(wall * 1e3 / steps, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic: seeded Ricker point source in a rigid hall at rest",
        "config": {"workload": "hall 385^3 nodes f64, CFL 0.5, Ricker source, audit every step",
                   "tile_x": args.tile_x or 64},
        "roofline": {"bound": "hbm", "achieved": round(64 * nodes / (ms / steps * 1e-3) / 1e9, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(64 * nodes / (ms / steps * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                     "traffic": None, "bytes_per_node": 64,
                     "note": "device loop time per step (fused step + audit finalize)"},
        "audit_drift": float(abs(totals[-1] - totals[0]) / max(abs(totals[0]), abs(h.supplied()))),
        "cpu_baseline": None,
    }


def bench_ensemble(args) -> dict:
    """512 uniform heat blocks, edges 16-96, 500 steps each in one graph launch."""
    from pantometry_amd import workloads as w
    ens, shapes = w.heat_ensemble(512)
    ens.set_steps_per_pass(args.steps_per_pass)
    by_policy = None
    if args.z_chunk > 0:
        ens.set_tiles(args.tiles)
        ens.set_schedule(0, args.z_chunk)
        report = f"fixed tiles={args.tiles} resident_edge=0 z_chunk={args.z_chunk}"
    else:
        report = ens.tune()
        # The tuner's best time per streamed tile policy, ms per step.
        by_policy = {}
        for line in report.splitlines():
            if line.startswith("tiles=") and line.endswith("ms/step"):
                pol = "tiles=" + line.split()[0].split("=")[1]
                ms_c = float(line.rsplit(":", 1)[1].split()[0])
                by_policy[pol] = min(by_policy.get(pol, ms_c), ms_c)
    cells = sum(e ** 3 for e, _ in shapes)
    ens.run(args.warmup)
    before = ens.ledgers()
    t0 = time.perf_counter()
    ms = ens.run(args.steps)
    wall = time.perf_counter() - t0
    after = ens.ledgers()
    return {
        "metric": "GCUPS, batched FP64 heat ensemble (512 grids, edges 16-96)",
        "value": round(cells * args.steps / wall / 1e9, 3), "unit": "GCUPS", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 5), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic: seeded edges, diffusivities, hot spots and noise",
        "config": {"workload": "ensemble of 512 uniform blocks, one graph launch per run",
                   "cells": cells, "schedule": report.strip().splitlines()[-1],
                   "tuner_best_ms_per_step_by_tile_policy": by_policy},
        "roofline": {"bound": "hbm", "achieved": round(16 * cells / (ms / args.steps * 1e-3) / 1e9, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(16 * cells / (ms / args.steps * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                     "traffic": _ensemble_traffic(args.steps_per_pass), "bytes_per_cell": 16,
                     "steps_per_launch": args.steps_per_pass,
                     "kernel": f"ens_tile<{args.steps_per_pass}> (whole-row tiles, z-marching, "
                               "LDS-staged planes)",
                     "note": "graph launch time per step (one launch per two steps: temporal "
                             "blocking, so 'achieved' counts 16 B per cell-step while 'traffic' "
                             "is the HBM bytes of one launch); resident members touch HBM only "
                             "at the ends"},
        "audit_drift": float((abs(after - before) / abs(before)).max()),
        "cpu_baseline": None,
    }


def bench_surrogate(args) -> dict:
    """The surrogate-assisted rollout: a heterogeneous 128^3 block advanced 1024 stencil steps by
    64 surrogate steps of 16, with the FP64 stencil run beside it on every fourth chunk as the
    judge (a checked chunk over 1e-3 in the error of the heat falls back to the stencil's
    result) and the one-step-change veto on the others, against the plain GPU stencil over the
    same 1024 steps.  `final_error_kelvin` is the relative L2 error ||T - T_h|| / ||T_h||,
    `final_error` the error of the heat ||T - T_h|| / ||T_h - 300 K||."""
    import numpy as np
    from pantometry_amd import surrogate as S
    s = S.Surrogate()
    s.load(S.WEIGHTS)
    block, k, dt = S.heterogeneous_pair_block(128, 424242)
    t0 = block.cells().copy()
    del block
    runs = {}
    for name, tol, every in (("checked_every_4th", 1e-3, 4), ("checked_every_chunk", 1e-3, 1),
                             ("unchecked", float("inf"), 0)):
        rep = S.rollout(s, k, t0, dt, steps=1024, tolerance=tol, check_every=every,
                        residual=tol != float("inf"))
        runs[name] = {"speedup": round(rep.speedup, 4), "rollout_s": round(rep.rollout_s, 4),
                      "stencil_s": round(rep.stencil_s, 4), "checked": rep.checked,
                      "fallbacks": rep.fallbacks, "step_change_vetoes": rep.rises,
                      "max_chunk_error": rep.max_error,
                      "final_error": rep.final_error,
                      "final_error_kelvin": rep.final_error_kelvin, "max_drift": rep.max_drift,
                      "surrogate_step_ms": round(rep.surrogate_ms, 4)}
    main_run = runs["checked_every_4th"]
    flops = S.forward_flops(128)
    tflops = flops / (runs["unchecked"]["surrogate_step_ms"] * 1e-3) / 1e12
    return {
        "metric": "surrogate-assisted rollout speedup vs the plain FP64 GPU stencil, heat 128^3, "
                  "1024 steps, veto at 1e-3 relative L2",
        "value": main_run["speedup"], "unit": "x", "n_gpus": 1, "steps": 1024, "warmup": 1,
        "ms_per_step": round(main_run["rollout_s"] * 1e3 / 1024, 5), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "bf16 (network), f64 (stencil, state)",
        "data": "synthetic: seeded heterogeneous block; network trained on GPU-stencil pairs "
                "and its own rollouts",
        "config": {"workload": "heat 128^3 heterogeneous, 64 surrogate steps of 16 stencil steps",
                   "weights": S.WEIGHTS.name},
        "roofline": {"bound": "mfma", "achieved": round(tflops, 2), "peak": 2500.0,
                     "unit": "TFLOP/s", "frac": round(tflops / 2500.0, 5), "traffic": None,
                     "note": "one surrogate step (network forward, features and conservative "
                             f"shift) over {flops / 1e9:.1f} GFLOP of convolutions"},
        "runs": runs,
        "cpu_baseline": None,
    }


def bench_film(args) -> dict:
    """The extended heat path: a copper block with an alumina layer, every face exposed to its own
    environment (convection and radiation) and two volumetric sources, at 128^3 and at a
    scene-sized 24^3.  Device-resident advance (limit, film, sweep, sources, books and audit every
    step on the device) against one call per step through the C ABI (each step's limit, loss and
    books read back on the host)."""
    import numpy as np
    import pantometry_amd as pm
    c0 = 273.15
    rows = {}
    for edge, steps, host_steps in ((128, max(args.steps, 200), 40), (24, max(args.steps, 500), 200)):
        counts, dx = (edge, edge, edge), 1.5e-3
        n = edge ** 3
        k, j, i = np.meshgrid(np.arange(edge), np.arange(edge), np.arange(edge), indexing="ij")
        layer = (k == 1).reshape(-1).astype(np.uint8)
        top = (k == edge - 1).reshape(-1).astype(np.uint8)
        core = ((abs(i - edge // 2) < edge // 8) & (abs(j - edge // 2) < edge // 8) &
                (abs(k - edge // 2) < edge // 8)).reshape(-1).astype(np.uint8)
        side = (edge * dx) ** 2

        def make():
            g = pm.Solid3D("module", pm.Substance.of("cu", 8.96e3, 401.0, 385.0, 0.04), counts, dx,
                           c0 + 60.0)
            g.fill(pm.Substance.of("alumina", 3720.0, 24.0, 880.0, 0.8), layer)
            for f, amb, h in ((0, 25.0, 40.0), (1, 30.0, 7.0), (2, 20.0, 150.0), (3, 22.0, 0.0),
                              (4, 40.0, 3000.0), (5, 35.0, 12.0)):
                g.losing_from(pm.Face(f), pm.Environment(c0 + amb, h, side))
            g.dissipating(45.0, top)
            g.dissipating(3.0, core)
            return g

        g = make()
        dt = 0.7 * g.max_stable_dt()
        g.advance(dt, args.warmup, tolerance=1e-9)
        t0 = time.perf_counter()
        totals, ms = g.advance(dt, steps, tolerance=1e-9)
        wall = time.perf_counter() - t0
        h = make()
        bus = pm.Exchange()
        for _ in range(min(args.warmup, 5)):
            h.step(0.0, dt, bus)
        t0 = time.perf_counter()
        for _ in range(host_steps):
            h.step(0.0, dt, bus)
        host = (time.perf_counter() - t0) / host_steps
        rows[f"{edge}^3"] = {
            "device_resident_us_per_step": round(wall * 1e6 / steps, 2),
            "device_loop_us_per_step": round(ms * 1e3 / steps, 2),
            "gcups": round(n * steps / wall / 1e9, 3),
            "one_call_per_step_us": round(host * 1e6, 2),
            "speedup_over_one_call_per_step": round(host / (wall / steps), 2),
            "steps": steps, "audit": "every step's ledger within 1e-9 of the step before",
            "roofline": {"bound": "hbm", "bytes_per_cell": 32,
                         "achieved": round(32 * n / (ms / steps * 1e-3) / 1e9, 1),
                         "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(32 * n / (ms / steps * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
        }
    big = rows["128^3"]
    return {
        "metric": "GCUPS, FP64 heat with exposed faces and sources (extended path), 128^3",
        "value": big["gcups"], "unit": "GCUPS", "n_gpus": 1, "steps": big["steps"],
        "warmup": args.warmup, "ms_per_step": round(big["device_resident_us_per_step"] / 1e3, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic: copper with an alumina layer, six exposed faces, two sources",
        "config": {"workload": "extended heat path, device-resident advance, audit every step"},
        "roofline": {**big["roofline"],
                     "note": "device loop time per step over 32 B a cell (T, k and mobility "
                             "read, T written); nine launches a step: film rates, limit, film "
                             "shed, sweep, sources, stored heat, books"},
        "rows": rows, "cpu_baseline": None,
    }


def bench_lines(args) -> dict:
    """The lower-dimensional domains at the reference's own sizes and at the largest line one
    workgroup holds in LDS: device-resident advance (audit every step) against the oracle's
    step on this host, both per step."""
    import math
    import numpy as np
    import pantometry_amd as pm
    from oracle.oracle import OracleBar1D, OracleRoom, OracleTube
    steps = max(args.steps, 100)
    rows = {}

    def cpu_time(step, budget=1.5):
        n, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < budget and n < 20000:
            step()
            n += 1
        return (time.perf_counter() - t0) / n * 1e6

    al = pm.Substance.aluminium_6061()
    for cells in (21, 2048):
        bar = pm.Bar1D("bar", al, cells, 1e-3, 1e-4, 293.15)
        c = np.full(cells, 293.15)
        c[cells // 2] += 40.0
        bar.set_cells(c)
        dt = 0.9 * bar.max_stable_dt()
        bar.advance(dt, args.warmup)
        _, ms = bar.advance(dt, steps)
        o = OracleBar1D(cells, 1e-3, 1e-4, al.density, al.conductivity, al.specific_heat, 293.15)
        o.cells = c.copy()
        rows[f"bar {cells} cells"] = {"gpu_us_per_step": round(ms * 1e3 / steps, 3),
                                      "cpu_us_per_step": round(cpu_time(lambda: o.step(dt)), 3)}
    for cells in (201, 2048):
        tube = pm.Tube.of_air("tube", 1.0, cells, 1e-4)
        tube.released_from(lambda x: 200.0 * math.exp(-((x - 0.5) / 0.08) ** 2))
        dt = tube.max_stable_dt()
        tube.advance(dt, args.warmup)
        _, ms = tube.advance(dt, steps)
        o = OracleTube(1.0, cells, 1e-4, 1.204, 343.0)
        o.release(np.array([200.0 * math.exp(-((i * o.dx - 0.5) / 0.08) ** 2) for i in range(o.n)]))
        rows[f"tube {cells} cells"] = {"gpu_us_per_step": round(ms * 1e3 / steps, 3),
                                       "cpu_us_per_step": round(cpu_time(lambda: o.step(dt)), 3)}
    for cells in (81, 241):
        room = pm.Room.of_air("room", 4.4, 3.1, cells)
        room.released_in_mode((1, 1), 1.0)
        dt = room.max_stable_dt()
        room.advance(dt, args.warmup)
        rsteps = min(steps, 500)
        _, ms = room.advance(dt, rsteps)
        o = OracleRoom(4.4, 3.1, cells, 1.0, 343.0, 1.204)
        o.release(o.mode2(1, 1, 1.0))
        rows[f"room {cells} nodes across"] = {
            "gpu_us_per_step": round(ms * 1e3 / rsteps, 3),
            "cpu_us_per_step": round(cpu_time(lambda: (o.step(dt), o.invariant())), 3)}
    return {"metric": "microseconds per audited step, lower-dimensional domains",
            "workload": "lines", "steps": steps, "rows": rows,
            "note": "GPU: device-resident advance with the audit every step (a line: one "
                    "launch for all steps; a room: three launches a step); CPU: the oracle's "
                    "step (plus the room's invariant) from Python on one core"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-secondary", dest="secondary", action="store_false",
                    help="skip the hall, ensemble, surrogate and film figures the headline line carries")
    ap.add_argument("--tile-y", type=int, default=0)
    ap.add_argument("--tile-x", type=int, default=0,
                    help="hall: width of the paired step's column tiles, 64/56/48 (0: 64)")
    ap.add_argument("--z-chunk", type=int, default=0)
    ap.add_argument("--tiles", type=int, default=0,
                    help="ensemble streamed tile policy 0-3 (with --z-chunk; else tuned)")
    ap.add_argument("--steps-per-pass", type=int, choices=[1, 2], default=2,
                    help="2 (default): pairs of steps per pass over HBM; 1: one step per pass")
    ap.add_argument("--workload", choices=["heat", "hall", "ensemble", "surrogate", "film", "copy",
                                           "lines"],
                    default="heat",
                    help="heat (the headline 512^3 block), or the hall / ensemble / surrogate / "
                         "film (extended heat path) secondaries")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: the heat block is not sharded across GPUs "
              "(it would shard by z-slabs with one plane of halo exchanged per step); "
              "run with --gpus 1", file=sys.stderr)
        return 2

    import pantometry_amd as pm

    if args.workload == "hall":
        print(json.dumps(bench_hall(args)), flush=True)
        return 0
    if args.workload == "ensemble":
        print(json.dumps(bench_ensemble(args)), flush=True)
        return 0
    if args.workload == "lines":
        print(json.dumps(bench_lines(args)), flush=True)
        return 0
    if args.workload == "film":
        print(json.dumps(bench_film(args)), flush=True)
        return 0
    if args.workload == "surrogate":
        print(json.dumps(bench_surrogate(args)), flush=True)
        return 0
    if args.workload == "copy":
        import ctypes as C
        ms = C.c_double(0.0)
        pm._capi.check(pm.lib().pm_stream_copy(N ** 3, args.steps, C.byref(ms)), "pm_stream_copy")
        gbs = BYTES_PER_CELL * N ** 3 / (ms.value * 1e-3) / 1e9
        print(json.dumps({"metric": "stream c=a+b over 512^3 doubles (counter calibration)",
                          "value": round(gbs, 1), "unit": "GB/s", "ms_per_launch": ms.value,
                          "bytes_read": 16 * N ** 3, "bytes_written": 8 * N ** 3}), flush=True)
        return 0

    out = bench_heat(args)
    if args.secondary:
        # The other configurations of the metric, measured in the same run.
        sec = {}
        for name, fn, steps in (("hall_385", bench_hall, max(args.steps, 200)),
                                ("ensemble_512", bench_ensemble, 500),
                                ("surrogate_128", bench_surrogate, 1024),
                                ("film_128", bench_film, 200)):
            print(f"[bench] {name}", file=sys.stderr, flush=True)
            sargs = argparse.Namespace(**vars(args))
            sargs.tile_y, sargs.z_chunk, sargs.steps_per_pass, sargs.steps = 0, 0, 2, steps
            sargs.warmup = min(args.warmup, 20)
            r = fn(sargs)
            sec[name] = {k: r[k] for k in ("metric", "value", "unit", "ms_per_step", "steps",
                                           "roofline", "config") if k in r}
            for k in ("runs", "rows", "audit_drift", "cpu_baseline"):
                if r.get(k) is not None:
                    sec[name][k] = r[k]
        out["secondary"] = sec
    if not args.no_cpu:
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline()
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)
    return 0


def bench_heat(args) -> dict:
    """The headline: the heterogeneous 512^3 block, audited every step."""
    from pantometry_amd import workloads as w
    block, dt = w.heterogeneous_block(N)
    if args.tile_y or args.z_chunk:
        block.set_tile(args.tile_y, args.z_chunk)
    block.set_steps_per_pass(args.steps_per_pass)
    if args.warmup:
        block.advance(dt, args.warmup, tolerance=1e-12, want_totals=False)
    # A profiled pass (HIP events around every launch) gives the finalize kernel's duration and
    # the evented sweep duration; the headline and the roofline come from the plain pass after it.
    block.set_profiling(True)
    totals, _ = block.advance(dt, args.steps, tolerance=1e-12)
    sweep_ms_evented, audit_ms, per_launch = block.kernel_times()
    block.set_profiling(False)
    t0 = time.perf_counter()
    _, loop_ms = block.advance(dt, args.steps, tolerance=1e-12, want_totals=False)
    wall = time.perf_counter() - t0

    cells = N ** 3
    ms_per_step = wall * 1e3 / args.steps
    gcups = cells * args.steps / wall / 1e9
    # The sweep's mean launch duration in the plain pass: the device loop time (HIP events on the
    # handle's stream around the whole loop) less the finalize launches, over the sweep launches.
    launches = args.steps // per_launch + args.steps % per_launch
    sweep_ms = (loop_ms - launches * audit_ms) / launches
    sweep_ms_per_step = sweep_ms / per_launch
    if sweep_ms_per_step > ms_per_step * (1 + 1e-9):
        raise RuntimeError(f"sweep time per step {sweep_ms_per_step:.5f} ms exceeds the measured "
                           f"{ms_per_step:.5f} ms per step")
    # Algorithmic bytes of one sweep launch: 24 B per cell-step, `per_launch` steps per launch.
    achieved = BYTES_PER_CELL * cells * per_launch / (sweep_ms * 1e-3) / 1e9
    traffic_file = "heat_sweep2_512" if per_launch == 2 else "heat_sweep_512"
    out = {
        "metric": METRIC,
        "value": round(gcups, 3),
        "unit": "GCUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: seeded lognormal(0, 0.5) conductivity and 8 Gaussian hot spots",
        "config": {
            "workload": "heat 512^3 f64, heterogeneous conductivity, harmonic faces, "
                        "audit every step (0.9 of the stability limit)",
            "cells": cells, "dt": dt, "tile_y": args.tile_y or 8, "z_chunk": args.z_chunk or 128,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": _traffic_from_profiles(per_launch),
            "traffic_source": f"profiles/{traffic_file}.md (numbers in {traffic_file}_counters.json): "
                              "rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE of this kernel at this "
                              "size, scaled by the same counters on a copy kernel of known size; "
                              "not measured in this run",
            "kernel": ("heat_sweep2<8> (per-cell k, uniform C, 64x8 tile, two steps per pass)"
                       if per_launch == 2 else "heat_sweep<1,0,8> (per-cell k, uniform C, 64x8 tile)"),
            "kernel_ms": round(sweep_ms, 5),
            "kernel_ms_method": "device loop time of the timed pass (HIP events on the sweep's "
                                "stream) less the finalize launches, per sweep launch",
            "kernel_ms_evented": round(sweep_ms_evented, 5),
            "steps_per_launch": per_launch,
            "audit_kernel_ms": round(audit_ms, 5),
            "bytes_per_cell_step": BYTES_PER_CELL,
            "note": ("temporal blocking: one pass over HBM advances two steps, so 'achieved' "
                     "(algorithmic 24 B per cell-step) can exceed the HBM traffic per launch "
                     "('traffic' ~ 24 B per cell)" if per_launch == 2 else
                     "one step per pass"),
        },
        "device_loop_ms_per_step": round(loop_ms / args.steps, 5),
        "audit_drift": float(abs(totals[-1] - totals[0]) / abs(totals[0])),
    }
    del block
    return out


if __name__ == "__main__":
    sys.exit(main())
