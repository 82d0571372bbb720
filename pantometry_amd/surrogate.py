"""The learned heat surrogate: a 3D U-Net that advances a heterogeneous heat block by k stencil
steps at once, trained and run on gfx950's bf16 matrix cores, with the FP64 stencil as its judge
and fallback.

Nothing in pantometry is learned; this is an extension beside the drop-in Solid3D path.  The
network maps ((T - 300 K) / 100 K, log k / 0.5) to (T_k - T_0) / 10 K; a step adds the
prediction to T in fp64 and shifts the result uniformly so the block's total heat is unchanged
(uniform capacity), which keeps the conservation audit's drift at rounding level by
construction.  The veto therefore rests on the error against the stencil: a rollout runs the
FP64 stencil beside the surrogate on the chunks it checks, and replaces any surrogate chunk whose
relative L2 error ||T_s - T_h|| / ||T_h - T_ref|| exceeds the tolerance by the stencil's result.
"""
from __future__ import annotations

import ctypes as C
import math
import time
from dataclasses import dataclass, field
from pathlib import Path
from typing import Callable, List, Optional, Sequence, Tuple

import numpy as np

from . import _capi
from ._capi import check, dptr, fptr, lib
from .rng import Rng
from .thermal import Solid3D, Substance

SURROGATE_SEED = 0x5EED0010
TRAJECTORY_SEED = 0x5EED0011
BATCH_SEED = 0x5EED0012
PUSHFORWARD_SEED = 0x5EED0013

WEIGHTS = Path(__file__).resolve().parent / "weights" / "heat_unet_k16.npz"


class Surrogate:
    """A device-resident U-Net with its Adam state and an optional device-resident training set."""

    def __init__(self, seed: int = SURROGATE_SEED):
        h = C.c_void_p()
        check(lib().pm_surrogate_create(int(seed), C.byref(h)), "pm_surrogate_create")
        self._h = h
        self._set: Tuple[int, int] = (0, 0)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().pm_surrogate_destroy(h)
            self._h = None

    # ---- parameters -------------------------------------------------------------------------
    @property
    def parameter_count(self) -> int:
        return int(lib().pm_surrogate_parameter_count(self._h))

    def parameters(self) -> np.ndarray:
        out = np.empty(self.parameter_count, dtype=np.float32)
        check(lib().pm_surrogate_get_parameters(self._h, fptr(out)), "pm_surrogate_get_parameters")
        return out

    def set_parameters(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != self.parameter_count:
            raise ValueError(f"expected {self.parameter_count} parameters, got {p.size}")
        check(lib().pm_surrogate_set_parameters(self._h, fptr(p)), "pm_surrogate_set_parameters")

    def save(self, path: Path, **meta) -> None:
        path = Path(path)
        path.parent.mkdir(parents=True, exist_ok=True)
        np.savez_compressed(path, parameters=self.parameters(),
                            meta=np.array(repr(sorted(meta.items()))))

    def load(self, path: Optional[Path] = None) -> None:
        """Parameters from `path`, by default the shipped weights."""
        with np.load(Path(path) if path is not None else WEIGHTS, allow_pickle=False) as f:
            self.set_parameters(f["parameters"])

    # ---- training set -----------------------------------------------------------------------
    def reserve(self, samples: int, edge: int) -> None:
        check(lib().pm_surrogate_reserve_set(self._h, int(samples), int(edge)),
              "pm_surrogate_reserve_set")
        self._set = (int(samples), int(edge))

    def set_sample(self, index: int, t0: np.ndarray, k: np.ndarray, tk: np.ndarray,
                   coupling: float) -> None:
        """Pair `index` from host arrays; `coupling` is dt * dx / C of the steps between them."""
        a = [np.ascontiguousarray(v, dtype=np.float64).reshape(-1) for v in (t0, k, tk)]
        check(lib().pm_surrogate_set_sample(self._h, int(index), dptr(a[0]), dptr(a[1]),
                                            dptr(a[2]), float(coupling)), "pm_surrogate_set_sample")

    def set_sample_from(self, index: int, before: Solid3D, after: Solid3D, dt: float) -> None:
        """Pair `index` from two blocks 16 stencil steps of `dt` apart, device to device."""
        check(lib().pm_surrogate_set_sample_from(self._h, int(index), before._h, after._h,
                                                 float(dt)), "pm_surrogate_set_sample_from")

    # ---- training ---------------------------------------------------------------------------
    def train_step(self, indices: Sequence[int], lr: float = 1e-3) -> float:
        idx = (C.c_int * len(indices))(*[int(i) for i in indices])
        loss = C.c_double(0.0)
        check(lib().pm_surrogate_train_step(self._h, idx, len(indices), float(lr), C.byref(loss)),
              "pm_surrogate_train_step")
        return loss.value

    def gradients(self, indices: Sequence[int]) -> Tuple[float, np.ndarray]:
        idx = (C.c_int * len(indices))(*[int(i) for i in indices])
        g = np.empty(self.parameter_count, dtype=np.float32)
        loss = C.c_double(0.0)
        check(lib().pm_surrogate_gradients(self._h, idx, len(indices), fptr(g), C.byref(loss)),
              "pm_surrogate_gradients")
        return loss.value, g

    # ---- inference --------------------------------------------------------------------------
    def forward(self, t: np.ndarray, k: np.ndarray, coupling: float) -> np.ndarray:
        """The normalised change (T_16 - T_0) / 10 K the network predicts for one cubic block
        whose stencil step has coupling dt * dx / C."""
        t = np.ascontiguousarray(t, dtype=np.float64)
        k = np.ascontiguousarray(k, dtype=np.float64)
        edge = round(t.size ** (1.0 / 3.0))
        y = np.empty(t.size, dtype=np.float32)
        check(lib().pm_surrogate_forward(self._h, edge, dptr(t), dptr(k), float(coupling),
                                         fptr(y)), "pm_surrogate_forward")
        return y.reshape(t.shape)

    def step(self, block: Solid3D, dt: float, reference: float) -> Tuple[float, float]:
        """One surrogate step of `block` in place, standing in for 16 stencil steps of `dt`;
        returns (audit drift, device ms)."""
        drift, ms = C.c_double(0.0), C.c_double(0.0)
        check(lib().pm_surrogate_step_solid(self._h, block._h, float(dt), float(reference),
                                            C.byref(drift), C.byref(ms)),
              "pm_surrogate_step_solid")
        return drift.value, ms.value


K_STEPS = 16   # stencil steps one surrogate step stands in for

# (cin, cout, level) of the fourteen 3x3x3 convolutions, encoder then decoder; level l runs at
# edge / 2^l.  The 1x1 head (16 -> 1) follows the last.
LAYERS = [(16, 16, 0), (16, 16, 0), (16, 32, 1), (32, 32, 1), (32, 64, 2), (64, 64, 2),
          (64, 128, 3), (128, 128, 3), (192, 64, 2), (64, 64, 2), (96, 32, 1), (32, 32, 1),
          (48, 16, 0), (16, 16, 0)]


def forward_flops(edge: int) -> int:
    """Multiply-add operations (x2) of one forward pass over an edge^3 block."""
    total = 2 * 16 * edge ** 3
    for cin, cout, level in LAYERS:
        total += 2 * 27 * cin * cout * (edge >> level) ** 3
    return total


def coupling(dt: float, dx: float = None, density: float = None, specific_heat: float = None) -> float:
    """dt * dx / C for a cubic cell of side dx: the stencil's per-step coupling scale."""
    from .workloads import CP, DX, RHO
    dx = DX if dx is None else dx
    rho = RHO if density is None else density
    cp = CP if specific_heat is None else specific_heat
    return dt * dx / ((rho * (dx * dx * dx)) * cp)


# ---- data from the GPU stencil ------------------------------------------------------------------

def heterogeneous_pair_block(edge: int, seed_index: int) -> Tuple[Solid3D, np.ndarray, float]:
    """A heterogeneous block of edge^3 cells (lognormal k, eight hot spots) seeded by
    `seed_index`, with its conductivity and 0.9 of its stability limit."""
    from .workloads import CP, DX, HETERO_K_SEED, HETERO_SPOT_SEED, RHO, heterogeneous_fields
    k, t = heterogeneous_fields((edge, edge, edge), 8, HETERO_K_SEED + 0x100 * (seed_index + 1),
                                HETERO_SPOT_SEED + 0x100 * (seed_index + 1))
    block = Solid3D("trajectory", Substance.of("heterogeneous", RHO, 1.0, CP), (edge,) * 3, DX,
                    300.0)
    block.with_conductivity(k)
    block.set_cells(t)
    return block, k, 0.9 * block.max_stable_dt(0.0)


def build_training_set(surr: Surrogate, samples: int, edge: int = 64, chunks: int = 64, seed: int = TRAJECTORY_SEED,
                       log: Optional[Callable[[str], None]] = None, capacity: int = 0) -> float:
    """Fills the surrogate's training set with `samples` pairs (T_0, T_k), one per seeded
    trajectory: trajectory i starts from its own heterogeneous field, runs 16 j stencil steps for
    a seeded j in [0, chunks), and records the pair across the next 16 steps.
    Every step is the FP64 GPU stencil with the audit on.  The set is sized for `capacity` pairs
    (at least `samples`), so more can be added later.  Returns the wall seconds taken."""
    t0 = time.perf_counter()
    surr.reserve(max(samples, capacity), edge)
    for i in range(samples):
        block, k, dt = heterogeneous_pair_block(edge, i)
        j = int(Rng.for_index(seed, i).unit() * chunks)
        if j:
            block.advance(dt, j * K_STEPS, tolerance=1e-9, want_totals=False)
        after = Solid3D("after", Substance.of("heterogeneous", 4000.0, 1.0, 1000.0), (edge,) * 3,
                        block.spacing(), 300.0)
        after.with_conductivity(k)
        after.copy_cells_from(block)
        after.advance(dt, K_STEPS, tolerance=1e-9, want_totals=False)
        surr.set_sample_from(i, block, after, dt)
        if log and (i + 1) % max(1, samples // 8) == 0:
            log(f"training set: {i + 1}/{samples} pairs ({time.perf_counter() - t0:.1f} s)")
    return time.perf_counter() - t0


def add_rollout_pairs(surr: Surrogate, first: int, count: int, edge: int = 64, chunks: int = 64,
                      unroll: int = 3, seed: int = PUSHFORWARD_SEED,
                      log: Optional[Callable[[str], None]] = None) -> float:
    """Writes pairs `first .. first + count` of the training set from the network's own
    rollouts: trajectory i (seeded like the stencil pairs, offset by `first`) runs 16 j stencil
    steps for a seeded j, then 1 .. `unroll` surrogate steps of the current network, and the pair
    is that state against the FP64 stencil 16 steps from it.  Training on these teaches the
    network the states it produces itself (its residual noise included), which the stencil
    pairs never show it.  Returns the wall seconds taken."""
    from .workloads import CP, RHO
    t0 = time.perf_counter()
    for n in range(count):
        i = first + n
        block, k, dt = heterogeneous_pair_block(edge, i)
        r = Rng.for_index(seed, i)
        j = int(r.unit() * chunks)
        m = 1 + int(r.unit() * unroll)
        if j:
            block.advance(dt, j * K_STEPS, tolerance=1e-9, want_totals=False)
        for _ in range(m):
            surr.step(block, dt, 300.0)
        after = Solid3D("after", Substance.of("heterogeneous", RHO, 1.0, CP), (edge,) * 3,
                        block.spacing(), 300.0)
        after.with_conductivity(k)
        after.copy_cells_from(block)
        after.advance(dt, K_STEPS, tolerance=1e-9, want_totals=False)
        surr.set_sample_from(i, block, after, dt)
        if log and (n + 1) % max(1, count // 8) == 0:
            log(f"rollout pairs: {n + 1}/{count} ({time.perf_counter() - t0:.1f} s)")
    return time.perf_counter() - t0


def train(surr: Surrogate, samples: int, steps: int, batch: int = 32, lr: float = 1e-3,
          seed: int = BATCH_SEED, log: Optional[Callable[[str], None]] = None,
          log_every: int = 50, cosine: bool = True, start: int = 0,
          total: Optional[int] = None) -> List[float]:
    """`steps` Adam steps on seeded random batches of the training set, with the learning rate
    decayed along a half cosine from `lr` towards zero over `total` steps (the run's own steps
    unless given; `start` offsets a run that continues another).  Returns the losses."""
    total = steps if total is None else total
    losses = []
    for s in range(start, start + steps):
        r = Rng.for_index(seed, s)
        idx = [int(r.unit() * samples) for _ in range(batch)]
        rate = lr * 0.5 * (1.0 + math.cos(math.pi * min(s, total) / total)) if cosine else lr
        losses.append(surr.train_step(idx, max(rate, lr * 1e-3)))
        if log and ((s + 1) % log_every == 0 or s == start):
            log(f"step {s + 1}/{start + steps}: loss {losses[-1]:.4e}")
    return losses


# ---- the audit-vetoed rollout -------------------------------------------------------------------

@dataclass
class RolloutReport:
    steps: int                     # stencil steps advanced
    k: int                         # stencil steps per surrogate step
    chunks: int                    # surrogate steps attempted
    checked: int                   # chunks checked against the stencil
    fallbacks: int                 # chunks vetoed and replaced by the stencil's result
    max_error: float               # worst relative L2 error seen on a checked chunk
    max_drift: float               # worst audit drift of a surrogate step
    final_error: float             # relative L2 error against a plain stencil run:
                                   # ||T - T_h|| / ||T_h - T_ref|| (of the heat above T_ref)
    rollout_s: float               # wall seconds of the rollout
    stencil_s: float               # wall seconds of the plain stencil over the same steps
    surrogate_ms: float            # mean device ms of one surrogate step
    errors: List[float] = field(default_factory=list)
    final_error_kelvin: float = 0.0  # ||T - T_h|| / ||T_h||, of the temperatures themselves
    rises: int = 0                   # chunks vetoed because their one-step change rose
    changes: List[Tuple[float, float]] = field(default_factory=list)
                                   # per chunk: one stencil step's change ||S T - T|| at its
                                   # start and after the surrogate step (when measured)

    @property
    def speedup(self) -> float:
        return self.stencil_s / self.rollout_s if self.rollout_s > 0 else float("nan")


def rollout(surr: Surrogate, k_field: np.ndarray, t0: np.ndarray, dt: float, steps: int = 1024,
            tolerance: float = 1e-3, check_every: int = 1, reference: float = 300.0,
            residual: bool = True) -> RolloutReport:
    """Advances a heterogeneous block `steps` stencil steps by surrogate steps of 16 each.  Every
    `check_every`-th chunk (all of them with 1) also runs the FP64 stencil from the same state;
    a checked chunk whose error exceeds `tolerance`, or any chunk whose audit drift does, is
    replaced by the stencil's result.  With `residual`, every chunk not checked is also vetoed
    when one FP64 stencil step moves the surrogate's result further than one step moved the
    chunk's start (||S T_s - T_s|| > ||S T_0 - T_0||).  The block is insulated with a uniform
    capacity, so a step is T <- S T with S = I + dt A symmetric, A negative semi-definite and
    every eigenvalue of S in [-1, 1] at a stable dt; S - I commutes with S, so along the exact
    trajectory that change never grows.  Noise the network leaves behind, which the stencil
    would have damped, makes it grow.  Returns the report, with the same steps of the plain
    stencil timed for comparison."""
    from .workloads import CP, DX, RHO
    k = K_STEPS
    edge = t0.shape[0]
    sub = Substance.of("heterogeneous", RHO, 1.0, CP)

    def block():
        b = Solid3D("rollout", sub, (edge,) * 3, DX, reference)
        b.with_conductivity(k_field)
        b.set_cells(t0)
        return b

    plain, state, judge, probe = block(), block(), block(), block()

    def change() -> float:
        """||S T - T|| of the rollout's present state: one stencil step on the probe."""
        probe.copy_cells_from(state)
        probe.advance(dt, 1, tolerance=1e-9, want_totals=False)
        return probe.l2_parts(state, reference)[0]

    # Warm both paths once (kernel loading, buffer sizing), then restore the start.
    surr.step(judge, dt, reference)
    judge.set_cells(t0)
    judge.advance(dt, 1, tolerance=1e-9, want_totals=False)
    judge.set_cells(t0)

    tp = time.perf_counter()
    plain.advance(dt, steps, tolerance=1e-9, want_totals=False)
    stencil_s = time.perf_counter() - tp

    chunks = steps // k
    checked = fallbacks = rises = 0
    max_err = max_drift = 0.0
    errors: List[float] = []
    changes: List[Tuple[float, float]] = []
    sms = 0.0
    start_change = None  # ||S T - T|| of the state the next chunk starts from, when known
    tr = time.perf_counter()
    for c in range(chunks):
        check_now = check_every > 0 and c % check_every == 0
        if check_now:
            judge.copy_cells_from(state)
            judge.advance(dt, k, tolerance=1e-9, want_totals=False)
        else:
            state.checkpoint()
            if residual and start_change is None:
                start_change = change()
        drift, ms = surr.step(state, dt, reference)
        sms += ms
        max_drift = max(max_drift, drift)
        veto = drift > tolerance
        if check_now:
            checked += 1
            err = state.relative_l2(judge, reference)
            errors.append(err)
            max_err = max(max_err, err)
            veto = veto or err > tolerance
            start_change = None
        elif residual:
            after = change()
            changes.append((start_change, after))
            if after > start_change:
                rises += 1
                veto = True
            start_change = after
        if veto:
            fallbacks += 1
            start_change = None
            if check_now:
                state.copy_cells_from(judge)
            else:
                state.restore()
                state.advance(dt, k, tolerance=1e-9, want_totals=False)
    rest = steps - chunks * k
    if rest:
        state.advance(dt, rest, tolerance=1e-9, want_totals=False)
    rollout_s = time.perf_counter() - tr
    final = state.relative_l2(plain, reference)
    return RolloutReport(steps, k, chunks, checked, fallbacks, max_err, max_drift, final,
                         rollout_s, stencil_s, sms / max(chunks, 1), errors,
                         state.relative_l2(plain, 0.0), rises, changes)
