"""The seeded workloads the stencil path is measured and checked on.

Every random draw comes from Rng::for_index(seed, index) and every transcendental from the host's
libm, so the same arrays can be handed to the GPU and to any CPU restatement.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import List, Tuple

import numpy as np

from ._capi import check, dptr, lib
from .acoustic import Hall
from .ensemble import Ensemble, Member
from .rng import Rng
from .thermal import Solid3D, Substance

UNIFORM_NOISE_SEED = 0x5EED0000
HETERO_K_SEED = 0x5EED0001
HETERO_SPOT_SEED = 0x5EED0002
HALL_SOURCE_SEED = 0x5EED0003
ENSEMBLE_SEED = 0x5EED0004

DX = 1e-3
RHO = 4000.0        # kg/m^3 (rho*c = 4e6 J/m^3/K)
CP = 1000.0         # J/kg/K


def hot_spot_cells(counts: Tuple[int, int, int], amp: float = 100.0, sigma: float = 6.0,
                   noise: float = 0.1, seed: int = UNIFORM_NOISE_SEED,
                   base: float = 300.0) -> np.ndarray:
    """base + amp*exp(-r^2/(2 sigma^2)) + range(-noise, noise), r in cells from the centre."""
    nx, ny, nz = counts
    out = np.empty(nx * ny * nz)
    check(lib().pm_workload_hot_spot(nx, ny, nz, base, amp, sigma, noise, seed, dptr(out)),
          "pm_workload_hot_spot")
    return out.reshape(nz, ny, nx)


def heterogeneous_fields(counts: Tuple[int, int, int], spots: int = 8,
                         k_seed: int = HETERO_K_SEED, spot_seed: int = HETERO_SPOT_SEED):
    """(conductivity, temperature): k lognormal(0, 0.5) W/m/K, T = 300 K + `spots` hot spots of
    amplitude 50-200 K and width 4-16 cells."""
    nx, ny, nz = counts
    k = np.empty(nx * ny * nz)
    t = np.empty(nx * ny * nz)
    check(lib().pm_workload_heterogeneous(nx, ny, nz, k_seed, spot_seed, spots, dptr(k), dptr(t)),
          "pm_workload_heterogeneous")
    return k.reshape(nz, ny, nx), t.reshape(nz, ny, nx)


def uniform_block(n: int = 64, alpha: float = 1e-4) -> Tuple[Solid3D, float]:
    """The uniform block: n^3 cells, alpha = k/(rho c) (k = alpha*4e6), dx = 1 mm, hot spot
    plus noise, stepped at 0.9 of its limit.  Returns (block, dt)."""
    sub = Substance.of("uniform", RHO, alpha * RHO * CP, CP)
    block = Solid3D("block", sub, (n, n, n), DX, 300.0)
    block.set_cells(hot_spot_cells((n, n, n)))
    return block, 0.9 * block.max_stable_dt(0.0)


def heterogeneous_block(n: int = 512, spots: int = 8) -> Tuple[Solid3D, float]:
    """The heterogeneous block: n^3 cells, per-cell k lognormal(0, 0.5), rho c = 4e6, dx = 1 mm,
    eight hot spots, stepped at 0.9 of its limit.  Returns (block, dt)."""
    k, t = heterogeneous_fields((n, n, n), spots)
    sub = Substance.of("heterogeneous", RHO, 1.0, CP)
    block = Solid3D("block", sub, (n, n, n), DX, 300.0)
    block.with_conductivity(k)
    block.set_cells(t)
    return block, 0.9 * block.max_stable_dt(0.0)


def ricker(t: float, f: float, t0: float, amplitude: float = 1.0) -> float:
    """amplitude * (1 - 2 pi^2 f^2 tau^2) exp(-pi^2 f^2 tau^2), tau = t - t0."""
    tau = t - t0
    a = math.pi * math.pi * f * f * tau * tau
    return amplitude * (1.0 - 2.0 * a) * math.exp(-a)


@dataclass
class HallCase:
    hall: Hall
    dt: float
    source_node: int
    increments: np.ndarray


def acoustic_hall(cells: int = 384, steps: int = 4000, cfl: float = 0.5,
                  amplitude: float = 1.0) -> HallCase:
    """The rigid hall: cells^3 cells (cells+1 nodes a side), air-like rho = 1.2, c = 343, dx =
    1 mm, CFL `cfl`, and a Ricker point source of peak frequency c/(20 dx) at a seeded interior
    node."""
    width = cells * DX
    hall = Hall("hall", width, width, width, cells + 1, 343.0, 1.2)
    dt = cfl * hall.max_stable_dt(0.0)
    nx, ny, nz = hall.nodes()
    r = Rng.for_index(HALL_SOURCE_SEED, 0)
    i = 1 + int(r.unit() * (nx - 2))
    j = 1 + int(r.unit() * (ny - 2))
    k = 1 + int(r.unit() * (nz - 2))
    f = 343.0 / (20.0 * hall.spacing())
    t0 = 1.0 / f
    inc = np.array([ricker((n + 1) * dt, f, t0, amplitude) for n in range(steps)])
    hall.released_from(np.zeros((nz, ny, nx)))
    hall.set_source(hall.node_index(i, j, k), inc)
    return HallCase(hall, dt, hall.node_index(i, j, k), inc)


def ensemble_members(count: int = 512, lo_edge: float = 16.0, hi_edge: float = 96.0,
                     lo_alpha: float = 1e-5, hi_alpha: float = 1e-3,
                     seed: int = ENSEMBLE_SEED) -> List[Tuple[int, float]]:
    """(edge, alpha) per member: edge = round(exp(U[ln lo, ln hi])), alpha = exp(U[...])."""
    out = []
    for i in range(count):
        r = Rng.for_index(seed, i)
        edge = int(round(math.exp(r.range(math.log(lo_edge), math.log(hi_edge)))))
        alpha = math.exp(r.range(math.log(lo_alpha), math.log(hi_alpha)))
        out.append((edge, alpha))
    return out


def uniform_limit(counts: Tuple[int, int, int], sub: Substance, dx: float = DX) -> float:
    """max_stable_dt of a uniform block (solid.rs:757-782, 1709-1715): 1/max(mob*face_sum) over
    one cell of each boundary class, faces summed x-, x+, y-, y+, z-, z+."""
    k = sub.conductivity
    kface = 2.0 * k * k / (k + k) if k + k > 0.0 else 0.0
    cap = (sub.density * (dx * dx * dx)) * sub.specific_heat
    mob = dx / cap

    def reps(n: int) -> List[int]:
        return sorted({0, min(1, n - 1), n - 1})

    nx, ny, nz = counts
    worst = 0.0
    for kk in reps(nz):
        for jj in reps(ny):
            for ii in reps(nx):
                f = [kface if ii > 0 else 0.0, kface if ii + 1 < nx else 0.0,
                     kface if jj > 0 else 0.0, kface if jj + 1 < ny else 0.0,
                     kface if kk > 0 else 0.0, kface if kk + 1 < nz else 0.0]
                total = 0.0
                for v in f:
                    total = total + v
                worst = max(worst, mob * total)
    return 1.0 / worst


def member_limit(edge: int, alpha: float) -> Tuple[Substance, float]:
    """The member's substance (k = alpha * rho c) and its max_stable_dt."""
    sub = Substance.of("member", RHO, alpha * RHO * CP, CP)
    return sub, uniform_limit((edge, edge, edge), sub)


def heat_ensemble(count: int = 512, **kw) -> Tuple[Ensemble, List[Tuple[int, float]]]:
    """The ensemble: `count` uniform blocks with log-uniform edges and diffusivities, each with
    its own seeded hot spot and noise and its own 0.9-of-limit step."""
    shapes = ensemble_members(count, **kw)
    members = []
    for edge, alpha in shapes:
        sub, limit = member_limit(edge, alpha)
        members.append(Member((edge, edge, edge), DX, sub, 300.0, 0.9 * limit))
    ens = Ensemble(members)
    for m, (edge, _) in enumerate(shapes):
        ens.set_cells(m, hot_spot_cells((edge, edge, edge), seed=ENSEMBLE_SEED + 1 + m))
    return ens, shapes
