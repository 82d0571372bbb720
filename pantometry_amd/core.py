"""The kernel-side types the stencil path is called through, mirrored from pantometry-core.

These are the caller's side of the drop-in boundary: the books (`Ledger`, `audit_with`,
`Violation`, `Tolerances`), the bus (`Exchange`) and the scheduler (`Simulation`).  In the
reference they are Rust (crates/pantometry-core/src/conserved.rs and sim.rs); a Rust caller keeps
its own and binds only the domains.  They are restated here so the GPU domains can be driven,
and tested, exactly as the reference drives its CPU ones.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from enum import Enum
from typing import Dict, Iterable, List, Optional, Tuple

ENERGY = "energy"  # quantity::ENERGY, conserved.rs:58
HEAT = ENERGY      # the thermal crate's channel, pantometry-thermal/src/lib.rs:74


class Violation(Exception):
    """A conservation law that did not hold, or a refused step (conserved.rs:128-215)."""

    def __init__(self, quantity: str, site: str, before: float, after: float, scale: float,
                 tolerance: float, step: int = -1):
        self.quantity = quantity
        self.site = site
        self.before = before
        self.after = after
        self.scale = scale
        self.tolerance = tolerance
        self.step = step
        super().__init__(str(self))

    @staticmethod
    def at(site: str, quantity: str, detail: float) -> "Violation":
        """Violation::at, conserved.rs:148-157."""
        return Violation(quantity, site, detail, detail, abs(detail), 0.0)

    def error(self) -> float:
        return abs(self.after - self.before)

    def relative_error(self) -> float:
        scale = self.scale if self.scale > 0.0 else max(abs(self.before), abs(self.after))
        return 0.0 if scale == 0.0 else self.error() / scale

    def __str__(self) -> str:  # Display, conserved.rs:180-215
        if self.tolerance == 0.0 and self.before == self.after:
            return f"at {self.site}: {self.quantity} ({self.before})"
        if self.before == self.after:
            return f"{self.quantity} is not conserved at {self.site}: {self.before}"
        verb = "created" if self.after > self.before else "destroyed"
        return (f"{self.quantity} {verb} at {self.site}: {self.before:.6e} became {self.after:.6e}, "
                f"a relative change of {self.relative_error():.3e} against a tolerance of "
                f"{self.tolerance:.3e}")


@dataclass
class _Entry:
    total: float = 0.0
    scale: float = 0.0


class Ledger:
    """What a process claims to hold, by quantity name (conserved.rs:45-123)."""

    def __init__(self):
        self._e: Dict[str, _Entry] = {}

    def with_(self, quantity: str, si_total: float) -> "Ledger":
        self.add(quantity, si_total)
        return self

    def add(self, quantity: str, si_total: float) -> None:
        e = self._e.setdefault(quantity, _Entry())
        e.total += si_total
        e.scale = _rmax(e.scale, abs(si_total))

    def get(self, quantity: str) -> Optional[float]:
        e = self._e.get(quantity)
        return None if e is None else e.total

    def scale_of(self, quantity: str) -> Optional[float]:
        e = self._e.get(quantity)
        return None if e is None else e.scale

    def is_empty(self) -> bool:
        return not self._e

    def quantities(self) -> List[Tuple[str, float]]:
        return [(k, self._e[k].total) for k in sorted(self._e)]

    def merged(self, other: "Ledger") -> "Ledger":
        out = Ledger()
        for k, e in self._e.items():
            out._e[k] = _Entry(e.total, e.scale)
        for k in sorted(other._e):
            e = other._e[k]
            mine = out._e.setdefault(k, _Entry())
            mine.total += e.total
            mine.scale = _rmax(mine.scale, e.scale)
        return out


def _rmax(a: float, b: float) -> float:
    """Rust's f64::max: the non-NaN operand when one is NaN."""
    if math.isnan(a):
        return b
    if math.isnan(b):
        return a
    return a if a >= b else b


class Tolerances:
    """Per-quantity relative tolerances, default 1e-9 (conserved.rs:254-308)."""

    def __init__(self, default: float = 1e-9, per_quantity: Optional[Dict[str, float]] = None):
        self.default = default
        self.per_quantity = dict(per_quantity or {})

    @staticmethod
    def uniform(tol: float) -> "Tolerances":
        return Tolerances(tol)

    def with_(self, quantity: str, tol: float) -> "Tolerances":
        t = Tolerances(self.default, self.per_quantity)
        t.per_quantity[quantity] = tol
        return t

    def for_quantity(self, quantity: str) -> float:
        return self.per_quantity.get(quantity, self.default)


def audit_with(site: str, before: Ledger, after: Ledger, tolerances: Tolerances) -> None:
    """audit_with, conserved.rs:323-363: raise on the first quantity (in name order) that moved."""
    names = sorted(set(before._e) | set(after._e))
    for name in names:
        b = before.get(name) or 0.0
        a = after.get(name) or 0.0
        scale = _rmax(_rmax(_rmax(abs(b), abs(a)), before.scale_of(name) or 0.0),
                      after.scale_of(name) or 0.0)
        if scale < 1e-300:
            continue
        tol = tolerances.for_quantity(name)
        if abs(a - b) / scale > tol:
            raise Violation(name, site, b, a, scale, tol)


def audit(site: str, before: Ledger, after: Ledger, rel_tol: float) -> None:
    audit_with(site, before, after, Tolerances.uniform(rel_tol))


class Exchange:
    """The bus of named amounts between domains (sim.rs:314-674), plain channels only."""

    def __init__(self):
        self.published: Dict[str, float] = {}
        self.consumed: Dict[str, float] = {}
        self.interval = 0.0
        self.unclaimed_time: Dict[str, float] = {}
        self.takers: Dict[str, int] = {}
        self.published_total: Dict[str, float] = {}
        self.taken_since_mark: Dict[str, float] = {}
        self.published_since_mark: Dict[str, float] = {}

    def publish(self, channel: str, amount: float) -> None:
        self.published[channel] = self.published.get(channel, 0.0) + amount
        self.published_total[channel] = self.published_total.get(channel, 0.0) + amount
        self.published_since_mark[channel] = self.published_since_mark.get(channel, 0.0) + amount

    def take(self, channel: str) -> float:
        amount = self.published.get(channel, 0.0)
        self.published[channel] = 0.0
        self.consumed[channel] = self.consumed.get(channel, 0.0) + amount
        self.taken_since_mark[channel] = self.taken_since_mark.get(channel, 0.0) + amount
        self.takers[channel] = self.takers.get(channel, 0) + 1
        return amount

    def peek(self, channel: str) -> float:
        return self.published.get(channel, 0.0)

    def take_share(self, channel: str, dt: float) -> float:
        """Exchange::take_share, sim.rs:413-434."""
        h = dt
        if self.interval <= 0.0 or not math.isfinite(h) or h <= 0.0:
            return self.take(channel)
        left = self.unclaimed_time.setdefault(channel, self.interval)
        if h >= left or left - h <= self.interval * 1e-12:
            self.unclaimed_time[channel] = 0.0
            return self.take(channel)
        amount = self.published.get(channel, 0.0)
        share = amount * h / left
        self.unclaimed_time[channel] = left - h
        self.published[channel] = self.published.get(channel, 0.0) - share
        self.consumed[channel] = self.consumed.get(channel, 0.0) + share
        self.taken_since_mark[channel] = self.taken_since_mark.get(channel, 0.0) + share
        return share

    def covering(self, dt: float) -> None:
        self.interval = max(dt, 0.0)
        self.unclaimed_time.clear()
        self.takers.clear()
        self.mark()

    def mark(self) -> None:
        self.taken_since_mark.clear()
        self.published_since_mark.clear()

    def clear_offers(self) -> None:
        self.published.clear()
        self.unclaimed_time.clear()
        self.takers.clear()
        self.mark()

    def traffic(self) -> List[Tuple[str, float, float]]:
        names = sorted(set(self.published_total) | set(self.consumed))
        return [(n, self.published_total.get(n, 0.0), self.consumed.get(n, 0.0)) for n in names]

    def audit_transfers(self, site: str, abs_tol: float) -> None:
        """sim.rs:553-583 (plain channels)."""
        for channel in sorted(self.published):
            left = self.published[channel]
            if abs(left) > abs_tol:
                raise Violation(channel, f"{site} (published but not consumed)", left, 0.0,
                                abs(left), abs_tol)


def substeps_for(dt: float, limit: float) -> int:
    """integrator.rs:186-199."""
    if not math.isfinite(limit) or limit <= 0.0 or dt <= 0.0:
        return 1
    n = math.ceil(dt / limit)
    if not math.isfinite(n) or n < 1.0:
        return 1
    return int(min(n, 4294967295))


class Schedule(Enum):
    ONE_WAY = "one_way"
    STAGGERED = "staggered"
    MULTIRATE = "multirate"
    ITERATIVE = "iterative"


class Kind(Enum):
    EVOLVING = "evolving"
    QUASI_STATIC = "quasi_static"


@dataclass
class Report:
    substeps: List[Tuple[str, int]] = field(default_factory=list)
    iterations: int = 1
    residual: float = 0.0


def _attribute(name: str, before: Ledger, after: Ledger, tb, ta, tolerances: Tolerances) -> None:
    """The per-domain books check of sim.rs:1169-1228."""
    def find(t, channel):
        for c, p, k in t:
            if c == channel:
                return p, k
        return 0.0, 0.0

    def moved(channel):
        pb, kb = find(tb, channel)
        pa, ka = find(ta, channel)
        return (ka - kb) - (pa - pb)

    names = sorted(set(n for n, _ in before.quantities()) | set(n for n, _ in after.quantities()))
    for q in names:
        hb = before.get(q) or 0.0
        ha = after.get(q) or 0.0
        expected = moved(q)
        discrepancy = (ha - hb) - expected
        scale = _rmax(_rmax(_rmax(_rmax(abs(hb), abs(ha)), before.scale_of(q) or 0.0),
                            after.scale_of(q) or 0.0), abs(expected))
        if scale < 1e-300:
            continue
        tol = tolerances.for_quantity(q)
        if abs(discrepancy) / scale > tol:
            raise Violation(q, f"{name} (its own books, against what it moved on the bus)",
                            hb + expected, ha, scale, tol)


class Simulation:
    """Simulation::advance / sweep / iterate (sim.rs:785-1153), for domains that implement the
    Python Domain protocol (name, kind, max_stable_dt, step, ledger, books_balance, checkpoint,
    restore, supports_restore, residual)."""

    def __init__(self, schedule: Schedule = Schedule.STAGGERED, max_iter: int = 8,
                 tol: float = 1e-9):
        self.schedule = schedule
        self.max_iter = max_iter
        self.iter_tol = tol
        self.domains: list = []
        self.bus = Exchange()
        self.t = 0.0
        self.transfer_tol = 1e-12
        self.conservation_tol = Tolerances()

    def with_(self, domain) -> "Simulation":
        self.domains.append(domain)
        return self

    def conservation_tolerance(self, tol: float) -> "Simulation":
        self.conservation_tol = Tolerances(tol, self.conservation_tol.per_quantity)
        return self

    def domain(self, name: str):
        for d in self.domains:
            if d.name() == name:
                return d
        return None

    def ledger(self) -> Ledger:
        total = Ledger()
        for d in self.domains:
            total = total.merged(d.ledger())
        return total

    def advance(self, dt: float) -> Report:
        before = self.ledger()
        self.bus.covering(dt)
        if self.schedule in (Schedule.ONE_WAY, Schedule.STAGGERED):
            report = self._sweep(dt, False)
        elif self.schedule == Schedule.MULTIRATE:
            report = self._sweep(dt, True)
        else:
            report = self._iterate(dt)
        self.bus.audit_transfers("bus", self.transfer_tol)
        after = self.ledger()
        if not before.is_empty() or not after.is_empty():
            audit_with("simulation", before, after, self.conservation_tol)
        self.t += dt
        return report

    def _sweep(self, dt: float, multirate: bool) -> Report:
        """Simulation::sweep, sim.rs:967-1105: each domain in declared order, its substeps, the
        attribution of its books, and the refusal of a domain that took nothing from a channel
        an earlier domain of this sweep had already emptied (sim.rs:1012-1098)."""
        now = self.t
        substeps = []
        # What earlier takers of this sweep moved on each channel, as a magnitude.
        moved: Dict[str, float] = {}
        for d in self.domains:
            kind = d.kind() if hasattr(d, "kind") else Kind.EVOLVING
            n = substeps_for(dt, d.max_stable_dt(now)) if (multirate and kind == Kind.EVOLVING) else 1
            h = dt / n
            t = now
            takes_before = dict(self.bus.takers)
            audited = d.books_balance() if hasattr(d, "books_balance") else False
            books_before = d.ledger() if audited else None
            traffic_before = self.bus.traffic() if audited else None
            self.bus.mark()
            for _ in range(n):
                d.step(t, h, self.bus)
                t += h
            if audited:
                _attribute(d.name(), books_before, d.ledger(), traffic_before, self.bus.traffic(),
                           self.conservation_tol)
            # Net amounts this turn (taken - published, summed from zero since the mark), for
            # each channel this domain took from, in channel-name order.
            for channel in sorted(self.bus.takers):
                if self.bus.takers[channel] <= takes_before.get(channel, 0):
                    continue
                net = (self.bus.taken_since_mark.get(channel, 0.0)
                       - self.bus.published_since_mark.get(channel, 0.0))
                earlier = moved.get(channel, 0.0)
                if earlier > 0.0 and net == 0.0:
                    raise Violation(channel, f"{d.name()} (a second domain took from a channel "
                                    "already emptied)", earlier, net, earlier, 0.0)
                moved[channel] = moved.get(channel, 0.0) + abs(net)
            substeps.append((d.name(), n))
        residual = 0.0
        for d in self.domains:
            r = d.residual() if hasattr(d, "residual") else 0.0
            residual = _rmax(residual, r)
        return Report(substeps, 1, residual)

    def _iterate(self, dt: float) -> Report:
        for d in self.domains:
            if not (hasattr(d, "supports_restore") and d.supports_restore()):
                raise Violation.at(d.name(), "iterative coupling needs a restorable domain", 0.0)
        for d in self.domains:
            d.checkpoint()
        last = Report()
        for it in range(1, self.max_iter + 1):
            if it > 1:
                for d in self.domains:
                    d.restore()
                self.bus.clear_offers()
            last = self._sweep(dt, True)
            last.iterations = it
            if last.residual <= self.iter_tol:
                return last
        raise Violation("coupling residual", f"simulation (after {self.max_iter} iterations)",
                        0.0, last.residual, abs(last.residual), self.iter_tol)
