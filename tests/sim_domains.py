"""Host-side domains for driving pantometry_amd.core.Simulation: the oracle's Solid3D and Hall
behind the Domain protocol (so a GPU-driven run can be compared with an oracle-driven one), and
small parts that force the Simulation's own paths: a coupling residual that settles or never does
(Schedule::Iterative, sim.rs:1114-1153) and a consumer that loses what it takes (the per-domain
books check, sim.rs:1169-1228)."""
from __future__ import annotations

import math

from pantometry_amd.core import ENERGY, HEAT, Kind, Ledger, Violation


class OracleSolidDomain:
    """Solid3D::step (solid.rs:1717-1901) on the oracle: the limit refused before the bus is
    touched, then the share of the heat channel, then the sweep."""

    def __init__(self, b, name="block"):
        self.b = b
        self._name = name
        self.absorbed = 0.0

    def name(self):
        return self._name

    def kind(self):
        return Kind.EVOLVING

    def max_stable_dt(self, now=0.0):
        return self.b.max_stable_dt()

    def books_balance(self):
        return True

    def supports_restore(self):
        return True

    def residual(self):
        return 0.0

    def step(self, t, dt, bus):
        ratio = dt * self.b.worst_rate_now()
        if math.isnan(ratio):
            raise Violation.at(self._name, "substance has no diffusivity", math.inf)
        if ratio > 1.0 + 6e-12:
            f = 1.0 / 6.0
            raise Violation("Fourier number", f"{self._name} (explicit 3D conduction)", f,
                            ratio * f, f, 1e-12)
        gained = bus.take_share(HEAT, dt)
        self.absorbed += gained
        self.b.step(dt, gained)

    def ledger(self):
        return (Ledger().with_(ENERGY, self.b.stored_heat()).with_(ENERGY, self.b.lost)
                .with_(ENERGY, -self.b.supplied))

    def checkpoint(self):
        # The cells and the books, not `absorbed`: a reading the reference does not rewind
        # (solid.rs:1935-1950).
        self.b.checkpoint()

    def restore(self):
        self.b.restore()


class OracleHallDomain:
    """Hall::step (hall.rs:408-530) on the oracle: the Courant guard, the leapfrog, and on the
    first step the startup check, with the Violation the reference builds for each."""

    def __init__(self, h, name="hall"):
        self.h = h
        self._name = name
        self._saved = None

    def name(self):
        return self._name

    def kind(self):
        return Kind.EVOLVING

    def max_stable_dt(self, now=0.0):
        return self.h.max_stable_dt()

    def books_balance(self):
        return True

    def supports_restore(self):
        return True

    def residual(self):
        return 0.0

    def step(self, t, dt, bus):
        c = self.h.courant(dt)
        if c > 1.0 + 1e-12:
            raise Violation("Courant number", f"{self._name} (three-dimensional wave equation)",
                            1.0, c, 1.0, 1e-12)
        try:
            self.h.step(dt)
        except ValueError:
            inv = self.h.invariant() - self.h.supplied
            scale = max(abs(self.h.datum), abs(inv))
            raise Violation("energy", f"{self._name} (starting the leapfrog)", self.h.datum, inv,
                            scale, 0.25)

    def ledger(self):
        led = Ledger().with_(ENERGY, self.h.energy())
        if self.h.supplied != 0.0:
            led = led.with_(ENERGY, -self.h.supplied)
        return led

    def checkpoint(self):
        """Hall::checkpoint, hall.rs:536-544: pressure, velocities and the staggered flag (with
        the source's cursor and books, which the reference's hall has no counterpart of)."""
        h = self.h
        self._saved = (h.p.copy(), h.vx.copy(), h.vy.copy(), h.vz.copy(), h.staggered,
                       h.supplied, h.cursor)

    def restore(self):
        """Hall::restore, hall.rs:546-555: pressure_prev becomes the saved pressure."""
        h = self.h
        p, vx, vy, vz, h.staggered, h.supplied, h.cursor = self._saved
        h.p, h.pp, h.vx, h.vy, h.vz = p.copy(), p.copy(), vx.copy(), vy.copy(), vz.copy()


class Coupler:
    """A quasi-static part whose coupling residual after the n-th sweep since the checkpoint is
    residuals[n-1] (the last one repeating), and which publishes nothing.  A restore does not
    reset the count: the n-th iteration of one advance sees residuals[n-1]."""

    def __init__(self, residuals, name="coupler"):
        self.residuals = list(residuals)
        self._name = name
        self.sweeps = 0
        self.checkpoints = 0
        self.restores = 0

    def name(self):
        return self._name

    def kind(self):
        return Kind.QUASI_STATIC

    def max_stable_dt(self, now=0.0):
        return math.inf

    def supports_restore(self):
        return True

    def books_balance(self):
        return False

    def step(self, t, dt, bus):
        self.sweeps += 1

    def residual(self):
        return self.residuals[min(self.sweeps, len(self.residuals)) - 1] if self.sweeps else 0.0

    def ledger(self):
        return Ledger()

    def checkpoint(self):
        self.checkpoints += 1
        self.sweeps = 0

    def restore(self):
        self.restores += 1


class Lossy:
    """A wrapper that takes a domain's share of the heat channel but hands the domain nothing,
    while claiming the domain's books: the per-domain check must refuse it."""

    def __init__(self, inner):
        self.inner = inner

    def name(self):
        return self.inner.name()

    def kind(self):
        return Kind.EVOLVING

    def max_stable_dt(self, now=0.0):
        return self.inner.max_stable_dt(now)

    def books_balance(self):
        return True

    def supports_restore(self):
        return False

    def residual(self):
        return 0.0

    def step(self, t, dt, bus):
        from pantometry_amd.core import Exchange
        bus.take_share(HEAT, dt)
        self.inner.step(t, dt, Exchange())

    def ledger(self):
        return self.inner.ledger()
