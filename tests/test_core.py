"""The kernel-side types (conserved.rs, sim.rs, integrator.rs) as restated in pantometry_amd.core:
the reference's own unit tests of the books, the bus and the substep count."""
import math

import pytest

from pantometry_amd.core import (ENERGY, Exchange, Ledger, Tolerances, Violation, audit,
                                 audit_with, substeps_for)


def test_zero_against_zero_is_not_a_violation():
    audit("s", Ledger().with_(ENERGY, 0.0), Ledger().with_(ENERGY, 0.0), 1e-9)


def test_an_absent_quantity_counts_as_zero():
    with pytest.raises(Violation) as err:
        audit("s", Ledger(), Ledger().with_("momentum", 1.0), 1e-9)
    assert err.value.quantity == "momentum" and err.value.before == 0.0


def test_names_are_checked_in_a_fixed_order():
    before = Ledger().with_("mass", 1.0).with_("energy", 1.0)
    after = Ledger().with_("mass", 2.0).with_("energy", 2.0)
    with pytest.raises(Violation) as err:
        audit("s", before, after, 1e-9)
    assert err.value.quantity == "energy"


def test_the_scale_is_the_largest_entry_not_the_net():
    # A net of zero from two large entries: rounding is judged against the entries.
    before = Ledger().with_(ENERGY, 1e6).with_(ENERGY, -1e6)
    after = Ledger().with_(ENERGY, 1e6).with_(ENERGY, -1e6 + 1e-6)
    audit_with("s", before, after, Tolerances.uniform(1e-9))
    with pytest.raises(Violation):
        audit_with("s", before, after, Tolerances.uniform(1e-13))


def test_per_quantity_tolerances():
    t = Tolerances.uniform(1e-9).with_("momentum", 1e-6)
    assert t.for_quantity("momentum") == 1e-6 and t.for_quantity("energy") == 1e-9


def test_take_share_hands_out_the_whole_channel_exactly():
    bus = Exchange()
    bus.covering(1.0)
    bus.publish(ENERGY, 10.0)
    got = [bus.take_share(ENERGY, 1.0 / 3.0) for _ in range(3)]
    assert sum(got) == pytest.approx(10.0, rel=1e-15)
    assert bus.peek(ENERGY) == 0.0
    bus.audit_transfers("bus", 1e-12)


def test_energy_that_arrives_nowhere_is_a_violation():
    bus = Exchange()
    bus.publish(ENERGY, 1.0)
    with pytest.raises(Violation) as err:
        bus.audit_transfers("bus", 1e-12)
    assert "published but not consumed" in err.value.site


def test_substeps_for():
    assert substeps_for(1.0, math.inf) == 1
    assert substeps_for(1.0, math.nan) == 1
    assert substeps_for(1.0, 0.3) == 4
    assert substeps_for(0.0, 0.3) == 1
    assert substeps_for(1.0, 1.0) == 1


def test_violation_display():
    v = Violation.at("plate", "substance has no heat capacity", math.inf)
    assert str(v) == "at plate: substance has no heat capacity (inf)"
