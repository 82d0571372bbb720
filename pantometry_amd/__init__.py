"""pantometry_amd — the MI355X (gfx950) backend of pantometry's explicit 3D grid-stencil path.

GPU domains (`Solid3D`, `Hall`, `Ensemble`) that drop into pantometry's Domain API, stepped by
hand-written HIP kernels through the C ABI of ``include/pantometry_amd.h``, and the kernel-side
types they are driven through (`Ledger`, `audit_with`, `Exchange`, `Simulation`).
"""
from ._capi import DeviceError, LIB_PATH, lib
from .acoustic import End, Hall, Room, Tube
from .core import (ENERGY, HEAT, Exchange, Kind, Ledger, Report, Schedule, Simulation,
                   Tolerances, Violation, audit, audit_with, substeps_for)
from .ensemble import Ensemble, Member
from .rng import Rng
from .thermal import STABLE_FOURIER_3D, Bar1D, Environment, Face, Fusion, GapPatch, Solid3D, Substance, celsius, mm

__all__ = [
    "DeviceError", "LIB_PATH", "lib", "Hall", "Room", "Tube", "End", "Bar1D", "ENERGY", "HEAT", "Exchange", "Kind", "Ledger",
    "Report", "Schedule", "Simulation", "Tolerances", "Violation", "audit", "audit_with",
    "substeps_for", "Ensemble", "Member", "Rng", "STABLE_FOURIER_3D", "Environment", "Face",
    "Solid3D", "Substance", "celsius", "mm",
]
