"""ctypes binding of ``include/pantometry_amd.h``.

Loads ``libpantometry_amd.so`` from this package directory and fails loudly when it is missing:
there is no Python or CPU stand-in for any entry point.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libpantometry_amd.so"

PM_OK, PM_VIOLATION, PM_ERR_DEVICE, PM_ERR_ARGUMENT, PM_ERR_STATE = 0, 1, 2, 3, 4


class pm_violation(C.Structure):
    _fields_ = [("quantity", C.c_char * 64), ("site", C.c_char * 160), ("before", C.c_double),
                ("after", C.c_double), ("scale", C.c_double), ("tolerance", C.c_double),
                ("step", C.c_int64)]


class pm_substance(C.Structure):
    _fields_ = [("density", C.c_double), ("conductivity", C.c_double),
                ("specific_heat", C.c_double), ("emissivity", C.c_double),
                ("melting_point", C.c_double), ("latent_heat", C.c_double),
                ("liquid_conductivity", C.c_double), ("liquid_specific_heat", C.c_double)]


class pm_environment(C.Structure):
    _fields_ = [("ambient", C.c_double), ("convection_w_per_m2_k", C.c_double),
                ("area", C.c_double)]


class pm_books(C.Structure):
    _fields_ = [("stored", C.c_double), ("lost", C.c_double), ("supplied", C.c_double)]


class pm_member(C.Structure):
    _fields_ = [("nx", C.c_size_t), ("ny", C.c_size_t), ("nz", C.c_size_t), ("dx", C.c_double),
                ("substance", pm_substance), ("initial", C.c_double), ("dt", C.c_double)]


_D = C.c_double
_DP = C.POINTER(C.c_double)
_FP = C.POINTER(C.c_float)
_VP = C.c_void_p
_S = C.c_size_t
_U64 = C.c_uint64
_I = C.c_int
_VIOL = C.POINTER(pm_violation)

# name: (restype, argtypes)
SIGNATURES = {
    "pm_last_error": (C.c_char_p, []),
    "pm_device_info": (_I, [C.c_char_p, _S, C.POINTER(_I), C.POINTER(_S)]),
    "pm_stream_copy": (_I, [_S, _I, _DP]),
    "pm_rng_new": (_U64, [_U64]),
    "pm_rng_for_index": (_U64, [_U64, _U64]),
    "pm_rng_unit": (_D, [C.POINTER(_U64)]),
    "pm_rng_range": (_D, [C.POINTER(_U64), _D, _D]),
    "pm_rng_gaussian": (_D, [C.POINTER(_U64)]),
    "pm_rng_split": (_U64, [C.POINTER(_U64)]),
    "pm_workload_hot_spot": (_I, [_S, _S, _S, _D, _D, _D, _D, _U64, _DP]),
    "pm_workload_heterogeneous": (_I, [_S, _S, _S, _U64, _U64, _I, _DP, _DP]),
    "pm_solid3d_create": (_I, [C.c_char_p, C.POINTER(pm_substance), _S, _S, _S, _D, _D,
                               C.POINTER(_VP)]),
    "pm_solid3d_destroy": (None, [_VP]),
    "pm_solid3d_counts": (_I, [_VP, C.POINTER(_S)]),
    "pm_solid3d_fill": (_I, [_VP, C.POINTER(pm_substance), C.POINTER(C.c_uint8)]),
    "pm_solid3d_set_conductivity": (_I, [_VP, _DP]),
    "pm_solid3d_set_cells": (_I, [_VP, _DP]),
    "pm_solid3d_get_cells": (_I, [_VP, _DP]),
    "pm_solid3d_set_temperature": (_I, [_VP, _S, _S, _S, _D]),
    "pm_solid3d_deposit": (_I, [_VP, _S, _S, _S, _D]),
    "pm_solid3d_max_stable_dt": (_D, [_VP]),
    "pm_solid3d_stability_ratio": (_D, [_VP, _D]),
    "pm_solid3d_refuses": (_I, [_VP, _D, _VIOL]),
    "pm_solid3d_step": (_I, [_VP, _D, _D, _VIOL]),
    "pm_solid3d_ledger": (_I, [_VP, C.POINTER(pm_books)]),
    "pm_solid3d_advance": (_I, [_VP, _D, _U64, _D, _DP, _DP, _VIOL]),
    "pm_solid3d_advance_static": (_I, [_VP, _D, _U64, _D, _DP, _DP, _VIOL]),
    "pm_solid3d_checkpoint": (_I, [_VP]),
    "pm_solid3d_losing_from": (_I, [_VP, _I, C.POINTER(pm_environment)]),
    "pm_solid3d_dissipating": (_I, [_VP, _D, C.POINTER(C.c_uint8)]),
    "pm_solid3d_generated_power": (_I, [_VP, _DP]),
    "pm_solid3d_sample": (_I, [_VP, _DP, _S, _DP]),
    "pm_solid3d_empty": (_I, [_VP, C.POINTER(C.c_uint8)]),
    "pm_solid3d_set_melted_fraction": (_I, [_VP, _S, _S, _S, _D]),
    "pm_solid3d_set_region": (_I, [_VP, C.POINTER(C.c_uint8), _D]),
    "pm_solid3d_melted": (_I, [_VP, _DP, _DP]),
    "pm_solid3d_melting_cells": (_S, [_VP]),
    "pm_solid3d_void_cells": (_S, [_VP]),
    "pm_solid3d_is_void": (_I, [_VP, _S, _S, _S]),
    "pm_solid3d_gap_pairs": (_S, [_VP, C.POINTER(C.c_int64), C.POINTER(C.c_int64), _DP, _S]),
    "pm_hall_sample": (_I, [_VP, _DP, _S, _DP]),
    "pm_solid3d_restore": (_I, [_VP]),
    "pm_solid3d_readings": (_I, [_VP, _DP]),
    "pm_solid3d_statistics": (_I, [_VP, _DP]),
    "pm_solid3d_heat_capacity": (_D, [_VP]),
    "pm_solid3d_set_tile": (_I, [_VP, _I, _I, _I]),
    "pm_solid3d_set_profiling": (_I, [_VP, _I]),
    "pm_solid3d_kernel_times": (_I, [_VP, _DP, _DP, C.POINTER(_I)]),
    "pm_solid3d_set_steps_per_pass": (_I, [_VP, _I]),
    "pm_solid3d_steps_per_pass": (_I, [_VP]),
    "pm_hall_create": (_I, [C.c_char_p, _D, _D, _D, _S, _D, _D, C.POINTER(_VP)]),
    "pm_hall_destroy": (None, [_VP]),
    "pm_hall_nodes": (_I, [_VP, C.POINTER(_S)]),
    "pm_hall_spacing": (_D, [_VP]),
    "pm_hall_release": (_I, [_VP, _DP]),
    "pm_hall_set_source": (_I, [_VP, _S, _DP, _U64]),
    "pm_hall_max_stable_dt": (_D, [_VP]),
    "pm_hall_courant": (_D, [_VP, _D]),
    "pm_hall_step": (_I, [_VP, _D, _VIOL]),
    "pm_hall_energy": (_I, [_VP, _DP, _DP, _DP]),
    "pm_hall_advance": (_I, [_VP, _D, _U64, _D, _DP, _DP, _VIOL]),
    "pm_hall_get_fields": (_I, [_VP, _DP, _DP, _DP, _DP, _DP]),
    "pm_hall_checkpoint": (_I, [_VP]),
    "pm_hall_restore": (_I, [_VP]),
    "pm_hall_peak_pressure": (_D, [_VP]),
    "pm_hall_set_tile": (_I, [_VP, _I, _I]),
    "pm_hall_set_tile_width": (_I, [_VP, _I]),
    "pm_hall_set_steps_per_pass": (_I, [_VP, _I]),
    "pm_room_create": (_I, [C.c_char_p, _D, _D, _S, _D, _D, _D, C.POINTER(_VP)]),
    "pm_hall_dimensions": (_I, [_VP]),
    "pm_bar1d_create": (_I, [C.c_char_p, C.POINTER(pm_substance), _S, _D, _D, _D, C.POINTER(_VP)]),
    "pm_bar1d_destroy": (None, [_VP]),
    "pm_bar1d_cells": (_S, [_VP]),
    "pm_bar1d_set_cells": (_I, [_VP, _DP]),
    "pm_bar1d_get_cells": (_I, [_VP, _DP]),
    "pm_bar1d_fourier_number": (_D, [_VP, _D]),
    "pm_bar1d_max_stable_dt": (_D, [_VP]),
    "pm_bar1d_refuses": (_I, [_VP, _D, _VIOL]),
    "pm_bar1d_step": (_I, [_VP, _D, _D, _DP, _VIOL]),
    "pm_bar1d_ledger": (_I, [_VP, _DP, _DP]),
    "pm_bar1d_advance": (_I, [_VP, _D, _U64, _D, _DP, _DP, _VIOL]),
    "pm_bar1d_sample": (_I, [_VP, _DP, _S, _DP]),
    "pm_bar1d_readings": (_I, [_VP, _DP]),
    "pm_bar1d_checkpoint": (_I, [_VP]),
    "pm_bar1d_restore": (_I, [_VP]),
    "pm_tube_create": (_I, [C.c_char_p, _D, _S, _D, _D, _D, C.POINTER(_VP)]),
    "pm_tube_destroy": (None, [_VP]),
    "pm_tube_cells": (_S, [_VP]),
    "pm_tube_spacing": (_D, [_VP]),
    "pm_tube_set_ends": (_I, [_VP, _I, _D, _I, _D]),
    "pm_tube_release": (_I, [_VP, _DP]),
    "pm_tube_max_stable_dt": (_D, [_VP]),
    "pm_tube_courant": (_D, [_VP, _D]),
    "pm_tube_step": (_I, [_VP, _D, _DP, _VIOL]),
    "pm_tube_energy": (_I, [_VP, _DP, _DP, _DP]),
    "pm_tube_advance": (_I, [_VP, _D, _U64, _D, _DP, _DP, _VIOL]),
    "pm_tube_get_fields": (_I, [_VP, _DP, _DP, _DP]),
    "pm_tube_checkpoint": (_I, [_VP]),
    "pm_tube_restore": (_I, [_VP]),
    "pm_ensemble_create": (_I, [C.POINTER(pm_member), _S, C.POINTER(_VP)]),
    "pm_ensemble_destroy": (None, [_VP]),
    "pm_ensemble_set_cells": (_I, [_VP, _S, _DP]),
    "pm_ensemble_get_cells": (_I, [_VP, _S, _DP]),
    "pm_ensemble_run": (_I, [_VP, _U64, _DP, _VIOL]),
    "pm_ensemble_ledgers": (_I, [_VP, _DP]),
    "pm_ensemble_set_schedule": (_I, [_VP, _I, _I]),
    "pm_ensemble_set_tiles": (_I, [_VP, _I]),
    "pm_ensemble_set_steps_per_pass": (_I, [_VP, _I]),
    "pm_ensemble_tune": (_I, [_VP, C.c_char_p, _S]),
    "pm_solid3d_copy_cells": (_I, [_VP, _VP]),
    "pm_surrogate_create": (_I, [_U64, C.POINTER(_VP)]),
    "pm_surrogate_destroy": (None, [_VP]),
    "pm_surrogate_parameter_count": (C.c_int64, [_VP]),
    "pm_surrogate_get_parameters": (_I, [_VP, _FP]),
    "pm_surrogate_set_parameters": (_I, [_VP, _FP]),
    "pm_surrogate_reserve_set": (_I, [_VP, C.c_int64, C.c_int64]),
    "pm_surrogate_set_sample": (_I, [_VP, C.c_int64, _DP, _DP, _DP, _D]),
    "pm_surrogate_set_sample_from": (_I, [_VP, C.c_int64, _VP, _VP, _D]),
    "pm_surrogate_train_step": (_I, [_VP, C.POINTER(_I), _I, C.c_float, _DP]),
    "pm_surrogate_gradients": (_I, [_VP, C.POINTER(_I), _I, _FP, _DP]),
    "pm_surrogate_forward": (_I, [_VP, C.c_int64, _DP, _DP, _D, _FP]),
    "pm_surrogate_step_solid": (_I, [_VP, _VP, _D, _D, _DP, _DP]),
    "pm_solid3d_relative_l2": (_I, [_VP, _VP, _D, _DP]),
    "pm_solid3d_l2_parts": (_I, [_VP, _VP, _D, _DP, _DP]),
    "pm_nn_conv3d": (_I, [_I, _I, _I, _I, _I, _I, _FP, _FP, _FP, _FP, _I]),
    "pm_nn_conv3d_backward": (_I, [_I, _I, _I, _I, _I, _I, _FP, _FP, _FP, _FP, _FP, _FP, _I]),
}

_lib = None


def lib() -> C.CDLL:
    """The loaded library, with every declared signature attached."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is not built; run `python -m pantometry_amd.build` (or "
                "__graft_entry__.build()) first. There is no CPU fallback.")
        handle = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


class DeviceError(RuntimeError):
    """A PM_ERR_* status: device, argument or state failure."""


def last_error() -> str:
    msg = lib().pm_last_error()
    return msg.decode() if msg else ""


def check(status: int, what: str, violation: pm_violation | None = None) -> None:
    """Raises Violation for PM_VIOLATION and DeviceError for any PM_ERR_* status."""
    if status == PM_OK:
        return
    if status == PM_VIOLATION and violation is not None:
        from .core import Violation
        raise Violation(violation.quantity.decode(), violation.site.decode(), violation.before,
                        violation.after, violation.scale, violation.tolerance, violation.step)
    raise DeviceError(f"{what}: status {status}: {last_error()}")


def dptr(a) -> "C.POINTER(C.c_double)":
    """A double* into a C-contiguous float64 numpy array."""
    return a.ctypes.data_as(_DP)


def fptr(a) -> "C.POINTER(C.c_float)":
    """A float* into a C-contiguous float32 numpy array."""
    return a.ctypes.data_as(_FP)
