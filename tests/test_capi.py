"""The C ABI without a GPU: the library loads, exports every symbol the header declares, refuses
to construct a domain without a gfx950 device (no CPU fallback), and its host logic matches."""
import ctypes as C
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def _declared():
    text = (ROOT / "include" / "pantometry_amd.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pm_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    import pantometry_amd as pm
    handle = C.CDLL(str(pm.LIB_PATH))
    names = _declared()
    assert len(names) > 40
    missing = [n for n in names if not hasattr(handle, n)]
    assert not missing, missing


def test_the_binding_covers_the_header():
    from pantometry_amd import _capi
    assert set(_declared()) == set(_capi.SIGNATURES)


def _has_device():
    import pantometry_amd as pm
    n = C.c_int(0)
    return pm.lib().pm_device_info(None, 0, C.byref(n), None) == 0


def test_no_device_means_no_domain():
    import pantometry_amd as pm
    if _has_device():
        pytest.skip("a device is present; the refusal path is for hosts without one")
    with pytest.raises(pm.DeviceError, match="no CPU fallback|no HIP device"):
        pm.Solid3D("b", pm.Substance.aluminium_6061(), (4, 4, 4), 1e-3, 300.0)
    with pytest.raises(pm.DeviceError):
        pm.Hall.of_air("h", 1.0, 1.0, 1.0, 5)


def test_null_arguments_are_refused():
    import pantometry_amd as pm
    lib = pm.lib()
    assert lib.pm_solid3d_create(b"x", None, 1, 1, 1, 1.0, 1.0, None) == 3
    assert lib.pm_solid3d_step(None, 1.0, 0.0, None) == 3
    assert lib.pm_hall_step(None, 1.0, None) == 3
    assert lib.pm_workload_hot_spot(0, 1, 1, 0.0, 0.0, 1.0, 0.0, 0, None) == 3


def test_the_c_binding_of_a_violation_round_trips():
    from pantometry_amd import _capi
    from pantometry_amd.core import Violation
    v = _capi.pm_violation()
    v.quantity = b"Fourier number"
    v.site = b"block (explicit 3D conduction)"
    v.before, v.after, v.scale, v.tolerance, v.step = 1 / 6, 1.05 / 6, 1 / 6, 1e-12, -1
    with pytest.raises(Violation) as err:
        _capi.check(1, "step", v)
    assert err.value.quantity == "Fourier number"
    assert "created" in str(err.value)
