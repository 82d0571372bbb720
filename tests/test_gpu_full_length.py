"""The three stencil configurations at full size and full length, against the oracle's digests.

`tests/golden/full_length.json` holds SHA-256 digests of the oracle's inputs and final fields and
`full_length_totals.npz` its ledger after every step (both written by
`tests/golden/make_digests.py`).  Each test first checks the digests of the inputs it generated on
this host (a host libm that differs would show up there, as input drift, not as a kernel
mismatch), then runs the whole configuration through the C ABI and checks the final fields bit for
bit and every step's ledger total:

  heat      512^3 x 2000 steps: every total within 1e-12 of the compensated sum and no farther
            from the reference's left fold than that fold's own rounding (plus 1e-12);
  hall      385^3 x 4000 steps with the Ricker source: every total within 1e-12 (of that
            step's ledger scale) of the compensated sum, and no farther from the reference's
            left fold than that fold's own rounding (plus 1e-12); the fold over 57 M nodes
            drifts by up to ~1e-10 of the scale;
  ensemble  512 members x 500 steps in one graph launch: every member's cells bit for bit and its
            stored heat within 1e-12.
"""
import hashlib
import json
from pathlib import Path

import numpy as np
import pytest

from tests.conftest import gpu_case

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"


def sha(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.float64).tobytes()).hexdigest()


@pytest.fixture(scope="module")
def digests():
    rec = json.loads((GOLDEN / "full_length.json").read_text())
    with np.load(GOLDEN / "full_length_totals.npz") as z:
        arrays = {k: z[k] for k in z.files}
    return rec, arrays


def _record(digests, name, keys):
    rec, arrays = digests
    if name not in rec or any(k not in arrays for k in keys):
        pytest.skip(f"no {name} digest in tests/golden/ "
                    f"(python tests/golden/make_digests.py {name})")
    return rec[name], arrays


@pytest.mark.timeout(600)
def test_heat_512_for_2000_steps(digests):
    d, arrays = _record(digests, "heat", ("heat_fold", "heat_exact"))
    gpu_case("heat 512^3 x 2000 against the oracle's digests")
    import pantometry_amd as pm
    from pantometry_amd import workloads as w
    n = d["counts"][0]
    k, t = w.heterogeneous_fields((n, n, n))
    assert sha(k) == d["k_sha256"] and sha(t) == d["t0_sha256"], "input drift (host libm?)"
    g = pm.Solid3D("block", pm.Substance("s", w.RHO, 1.0, w.CP), (n, n, n), w.DX, 300.0)
    g.with_conductivity(k)
    g.set_cells(t)
    del k, t
    dt = 0.9 * g.max_stable_dt()
    assert dt.hex() == d["dt"]
    totals, _ = g.advance(dt, d["steps"], tolerance=1e-12)
    assert sha(g.cells()) == d["final_sha256"]
    fold, exact = arrays["heat_fold"], arrays["heat_exact"]
    assert np.all(np.abs(totals - exact) <= 1e-12 * np.abs(exact))
    assert np.all(np.abs(totals - fold) <= np.abs(fold - exact) + 1e-12 * np.abs(exact))
    print(f"largest |gpu - exact| / exact {np.max(np.abs(totals - exact) / np.abs(exact)):.3e}, "
          f"largest |fold - exact| / exact {np.max(np.abs(fold - exact) / np.abs(exact)):.3e}",
          flush=True)


@pytest.mark.timeout(600)
def test_hall_385_for_4000_steps(digests):
    d, arrays = _record(digests, "hall", ("hall_totals", "hall_exact", "hall_scale"))
    gpu_case("hall 385^3 x 4000 against the oracle's digests")
    from pantometry_amd import workloads as w
    case = w.acoustic_hall(d["nodes"][0] - 1, d["steps"])
    assert case.source_node == d["source_node"] and case.dt.hex() == d["dt"]
    assert sha(case.increments) == d["increments_sha256"], "input drift (host libm?)"
    h = case.hall
    totals, _ = h.advance(case.dt, d["steps"], tolerance=1e-12)
    for name, field in zip(("p", "pp", "vx", "vy", "vz"), h.fields()):
        assert sha(field) == d["final_sha256"][name], name
    assert h.supplied() == float.fromhex(d["supplied"])
    fold, exact, scale = arrays["hall_totals"], arrays["hall_exact"], arrays["hall_scale"]
    assert np.all(np.abs(totals - exact) <= 1e-12 * scale)
    assert np.all(np.abs(totals - fold) <= np.abs(fold - exact) + 1e-12 * scale)
    print(f"largest |gpu - exact| / scale {np.max(np.abs(totals - exact) / np.maximum(scale, 1e-300)):.3e}, "
          f"largest |fold - exact| / scale {np.max(np.abs(fold - exact) / np.maximum(scale, 1e-300)):.3e}",
          flush=True)


@pytest.mark.timeout(600)
def test_ensemble_of_512_for_500_steps(digests):
    d, arrays = _record(digests, "ensemble", ("ensemble_stored",))
    gpu_case("ensemble 512 x 500 against the oracle's digests")
    from pantometry_amd import workloads as w
    ens, shapes = w.heat_ensemble(d["count"])
    for m, ((edge, alpha), want) in enumerate(zip(shapes, d["members"])):
        assert edge == want["edge"] and alpha.hex() == want["alpha"], m
        assert ens.members[m].dt.hex() == want["dt"], m
        assert sha(ens.cells(m)) == want["t0_sha256"], ("input drift", m)
    print(ens.tune(), flush=True)
    ens.run(d["steps"])
    stored = ens.ledgers()
    for m, want in enumerate(d["members"]):
        assert sha(ens.cells(m)) == want["final_sha256"], (m, want["edge"])
    ref = arrays["ensemble_stored"]
    assert np.all(np.abs(stored - ref) <= 1e-12 * np.abs(ref))
