import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and runs the HIP kernels")


def gpu_case(msg: str) -> None:
    """Names the case about to run on the GPU, flushed, so a run that ends the box still says
    which case it was."""
    print(f"[gpu case] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as o
    o.lib()
    return o
