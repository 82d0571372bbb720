"""Reading tests/golden/stencil_fixtures.json (written by tests/golden/make_fixtures.py)."""
from __future__ import annotations

import json
import struct
from pathlib import Path

import numpy as np

FIXTURES = Path(__file__).resolve().parent / "golden" / "stencil_fixtures.json"


def load() -> dict:
    return json.loads(FIXTURES.read_text())


def arr(h: str) -> np.ndarray:
    """float64 array from big-endian binary64 hex."""
    return np.frombuffer(bytes.fromhex(h), dtype=">f8").astype(np.float64)


def num(h: str) -> float:
    return struct.unpack(">d", bytes.fromhex(h))[0]


def bits(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64).reshape(-1).view(np.uint64)
