"""Shared fixtures. Tests marked ``gpu`` need an MI355X and the built libgbsk.so;
everything else runs on the CPU (oracle, host logic, ABI surface)."""
from __future__ import annotations

import json
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

GOLDEN = Path(__file__).resolve().parent / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU and the built HIP library")
    config.addinivalue_line("markers", "slow: long-running case")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.lib()
    return O


@pytest.fixture(scope="session")
def ref_checksums():
    return json.loads((GOLDEN / "ref_checksums.json").read_text())["rows"]


def rng_complex(shape, seed, scale=1.0):
    g = np.random.default_rng(seed)
    return scale * (g.standard_normal(shape) + 1j * g.standard_normal(shape))


def rng_symmetric(B, N, seed, scale=1.0):
    a = rng_complex((B, N, N), seed, scale)
    return np.ascontiguousarray(a + a.transpose(0, 2, 1))


def rel_err(a, b):
    a = np.asarray(a, np.complex128)
    b = np.asarray(b, np.complex128)
    return np.abs(a - b) / np.maximum(np.abs(b), 1e-300)
