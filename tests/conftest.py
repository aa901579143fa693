"""Shared test setup: import paths, the `gpu` marker, and common scene fixtures."""
import os
import sys
from pathlib import Path

import pytest

# Under pytest-xdist every worker would give the oracle's OpenMP loops all the host's cores, and
# the workers' spinning threads then starve one another (a 40 s suite took 28 minutes at -n 4).
if os.environ.get("PYTEST_XDIST_WORKER") and "OMP_NUM_THREADS" not in os.environ:
    os.environ["OMP_NUM_THREADS"] = "1"

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT / "break-that-denim_amd"))
sys.path.insert(0, str(ROOT / "oracle"))
sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built libbd.so")


def lcg(seed):
    """Deterministic LCG of the reference tests (newtonMultigrid.test.ts:19-25): values in [-0.5, 0.5)."""
    s = seed & 0xFFFFFFFF

    def rnd():
        nonlocal s
        s = (s * 1664525 + 1013904223) & 0xFFFFFFFF
        return s / 4294967296 - 0.5

    return rnd
