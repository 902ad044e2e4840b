"""Shared fixtures. Tests marked `gpu` need a HIP device and the built
liblucida.so; everything else runs on the CPU (oracle, host logic, ABI)."""
from __future__ import annotations

import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))
GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built liblucida.so")


@pytest.fixture(scope="session")
def gpu():
    import lucida_amd as L
    g = L.Gpu(0)
    yield g
    g.close()


@pytest.fixture(scope="session")
def oracle():
    import pyoracle
    pyoracle.lib()
    return pyoracle
