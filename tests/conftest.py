"""Shared test setup: the `gpu` marker, import paths and fixtures.

Tests marked `gpu` need an MI355X (gfx950) and the built libswmi.so; they fail (not skip) when either
is missing, so a silent fallback can never pass. Everything else runs on the CPU.
"""
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as o
    o.lib()
    return o


@pytest.fixture(scope="session")
def swmi():
    import computational_physics_amd as cpa
    info = cpa.device_info()  # raises if the library or the device is missing
    assert info["arch"].startswith("gfx950"), info
    return cpa
