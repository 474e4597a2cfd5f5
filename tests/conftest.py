import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs an MI355X (gfx950) and the built HIP library")


@pytest.fixture(scope="session")
def orc():
    """The CPU oracle (built from oracle/ on first use)."""
    from oracle import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def hip():
    """The HIP backend binding; fails loudly when the library or GPU is
    missing (GPU tests never fall back to the CPU)."""
    from gpumetropolis_amd import _native
    if not _native.available():
        raise RuntimeError("HIP backend unavailable: " + _native.last_error())
    return _native
