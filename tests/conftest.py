import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built libgdem.so")


def _ensure_oracle():
    so = os.path.join(ROOT, "oracle", "liboracle_dem.so")
    src = os.path.join(ROOT, "oracle", "dem_oracle.cpp")
    if not os.path.exists(so) or os.path.getmtime(so) < os.path.getmtime(src):
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])


_ensure_oracle()


@pytest.fixture(scope="session")
def ctx():
    """One gdem context on device 0 for all GPU tests."""
    from gpudem_amd import Context
    c = Context(0)
    yield c
    c.close()


def golden(name):
    return os.path.join(GOLDEN, name)
