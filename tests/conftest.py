"""Shared fixtures. GPU tests carry @pytest.mark.gpu; everything else runs on CPU."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from _pkgload import load_package  # noqa: E402

P = load_package()


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def pkg():
    return P


@pytest.fixture(scope="session")
def orc():
    import pyoracle
    return pyoracle


@pytest.fixture(scope="session")
def gpu_ready():
    n = P.lib().cvp9_hip_device_count()
    if n < 1:
        pytest.fail("no HIP device visible to a GPU test")
    return n
