import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "graphgolf-cuda_amd")
for p in (REPO, PKG, os.path.join(REPO, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs on the GPU box)")
    config.addinivalue_line("markers", "slow: long CPU test")


@pytest.fixture(scope="session")
def gg():
    import graphgolf_mi355x as m
    return m


@pytest.fixture(scope="session")
def gpu(gg):
    """The library with a visible device; a GPU test without one fails loudly."""
    n = gg.device_count()
    assert n >= 1, "no HIP device visible: GPU tests must run on the MI355X box"
    return gg
