import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "nbodysimulator.jl_amd")
for p in (ROOT, PKG, os.path.dirname(os.path.abspath(__file__))):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libnbx on cuda:0)")


@pytest.fixture(scope="session")
def oracle():
    import oracle_lib
    return oracle_lib


@pytest.fixture(scope="session")
def nbx():
    """The product library; GPU tests fail loudly (no fallback) if it cannot run."""
    import nbodysimulator
    nbodysimulator.load()
    if nbodysimulator.device_count() < 1:
        pytest.fail("no HIP device visible: GPU tests need an MI355X")
    return nbodysimulator
