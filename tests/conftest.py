import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "mini-fem_amd"), os.path.join(ROOT, "oracle"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU")
    config.addinivalue_line("markers", "slow: larger CPU cases")


@pytest.fixture(scope="session")
def orc():
    import pyoracle
    pyoracle.lib()
    return pyoracle


@pytest.fixture(scope="session")
def mf():
    import minifem
    minifem.lib()
    return minifem


@pytest.fixture(scope="session")
def gpu(mf):
    import minifem.gpu as g
    if g.device_count() < 1:
        pytest.fail("GPU test collected without a GPU")
    g.init(0)
    return g
