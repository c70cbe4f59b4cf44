import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "parareal-cc-cfv-split-gpu_amd")
for p in (PKG, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def oracle():
    import pyoracle
    pyoracle.lib()
    pyoracle.set_threads(min(16, os.cpu_count() or 1))
    return pyoracle


@pytest.fixture(scope="session")
def gpu():
    import ccfv
    ccfv.load()
    if ccfv.device_count() < 1:
        pytest.fail("no HIP device visible to the GPU tests")
    return ccfv
