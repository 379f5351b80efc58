import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "n-body-simulation_amd")
for p in (PKG, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


@pytest.fixture(scope="session")
def oracle():
    import oracle as O
    return O


@pytest.fixture(scope="session")
def nb():
    import nbody
    return nbody


def _gpu_present():
    try:
        import nbody
        nbody.device_name()
        return True
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu(nb):
    if not _gpu_present():
        pytest.fail("no HIP device visible: gpu tests need an MI355X")
    return nb
