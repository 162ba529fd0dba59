import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built libsu2mc.so")
    config.addinivalue_line("markers", "slow: long CPU test")


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle as O

    O.load()
    return O


@pytest.fixture(scope="session")
def gpu():
    """The product package with a usable device; GPU tests fail loudly (not
    skip) when the library or the device is missing."""
    import qcd_2d_gpu_amd as q

    q.load_library()
    n = q.device_count()
    assert n >= 1, "no HIP device visible"
    return q
