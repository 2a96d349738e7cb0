import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library on the device)")


@pytest.fixture(scope="session")
def ws():
    """The device workspace. A GPU test fails (not skips) when the library
    is missing or the device is unusable."""
    import csharp_linearalgebra_amd as T

    w = T.Workspace(0)
    yield w
    w.dispose()


@pytest.fixture
def rng():
    return np.random.default_rng(12345)
