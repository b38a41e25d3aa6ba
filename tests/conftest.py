import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O

    O.lib()
    return O


@pytest.fixture(scope="session")
def umapx():
    """The C-ABI library (loaded, never rebuilt here)."""
    from umapers_amd import _lib as L

    L.lib()
    return L


@pytest.fixture(scope="session")
def gpu(umapx):
    n = umapx.device_count()
    if n < 1:
        raise RuntimeError("GPU test collected but no GPU is visible to HIP")
    return n


@pytest.fixture
def rng():
    return np.random.default_rng(1234)
