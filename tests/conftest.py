import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

WIDTHS = (8, 16, 32, 64)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")


@pytest.fixture(scope="session")
def orc():
    from oracle import load
    return load()


def rand_limbs(rng, k, shape):
    """Uniform limbs of Z/2^k as uint64."""
    x = rng.integers(0, 2**63, size=shape, dtype=np.uint64) * np.uint64(2) + rng.integers(0, 2, size=shape, dtype=np.uint64)
    if k < 64:
        x &= np.uint64((1 << k) - 1)
    return x


@pytest.fixture
def rng():
    return np.random.default_rng(1234)
