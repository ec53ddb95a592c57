import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a real MI355X (runs through the C ABI)")


def gmm_int(N, D, ncomp, seed, sigma=20.0):
    """Small integer-valued clipped Gaussian mixture (the C1/C2 data family) from numpy."""
    rng = np.random.default_rng(seed)
    C = rng.uniform(0, 255, (ncomp, D))
    X = C[rng.integers(0, ncomp, N)] + sigma * rng.standard_normal((N, D))
    return np.clip(np.rint(X), 0, 255).astype(np.float32)


def gmm_unit(N, D, ncomp, seed):
    rng = np.random.default_rng(seed)
    C = rng.standard_normal((ncomp, D))
    X = C[rng.integers(0, ncomp, N)] + rng.standard_normal((N, D))
    X /= np.linalg.norm(X, axis=1, keepdims=True)
    return X.astype(np.float32)


@pytest.fixture(scope="session")
def small_int_data():
    return gmm_int(3000, 128, 32, 7)


@pytest.fixture(scope="session")
def small_unit_data():
    return gmm_unit(2000, 96, 16, 11)
