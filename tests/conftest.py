import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP sweep through the C-ABI)")


@pytest.fixture(scope="session")
def small_problem():
    import umt_amd as U
    return U.TiledProblem(2, 2, 2, 2, 2)


@pytest.fixture(scope="session")
def config1_problem():
    """4x4x4 tiles (1536 zones), 2x2 product quadrature (32 angles)."""
    import umt_amd as U
    return U.TiledProblem(4, 4, 4, 2, 2)


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = np.maximum(np.abs(b), 1e-300)
    return float(np.max(np.abs(a - b) / scale))
