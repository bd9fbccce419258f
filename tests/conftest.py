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


def hex_grid(n, amount, seed, twist=0.0):
    """Points (n+1)^3 of the unit cube and n^3 hexes (VTK node order). Interior
    nodes move by up to amount/n in each direction; twist rotates interior
    nodes about the z axis by twist * z radians. Strong distortions give
    warped faces, zones whose corners depend on each other in a cycle and
    cycles among zones."""
    r = np.arange(n + 1) / n
    X, Y, Z = np.meshgrid(r, r, r, indexing="ij")
    pts = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)

    def idx(i, j, k):
        return (i * (n + 1) + j) * (n + 1) + k

    hexes = np.array([[idx(i, j, k), idx(i + 1, j, k), idx(i + 1, j + 1, k), idx(i, j + 1, k),
                       idx(i, j, k + 1), idx(i + 1, j, k + 1), idx(i + 1, j + 1, k + 1),
                       idx(i, j + 1, k + 1)]
                      for i in range(n) for j in range(n) for k in range(n)], np.int32)
    rng = np.random.default_rng(seed)
    inside = np.all((pts > 1e-12) & (pts < 1 - 1e-12), axis=1)
    pts = pts.copy()
    pts[inside] += rng.uniform(-amount / n, amount / n, (int(inside.sum()), 3))
    if twist:
        c = pts[:, :2] - 0.5
        th = twist * pts[:, 2] * inside
        pts[:, 0] = 0.5 + np.cos(th) * c[:, 0] - np.sin(th) * c[:, 1]
        pts[:, 1] = 0.5 + np.sin(th) * c[:, 0] + np.cos(th) * c[:, 1]
    return pts, hexes


# Meshes whose S_N schedules have cycles: (n, amount, seed, twist).
CYCLIC_MESHES = [(4, 0.3, 1, 0.0), (6, 0.45, 3, 0.0), (6, 0.2, 4, 1.5)]


def cyclic_problem(n, amount, seed, twist, npolar=2, nazimuthal=2):
    import umt_amd as U
    pts, hexes = hex_grid(n, amount, seed, twist)
    return U.TiledProblem.from_hex_mesh(pts, hexes, npolar, nazimuthal, kdtree=False)
