"""Corner geometry: the oracle's restatement of rt/geometryUCBxyz.F90 against
what the mathematics fixes (closure, exact volumes, equal-and-opposite
areas, outward normals, closed forms on box zones), and the product
library's host geometry against the oracle."""
import numpy as np
import pytest

import umt_amd as U
from oracle import oracle as O


@pytest.fixture(scope="module", params=[(1, 1, 1), (2, 3, 2)])
def prob(request):
    p = U.TiledProblem(*request.param, 2, 2)
    m = p.mesh()
    A_fp, A_ez, V = O.geometry(m)
    return p, m, A_fp, A_ez, V


def test_product_geometry_equals_oracle(prob):
    p, m, A_fp, A_ez, V = prob
    g = p.geometry()
    np.testing.assert_array_equal(g["A_fp"], A_fp)
    np.testing.assert_array_equal(g["A_ez"], A_ez)
    np.testing.assert_array_equal(g["Volume"], V)
    np.testing.assert_allclose(g["VolumeZone"], V.reshape(-1, 8).sum(axis=1), rtol=1e-15)


def test_corner_closure(prob):
    """Each corner is a closed polyhedron: sum of its face area vectors = 0."""
    _, _, A_fp, A_ez, _ = prob
    s = A_fp.sum(axis=1) + A_ez.sum(axis=1)
    assert np.abs(s).max() < 1e-16 * 1e2


def test_zone_and_total_volume(prob):
    """Sum of corner volumes = exact volume of the extruded quadrilateral."""
    p, m, _, _, V = prob
    px = m["px"].reshape(-1, 8, 3)
    x, y = px[:, :4, 0], px[:, :4, 1]
    area = 0.5 * np.abs(np.sum(x * np.roll(y, -1, axis=1) - np.roll(x, -1, axis=1) * y, axis=1))
    dz = px[:, 4, 2] - px[:, 0, 2]
    np.testing.assert_allclose(V.reshape(-1, 8).sum(axis=1), area * dz, rtol=1e-13)
    assert abs(V.sum() - 1.0) < 1e-13


def test_equal_and_opposite(prob):
    p, m, A_fp, A_ez, _ = prob
    Cn = p.ncorner
    cFP, cEZ = m["cFP"], m["cEZ"]
    for c in range(Cn):
        for f in range(3):
            n = cFP[c, f]
            if n <= Cn:
                ff = np.where(cFP[n - 1] == c + 1)[0]
                assert len(ff) == 1
                np.testing.assert_array_equal(A_fp[n - 1, ff[0]], -A_fp[c, f])
            z0 = 8 * (c // 8)
            e = z0 + cEZ[c, f] - 1
            fe = np.where(cEZ[e] == c - z0 + 1)[0]
            assert len(fe) == 1
            np.testing.assert_array_equal(A_ez[e, fe[0]], -A_ez[c, f])


def test_fp_areas_point_out_of_the_zone(prob):
    p, m, A_fp, _, _ = prob
    px = m["px"]
    zc = px.reshape(-1, 8, 3).mean(axis=1)
    for c in range(p.ncorner):
        z = c // 8
        for f in range(3):
            face = m["CToFace"][c, f]
            corners = [8 * z + k for k in range(8) if face in m["CToFace"][8 * z + k]]
            fc = px[corners].mean(axis=0)
            assert np.dot(fc - zc[z], A_fp[c, f]) > 0


def test_box_zone_closed_forms(prob):
    """Axis-aligned box zones (the tile's corner and centre squares): every
    corner holds 1/8 of the box; each FP corner face is a quarter of its box
    face; each EZ face is a quarter of the mid-plane section."""
    p, m, A_fp, A_ez, V = prob
    px = m["px"].reshape(-1, 8, 3)
    nbox = 0
    for z in range(p.nzones):
        lo, hi = px[z].min(axis=0), px[z].max(axis=0)
        corners_on_box = np.all(np.isclose(px[z], lo) | np.isclose(px[z], hi))
        if not corners_on_box:
            continue
        nbox += 1
        d = hi - lo
        np.testing.assert_allclose(V[8 * z:8 * z + 8], np.prod(d) / 8, rtol=1e-14)
        quarter = np.array([d[1] * d[2], d[0] * d[2], d[0] * d[1]]) / 4
        for c in range(8):
            for f in range(3):
                for A in (A_fp[8 * z + c, f], A_ez[8 * z + c, f]):
                    axis = int(np.argmax(np.abs(A)))
                    np.testing.assert_allclose(abs(A[axis]), quarter[axis], rtol=1e-14)
                    assert np.abs(np.delete(A, axis)).max() < 1e-14 * abs(A[axis])
    assert nbox >= 8 * p.tiles[0] * p.tiles[1] * p.tiles[2]
