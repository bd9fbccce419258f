"""Quadrature: the oracle's restatement of rt/quadProduct.F90 fed with the
reference's own QRS45 tables (fixture parsed from
mods/QuadratureData_mod.F90), the product library's compiled-in tables,
and the identities a product quadrature must satisfy."""
import json
import os

import numpy as np
import pytest

import umt_amd as U
from oracle import oracle as O

FIX = os.path.join(os.path.dirname(__file__), "golden", "qrs45_n1to4.json")


def _tables(n_polar, n_azi):
    with open(FIX) as f:
        t = json.load(f)["orders"]
    p, a = t[str(n_polar)], t[str(n_azi)]
    return p["cosTheta"], p["weightTheta"], a["cosPhiXYZ"], a["weightPhiXYZ"]


@pytest.mark.parametrize("npol,nazi", [(1, 1), (2, 2), (3, 3), (4, 4), (2, 4), (4, 1)])
def test_product_quadrature_matches_reference_tables(npol, nazi):
    omega, weight = O.quadrature(npol, nazi, *_tables(npol, nazi), polaraxis=1)
    p = U.TiledProblem(1, 1, 1, npol, nazi)
    q = p.quadrature()
    np.testing.assert_array_equal(q["omega"], omega)
    np.testing.assert_array_equal(q["weight"], weight)


@pytest.mark.parametrize("npol,nazi", [(2, 2), (3, 3), (4, 4)])
def test_quadrature_identities(npol, nazi):
    omega, w = O.quadrature(npol, nazi, *_tables(npol, nazi), polaraxis=1)
    assert omega.shape == (8 * npol * nazi, 3)
    assert abs(w.sum() - 4 * np.pi) < 1e-13
    np.testing.assert_allclose(np.linalg.norm(omega, axis=1), 1.0, atol=1e-15)
    # Octant symmetry: odd moments and mixed second moments vanish.
    for i in range(3):
        assert abs(np.sum(w * omega[:, i])) < 1e-14
        for j in range(i + 1, 3):
            assert abs(np.sum(w * omega[:, i] * omega[:, j])) < 1e-14
    # Angles are numbered octant-fastest: angle 8k+o is base direction k in octant o.
    signs = np.sign(omega[:8])
    expect = np.array([[1, 1, 1], [-1, 1, 1], [-1, -1, 1], [1, -1, 1],
                       [1, 1, -1], [-1, 1, -1], [-1, -1, -1], [1, -1, -1]])
    np.testing.assert_array_equal(signs, expect)


def test_polar_axis_choice_permutes_components():
    t = _tables(3, 2)
    o1, w1 = O.quadrature(3, 2, *t, polaraxis=1)
    o3, w3 = O.quadrature(3, 2, *t, polaraxis=3)
    np.testing.assert_array_equal(w1, w3)
    # polar cosine sits on the chosen axis
    np.testing.assert_allclose(np.sort(np.abs(o1[:, 0])), np.sort(np.abs(o3[:, 2])), atol=0)


def test_fixture_tables_are_consistent():
    """Each tabulated order: polar weights sum to 1, azimuthal weights to
    pi/2 (one quadrant), cosines lie in (0,1) and are ordered, as a product
    rule over one octant requires (catches a misread or shifted table)."""
    with open(FIX) as f:
        t = json.load(f)["orders"]
    for n in range(1, 5):
        o = t[str(n)]
        assert abs(sum(o["weightTheta"]) - 1.0) < 1e-15
        assert abs(sum(o["weightPhiXYZ"]) - np.pi / 2) < 1e-15
        for key in ("cosTheta", "cosPhiXYZ"):
            v = np.array(o[key])
            assert np.all((v > 0) & (v < 1))
        assert np.all(np.diff(o["cosTheta"]) < 0)
        assert np.all(np.diff(o["cosPhiXYZ"]) > 0)
