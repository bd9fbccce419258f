"""The oracle's polyhedral build (liboracle_poly.so: the same C restatement of
rt/geometryUCBxyz.F90, the schedule and snac/SweepUCBxyz.F90 with corner-face,
corner and zone-face arrays of leading dimensions 6, 12 and 8), pinned
before it checks the GPU's polyhedral sweep:

- on hex meshes laid out with the wider dimensions it gives the hex build's
  geometry, schedule and sweep bit for bit;
- on a hex-pyramid mesh (pyramid apexes are corners with four faces) its
  geometry fills the unit cube, closes every corner's surface, sums each zone
  face's corner areas to the face's own area vector (computed here from the
  points) and is antisymmetric across faces;
- its sweep keeps an infinite-medium solution psi = Q / sigma exactly (every
  coupling term, the mean over the faces opposite an EZ face included, must
  vanish for that) and balances every zone: the EZ terms cancel in the sum
  over a zone's corners, leaving outflow + absorption = volume source +
  inflow.

Not pinned beyond agreement with the GPU kernel, which restates the same
reference lines independently: which of a many-faced corner's incoming FP
faces count as opposite an EZ face (snac/SweepUCBxyz.F90:201-208)."""
import numpy as np
import pytest

import umt_amd as U
from oracle import oracle as O
from polymesh import PolyProblem, hex_pyramid_mesh, widen


@pytest.fixture(scope="module")
def quad():
    return U.TiledProblem(1, 1, 1, 2, 2).quadrature()


@pytest.fixture(scope="module")
def pyr(quad):
    return PolyProblem(hex_pyramid_mesh(3, 3, 2), quad)


def _sweep(p, a, S, sigt, psib=None, phi=None):
    G = S.shape[1]
    s = p.schedule()
    start = int(np.sum(s["nHyperPlanes"][:a]))
    zip_ = s["zonesInPlane"][start:start + int(s["nHyperPlanes"][a])]
    Psi1 = np.zeros((p.ncorner + p.nbelem, G))
    PsiB = np.zeros((p.nbelem, G)) if psib is None else psib.copy()
    Phi = np.zeros((p.ncorner, G)) if phi is None else phi.copy()
    O.sweep_angle(p.mesh(), p.geometry(), S, sigt, p.quadrature()["omega"][a],
                  p.quadrature()["weight"][a], s["nextZ"][a], s["nextC"][a], zip_, Psi1, PsiB, Phi)
    return Psi1, PsiB, Phi


def test_wide_build_reproduces_the_hex_build(quad):
    tp = U.TiledProblem(2, 2, 1, 2, 2)
    m = tp.mesh()
    wide = widen(m)
    A_fp, A_ez, V = O.geometry(m)
    W_fp, W_ez, WV = O.geometry(wide)
    assert np.array_equal(W_fp[:, :3], A_fp) and not W_fp[:, 3:].any()
    assert np.array_equal(W_ez[:, :3], A_ez) and np.array_equal(WV, V)
    omega = quad["omega"]
    for a in (0, 5, 17):
        r1 = O.schedule_angle(m, A_fp, A_ez, omega[a])
        r2 = O.schedule_angle(wide, W_fp, W_ez, omega[a])
        for x, y in zip(r1, r2):
            assert np.array_equal(x, y)
    G = 3
    st, _, q = U.draw_cross_sections(4, tp.nzones, G, (0.1, 10.0), (0.0, 0.0), (0.0, 1.0))
    S = np.ascontiguousarray(np.repeat(q, 8, axis=0))
    hexp = PolyProblem(m, quad)
    widep = PolyProblem(wide, quad)
    for a in (0, 9, 30):
        for x, y in zip(_sweep(hexp, a, S, st), _sweep(widep, a, S, st)):
            assert np.array_equal(x, y)


def _newell(p):
    return 0.5 * sum(np.cross(p[i], p[(i + 1) % len(p)]) for i in range(len(p)))


def test_pyramid_geometry_invariants(pyr):
    m, g = pyr.mesh(), pyr.geometry()
    A_fp, A_ez, V = g["A_fp"], g["A_ez"], g["Volume"]
    assert (m["nCFaces"] == 4).sum() == 9 * 6  # one apex per pyramid
    assert abs(V.sum() - 1.0) < 1e-14 and V.min() > 0
    scale = np.abs(A_fp).max()
    # A corner's FP and EZ faces close its surface.
    assert np.abs(A_fp.sum(axis=1) + A_ez.sum(axis=1)).max() < 1e-15 * max(1.0, scale) * 10
    # Each zone face: the corner pieces sum to the face's outward area vector
    # (cEZ on a face points to the previous corner of its inward winding, so
    # following it walks the face outward); across a face the pieces are
    # equal and opposite.
    Z = len(m["numCorner"])
    for z in range(Z):
        c0, n = m["cOffSet"][z], m["numCorner"][z]
        for f in range(m["zoneFaces"][z]):
            cs = [(c, k) for c in range(c0, c0 + n) for k in range(m["nCFaces"][c])
                  if m["CToFace"][c, k] == f + 1]
            tot = sum(A_fp[c, k] for c, k in cs)
            # The face's corners in winding order: chained through cEZ on it.
            corners = [c for c, _ in cs]
            ring = [corners[0]]
            while len(ring) < len(corners):
                c = ring[-1]
                k = [k for cc, k in cs if cc == c][0]
                ring.append(c0 + m["cEZ"][c, k] - 1)
            area = _newell(m["px"][ring])
            assert np.abs(tot - area).max() < 1e-15 * 10, (z, f)
            for c, k in cs:
                cfp = m["cFP"][c, k] - 1
                if cfp < len(V):
                    kk = [j for j in range(m["nCFaces"][cfp]) if m["cFP"][cfp, j] == c + 1]
                    assert len(kk) == 1 and np.array_equal(A_fp[cfp, kk[0]], -A_fp[c, k])


def test_pyramid_sweep_keeps_the_infinite_medium_solution(pyr):
    G = 4
    sig = np.array([0.3, 1.0, 7.0, 50.0])
    Q = np.array([1.0, 0.25, 3.0, 0.5])
    S = np.ascontiguousarray(np.tile(Q, (pyr.ncorner, 1)))
    st = np.ascontiguousarray(np.tile(sig, (pyr.nzones, 1)))
    psib = np.ascontiguousarray(np.tile(Q / sig, (pyr.nbelem, 1)))
    for a in range(pyr.nangles):
        Psi1, PsiB, _ = _sweep(pyr, a, S, st, psib=psib)
        assert np.abs(Psi1[:pyr.ncorner] / (Q / sig) - 1).max() < 1e-13, a


def test_pyramid_sweep_balances_every_zone(pyr):
    m, g = pyr.mesh(), pyr.geometry()
    G = 3
    st, _, q = U.draw_cross_sections(8, pyr.nzones, G, (0.05, 20.0), (0.0, 0.0), (0.0, 1.0))
    rng = np.random.default_rng(5)
    S = rng.uniform(0.0, 1.0, (pyr.ncorner, G))
    psib = rng.uniform(0.0, 0.5, (pyr.nbelem, G))
    omega = pyr.quadrature()["omega"]
    s = pyr.schedule()
    assert s["numCycles"].sum() == 0
    for a in (0, 7, 21):
        Psi1, _, _ = _sweep(pyr, a, S, st, psib=psib)
        # The boundary tail holds the incident fluxes the sweep read.
        Psi1[pyr.ncorner:] = psib
        for z in range(pyr.nzones):
            c0, n = m["cOffSet"][z], m["numCorner"][z]
            lhs = np.zeros(G)
            rhs = np.zeros(G)
            for c in range(c0, c0 + n):
                rhs += g["Volume"][c] * S[c]
                lhs += st[z] * g["Volume"][c] * Psi1[c]
                for k in range(m["nCFaces"][c]):
                    afp = float(np.dot(omega[a], g["A_fp"][c, k]))
                    lhs += afp * (Psi1[c] if afp > 0 else Psi1[m["cFP"][c, k] - 1])
            assert np.abs(lhs - rhs).max() <= 1e-12 * np.abs(rhs).max(), (a, z)
