"""Sweep schedule: the oracle's restatement of snneed/findseeds/
getDownStreamData/snnext against brute force (longest-path layering of the
zone dependency graph, lexicographically smallest corner order by
enumeration), the product library's schedule against the oracle, and the
cycle guard."""
import itertools

import numpy as np
import pytest

import umt_amd as U
from oracle import oracle as O


@pytest.fixture(scope="module", params=[((2, 2, 1), (2, 2)), ((2, 1, 2), (3, 3))])
def prob(request):
    tiles, quad = request.param
    p = U.TiledProblem(*tiles, *quad)
    return p, p.mesh(), p.geometry(), p.quadrature()


def test_product_schedule_equals_oracle(prob):
    p, m, g, q = prob
    ref = O.schedule(m, g["A_fp"], g["A_ez"], q["omega"])
    s = p.schedule()
    for k in ("nHyperPlanes", "zonesInPlane", "nextZ", "nextC"):
        np.testing.assert_array_equal(s[k], ref[k], err_msg=k)


def _zone_edges(m, A_fp, omega):
    """Upwind edges (u -> d) from the sign of Omega . A_fp summed over the
    corner faces a zone has on each shared face."""
    Z = m["numCorner"].shape[0]
    afp = A_fp @ omega  # (C, 3)
    edges = set()
    for z in range(Z):
        for face in range(1, 7):
            zo = m["zoneOpp"][z, face - 1]
            if zo <= 0:
                continue
            tot = 0.0
            for c in range(8 * z, 8 * z + 8):
                for f in range(3):
                    if m["CToFace"][c, f] == face:
                        tot += afp[c, f]
            if tot > 0:
                edges.add((z, zo - 1))
            elif tot < 0:
                edges.add((zo - 1, z))
    return edges


def test_hyperplanes_are_longest_path_layers(prob):
    p, m, g, q = prob
    for a in range(0, p.nangles, 5):
        rc, nextZ, nextC, zip_ = O.schedule_angle(m, g["A_fp"], g["A_ez"], q["omega"][a])
        assert rc == 0
        edges = _zone_edges(m, g["A_fp"], q["omega"][a])
        level = np.zeros(p.nzones, dtype=int)
        for _ in range(p.nzones):
            changed = False
            for u, d in edges:
                if level[d] < level[u] + 1:
                    level[d] = level[u] + 1
                    changed = True
            if not changed:
                break
        plane = np.repeat(np.arange(len(zip_)), zip_)
        order = nextZ - 1
        assert sorted(order.tolist()) == list(range(p.nzones))
        got = np.empty(p.nzones, dtype=int)
        got[order] = plane
        np.testing.assert_array_equal(got, level)


def test_corner_order_is_lexicographically_smallest_topological_order(prob):
    p, m, g, q = prob
    rng = np.random.default_rng(3)
    perms = np.array(list(itertools.permutations(range(8))))
    pos = np.argsort(perms, axis=1)  # position of each corner in each permutation
    for a in rng.choice(p.nangles, 4, replace=False):
        rc, nextZ, nextC, _ = O.schedule_angle(m, g["A_fp"], g["A_ez"], q["omega"][a])
        assert rc == 0
        for z in rng.choice(p.nzones, 6, replace=False):
            before = []
            for c in range(8):
                for f in range(3):
                    n = m["cEZ"][8 * z + c, f] - 1
                    if n > c:
                        aez = g["A_ez"][8 * z + c, f] @ q["omega"][a]
                        if aez > 0:
                            before.append((c, n))
                        elif aez < 0:
                            before.append((n, c))
            ok = np.ones(len(perms), dtype=bool)
            for u, d in before:
                ok &= pos[:, u] < pos[:, d]
            best = perms[ok][0]  # permutations() yields lexicographic order
            np.testing.assert_array_equal(nextC[8 * z:8 * z + 8] - 1, best)


def test_warped_face_puts_downstream_zone_on_cycle_list(prob):
    """Flip one corner-face area so a face has mixed Omega.A signs: snneed
    (snac/snneed.F90:165-186) keeps the face-summed dependency and puts the
    downstream zone's corners on the cycle list, since one of its corners
    feeds the upstream zone."""
    p, m, g, q = prob
    A_fp = g["A_fp"].copy()
    om = q["omega"][0]
    for c in range(p.ncorner):
        z = c // 8
        for f in range(3):
            face = m["CToFace"][c, f]
            zo = m["zoneOpp"][z, face - 1]
            # snneed reads the face from its lower-numbered zone
            if zo > z + 1 and abs(A_fp[c, f] @ om) > 0:
                out = A_fp[c, f] @ om > 0  # zone z is upstream across this face
                A_fp[c, f] *= -0.5  # opposite sign, face sum keeps its sign
                rc, nextZ, _, _ = O.schedule_angle(m, A_fp, g["A_ez"], om)
                assert rc == 2  # a cycle list: the plain call refuses
                rc, nextZ2, _, _, cyc = O.schedule_angle(m, A_fp, g["A_ez"], om, cycles=True)
                assert rc == 0
                down = zo - 1 if out else z
                np.testing.assert_array_equal(cyc, 8 * down + np.arange(1, 9))
                return
    pytest.fail("no interior corner face found")
