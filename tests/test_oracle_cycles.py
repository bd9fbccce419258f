"""Dependency cycles on distorted hex meshes: the oracle's restatement of the
reference's cycle handling (snac/snneed.F90 mixed faces, snac/findseeds.F90,
snac/getDownStreamData.F90 + snac/fixZone.F90 cyclic zones,
snac/cyclebreaker.F90 + snac/sccsearch.F90 zone cycles; lagged cycle psi,
control/constructDynMemory.F90:118-189) against what the mathematics fixes,
and the product's schedule builder against the oracle.

The reference has no test or fixture for this path. Pins:
  - every psi a zone reads before the serial sweep has computed it belongs
    to a corner on that angle's cycle list, and cyclic zones have all their
    corners listed (otherwise the sweep would read another angle's values);
  - the zone dependency graphs of these meshes do have strongly connected
    components (scipy finds them independently), so the Tarjan breaker runs;
  - repeated lagged sweeps converge to the solution of the coupled corner
    equations of all zones, assembled and solved densely here without any
    sweep order: the fixed point does not depend on which links were broken,
    and a planted mistake (lagged values never updated) misses it.
The exact break choices (which links, in which order) follow the reference's
search order and are pinned only by the product agreeing with the oracle;
a verbatim run of the reference (unbuildable here) would pin them further.
"""
import numpy as np
import pytest
from scipy.sparse import csr_matrix
from scipy.sparse.csgraph import connected_components

from conftest import CYCLIC_MESHES, cyclic_problem
from oracle import oracle as O


@pytest.fixture(scope="module", params=CYCLIC_MESHES, ids=lambda p: f"n{p[0]}-a{p[1]}-t{p[3]}")
def cyc(request):
    p = cyclic_problem(*request.param)
    return p, p.mesh(), p.geometry(), p.quadrature(), p.schedule()


def test_product_schedule_equals_oracle(cyc):
    p, m, g, q, s = cyc
    ref = O.schedule(m, g["A_fp"], g["A_ez"], q["omega"])
    for k, v in ref.items():
        np.testing.assert_array_equal(s[k], v, err_msg=k)
    assert len(s["cycleList"]) > 0


def _plane_and_position(s, a, Z):
    poff = np.concatenate([[0], np.cumsum(s["nHyperPlanes"])])
    nz = np.abs(s["nextZ"][a]) - 1
    pos = np.empty(Z, int)
    pos[nz] = np.arange(Z)
    plane = np.empty(Z, int)
    k = 0
    for h, cnt in enumerate(s["zonesInPlane"][poff[a]:poff[a + 1]]):
        plane[nz[k:k + cnt]] = h
        k += cnt
    return pos, plane, nz


def test_every_lagged_read_is_on_the_cycle_list(cyc):
    p, m, g, q, s = cyc
    Z, Cn = p.nzones, p.ncorner
    off = np.concatenate([[0], np.cumsum(s["numCycles"])])
    lagged = 0
    for a in range(p.nangles):
        pos, plane, nz = _plane_and_position(s, a, Z)
        assert sorted(nz.tolist()) == list(range(Z))  # every zone once
        assert s["zonesInPlane"].min() > 0
        cyc_set = set((s["cycleList"][off[a]:off[a + 1]] - 1).tolist())
        dots = np.einsum("cfk,k->cf", g["A_fp"], q["omega"][a])
        for c in range(Cn):
            for f in range(3):
                up = m["cFP"][c, f] - 1
                if up < Cn and dots[c, f] < 0 and pos[up // 8] > pos[c // 8]:
                    assert up in cyc_set, (a, c, up)
                    lagged += 1
        # cyclic zones (nextZ < 0) keep the natural corner order, all corners listed
        for z in np.where(s["nextZ"][a] < 0)[0]:
            zone = -s["nextZ"][a][z] - 1
            assert set(range(8 * zone, 8 * zone + 8)) <= cyc_set
            np.testing.assert_array_equal(s["nextC"][a][8 * zone:8 * zone + 8], np.arange(1, 9))
    assert lagged > 0


def test_zone_graphs_have_strong_components(cyc):
    """The face-summed zone dependency graph (snneed's exit faces) has cycles
    in some direction, found here by scipy's SCC search."""
    p, m, g, q, s = cyc
    Z = p.nzones
    found = 0
    for a in range(p.nangles):
        dots = np.einsum("cfk,k->cf", g["A_fp"], q["omega"][a])
        flow = {}
        for c in range(p.ncorner):
            for f in range(3):
                up = m["cFP"][c, f] - 1
                if up < p.ncorner:
                    key = (c // 8, up // 8)
                    flow[key] = flow.get(key, 0.0) + dots[c, f]  # > 0: leaves c's zone
        rows, cols = [], []
        for (z, zn), v in flow.items():
            if v > 0:
                rows.append(z)
                cols.append(zn)
        G = csr_matrix((np.ones(len(rows)), (rows, cols)), shape=(Z, Z))
        ncomp, lab = connected_components(G, directed=True, connection="strong")
        found += int((np.bincount(lab) > 1).sum())
    if s["nextZ"].min() > 0 and found == 0:
        pytest.skip("this mesh has only warped faces")
    assert found > 0


def dense_psi(m, g, omega, S, sig):
    """Corner equations of every zone coupled, for one direction and group,
    vacuum boundaries: psi_c (outgoing areas + sigma V_c) = V_c S_c + incoming
    FP-face fluxes + the EZ coupling terms + upwind EZ neighbours
    (snac/SweepUCBxyz.F90:119-306 written as one linear system, no sweep)."""
    Cn = m["nCFaces"].shape[0]
    M = np.zeros((Cn, Cn))
    b = np.zeros(Cn)
    Afp, Aez, V = g["A_fp"], g["A_ez"], g["Volume"]
    cFP, cEZ = m["cFP"] - 1, m["cEZ"] - 1
    for c in range(Cn):
        z = c // 8
        s_, v = sig[z], V[c]
        afp, aez = Afp[c] @ omega, Aez[c] @ omega
        M[c, c] += s_ * v + afp[afp > 0].sum() + aez[aez > 0].sum()
        b[c] += v * S[c]
        for f in range(3):
            if afp[f] < 0 and cFP[c, f] < Cn:
                M[c, cFP[c, f]] += afp[f]
        for f in range(3):
            nb = 8 * z + cEZ[c, f]
            if aez[f] < 0:
                M[c, nb] += aez[f]
            elif aez[f] > 0:
                A, fo = aez[f], (f + 1) % 3
                Qu, Qd = S[c], S[nb]
                if afp[fo] < 0:
                    sv = s_ * v
                    gnum = A * A * (1.82 * sv * sv + A * (4 * sv + 3 * A))
                    gden = v * (4 * sv ** 3 + A * (6 * sv * sv + 2 * A * (2 * sv + A)))
                    den = gnum + gden * s_
                    const = (-v * gnum * Qu + 0.5 * A * gden * (Qu - Qd)) / den
                    b[c] += const
                    b[nb] -= const
                    up = cFP[c, fo]
                    if up < Cn:  # sez carries v gnum sig psi_opp / den
                        M[c, up] -= v * gnum * s_ / den
                        M[nb, up] += v * gnum * s_ / den
                else:
                    const = 0.5 * A / s_ * (Qu - Qd)
                    b[c] += const
                    b[nb] -= const
    return np.linalg.solve(M, b)


@pytest.mark.parametrize("mesh", [(3, 0.45, 3, 0.0), (4, 0.2, 4, 1.5)])
def test_lagged_sweeps_converge_to_coupled_solution(mesh):
    p = cyclic_problem(*mesh)
    m, g, q, s = p.mesh(), p.geometry(), p.quadrature(), p.schedule()
    assert len(s["cycleList"]) > 0
    rng = np.random.default_rng(0)
    n = mesh[0]
    sig = rng.uniform(2.0, 8.0, p.nzones) * n
    S = rng.uniform(0.1, 1.0, p.ncorner)
    ref = sum(q["weight"][a] * dense_psi(m, g, q["omega"][a], S, sig) for a in range(p.nangles))
    St, sg = S[:, None].copy(), sig[:, None].copy()
    cp = np.zeros((len(s["cycleList"]), 1))
    prev = None
    for it in range(200):
        phi, _ = O.sweep_all(m, g, q, s, St, sg, cycle_psi=cp)
        if prev is not None and np.abs(phi - prev).max() <= 1e-15 * np.abs(phi).max():
            break
        prev = phi
    assert it < 199
    err = np.abs(phi[:, 0] - ref).max() / np.abs(ref).max()
    assert err < 1e-13, err
    # planted mistake: the lagged values are never carried to the next sweep
    for _ in range(3):
        bad, _ = O.sweep_all(m, g, q, s, St, sg, cycle_psi=np.zeros_like(cp))
    assert np.abs(bad[:, 0] - ref).max() / np.abs(ref).max() > 1e-4


def test_source_iteration_with_cycles_reaches_balance():
    """Source iteration on a cyclic mesh converges, and its fixed point is
    the coupled solution with the scattering source S = (sigs phi + q)/4pi."""
    p = cyclic_problem(3, 0.45, 3, 0.0)
    m, g, q, s = p.mesh(), p.geometry(), p.quadrature(), p.schedule()
    rng = np.random.default_rng(1)
    sig = rng.uniform(6.0, 20.0, (p.nzones, 1))
    sigs = 0.5 * sig
    qq = rng.uniform(0.1, 1.0, (p.nzones, 1))
    phi, it, res, _ = O.source_iteration(m, g, q, s, sig, sigs, qq, 1e-13, 500)
    assert res < 1e-13
    S = (np.repeat(sigs, 8, 0) * phi + np.repeat(qq, 8, 0)) / (4 * np.pi)
    ref = sum(q["weight"][a] * dense_psi(m, g, q["omega"][a], S[:, 0], sig[:, 0])
              for a in range(p.nangles))
    assert np.abs(phi[:, 0] - ref).max() / np.abs(ref).max() < 1e-11
