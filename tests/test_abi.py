"""C-ABI of libumt_amd.so without a GPU: the library loads, exports every
function include/umt_sweep.h declares, builds problems on the host, and
rejects bad input with a status code and message before any device call."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import umt_amd as U
from umt_amd import _lib

INCLUDE = os.path.join(os.path.dirname(os.path.dirname(__file__)), "include")


def _declared_functions():
    text = "".join(open(os.path.join(INCLUDE, f)).read() for f in sorted(os.listdir(INCLUDE))
                   if f.endswith(".h"))
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(\w+)\s*\(", text, flags=re.M)
    return sorted(set(n for n in names if n not in ("if", "while")))


def test_every_declared_symbol_is_exported():
    lib = _lib.load()
    names = _declared_functions()
    assert "gpu_sweepucbxyz" in names and "umt_ctx_create" in names
    assert "umt_rz_ctx_create" in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python binding covers the whole header
    assert set(names) <= set(_lib.SIGNATURES), set(names) - set(_lib.SIGNATURES)


def test_problem_dims_match_tile_counts():
    p = U.TiledProblem(3, 2, 1, 2, 3)
    assert p.nzones == 24 * 6
    assert p.ncorner == 8 * p.nzones
    assert p.nangles == 8 * 2 * 3
    # boundary corner faces: 4 per boundary face; top/bottom 24 faces per tile,
    # each tile edge of length 20 carries 4 boundary faces per layer
    bfaces = 2 * 24 * 3 * 2 + 2 * (3 + 2) * 4 * 1
    assert p.nbelem == 4 * bfaces
    m = p.mesh()
    assert np.count_nonzero(m["cFP"] > p.ncorner) == p.nbelem
    assert sorted(m["cFP"][m["cFP"] > p.ncorner].tolist()) == list(
        range(p.ncorner + 1, p.ncorner + p.nbelem + 1))


def test_problem_rejects_bad_quadrature_order():
    lib = _lib.load()
    h = C.c_void_p()
    rc = lib.umt_problem_tiled(1, 1, 1, 7, 2, 1, C.byref(h))
    assert rc != 0
    assert b"npolar" in lib.umt_last_error()


def test_cross_sections_are_reproducible_and_in_range():
    a = U.draw_cross_sections(5, 100, 4, (1e-2, 1e3), (0.0, 0.95), (0.0, 1.0))
    b = U.draw_cross_sections(5, 100, 4, (1e-2, 1e3), (0.0, 0.95), (0.0, 1.0))
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)
    st, ss, q = a
    assert st.min() >= 1e-2 and st.max() <= 1e3
    c = ss / st
    assert c.min() >= 0 and c.max() <= 0.95
    # log-uniform: log10(sigma_t) spread over five decades
    assert np.log10(st).std() > 1.0


def _ctx_args(p, **over):
    m, g, q, s = p.mesh(), p.geometry(), p.quadrature(), p.schedule()
    arr = dict(numCorner=m["numCorner"], cOffSet=m["cOffSet"], nCFaces=m["nCFaces"],
               cFP=m["cFP"].copy(), cEZ=m["cEZ"].copy(), A_fp=g["A_fp"], A_ez=g["A_ez"],
               Volume=g["Volume"], VolumeZone=g["VolumeZone"], omega=q["omega"],
               weight=q["weight"], nHyperPlanes=s["nHyperPlanes"],
               zonesInPlane=s["zonesInPlane"], nextZ=s["nextZ"].copy())
    arr.update(over)
    return arr


def _create(p, arr, G=2):
    lib = _lib.load()
    ip = lambda a: np.ascontiguousarray(a, np.int32).ctypes.data_as(C.POINTER(C.c_int))
    dp = lambda a: np.ascontiguousarray(a, np.float64).ctypes.data_as(C.POINTER(C.c_double))
    h = C.c_void_p()
    rc = lib.umt_ctx_create(p.nzones, p.ncorner, p.nbelem, G, p.nangles, 64, ip(arr["numCorner"]),
                            ip(arr["cOffSet"]), ip(arr["nCFaces"]), ip(arr["cFP"]), ip(arr["cEZ"]),
                            dp(arr["A_fp"]), dp(arr["A_ez"]), dp(arr["Volume"]),
                            dp(arr["VolumeZone"]), dp(arr["omega"]), dp(arr["weight"]),
                            ip(arr["nHyperPlanes"]), ip(arr["zonesInPlane"]), ip(arr["nextZ"]),
                            C.byref(h))
    return rc, lib.umt_last_error().decode()


@pytest.fixture(scope="module")
def prob():
    return U.TiledProblem(1, 1, 1, 2, 2)


def test_ctx_rejects_non_hex_corner_order(prob):
    cEZ = prob.mesh()["cEZ"].copy()
    cEZ[5, [0, 1]] = cEZ[5, [1, 0]]
    rc, msg = _create(prob, _ctx_args(prob, cEZ=cEZ))
    assert rc == 13 and "corner-face order" in msg


def test_ctx_rejects_out_of_range_zone_in_schedule(prob):
    """Negative nextZ entries are cyclic zones (snnext); beyond -Z they are invalid."""
    nextZ = prob.schedule()["nextZ"].copy()
    nextZ[3, 10] = -(prob.nzones + 1)
    rc, msg = _create(prob, _ctx_args(prob, nextZ=nextZ))
    assert rc == 15 and "angle 4" in msg


def test_ctx_rejects_out_of_range_connectivity(prob):
    cFP = prob.mesh()["cFP"].copy()
    cFP[7, 2] = prob.ncorner + prob.nbelem + 1
    rc, msg = _create(prob, _ctx_args(prob, cFP=cFP))
    assert rc == 14


def test_dropin_rejects_bad_schedule_and_cycle_list_without_touching_outputs(prob):
    """gpu_sweepucbxyz returns through its status code (the reference exits)."""
    p = prob
    G = 2
    st = np.ones((p.nzones, G))
    S = np.ones((p.ncorner, G))
    Psi1 = np.zeros((p.ncorner + p.nbelem, G))
    PsiB = np.zeros((p.nbelem, G))
    Phi = np.full((p.ncorner, G), 3.0)
    s = p.schedule()
    saved = s["nextZ"][0, 0]
    s["nextZ"][0, 0] = -(p.nzones + 1)
    try:
        status = U.sweep_ucbxyz(p, 1, S, st, Psi1, PsiB, Phi)
    finally:
        s["nextZ"][0, 0] = saved
    assert status == 45
    # a cycle list naming a corner outside the mesh
    s["numCycles"] = np.array([1] + [0] * (p.nangles - 1), np.int32)
    s["cycleList"] = np.array([p.ncorner + 1], np.int32)
    try:
        status = U.sweep_ucbxyz(p, 1, S, st, Psi1, PsiB, Phi)
    finally:
        s["numCycles"] = np.zeros(p.nangles, np.int32)
        s["cycleList"] = np.zeros(0, np.int32)
    assert status == 42
    assert np.all(Phi == 3.0)


def test_dropin_refuses_polyhedral_arrays_it_cannot_take():
    """A mesh that is not the canonical hex goes to the polyhedral kernel
    after its arrays are checked; arrays it cannot take leave the outputs
    alone and set the status (before any device call)."""
    from polymesh import PolyProblem, hex_pyramid_mesh

    quad = U.TiledProblem(1, 1, 1, 2, 2).quadrature()
    p = PolyProblem(hex_pyramid_mesh(2, 2, 1), quad)
    G = 2
    st = np.ones((p.nzones, G))
    S = np.ones((p.ncorner, G))
    m, sched = p.mesh(), p.schedule()

    class Patched:
        ncorner, nbelem, nangles = p.ncorner, p.nbelem, p.nangles

        def __init__(self, mesh=None, schedule=None):
            self._m = dict(m, **(mesh or {}))
            self._s = dict(sched, **(schedule or {}))

        def mesh(self):
            return self._m

        def schedule(self):
            return self._s

        geometry = staticmethod(p.geometry)
        quadrature = staticmethod(p.quadrature)

    def status(q):
        Psi1 = np.zeros((p.ncorner + p.nbelem, G))
        Phi = np.full((p.ncorner, G), 7.0)
        rc = U.sweep_ucbxyz(q, 1, S, st, Psi1, np.zeros((p.nbelem, G)), Phi)
        assert np.all(Phi == 7.0)
        return rc, U.load().umt_last_error().decode()

    nextC = sched["nextC"].copy()
    nextC[0, 1] = nextC[0, 0]  # not an order of the zone's corners
    rc, msg = status(Patched(schedule=dict(nextC=nextC)))
    assert rc == 47 and "nextC" in msg
    cEZ = m["cEZ"].copy()
    cEZ[3, 0] = m["numCorner"][0] + 1  # EZ neighbour outside its zone
    rc, msg = status(Patched(mesh=dict(cEZ=cEZ)))
    assert rc == 13 and "EZ neighbour" in msg
    nCF = m["nCFaces"].copy()
    nCF[2] = 7  # more corner faces than maxcf
    rc, msg = status(Patched(mesh=dict(nCFaces=nCF)))
    assert rc == 12 and "corner faces" in msg
    cFP = m["cFP"].copy()
    cFP[5, 1] = p.ncorner + p.nbelem + 1
    rc, msg = status(Patched(mesh=dict(cFP=cFP)))
    assert rc == 46
