"""ORACLE — test infrastructure only. ctypes binding of the C restatement of
the reference sweep path (oracle.c). Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg may import this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_pi = C.POINTER(C.c_int)
_pd = C.POINTER(C.c_double)
_i, _d = C.c_int, C.c_double

_SIG = {
    "oracle_geometry": (_i, [_i, _i, _pi, _pi, _pi, _pi, _pi, _pi, _pi, _pi, _pd, _pd, _pd, _pd]),
    "oracle_quadrature": (_i, [_i, _i, _i, _pd, _pd, _pd, _pd, _pd, _pd]),
    "oracle_schedule_angle": (_i, [_i, _i, _pi, _pi, _pi, _pi, _pi, _pi, _pi, _pi, _pd, _pd, _pd,
                                   _pi, _pi, _pi, _pi]),
    "oracle_schedule_angle_tol": (_i, [_i, _i, _pi, _pi, _pi, _pi, _pi, _pi, _pi, _pi, _pd, _pd,
                                       _pd, _d, _pi, _pi, _pi, _pi]),
    "oracle_sweep_angle": (None, [_i, _i, _i, _i, _pi, _pi, _pi, _pi, _pi, _pd, _pd, _pd, _pd,
                                  _pd, _d, _pd, _pd, _d, _i, _pi, _pi, _pi, _pd, _pd, _pd]),
    "oracle_sweep_all": (None, [_i, _i, _i, _i, _i, _pi, _pi, _pi, _pi, _pi, _pd, _pd, _pd, _pd,
                                _pd, _pd, _pd, _pi, _pi, _pi, _pi, _i, _pd, _pd]),
    "oracle_sweep_all_tol": (None, [_i, _i, _i, _i, _i, _pi, _pi, _pi, _pi, _pi, _pd, _pd, _pd,
                                    _pd, _pd, _pd, _pd, _pi, _pi, _pi, _pi, _i, _d, _pd, _pd]),
    "oracle_source_iteration": (_i, [_i, _i, _i, _i, _i, _pi, _pi, _pi, _pi, _pi, _pd, _pd, _pd,
                                     _pd, _pd, _pd, _pd, _pd, _pd, _pi, _pi, _pi, _pi, _i, _d, _i,
                                     _pd, _pd, _pd]),
    "oracle_sweep_angle_cyc": (None, [_i, _i, _i, _i, _pi, _pi, _pi, _pi, _pi, _pd, _pd, _pd,
                                      _pd, _pd, _d, _pd, _pd, _d, _i, _pi, _pi, _pi, _i, _pi, _pd,
                                      _pd, _pd, _pd]),
    "oracle_schedule_angle_cyc": (_i, [_i, _i, _pi, _pi, _pi, _pi, _pi, _pi, _pi, _pi, _pd, _pd,
                                       _pd, _d, _pi, _pi, _pi, _pi, _pi, _pi]),
    "oracle_sweep_all_cyc": (None, [_i, _i, _i, _i, _i, _pi, _pi, _pi, _pi, _pi, _pd, _pd, _pd,
                                    _pd, _pd, _pd, _pd, _pi, _pi, _pi, _pi, _i, _d, _pi, _pi, _pd,
                                    _pd, _pd]),
    "oracle_source_iteration_cyc": (_i, [_i, _i, _i, _i, _i, _pi, _pi, _pi, _pi, _pi, _pd, _pd,
                                         _pd, _pd, _pd, _pd, _pd, _pd, _pd, _pi, _pi, _pi, _pi,
                                         _pi, _pi, _pd, _i, _d, _i, _pd, _pd, _pd]),
}
_libs = {}


# Leading dimensions (corner faces, zone faces) of the polyhedral build.
POLY_MAXCF, POLY_MAXCORNER, POLY_MAXFACES = 6, 12, 8


def load(fast=False, poly=False):
    """liboracle.so (parity checker), liboracle_fast.so (timed CPU baseline)
    or liboracle_poly.so (the checker for polyhedral zones: corner-face and
    zone-face arrays with leading dimensions POLY_MAXCF and POLY_MAXFACES)."""
    name = "liboracle_poly.so" if poly else ("liboracle_fast.so" if fast else "liboracle.so")
    if name in _libs:
        return _libs[name]
    path = os.path.join(_HERE, name)
    if not os.path.exists(path):
        subprocess.run(["make", "-C", _HERE, name], check=True, capture_output=True)
    lib = C.CDLL(path)
    for fn, (res, args) in _SIG.items():
        f = getattr(lib, fn)
        f.restype, f.argtypes = res, args
    _libs[name] = lib
    return lib


def _ip(a):
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_pi)


def _dp(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_pd)


def _lib_for(mesh):
    """The build whose array dimensions the mesh arrays have: hex (corner
    faces 3, zone faces 6) or polyhedral (POLY_MAXCF, POLY_MAXFACES)."""
    maxcf, maxfaces = mesh["cFP"].shape[1], mesh["zoneOpp"].shape[1]
    if (maxcf, maxfaces) == (3, 6):
        return load()
    assert (maxcf, maxfaces) == (POLY_MAXCF, POLY_MAXFACES), (maxcf, maxfaces)
    assert int(mesh["numCorner"].max()) <= POLY_MAXCORNER
    return load(poly=True)


def geometry(mesh):
    """A_fp (C,maxcf,3), A_ez (C,maxcf,3), Volume (C) from mesh arrays."""
    Cn = mesh["nCFaces"].shape[0]
    Z = mesh["numCorner"].shape[0]
    maxcf = mesh["cFP"].shape[1]
    A_fp, A_ez, V = np.zeros((Cn, maxcf, 3)), np.zeros((Cn, maxcf, 3)), np.zeros(Cn)
    _lib_for(mesh).oracle_geometry(Z, Cn, _ip(mesh["numCorner"]), _ip(mesh["cOffSet"]),
                           _ip(mesh["zoneFaces"]), _ip(mesh["nCFaces"]), _ip(mesh["cFP"]),
                           _ip(mesh["cEZ"]), _ip(mesh["CToFace"]), _ip(mesh["zoneOpp"]),
                           _dp(mesh["px"]), _dp(A_fp), _dp(A_ez), _dp(V))
    return A_fp, A_ez, V


def quadrature(npolar, nazimuthal, cos_theta, w_theta, cos_phi, w_phi, polaraxis=1):
    A = 8 * npolar * nazimuthal
    omega, weight = np.zeros((A, 3)), np.zeros(A)
    rc = load().oracle_quadrature(npolar, nazimuthal, polaraxis,
                                  _dp(np.ascontiguousarray(cos_theta, np.float64)),
                                  _dp(np.ascontiguousarray(w_theta, np.float64)),
                                  _dp(np.ascontiguousarray(cos_phi, np.float64)),
                                  _dp(np.ascontiguousarray(w_phi, np.float64)), _dp(omega),
                                  _dp(weight))
    assert rc == 0
    return omega, weight


def schedule_angle(mesh, A_fp, A_ez, omega, par_tol=0.0, cycles=False):
    """(rc, nextZ (Z), nextC (C), zonesInPlane (nHP)) for one direction; with
    cycles=True also the cycle list (1-based corners) as a fifth entry, and
    rc is 0 for cyclic schedules too."""
    Z = mesh["numCorner"].shape[0]
    Cn = mesh["nCFaces"].shape[0]
    nextZ, nextC = np.zeros(Z, np.int32), np.zeros(Cn, np.int32)
    zip_ = np.zeros(Z, np.int32)
    nhp = C.c_int(0)
    args = (Z, Cn, _ip(mesh["numCorner"]), _ip(mesh["cOffSet"]), _ip(mesh["zoneFaces"]),
            _ip(mesh["nCFaces"]), _ip(mesh["cEZ"]), _ip(mesh["CToFace"]), _ip(mesh["zoneOpp"]),
            _ip(mesh["faceOpp"]), _dp(A_fp), _dp(A_ez), _dp(np.ascontiguousarray(omega, np.float64)),
            par_tol, _ip(nextZ), _ip(nextC), _ip(zip_), C.byref(nhp))
    lib = _lib_for(mesh)
    if not cycles:
        rc = lib.oracle_schedule_angle_tol(*args)
        return rc, nextZ, nextC, zip_[:nhp.value].copy()
    cyc = np.zeros(Cn, np.int32)
    nc = C.c_int(0)
    rc = lib.oracle_schedule_angle_cyc(*args, _ip(cyc), C.byref(nc))
    return rc, nextZ, nextC, zip_[:nhp.value].copy(), cyc[:nc.value].copy()


def schedule(mesh, A_fp, A_ez, omega_all, par_tol=0.0):
    """Per-angle schedule for every direction: dict like TiledProblem.schedule(),
    with the cycle lists (numCycles (A), cycleList concatenated, 1-based)."""
    A = omega_all.shape[0]
    nh, zips, nz, nc, cyc = [], [], [], [], []
    for a in range(A):
        rc, nextZ, nextC, zip_, cl = schedule_angle(mesh, A_fp, A_ez, omega_all[a], par_tol,
                                                    cycles=True)
        if rc:
            raise RuntimeError(f"oracle schedule failed for angle {a + 1}: code {rc}")
        nh.append(len(zip_))
        zips.append(zip_)
        nz.append(nextZ)
        nc.append(nextC)
        cyc.append(cl)
    return dict(nHyperPlanes=np.array(nh, np.int32), zonesInPlane=np.concatenate(zips),
                nextZ=np.stack(nz), nextC=np.stack(nc),
                numCycles=np.array([len(c) for c in cyc], np.int32),
                cycleList=np.concatenate(cyc).astype(np.int32))


def _cycle_args(sched, G, cycle_psi):
    """(numCycles, cycleList, cyclePsi) pointers; cyclePsi (total, G) is
    created zero when not given and updated in place."""
    ncyc = sched.get("numCycles")
    if ncyc is None or int(np.sum(ncyc)) == 0:
        return None, None, None, None
    total = int(np.sum(ncyc))
    if cycle_psi is None:
        cycle_psi = np.zeros((total, G))
    assert cycle_psi.shape == (total, G) and cycle_psi.flags["C_CONTIGUOUS"]
    return (_ip(np.ascontiguousarray(ncyc, np.int32)),
            _ip(np.ascontiguousarray(sched["cycleList"], np.int32)), _dp(cycle_psi), cycle_psi)


def sweep_angle(mesh, geom, STotal, Sigt, omega, quadwt, nextZ, nextC, zonesInPlane, Psi1, PsiB,
                Phi, tau=0.0, Psi=None, cycle_list=None, cycle_psi=None):
    """One angle, in place on Psi1 (C+nb,G), PsiB (nb,G), Phi (C,G); with
    this angle's cycle list (1-based corners) and its lagged psi cycle_psi
    (n, G), updated in place."""
    Z = mesh["numCorner"].shape[0]
    Cn = mesh["nCFaces"].shape[0]
    G = STotal.shape[1]
    nb = PsiB.shape[0]
    n = 0 if cycle_list is None else len(cycle_list)
    if n:
        assert cycle_psi.shape == (n, G) and cycle_psi.flags["C_CONTIGUOUS"]
    _lib_for(mesh).oracle_sweep_angle_cyc(
        Z, Cn, nb, G, _ip(mesh["numCorner"]), _ip(mesh["cOffSet"]), _ip(mesh["nCFaces"]),
        _ip(mesh["cFP"]), _ip(mesh["cEZ"]), _dp(geom["A_fp"]), _dp(geom["A_ez"]),
        _dp(geom["Volume"]), _dp(STotal), _dp(Sigt), tau, _dp(Psi),
        _dp(np.ascontiguousarray(omega, np.float64)), quadwt, len(zonesInPlane),
        _ip(np.ascontiguousarray(zonesInPlane, np.int32)), _ip(np.ascontiguousarray(nextZ, np.int32)),
        _ip(np.ascontiguousarray(nextC, np.int32)), n,
        _ip(np.ascontiguousarray(cycle_list, np.int32)) if n else None,
        _dp(cycle_psi) if n else None, _dp(Psi1), _dp(PsiB), _dp(Phi))


def sweep_all(mesh, geom, quad, sched, STotal, Sigt, nthreads=1, PsiB=None, fast=False,
              par_tol=0.0, cycle_psi=None):
    """All angles (SetSweep semantics). Returns (Phi (C,G), PsiB (A,nb,G)).
    par_tol > 0: corner faces within par_tol of parallel to Omega count as
    parallel in the sweep (use with a schedule built with the same par_tol).
    A schedule with cycle lists reads and updates cycle_psi (total, G) in
    place (zero when not given: the first sweep's lagged values)."""
    Z = mesh["numCorner"].shape[0]
    Cn = mesh["nCFaces"].shape[0]
    G = STotal.shape[1]
    A = quad["weight"].shape[0]
    nb = int(np.count_nonzero(mesh["cFP"] > Cn))
    if PsiB is None:
        PsiB = np.zeros((A, nb, G))
    Phi = np.zeros((Cn, G))
    ncp, clp, cpp, _ = _cycle_args(sched, G, cycle_psi)
    load(fast).oracle_sweep_all_cyc(
        Z, Cn, nb, G, A, _ip(mesh["numCorner"]), _ip(mesh["cOffSet"]), _ip(mesh["nCFaces"]),
        _ip(mesh["cFP"]), _ip(mesh["cEZ"]), _dp(geom["A_fp"]), _dp(geom["A_ez"]),
        _dp(geom["Volume"]), _dp(np.ascontiguousarray(STotal)), _dp(np.ascontiguousarray(Sigt)),
        _dp(quad["omega"]), _dp(quad["weight"]), _ip(sched["nHyperPlanes"]),
        _ip(sched["zonesInPlane"]), _ip(sched["nextZ"]), _ip(sched["nextC"]), nthreads, par_tol,
        ncp, clp, cpp, _dp(PsiB), _dp(Phi))
    return Phi, PsiB


def source_iteration(mesh, geom, quad, sched, Sigt, sigs, q, tol, max_iters, phi0=None,
                     nthreads=1, fast=False):
    """Returns (Phi (C,G), iterations, residual, history)."""
    Z = mesh["numCorner"].shape[0]
    Cn = mesh["nCFaces"].shape[0]
    G = Sigt.shape[1]
    A = quad["weight"].shape[0]
    nb = int(np.count_nonzero(mesh["cFP"] > Cn))
    Phi = np.zeros((Cn, G)) if phi0 is None else np.array(phi0, dtype=np.float64, order="C")
    res = C.c_double(0.0)
    hist = np.zeros(max(max_iters, 1))
    ncp, clp, cpp, _ = _cycle_args(sched, G, None)
    it = load(fast).oracle_source_iteration_cyc(
        Z, Cn, nb, G, A, _ip(mesh["numCorner"]), _ip(mesh["cOffSet"]), _ip(mesh["nCFaces"]),
        _ip(mesh["cFP"]), _ip(mesh["cEZ"]), _dp(geom["A_fp"]), _dp(geom["A_ez"]),
        _dp(geom["Volume"]), _dp(geom["VolumeZone"]), _dp(np.ascontiguousarray(Sigt)),
        _dp(np.ascontiguousarray(sigs)), _dp(np.ascontiguousarray(q)), _dp(quad["omega"]),
        _dp(quad["weight"]), _ip(sched["nHyperPlanes"]), _ip(sched["zonesInPlane"]),
        _ip(sched["nextZ"]), _ip(sched["nextC"]), ncp, clp, cpp, nthreads, tol, max_iters,
        _dp(Phi), C.byref(res), _dp(hist))
    return Phi, it, res.value, hist[:it].copy()


def time_step(mesh, geom, quad, sched, sigt, sigs, q, tau, psi_prev, tol, max_iters, phi0=None):
    """One implicit time step: source iteration with the sweep of
    snac/SweepUCBxyz.F90:119-130 for tau > 0 (total Sigt + tau,
    control/setTotalOpacity.F90; source STotal + tau Psi(:,:,Angle) with Psi
    the previous cycle's psi_prev (A, C, G)), residual as in
    source_iteration, then the cycle's final sweep with savePsi
    (rt/rtmainsn.F90:238-246). Angles are swept one after another.
    Returns (phi (C, G), psi_new (A, C, G), iterations, residual, history)."""
    Cn = mesh["nCFaces"].shape[0]
    G = sigt.shape[1]
    A = quad["weight"].shape[0]
    nb = int(np.count_nonzero(mesh["cFP"] > Cn))
    V, VZ = geom["Volume"], geom["VolumeZone"]
    st = np.ascontiguousarray(sigt + tau)
    sigs_c, q_c = np.repeat(sigs, 8, axis=0), np.repeat(q, 8, axis=0)
    starts = np.concatenate([[0], np.cumsum(sched["nHyperPlanes"])])
    psi_prev = np.ascontiguousarray(psi_prev, dtype=np.float64)
    ncyc = sched.get("numCycles", np.zeros(A, np.int32))
    coff = np.concatenate([[0], np.cumsum(ncyc)])
    clist = sched.get("cycleList", np.zeros(0, np.int32))
    # initCyclePsi (control/constructDynMemory.F90:75-114): lagged values
    # start from the previous cycle's psi
    cyc_psi = [np.ascontiguousarray(psi_prev[a][clist[coff[a]:coff[a + 1]] - 1]) for a in range(A)]

    def sweep(phi, save):
        S = np.ascontiguousarray((sigs_c * phi + q_c) / (4.0 * np.pi))
        Phi = np.zeros((Cn, G))
        new = np.zeros((A, Cn, G)) if save else None
        for a in range(A):
            Psi1 = np.zeros((Cn + nb, G))
            PsiB = np.zeros((nb, G))
            sweep_angle(mesh, geom, S, st, quad["omega"][a], quad["weight"][a], sched["nextZ"][a],
                        sched["nextC"][a], sched["zonesInPlane"][starts[a]:starts[a + 1]], Psi1,
                        PsiB, Phi, tau=tau, Psi=np.ascontiguousarray(psi_prev[a]),
                        cycle_list=clist[coff[a]:coff[a + 1]], cycle_psi=cyc_psi[a])
            if save:
                new[a] = Psi1[:Cn]
        return Phi, new

    def zavg(x):
        return (x * V[:, None]).reshape(-1, 8, G).sum(1) / VZ[:, None]

    phi = np.zeros((Cn, G)) if phi0 is None else np.array(phi0, dtype=np.float64)
    old = zavg(phi)
    hist, res, it = [], 1.0, 0
    while it < max_iters:
        phi, _ = sweep(phi, False)
        cur = zavg(phi)
        mask = cur > 1e-6 * cur.max()
        res = float(np.max(np.abs((cur[mask] - old[mask]) / cur[mask]))) if mask.any() else 0.0
        old = cur
        hist.append(res)
        it += 1
        if res < tol:
            break
    phi, new = sweep(phi, True)
    return phi, new, it, res, np.array(hist)
