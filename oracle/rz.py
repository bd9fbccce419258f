"""ORACLE — test infrastructure only. The 2D cylindrical (RZ) sweep path:
the corner topology of quad zones (aux/setTetonZone.F90:122-142 for
ndim = 2, aux/setOppositeFace.F90:32-53) in numpy, and ctypes bindings of
the C restatement in oracle_rz.c. Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg may import this module.

Arrays are numpy views of Teton's column-major layout: (C, 2) here is
Fortran (2, C), (C, 2, 2) is A_fp(2, 2, C), 1-based index contents.
"""
import ctypes as C
import json
import os

import numpy as np

from .oracle import load

_HERE = os.path.dirname(os.path.abspath(__file__))
_pi = C.POINTER(C.c_int)
_pd = C.POINTER(C.c_double)
_i, _d = C.c_int, C.c_double

_SIG = {
    "oracle_rz_geometry": (_i, [_i, _i, _pi, _pi, _pi, _pi, _pd, _pd, _pd, _pd, _pd, _pd, _pd]),
    "oracle_rz_quadrature": (_i, [_i, _i, _pd, _pd, _pd, _pd, _pd, _pd, _pi, _pi]),
    "oracle_rz_angle_coef": (_i, [_i, _pd, _pd, _pi, _pi, _pd, _pd, _pd]),
    "oracle_rz_schedule_angle": (_i, [_i, _i, _pi, _pi, _pi, _pi, _pi, _pi, _pi, _pd, _pd, _pd,
                                      _pi, _pi, _pi, _pi]),
    "oracle_rz_sweep_all": (None, [_i, _i, _i, _i, _i, _pi, _pi, _pi, _pi, _pd, _pd, _pd, _pd,
                                   _pd, _pd, _pd, _pd, _pd, _pd, _pi, _pi, _pd, _pd, _pd, _pi,
                                   _pi, _pi, _pi, _d, _pd, _i, _pd, _pd, _pd]),
    "oracle_rz_source_iteration": (_i, [_i, _i, _i, _i, _i, _pi, _pi, _pi, _pi, _pd, _pd, _pd,
                                        _pd, _pd, _pd, _pd, _pd, _pd, _pd, _pd, _pd, _pi, _pi,
                                        _pd, _pd, _pd, _pi, _pi, _pi, _pi, _i, _d, _i, _pd, _pd,
                                        _pd]),
}


def lib(fast=False):
    L = load(fast)
    if not getattr(L, "_rz_bound", False):
        for fn, (res, args) in _SIG.items():
            f = getattr(L, fn)
            f.restype, f.argtypes = res, args
        L._rz_bound = True
    return L


def _ip(a):
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_pi)


def _dp(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_pd)


def structured(nr, nz, r0=0.0, r1=1.0, z0=0.0, z1=1.0, jitter=0.0, seed=0):
    """Points (N, 2) as (r, z) of an nr x nz grid on [r0, r1] x [z0, z1] and
    quads (Z, 4) listing each zone's nodes clockwise in the (r, z) plane,
    zones r-fastest. Interior nodes move by up to jitter times the cell size."""
    r = np.linspace(r0, r1, nr + 1)
    z = np.linspace(z0, z1, nz + 1)
    R, Zg = np.meshgrid(r, z, indexing="xy")  # (nz+1, nr+1), r fastest
    pts = np.stack([R.ravel(), Zg.ravel()], 1)
    if jitter:
        rng = np.random.default_rng(seed)
        i, j = np.meshgrid(np.arange(nr + 1), np.arange(nz + 1), indexing="xy")
        inside = ((i > 0) & (i < nr) & (j > 0) & (j < nz)).ravel()
        d = np.array([(r1 - r0) / nr, (z1 - z0) / nz])
        pts[inside] += rng.uniform(-jitter, jitter, (int(inside.sum()), 2)) * d
    node = lambda i, j: j * (nr + 1) + i  # noqa: E731
    quads = [[node(i, j), node(i, j + 1), node(i + 1, j + 1), node(i + 1, j)]
             for j in range(nz) for i in range(nr)]  # clockwise: up, right, down
    return pts, np.array(quads, np.int32)


def quad_mesh(points, quads):
    """Teton corner topology of a quad mesh: setTetonZone's 2D branch with
    faces (edges) k = (node k, node k+1) of each zone's clockwise node list,
    boundary elements numbered per boundary edge in zone order (one vacuum
    boundary), setOppositeFace's opposite faces."""
    points = np.asarray(points, np.float64)
    quads = np.asarray(quads, np.int32)
    Z = quads.shape[0]
    Cn = 4 * Z
    edges = {}
    for z in range(Z):
        for f in range(4):
            a, b = quads[z, f], quads[z, (f + 1) % 4]
            edges.setdefault((min(a, b), max(a, b)), []).append((z, f))
    zoneOpp = np.full((Z, 4), -1, np.int32)
    for key, lst in edges.items():
        if len(lst) == 2:
            (z1, f1), (z2, f2) = lst
            zoneOpp[z1, f1] = z2 + 1
            zoneOpp[z2, f2] = z1 + 1
        elif len(lst) > 2:
            raise ValueError("edge shared by more than two zones")
    faceOpp = np.zeros((Z, 4), np.int32)
    for z in range(Z):
        for f in range(4):
            zo = zoneOpp[z, f]
            if zo < 0:
                faceOpp[z, f] = -1
            else:
                for f2 in range(4):
                    if zoneOpp[zo - 1, f2] == z + 1:
                        faceOpp[z, f] = f2 + 1
    cFP = np.zeros((Cn, 2), np.int32)
    cEZ = np.zeros((Cn, 2), np.int32)
    CToFace = np.zeros((Cn, 2), np.int32)
    nb = 0
    corner_at = lambda z, node: 4 * z + int(np.where(quads[z] == node)[0][0]) + 1  # noqa: E731
    for z in range(Z):
        for f in range(4):
            l1, l2 = f + 1, (f + 1) % 4 + 1  # local corners (1-based) of the face
            a, b = quads[z, f], quads[z, (f + 1) % 4]
            if zoneOpp[z, f] < 0:
                opp1, opp2 = Cn + nb + 1, Cn + nb + 2
                nb += 2
            else:
                zo = zoneOpp[z, f] - 1
                opp1, opp2 = corner_at(zo, a), corner_at(zo, b)
            cFP[4 * z + l1 - 1, 1] = opp1
            cFP[4 * z + l2 - 1, 0] = opp2
            cEZ[4 * z + l1 - 1, 0] = l2
            cEZ[4 * z + l2 - 1, 1] = l1
            CToFace[4 * z + l1 - 1, 1] = f + 1
            CToFace[4 * z + l2 - 1, 0] = f + 1
    px = np.ascontiguousarray(points[quads.ravel()])
    return dict(numCorner=np.full(Z, 4, np.int32), cOffSet=(4 * np.arange(Z)).astype(np.int32),
                zoneFaces=np.full(Z, 4, np.int32), cFP=cFP, cEZ=cEZ, CToFace=CToFace,
                zoneOpp=zoneOpp, faceOpp=faceOpp, px=px, nbelem=nb)


def geometry(mesh):
    Z = mesh["numCorner"].shape[0]
    Cn = mesh["cFP"].shape[0]
    A_fp, A_ez = np.zeros((Cn, 2, 2)), np.zeros((Cn, 2, 2))
    RFP, REZ = np.zeros((Cn, 2)), np.zeros((Cn, 2))
    Area, V = np.zeros(Cn), np.zeros(Cn)
    rc = lib().oracle_rz_geometry(Z, Cn, _ip(mesh["numCorner"]), _ip(mesh["cOffSet"]),
                                  _ip(mesh["cFP"]), _ip(mesh["cEZ"]), _dp(mesh["px"]), _dp(A_fp),
                                  _dp(A_ez), _dp(RFP), _dp(REZ), _dp(Area), _dp(V))
    assert rc == 0
    return dict(A_fp=A_fp, A_ez=A_ez, RadiusFP=RFP, RadiusEZ=REZ, Area=Area, Volume=V,
                VolumeZone=V.reshape(Z, 4).sum(1))


def tables(order):
    """One order's polar and RZ azimuthal QRS45 tables from the fixture."""
    path = os.path.join(os.path.dirname(_HERE), "tests", "golden", "qrs45_n1to4.json")
    return json.load(open(path))["orders"][str(order)]


def quadrature(npolar, nazimuthal):
    tp, ta = tables(npolar), tables(nazimuthal)
    A = npolar * 2 * (2 * nazimuthal + 2)
    omega, weight = np.zeros((A, 2)), np.zeros(A)
    st, fi = np.zeros(A, np.int32), np.zeros(A, np.int32)
    arr = lambda x: np.ascontiguousarray(x, np.float64)  # noqa: E731
    rc = lib().oracle_rz_quadrature(npolar, nazimuthal, _dp(arr(tp["cosTheta"])),
                                    _dp(arr(tp["weightTheta"])), _dp(arr(ta["cosPhiRZ"])),
                                    _dp(arr(ta["weightPhiRZ"])), _dp(omega), _dp(weight), _ip(st),
                                    _ip(fi))
    assert rc == 0
    fac, w1, w2 = np.zeros(A), np.zeros(A), np.zeros(A)
    rc = lib().oracle_rz_angle_coef(A, _dp(omega), _dp(weight), _ip(st), _ip(fi), _dp(fac),
                                    _dp(w1), _dp(w2))
    if rc:
        raise RuntimeError(f"angular coefficients failed: code {rc}")
    return dict(omega=omega, weight=weight, starting=st, finishing=fi, angDerivFac=fac,
                quadTauW1=w1, quadTauW2=w2)


def schedule(mesh, geom, quad):
    """Per-direction schedule; finishing directions are not swept and have
    no hyperplanes."""
    Z = mesh["numCorner"].shape[0]
    Cn = mesh["cFP"].shape[0]
    A = quad["weight"].shape[0]
    nh, zips = np.zeros(A, np.int32), []
    nz, nc = np.zeros((A, Z), np.int32), np.zeros((A, Cn), np.int32)
    for a in range(A):
        if quad["finishing"][a]:
            continue
        zp = np.zeros(Z, np.int32)
        n = C.c_int(0)
        rc = lib().oracle_rz_schedule_angle(
            Z, Cn, _ip(mesh["numCorner"]), _ip(mesh["cOffSet"]), _ip(mesh["zoneFaces"]),
            _ip(mesh["cEZ"]), _ip(mesh["CToFace"]), _ip(mesh["zoneOpp"]), _ip(mesh["faceOpp"]),
            _dp(geom["A_fp"]), _dp(geom["A_ez"]), _dp(np.ascontiguousarray(quad["omega"][a])),
            _ip(nz[a]), _ip(nc[a]), _ip(zp), C.byref(n))
        if rc:
            raise RuntimeError(f"oracle RZ schedule failed for angle {a + 1}: code {rc}")
        nh[a] = n.value
        zips.append(zp[:n.value])
    return dict(nHyperPlanes=nh, zonesInPlane=np.concatenate(zips).astype(np.int32),
                nextZ=nz, nextC=nc)


def sweep_all(mesh, geom, quad, sched, STotal, Sigt, nthreads=1, PsiB=None, fast=False,
              want_psi=False, tau=0.0, psi_prev=None):
    """All directions (SetSweep semantics). With psi_prev (A, C, G) each
    direction's source is STotal + tau psi_prev[a] (Sigt is then sigma_t +
    tau). Returns (Phi (C,G), PsiB (A,nb,G), Psi (A,C,G) or None)."""
    Z = mesh["numCorner"].shape[0]
    Cn = mesh["cFP"].shape[0]
    G = STotal.shape[1]
    A = quad["weight"].shape[0]
    nb = mesh["nbelem"]
    if PsiB is None:
        PsiB = np.zeros((A, nb, G))
    Phi = np.zeros((Cn, G))
    Psi = np.zeros((A, Cn, G)) if want_psi else None
    prev = None if psi_prev is None else np.ascontiguousarray(psi_prev, dtype=np.float64)
    if prev is not None:
        assert prev.shape == (A, Cn, G)
    lib(fast).oracle_rz_sweep_all(
        Z, Cn, nb, G, A, _ip(mesh["numCorner"]), _ip(mesh["cOffSet"]), _ip(mesh["cFP"]),
        _ip(mesh["cEZ"]), _dp(geom["A_fp"]), _dp(geom["A_ez"]), _dp(geom["RadiusFP"]),
        _dp(geom["RadiusEZ"]), _dp(geom["Area"]), _dp(geom["Volume"]),
        _dp(np.ascontiguousarray(STotal)), _dp(np.ascontiguousarray(Sigt)), _dp(quad["omega"]),
        _dp(quad["weight"]), _ip(quad["starting"]), _ip(quad["finishing"]),
        _dp(quad["angDerivFac"]), _dp(quad["quadTauW1"]), _dp(quad["quadTauW2"]),
        _ip(sched["nHyperPlanes"]), _ip(sched["zonesInPlane"]), _ip(sched["nextZ"]),
        _ip(sched["nextC"]), float(tau), _dp(prev), nthreads, _dp(PsiB), _dp(Phi), _dp(Psi))
    return Phi, PsiB, Psi


def source_iteration(mesh, geom, quad, sched, Sigt, sigs, q, tol, max_iters, phi0=None,
                     nthreads=1, fast=False):
    """Returns (Phi (C,G), iterations, residual, history)."""
    Z = mesh["numCorner"].shape[0]
    Cn = mesh["cFP"].shape[0]
    G = Sigt.shape[1]
    A = quad["weight"].shape[0]
    Phi = np.zeros((Cn, G)) if phi0 is None else np.array(phi0, dtype=np.float64, order="C")
    res = C.c_double(0.0)
    hist = np.zeros(max(max_iters, 1))
    it = lib(fast).oracle_rz_source_iteration(
        Z, Cn, mesh["nbelem"], G, A, _ip(mesh["numCorner"]), _ip(mesh["cOffSet"]),
        _ip(mesh["cFP"]), _ip(mesh["cEZ"]), _dp(geom["A_fp"]), _dp(geom["A_ez"]),
        _dp(geom["RadiusFP"]), _dp(geom["RadiusEZ"]), _dp(geom["Area"]), _dp(geom["Volume"]),
        _dp(geom["VolumeZone"]), _dp(np.ascontiguousarray(Sigt)), _dp(np.ascontiguousarray(sigs)),
        _dp(np.ascontiguousarray(q)), _dp(quad["omega"]), _dp(quad["weight"]),
        _ip(quad["starting"]), _ip(quad["finishing"]), _dp(quad["angDerivFac"]),
        _dp(quad["quadTauW1"]), _dp(quad["quadTauW2"]), _ip(sched["nHyperPlanes"]),
        _ip(sched["zonesInPlane"]), _ip(sched["nextZ"]), _ip(sched["nextC"]), nthreads, tol,
        max_iters, _dp(Phi), C.byref(res), _dp(hist))
    return Phi, it, res.value, hist[:it].copy()


def time_step(mesh, geom, quad, sched, sigt, sigs, q, tau, psi_prev, tol, max_iters, phi0=None):
    """One implicit time step of the RZ sweep: source iteration with the
    sweep of snac/SweepUCBrz.F90:130-137 for tau > 0 (total Sigt + tau,
    control/setTotalOpacity.F90; source STotal + tau Psi(:,:,Angle) with Psi
    the previous cycle's psi_prev (A, C, G)), residual as in
    source_iteration, then the cycle's final sweep with savePsi
    (rt/rtmainsn.F90:238-246, SweepUCBrz.F90:268-276: every swept direction's
    psi, finishing directions the last half-angle flux).
    Returns (phi (C, G), psi_new (A, C, G), iterations, residual, history)."""
    Cn = mesh["cFP"].shape[0]
    G = sigt.shape[1]
    V, VZ = geom["Volume"], geom["VolumeZone"]
    st = np.ascontiguousarray(sigt + tau)
    sigs_c, q_c = np.repeat(sigs, 4, axis=0), np.repeat(q, 4, axis=0)

    def sweep(phi, save):
        S = (sigs_c * phi + q_c) / (2.0 * np.pi)
        Phi, _, Psi = sweep_all(mesh, geom, quad, sched, S, st, want_psi=save, tau=tau,
                                psi_prev=psi_prev)
        return Phi, Psi

    def zavg(x):
        return (x * V[:, None]).reshape(-1, 4, G).sum(1) / VZ[:, None]

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


def problem(nr, nz, npolar, nazimuthal, **kw):
    """(mesh, geometry, quadrature, schedule) of a structured RZ problem."""
    pts, quads = structured(nr, nz, **kw)
    m = quad_mesh(pts, quads)
    g = geometry(m)
    q = quadrature(npolar, nazimuthal)
    return m, g, q, schedule(m, g, q)
