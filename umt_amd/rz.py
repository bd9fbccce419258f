"""2D cylindrical (RZ) sweep: host side of libumt_amd.so's umt_rz_* entry
points (include/umt_rz.h). Problems are built on the host; RzContext holds
everything on the GPU and sweeps every direction of every xi level in one
pass of level launches (snac/SweepUCBrz.F90 semantics, directions of a xi
level chained through the half-angle flux PsiM)."""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check

_pd = C.POINTER(C.c_double)
_pi = C.POINTER(C.c_int)


def _dp(a):
    return a.ctypes.data_as(_pd)


def _ip(a):
    return a.ctypes.data_as(_pi)


class RzProblem:
    """A quad mesh in the (r, z) plane with the RZ product quadrature
    (npolar x nazimuthal, each 1..4) and its schedule."""

    def __init__(self, nr, nz, npolar, nazimuthal, r0=0.0, r1=1.0, z0=0.0, z1=1.0, jitter=0.0,
                 seed=0):
        h = C.c_void_p()
        check("umt_rz_problem_structured",
              _lib.load().umt_rz_problem_structured(nr, nz, r0, r1, z0, z1, jitter, seed, npolar,
                                                    nazimuthal, C.byref(h)))
        self._init(h)

    @classmethod
    def from_quad_mesh(cls, points, quads, npolar, nazimuthal):
        """points (N, 2) as (r, z), quads (Z, 4) 0-based node ids."""
        pts = np.ascontiguousarray(points, dtype=np.float64)
        qd = np.ascontiguousarray(quads, dtype=np.int32)
        h = C.c_void_p()
        check("umt_rz_problem_from_quad_mesh",
              _lib.load().umt_rz_problem_from_quad_mesh(pts.shape[0], _dp(pts), qd.shape[0],
                                                        _ip(qd), npolar, nazimuthal, C.byref(h)))
        self = cls.__new__(cls)
        self._init(h)
        return self

    def _init(self, h):
        self._h = h
        d = np.zeros(8, dtype=np.int64)
        check("umt_rz_problem_dims",
              _lib.load().umt_rz_problem_dims(h, d.ctypes.data_as(C.POINTER(C.c_int64))))
        (self.nzones, self.ncorner, self.nbelem, self.nangles, self.npoints, self.total_planes,
         self.nlevels, self.sweep_levels) = (int(x) for x in d)
        self._cache = {}

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                _lib.load().umt_rz_problem_destroy(h)
            except Exception:
                pass
            self._h = None

    @property
    def handle(self):
        return self._h

    def mesh(self):
        if "mesh" not in self._cache:
            Z, Cn = self.nzones, self.ncorner
            m = dict(numCorner=np.zeros(Z, np.int32), cOffSet=np.zeros(Z, np.int32),
                     cFP=np.zeros((Cn, 2), np.int32), cEZ=np.zeros((Cn, 2), np.int32),
                     CToFace=np.zeros((Cn, 2), np.int32), zoneOpp=np.zeros((Z, 4), np.int32),
                     faceOpp=np.zeros((Z, 4), np.int32), px=np.zeros((Cn, 2)))
            check("umt_rz_problem_mesh", _lib.load().umt_rz_problem_mesh(
                self._h, *(_ip(m[k]) for k in ("numCorner", "cOffSet", "cFP", "cEZ", "CToFace",
                                               "zoneOpp", "faceOpp")), _dp(m["px"])))
            m["zoneFaces"] = np.full(Z, 4, np.int32)
            m["nbelem"] = self.nbelem
            self._cache["mesh"] = m
        return self._cache["mesh"]

    def points(self):
        """(points (N, 2), clockwise quads (Z, 4))."""
        pts = np.zeros((self.npoints, 2))
        qd = np.zeros((self.nzones, 4), np.int32)
        check("umt_rz_problem_points", _lib.load().umt_rz_problem_points(self._h, _dp(pts), _ip(qd)))
        return pts, qd

    def geometry(self):
        if "geom" not in self._cache:
            Cn = self.ncorner
            g = dict(A_fp=np.zeros((Cn, 2, 2)), A_ez=np.zeros((Cn, 2, 2)),
                     RadiusFP=np.zeros((Cn, 2)), RadiusEZ=np.zeros((Cn, 2)), Area=np.zeros(Cn),
                     Volume=np.zeros(Cn), VolumeZone=np.zeros(self.nzones))
            check("umt_rz_problem_geometry", _lib.load().umt_rz_problem_geometry(
                self._h, *(_dp(g[k]) for k in ("A_fp", "A_ez", "RadiusFP", "RadiusEZ", "Area",
                                               "Volume", "VolumeZone"))))
            self._cache["geom"] = g
        return self._cache["geom"]

    def quadrature(self):
        if "quad" not in self._cache:
            A = self.nangles
            q = dict(omega=np.zeros((A, 2)), weight=np.zeros(A), starting=np.zeros(A, np.int32),
                     finishing=np.zeros(A, np.int32), angDerivFac=np.zeros(A),
                     quadTauW1=np.zeros(A), quadTauW2=np.zeros(A))
            check("umt_rz_problem_quadrature", _lib.load().umt_rz_problem_quadrature(
                self._h, _dp(q["omega"]), _dp(q["weight"]), _ip(q["starting"]),
                _ip(q["finishing"]), _dp(q["angDerivFac"]), _dp(q["quadTauW1"]),
                _dp(q["quadTauW2"])))
            self._cache["quad"] = q
        return self._cache["quad"]

    def schedule(self):
        if "sched" not in self._cache:
            A, Z, Cn = self.nangles, self.nzones, self.ncorner
            s = dict(nHyperPlanes=np.zeros(A, np.int32),
                     zonesInPlane=np.zeros(self.total_planes, np.int32),
                     nextZ=np.zeros((A, Z), np.int32), nextC=np.zeros((A, Cn), np.int32))
            check("umt_rz_problem_schedule", _lib.load().umt_rz_problem_schedule(
                self._h, _ip(s["nHyperPlanes"]), _ip(s["zonesInPlane"]), _ip(s["nextZ"]),
                _ip(s["nextC"])))
            self._cache["sched"] = s
        return self._cache["sched"]

    def unknowns(self, groups):
        """Corner-direction-group unknowns of one sweep (swept directions)."""
        return self.ncorner * int(np.sum(self.quadrature()["finishing"] == 0)) * groups


class RzContext:
    """Device-resident RZ sweep of one problem with `groups` groups."""

    def __init__(self, problem, groups, precision=64):
        h = C.c_void_p()
        check("umt_rz_ctx_create",
              _lib.load().umt_rz_ctx_create(problem.handle, groups, precision, C.byref(h)))
        self._h = h
        self.problem = problem
        self.G = groups
        self.precision = precision

    def close(self):
        h = getattr(self, "_h", None)
        if h:
            _lib.load().umt_rz_ctx_destroy(h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def dims(self):
        d = np.zeros(8, dtype=np.int64)
        check("umt_rz_ctx_dims",
              _lib.load().umt_rz_ctx_dims(self._h, d.ctypes.data_as(C.POINTER(C.c_int64))))
        return dict(zip(("zones", "corners", "nbelem", "groups", "angles", "levels", "items",
                         "tiled"), (int(x) for x in d)))

    def set_sigt(self, sigt):
        a = np.ascontiguousarray(sigt, dtype=np.float64)
        check("umt_rz_ctx_set_sigt", _lib.load().umt_rz_ctx_set_sigt(self._h, _dp(a)))

    def set_source(self, S):
        a = np.ascontiguousarray(S, dtype=np.float64)
        check("umt_rz_ctx_set_source", _lib.load().umt_rz_ctx_set_source(self._h, _dp(a)))

    def set_scattering(self, sigs, q):
        a, b = (np.ascontiguousarray(x, dtype=np.float64) for x in (sigs, q))
        check("umt_rz_ctx_set_scattering",
              _lib.load().umt_rz_ctx_set_scattering(self._h, _dp(a), _dp(b)))

    def sweep(self):
        check("umt_rz_ctx_sweep", _lib.load().umt_rz_ctx_sweep(self._h))

    def phi(self):
        out = np.zeros((self.problem.ncorner, self.G))
        check("umt_rz_ctx_get_phi", _lib.load().umt_rz_ctx_get_phi(self._h, _dp(out)))
        return out

    def set_phi(self, phi):
        a = np.ascontiguousarray(phi, dtype=np.float64)
        check("umt_rz_ctx_set_phi", _lib.load().umt_rz_ctx_set_phi(self._h, _dp(a)))

    def psi(self):
        """(A, C + nbelem, G)."""
        p = self.problem
        out = np.zeros((p.nangles, p.ncorner + p.nbelem, self.G))
        check("umt_rz_ctx_get_psi", _lib.load().umt_rz_ctx_get_psi(self._h, _dp(out)))
        return out

    def set_tau(self, tau):
        """Time-dependent sweeps with tau = 1/(c dt): the sweep sees sigma_t +
        tau and the source S + tau psi_prev[direction]."""
        check("umt_rz_ctx_set_tau", _lib.load().umt_rz_ctx_set_tau(self._h, float(tau)))

    def set_psi_prev(self, psi):
        """The previous cycle's psi, (A, C, G)."""
        p = self.problem
        a = np.ascontiguousarray(psi, dtype=np.float64)
        if a.shape != (p.nangles, p.ncorner, self.G):
            raise ValueError(f"psi_prev must be {(p.nangles, p.ncorner, self.G)}, got {a.shape}")
        check("umt_rz_ctx_set_psi_prev", _lib.load().umt_rz_ctx_set_psi_prev(self._h, _dp(a)))

    def psi_prev(self):
        p = self.problem
        out = np.zeros((p.nangles, p.ncorner, self.G))
        check("umt_rz_ctx_get_psi_prev", _lib.load().umt_rz_ctx_get_psi_prev(self._h, _dp(out)))
        return out

    def save_psi(self):
        """The cycle's final sweep: source from the current phi, one sweep,
        psi_prev = its psi of every direction."""
        check("umt_rz_ctx_save_psi", _lib.load().umt_rz_ctx_save_psi(self._h))

    def time_step(self, tol=1e-8, max_iters=1000):
        """Source iteration with the current tau and psi_prev, then the
        savePsi sweep. Returns (converged, iterations, residual, history)."""
        out = self.source_iteration(tol, max_iters)
        self.save_psi()
        return out

    def set_tiles(self, on):
        """Sweep in tile levels or one launch per zone level (default); both
        give the same bits."""
        check("umt_rz_ctx_set_tiles", _lib.load().umt_rz_ctx_set_tiles(self._h, int(bool(on))))

    def set_graphs(self, on):
        """Replay sweeps from a captured graph (default) or launch them eagerly."""
        check("umt_rz_ctx_set_graphs", _lib.load().umt_rz_ctx_set_graphs(self._h, int(bool(on))))

    def set_groups_per_lane(self, K):
        check("umt_rz_ctx_set_groups_per_lane", _lib.load().umt_rz_ctx_set_groups_per_lane(self._h, K))

    def autotune(self, reps=3):
        """Times groups per lane K = 1, 2, 4 on the current source and keeps
        the fastest. Returns (best K, {K: ms per sweep})."""
        k = C.c_int(0)
        ms = np.zeros(3)
        check("umt_rz_ctx_autotune", _lib.load().umt_rz_ctx_autotune(self._h, reps, C.byref(k), _dp(ms)))
        return k.value, {K: float(t) for K, t in zip((1, 2, 4), ms) if t >= 0}

    def source_iteration(self, tol=1e-8, max_iters=1000):
        """Returns (converged, iterations, residual, residual_history)."""
        it = C.c_int(0)
        res = C.c_double(0.0)
        hist = np.zeros(max(max_iters, 1))
        rc = _lib.load().umt_rz_ctx_source_iteration(self._h, tol, max_iters, C.byref(it),
                                                     C.byref(res), _dp(hist))
        if rc not in (0, 1):
            check("umt_rz_ctx_source_iteration", rc)
        return rc == 0, it.value, res.value, hist[:it.value].copy()

    def time_sweeps(self, warmup=2, reps=10):
        """(ms per sweep, ms of level kernels in one sweep)."""
        a, b = C.c_double(0.0), C.c_double(0.0)
        check("umt_rz_ctx_time_sweeps",
              _lib.load().umt_rz_ctx_time_sweeps(self._h, warmup, reps, C.byref(a), C.byref(b)))
        return a.value, b.value
