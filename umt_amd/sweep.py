"""Python host side of the MI355X UCB sweep: problem construction, the
device-resident sweep context and the per-angle drop-in entry point.

This mirrors the reference's sweep interface for the parity tests:
``SweepContext.sweep`` is one pass of snac/SetSweep.F90:100-170 (all angles
of a single domain, vacuum boundaries), ``sweep_ucbxyz`` is
snac/SweepUCBxyz.F90:11 for one angle through the C-ABI that replaces
gpu/GPU_SweepUCBxyz.cu:532, and ``SweepContext.source_iteration`` is the
steady source iteration of rt/rtmainsn.F90:167-235 with the convergence
measure of rt/ConvergenceTest.F90:150-168.

Numpy shapes are the Teton Fortran shapes reversed (row-major): Teton
``Phi(G, C)`` is a numpy array of shape ``(C, G)``; index contents stay
1-based.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check

_pi = C.POINTER(C.c_int)
_pd = C.POINTER(C.c_double)

# Layout code of the zone layout for SweepContext.set_groups_per_lane: one
# lane per group of a zone carrying its eight corners (include/umt_sweep.h
# UMT_ZONE_LAYOUT); 1, 2, 4, 8 are the corner layout's groups per lane.
ZONE_LAYOUT = -1


def _ip(a):
    return a.ctypes.data_as(_pi) if a is not None else None


def _dp(a):
    return a.ctypes.data_as(_pd) if a is not None else None


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class TiledProblem:
    """Hex mesh (UMT's tiled mesh of [0,1]^3, or any hex mesh through
    from_hex_mesh) with geometry, quadrature and schedules."""

    def __init__(self, nx, ny, nz, npolar, nazimuthal, kdtree=True):
        lib = _lib.load()
        h = C.c_void_p()
        check("umt_problem_tiled",
              lib.umt_problem_tiled(nx, ny, nz, npolar, nazimuthal, int(kdtree), C.byref(h)))
        self._init(h, (nx, ny, nz), npolar, nazimuthal)

    @classmethod
    def from_hex_mesh(cls, points, hexes, npolar, nazimuthal, kdtree=True):
        """Problem on any hex mesh: points (npoints, 3), hexes (nzones, 8)
        0-based node ids in Conduit Blueprint / VTK hexahedron order."""
        pts = np.ascontiguousarray(points, dtype=np.float64)
        hx = np.ascontiguousarray(hexes, dtype=np.int32)
        h = C.c_void_p()
        check("umt_problem_from_hex_mesh", _lib.load().umt_problem_from_hex_mesh(
            pts.shape[0], _dp(pts), hx.shape[0], _ip(hx), npolar, nazimuthal, int(kdtree),
            C.byref(h)))
        self = cls.__new__(cls)
        self._init(h, None, npolar, nazimuthal)
        return self

    def _init(self, h, tiles, npolar, nazimuthal):
        lib = _lib.load()
        self._h = h
        d = np.zeros(8, dtype=np.int64)
        check("umt_problem_dims", lib.umt_problem_dims(h, d.ctypes.data_as(C.POINTER(C.c_int64))))
        (self.nzones, self.ncorner, self.nbelem, self.nangles, self.max_planes, self.npoints,
         self.total_planes, self.maxcf) = (int(x) for x in d)
        self.tiles = tiles
        self.npolar, self.nazimuthal = npolar, nazimuthal
        self._mesh = self._geom = self._quad = self._sched = None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                _lib.load().umt_problem_destroy(h)
            except Exception:
                pass
            self._h = None

    @property
    def handle(self):
        return self._h

    def mesh(self):
        if self._mesh is None:
            Z, Cn = self.nzones, self.ncorner
            m = dict(numCorner=np.zeros(Z, np.int32), cOffSet=np.zeros(Z, np.int32),
                     nCFaces=np.zeros(Cn, np.int32), cFP=np.zeros((Cn, 3), np.int32),
                     cEZ=np.zeros((Cn, 3), np.int32), CToFace=np.zeros((Cn, 3), np.int32),
                     zoneOpp=np.zeros((Z, 6), np.int32), faceOpp=np.zeros((Z, 6), np.int32),
                     px=np.zeros((Cn, 3), np.float64))
            check("umt_problem_mesh", _lib.load().umt_problem_mesh(
                self._h, _ip(m["numCorner"]), _ip(m["cOffSet"]), _ip(m["nCFaces"]), _ip(m["cFP"]),
                _ip(m["cEZ"]), _ip(m["CToFace"]), _ip(m["zoneOpp"]), _ip(m["faceOpp"]),
                _dp(m["px"])))
            m["zoneFaces"] = np.full(Z, 6, np.int32)
            self._mesh = m
        return self._mesh

    def geometry(self):
        if self._geom is None:
            Z, Cn = self.nzones, self.ncorner
            g = dict(A_fp=np.zeros((Cn, 3, 3)), A_ez=np.zeros((Cn, 3, 3)), Volume=np.zeros(Cn),
                     VolumeZone=np.zeros(Z))
            check("umt_problem_geometry", _lib.load().umt_problem_geometry(
                self._h, _dp(g["A_fp"]), _dp(g["A_ez"]), _dp(g["Volume"]), _dp(g["VolumeZone"])))
            self._geom = g
        return self._geom

    def quadrature(self):
        if self._quad is None:
            q = dict(omega=np.zeros((self.nangles, 3)), weight=np.zeros(self.nangles))
            check("umt_problem_quadrature", _lib.load().umt_problem_quadrature(
                self._h, _dp(q["omega"]), _dp(q["weight"])))
            self._quad = q
        return self._quad

    def schedule(self):
        if self._sched is None:
            A, Z, Cn = self.nangles, self.nzones, self.ncorner
            s = dict(nHyperPlanes=np.zeros(A, np.int32),
                     zonesInPlane=np.zeros(self.total_planes, np.int32),
                     nextZ=np.zeros((A, Z), np.int32), nextC=np.zeros((A, Cn), np.int32))
            check("umt_problem_schedule", _lib.load().umt_problem_schedule(
                self._h, _ip(s["nHyperPlanes"]), _ip(s["zonesInPlane"]), _ip(s["nextZ"]),
                _ip(s["nextC"])))
            total = C.c_int(0)
            lib = _lib.load()
            check("umt_problem_cycles", lib.umt_problem_cycles(self._h, C.byref(total), None, None))
            s["numCycles"] = np.zeros(A, np.int32)
            s["cycleList"] = np.zeros(total.value, np.int32)
            check("umt_problem_cycles", lib.umt_problem_cycles(
                self._h, None, _ip(s["numCycles"]), _ip(s["cycleList"])))
            self._sched = s
        return self._sched

    def gta(self):
        """Grey-acceleration S2 quadrature and schedule (dict like schedule())."""
        if getattr(self, "_gta", None) is None:
            A = 8
            n = C.c_int(0)
            om, w = np.zeros((A, 3)), np.zeros(A)
            nh = np.zeros(A, np.int32)
            zp = np.zeros(A * self.nzones, np.int32)  # upper bound on total planes
            nz = np.zeros((A, self.nzones), np.int32)
            check("umt_problem_gta", _lib.load().umt_problem_gta(
                self._h, C.byref(n), _dp(om), _dp(w), _ip(nh), _ip(zp), _ip(nz)))
            self._gta = dict(omega=om, weight=w, nHyperPlanes=nh,
                             zonesInPlane=zp[:int(nh.sum())].copy(), nextZ=nz)
        return self._gta

    def unknowns(self, groups):
        """Unknowns per sweep: corners x angles x groups (driver/test_driver.cc:1895)."""
        return self.ncorner * self.nangles * groups


def draw_cross_sections(seed, nzones, groups, sigt=(0.1, 10.0), c=(0.0, 0.9), q=(1.0, 1.0)):
    """Seeded (zone, group) draws: returns sigt, sigs, q as (Z, G) arrays."""
    st = np.zeros((nzones, groups))
    ss = np.zeros((nzones, groups))
    qq = np.zeros((nzones, groups))
    check("umt_draw_cross_sections", _lib.load().umt_draw_cross_sections(
        seed, nzones, groups, sigt[0], sigt[1], c[0], c[1], q[0], q[1], _dp(st), _dp(ss), _dp(qq)))
    return st, ss, qq


class SweepContext:
    """Device-resident sweep of all angles (precision 64 or 32)."""

    def __init__(self, problem, groups, precision=64):
        lib = _lib.load()
        h = C.c_void_p()
        check("umt_ctx_create_from_problem",
              lib.umt_ctx_create_from_problem(problem.handle, groups, precision, C.byref(h)))
        self._h = h
        self.problem = problem
        self.G = groups
        self.precision = precision
        d = np.zeros(7, dtype=np.int64)
        check("umt_ctx_dims", lib.umt_ctx_dims(h, d.ctypes.data_as(C.POINTER(C.c_int64))))
        self.levels = int(d[5])
        self.items = int(d[6])

    def close(self):
        h = getattr(self, "_h", None)
        if h:
            _lib.load().umt_ctx_destroy(h)
            self._h = None

    def __del__(self):
        self.close()

    def set_sigt(self, sigt):
        a = _f64(sigt)
        check("umt_ctx_set_sigt", _lib.load().umt_ctx_set_sigt(self._h, _dp(a)))

    def set_source(self, S):
        a = _f64(S)
        check("umt_ctx_set_source", _lib.load().umt_ctx_set_source(self._h, _dp(a)))

    def set_scattering(self, sigs, q):
        a, b = _f64(sigs), _f64(q)
        check("umt_ctx_set_scattering", _lib.load().umt_ctx_set_scattering(self._h, _dp(a), _dp(b)))

    def sweep(self):
        check("umt_ctx_sweep", _lib.load().umt_ctx_sweep(self._h))

    def sweep_n(self, n):
        check("umt_ctx_sweep_n", _lib.load().umt_ctx_sweep_n(self._h, n))

    def set_groups_per_lane(self, K):
        check("umt_ctx_set_groups_per_lane", _lib.load().umt_ctx_set_groups_per_lane(self._h, K))

    def groups_per_lane(self):
        k = C.c_int(0)
        check("umt_ctx_get_groups_per_lane", _lib.load().umt_ctx_get_groups_per_lane(self._h, C.byref(k)))
        return k.value

    def autotune(self, reps=2):
        """Time each supported layout (groups per lane 1, 2, 4, 8 and
        ZONE_LAYOUT); keep the fastest. Returns (best K, {K: ms per sweep})."""
        k = C.c_int(0)
        ms = np.zeros(5)
        check("umt_ctx_autotune_layouts",
              _lib.load().umt_ctx_autotune_layouts(self._h, reps, C.byref(k), _dp(ms)))
        return k.value, {kk: float(t) for kk, t in zip((1, 2, 4, 8, ZONE_LAYOUT), ms) if t >= 0}

    def autotune_grey(self, reps=2):
        """The same for the S2 grey sweeps of GTA. Returns (best K, {K: ms per S2 sweep})."""
        k = C.c_int(0)
        ms = np.zeros(5)
        check("umt_ctx_autotune_grey_layouts",
              _lib.load().umt_ctx_autotune_grey_layouts(self._h, reps, C.byref(k), _dp(ms)))
        return k.value, {kk: float(t) for kk, t in zip((1, 2, 4, 8, ZONE_LAYOUT), ms) if t >= 0}

    def incoming_faces(self):
        v = C.c_double(0.0)
        check("umt_ctx_incoming_faces", _lib.load().umt_ctx_incoming_faces(self._h, C.byref(v)))
        return v.value

    def level_stats(self):
        """{launches, items, zone_visits, blocks} of the S_N sweep as launched."""
        d = np.zeros(4, dtype=np.int64)
        check("umt_ctx_level_stats", _lib.load().umt_ctx_level_stats(
            self._h, d.ctypes.data_as(C.POINTER(C.c_int64))))
        return dict(launches=int(d[0]), items=int(d[1]), zone_visits=int(d[2]), blocks=int(d[3]))

    def phi(self):
        out = np.zeros((self.problem.ncorner, self.G))
        check("umt_ctx_get_phi", _lib.load().umt_ctx_get_phi(self._h, _dp(out)))
        return out

    def set_phi(self, phi):
        a = _f64(phi)
        check("umt_ctx_set_phi", _lib.load().umt_ctx_set_phi(self._h, _dp(a)))

    def psi(self):
        """psi(G, C+nbelem, A) as a numpy array of shape (A, C+nbelem, G)."""
        p = self.problem
        out = np.zeros((p.nangles, p.ncorner + p.nbelem, self.G))
        check("umt_ctx_get_psi", _lib.load().umt_ctx_get_psi(self._h, _dp(out)))
        return out

    def set_acceleration(self, mode):
        """"none" (plain source iteration), "gta" (grey-transport acceleration,
        kernel 1: grey and stretched-TSA sweeps), "gta2" (GTA kernel 2: the
        transfer-matrix solver), "learned" (the network given to set_learner),
        "gta+learned" or "gta2+learned"."""
        m = {"none": 0, "gta": 1, "gta2": 1, "learned": 2, "gta+learned": 3, "gta2+learned": 3}[mode]
        if mode.startswith("gta"):
            self.set_gta_solver(2 if mode.startswith("gta2") else 1)
        check("umt_ctx_set_acceleration", _lib.load().umt_ctx_set_acceleration(self._h, m))

    def set_learner(self, learner, every=2, guard=True):
        """Network for mode "learned" (None detaches); corrections every
        `every` sweeps; guard stops them once one raises the residual."""
        self._learner = learner  # keep it alive while attached
        h = learner.handle if learner is not None else None
        check("umt_ctx_set_learner", _lib.load().umt_ctx_set_learner(self._h, h, every, int(bool(guard))))

    def learn_collect(self, on=True, stride=1):
        """Record every stride-th state source iteration could correct (training data)."""
        check("umt_ctx_learn_collect",
              _lib.load().umt_ctx_learn_collect(self._h, int(bool(on)), stride))

    def learn_commit(self, learner, min_residual=1e-9):
        """Turn the recorded states into samples of learner, with targets
        against the current (converged) zone averages. Returns the state count."""
        n = C.c_int(0)
        check("umt_ctx_learn_commit", _lib.load().umt_ctx_learn_commit(
            self._h, learner.handle, min_residual, C.byref(n)))
        return n.value

    def learn_discard(self):
        check("umt_ctx_learn_discard", _lib.load().umt_ctx_learn_discard(self._h))

    def learn_features(self, ph, pp, dp=None):
        """Features (Z, G, 16) for zone averages ph, pp, dp given as (Z, G)."""
        a, b = _f64(ph), _f64(pp)
        c = _f64(dp) if dp is not None else None
        X = np.zeros((self.problem.nzones, self.G, 16), dtype=np.float32)
        check("umt_ctx_learn_features", _lib.load().umt_ctx_learn_features(
            self._h, _dp(a), _dp(b), _dp(c) if c is not None else None,
            X.ctypes.data_as(C.POINTER(C.c_float))))
        return X

    def learn_stats(self):
        """(corrections applied so far, iteration the last solve stopped correcting at or 0)."""
        a, b = C.c_int64(0), C.c_int(0)
        check("umt_ctx_learn_stats", _lib.load().umt_ctx_learn_stats(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def set_graphs(self, on):
        """Replay each sweep's level launches from a captured hipGraph."""
        check("umt_ctx_set_graphs", _lib.load().umt_ctx_set_graphs(self._h, int(bool(on))))

    def grey_sweep(self, P=None, gsrc=None):
        """The grey operator M(P) (+ the response to gsrc), arrays (C, G)."""
        a = _f64(P) if P is not None else None
        b = _f64(gsrc) if gsrc is not None else None
        out = np.zeros((self.problem.ncorner, self.G))
        check("umt_ctx_grey_sweep", _lib.load().umt_ctx_grey_sweep(self._h, _dp(a), _dp(b), _dp(out)))
        return out

    def grey_solve(self, gsrc, phi_half, eps_grey):
        """(F, BiCGStab iterations) of one grey solve."""
        a, b = _f64(gsrc), _f64(phi_half)
        F = np.zeros((self.problem.ncorner, self.G))
        n = C.c_int64(0)
        check("umt_ctx_grey_solve", _lib.load().umt_ctx_grey_solve(
            self._h, _dp(a), _dp(b), eps_grey, _dp(F), C.byref(n)))
        return F, n.value

    def set_incident_flux(self, PsiB):
        """Incident boundary psi as (A, nbelem, G) = Teton PsiB(G, nbelem, A); None: vacuum."""
        a = _f64(PsiB) if PsiB is not None else None
        check("umt_ctx_set_incident_flux", _lib.load().umt_ctx_set_incident_flux(self._h, _dp(a)))

    def boundary_flux(self):
        """PsiB (A, nbelem, G) after a sweep: exiting fluxes, incident ones elsewhere."""
        p = self.problem
        out = np.zeros((p.nangles, p.nbelem, self.G))
        check("umt_ctx_get_boundary_flux", _lib.load().umt_ctx_get_boundary_flux(self._h, _dp(out)))
        return out

    def set_gta_solver(self, kernel):
        """GTA kernel 1 (grey + stretched-TSA sweeps) or 2 (transfer matrices)."""
        check("umt_ctx_set_gta_solver", _lib.load().umt_ctx_set_gta_solver(self._h, int(kernel)))

    def set_grey_controls(self, rel_tol=None, max_its=None):
        """Grey iteration tolerance and maximum (teton_adjust_grey_reltol,
        teton_adjust_grey_maxits); None keeps the current value."""
        t, m, _ = self.grey_controls()
        check("umt_ctx_set_grey_controls", _lib.load().umt_ctx_set_grey_controls(
            self._h, float(t if rel_tol is None else rel_tol), int(m if max_its is None else max_its)))

    def grey_controls(self):
        """(relative tolerance, maximum grey iterations, hard maximum on)."""
        t, m, h = C.c_double(0.0), C.c_int(0), C.c_int(0)
        check("umt_ctx_get_grey_controls", _lib.load().umt_ctx_get_grey_controls(
            self._h, C.byref(t), C.byref(m), C.byref(h)))
        return t.value, m.value, bool(h.value)

    def set_gta_options(self, enforce_hard_iter_max):
        """Stop the grey iteration at its maximum whatever its error
        (teton_setgtaoptions; default on for kernel 2, off for kernel 1)."""
        check("umt_ctx_set_gta_options",
              _lib.load().umt_ctx_set_gta_options(self._h, int(bool(enforce_hard_iter_max))))

    def grey_transfer(self):
        """Kernel 2's matrices as (Z, G, 8, 8) arrays: T[z, g, s, c] (flux at
        corner c from a unit source at corner s) and the LU factors
        LU[z, g, c, s] of I - wtiso sigma_s T."""
        p, G = self.problem, self.G
        T = np.zeros((8, p.ncorner, G))
        LU = np.zeros((8, p.ncorner, G))
        check("umt_ctx_grey_transfer", _lib.load().umt_ctx_grey_transfer(self._h, _dp(T), _dp(LU)))
        T = T.reshape(8, p.nzones, 8, G).transpose(1, 3, 0, 2)    # [z, g, s, c]
        LU = LU.reshape(8, p.nzones, 8, G).transpose(1, 3, 2, 0)  # [z, g, c, s]
        return np.ascontiguousarray(T), np.ascontiguousarray(LU)

    def set_angles_per_block(self, n):
        """Sweep angles in blocks of n (psi memory for n angles)."""
        check("umt_ctx_set_angles_per_block", _lib.load().umt_ctx_set_angles_per_block(self._h, n))

    def angles_per_block(self):
        """(angles per block, blocks)."""
        n, b = C.c_int(0), C.c_int(0)
        check("umt_ctx_get_angles_per_block",
              _lib.load().umt_ctx_get_angles_per_block(self._h, C.byref(n), C.byref(b)))
        return n.value, b.value

    def copy_phi_from(self, other):
        """phi = other's phi (device to device), with the lagged cycle-list psi."""
        check("umt_ctx_copy_phi", _lib.load().umt_ctx_copy_phi(self._h, other._h))

    def set_cycles(self, num_cycles, cycle_list):
        """Cycle lists of the schedule (numCycles (A), cycleList 1-based);
        contexts made from a problem already hold the problem's."""
        nc = np.ascontiguousarray(num_cycles, dtype=np.int32)
        cl = np.ascontiguousarray(cycle_list, dtype=np.int32)
        check("umt_ctx_set_cycles", _lib.load().umt_ctx_set_cycles(self._h, _ip(nc), _ip(cl)))

    def gta_stats(self):
        """(BiCGStab iterations, S2 sweeps) done by GTA so far."""
        a, b = C.c_int64(0), C.c_int64(0)
        check("umt_ctx_gta_stats", _lib.load().umt_ctx_gta_stats(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def source_iteration(self, tol=1e-8, max_iters=1000):
        """Returns (converged, iterations, residual, residual_history)."""
        it = C.c_int(0)
        res = C.c_double(0.0)
        hist = np.zeros(max(max_iters, 1))
        rc = _lib.load().umt_ctx_source_iteration(self._h, tol, max_iters, C.byref(it),
                                                  C.byref(res), _dp(hist))
        if rc not in (0, 1, 2):
            check("umt_ctx_source_iteration", rc)
        return rc == 0, it.value, res.value, hist[:it.value].copy()

    def set_tau(self, tau):
        """Time-dependent sweeps with tau = 1/(c dt) (0: steady)."""
        check("umt_ctx_set_tau", _lib.load().umt_ctx_set_tau(self._h, tau))

    def set_psi_prev(self, psi):
        """Previous cycle's Psi as (A, C, G) = Teton Psi(G, C, A)."""
        a = _f64(psi)
        check("umt_ctx_set_psi_prev", _lib.load().umt_ctx_set_psi_prev(self._h, _dp(a)))

    def psi_prev(self):
        p = self.problem
        out = np.zeros((p.nangles, p.ncorner, self.G))
        check("umt_ctx_get_psi_prev", _lib.load().umt_ctx_get_psi_prev(self._h, _dp(out)))
        return out

    def save_psi(self):
        """The cycle's final sweep: Psi = psi of a sweep from the current phi."""
        check("umt_ctx_save_psi", _lib.load().umt_ctx_save_psi(self._h))

    def set_stall_window(self, window):
        """Stop source iteration after `window` iterations without a new best residual."""
        check("umt_ctx_set_stall_window", _lib.load().umt_ctx_set_stall_window(self._h, window))

    def time_sweeps(self, warmup=2, reps=10):
        """Returns (ms per sweep, ms of level kernels per sweep, level launches per sweep)."""
        ms = C.c_double(0.0)
        msl = C.c_double(0.0)
        n = C.c_int(0)
        check("umt_ctx_time_sweeps", _lib.load().umt_ctx_time_sweeps(
            self._h, warmup, reps, C.byref(ms), C.byref(msl), C.byref(n)))
        return ms.value, msl.value, n.value


def device_memory():
    """(free, total) bytes of device memory on the current GPU."""
    f, t = C.c_int64(0), C.c_int64(0)
    check("umt_device_memory_info", _lib.load().umt_device_memory_info(C.byref(f), C.byref(t)))
    return f.value, t.value


def hbm_copy_probe(n_doubles, reps=10):
    """Mean ms of one device-to-device copy of n_doubles FP64 values."""
    ms = C.c_double(0.0)
    check("umt_hbm_copy_probe", _lib.load().umt_hbm_copy_probe(n_doubles, reps, C.byref(ms)))
    return ms.value


def sweep_ucbxyz(problem, angle, STotal, Sigt, Psi1, PsiB_angle, Phi, tau=0.0, Psi_angle=None,
                 save_psi=False, stream_id=0, cycle_psi=None, noalloc=False):
    """One angle through the drop-in C-ABI ``gpu_sweepucbxyz``.

    ``angle`` is 1-based. Psi1 (C+nbelem, G), PsiB_angle (nbelem, G) and Phi
    (C, G) are updated in place, as the Fortran caller's arrays would be.
    When the problem's schedule has cycle lists, cycle_psi (total, G) holds
    the lagged psi of every angle's cycle corners (Teton's cyclePsi) and is
    updated in place; None uses zeros. noalloc selects ``gpu_sweepucbxyz_noalloc``
    (the same sweep with no device allocation). Returns the library status
    (0 on success).
    """
    lib = _lib.load()
    m, g, q, s = problem.mesh(), problem.geometry(), problem.quadrature(), problem.schedule()
    a = angle - 1
    start = int(np.sum(s["nHyperPlanes"][:a]))
    nhp = int(s["nHyperPlanes"][a])
    zip_ = _i32(s["zonesInPlane"][start:start + nhp])
    nextZ = _i32(s["nextZ"][a])
    nextC = _i32(s["nextC"][a])
    G = STotal.shape[1]
    omega = _f64(q["omega"][a])
    ncyc = s["numCycles"]
    cyc_list = _i32(s["cycleList"]) if len(s["cycleList"]) else np.zeros(1, np.int32)
    if cycle_psi is None:
        cycle_psi = np.zeros((max(len(s["cycleList"]), 1), STotal.shape[1]))
    assert cycle_psi.dtype == np.float64 and cycle_psi.flags["C_CONTIGUOUS"]
    # Teton's maxcf and maxCorner: the corner-face arrays' leading dimension
    # and the most corners of a zone (3 and 8 on hex meshes).
    maxcf = int(m["cFP"].shape[1])
    max_corner = int(np.max(m["numCorner"]))
    ints = {k: C.c_int(v) for k, v in dict(
        Angle=angle, nHP=nhp, Groups=G, ndim=3, maxcf=maxcf, ncorner=problem.ncorner,
        nbelem=problem.nbelem, NumAngles=problem.nangles, maxCorner=max_corner, mem0solve1=0,
        streamId=stream_id, totalStreams=1, savePsi=int(save_psi), numCycles=int(ncyc[a]),
        cycleOffSet=int(np.sum(ncyc[:a])), b0=0, nBdyElem=0, Mref=-1).items()}
    tau_c = C.c_double(tau)
    wt = C.c_double(q["weight"][a])
    Psi = Psi_angle if Psi_angle is not None else np.zeros((problem.ncorner, G))
    dummy_d = np.zeros(1)
    for arr in (STotal, Sigt, Psi1, PsiB_angle, Phi, Psi):
        assert arr.dtype == np.float64 and arr.flags["C_CONTIGUOUS"]
    fn = lib.gpu_sweepucbxyz_noalloc if noalloc else lib.gpu_sweepucbxyz
    fn(
        C.byref(ints["Angle"]), C.byref(ints["nHP"]), _ip(zip_), _ip(nextZ), _ip(nextC),
        _dp(STotal), C.byref(tau_c), _dp(Psi), C.byref(ints["Groups"]), _dp(g["Volume"]),
        _dp(Sigt), _ip(m["nCFaces"]), C.byref(ints["ndim"]), C.byref(ints["maxcf"]),
        C.byref(ints["ncorner"]), _dp(g["A_fp"]), _dp(omega), _ip(m["cFP"]), _dp(Psi1),
        C.byref(ints["nbelem"]), _dp(g["A_ez"]), _ip(m["cEZ"]), C.byref(ints["NumAngles"]),
        C.byref(wt), _dp(Phi), _dp(PsiB_angle), C.byref(ints["maxCorner"]),
        C.byref(ints["mem0solve1"]), C.byref(ints["streamId"]), C.byref(ints["totalStreams"]),
        C.byref(ints["savePsi"]), C.byref(ints["numCycles"]), C.byref(ints["cycleOffSet"]),
        _dp(cycle_psi), _ip(cyc_list), C.byref(ints["b0"]), C.byref(ints["nBdyElem"]),
        _dp(dummy_d), C.byref(ints["Mref"]), _ip(m["numCorner"]), _ip(m["cOffSet"]))
    sid = C.c_int(stream_id)
    lib.gpu_streamsynchronize(C.byref(sid))
    return lib.gpu_sweepucbxyz_status()


# Grey iterations per grey solve in the FP32 phase of solve_mixed: that
# phase needs its correction only to the FP32 switch tolerance. With the
# reference's 21, GTA kernel 2 solves on 32^3 tiles took 14.6 and 14.5 s
# (two seeds) against 11.9 and 11.6 s with 9, and 0.71 against 0.66 s on
# 20^3 tiles, same outer iterations; 7, 5 and 3 were slower.
FP32_PHASE_GREY_MAX_ITS = 9


def solve_mixed(fast, exact, tol=1e-8, switch_tol=1e-6, max_iters=1000, stall_window=8,
                fast_grey_max_its=FP32_PHASE_GREY_MAX_ITS):
    """Source iteration with FP32 sweeps (context `fast`) until the residual
    falls below switch_tol (or stops improving at the FP32 rounding floor),
    then FP64 sweeps (context `exact`, same problem and cross sections) from
    that phi to tol; the FP64 residual decides convergence. Both contexts keep
    their own acceleration; the FP32 phase's grey solves stop after
    fast_grey_max_its grey iterations (None: fast's own controls, restored
    afterwards either way). fast's phi is set to the result. Returns
    (converged, FP32 iterations, FP64 iterations, residual, history)."""
    fast.set_stall_window(stall_window)
    controls = fast.grey_controls()
    if fast_grey_max_its is not None:
        fast.set_grey_controls(max_its=fast_grey_max_its)
    try:
        _, it1, _, h1 = fast.source_iteration(switch_tol, max_iters)
    finally:
        fast.set_grey_controls(rel_tol=controls[0], max_its=controls[1])
    exact.copy_phi_from(fast)
    ok, it2, res, h2 = exact.source_iteration(tol, max(1, max_iters - it1))
    fast.copy_phi_from(exact)
    return ok, it1, it2, res, np.concatenate([h1, h2])
