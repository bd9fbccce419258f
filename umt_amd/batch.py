"""Batches of uneven domains (host side of libumt_amd.so's umt_batch_*):
many independent tiled problems, each with its own quadrature, group count
and cross sections, swept together one level launch per depth."""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check

_pd = C.POINTER(C.c_double)
_pi = C.POINTER(C.c_int)


def _dp(a):
    return a.ctypes.data_as(_pd)


class SweepBatch:
    """Device-resident batch of domains (problems: TiledProblem list)."""

    def __init__(self, problems, groups, precision=64):
        lib = _lib.load()
        self.problems = list(problems)
        self.groups = [int(g) for g in groups]
        n = len(self.problems)
        hs = (C.c_void_p * n)(*[p.handle.value for p in self.problems])
        gs = np.ascontiguousarray(self.groups, dtype=np.int32)
        h = C.c_void_p()
        check("umt_batch_create", lib.umt_batch_create(n, hs, gs.ctypes.data_as(_pi), precision,
                                                       C.byref(h)))
        self._h = h
        self.precision = precision

    def close(self):
        h = getattr(self, "_h", None)
        if h:
            _lib.load().umt_batch_destroy(h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def dims(self):
        d = np.zeros(6, dtype=np.int64)
        check("umt_batch_dims",
              _lib.load().umt_batch_dims(self._h, d.ctypes.data_as(C.POINTER(C.c_int64))))
        keys = ("domains", "unknowns", "launches", "groups_per_lane", "angles_per_block", "blocks")
        return dict(zip(keys, (int(x) for x in d)))

    def set_cross_sections(self, domain, sigt, sigs, q):
        a, b, c = (np.ascontiguousarray(x, dtype=np.float64) for x in (sigt, sigs, q))
        check("umt_batch_set_cross_sections",
              _lib.load().umt_batch_set_cross_sections(self._h, domain, _dp(a), _dp(b), _dp(c)))

    def reset(self):
        check("umt_batch_reset", _lib.load().umt_batch_reset(self._h))

    def sweep(self):
        """One step of every domain: S from phi, then phi from one sweep."""
        check("umt_batch_sweep", _lib.load().umt_batch_sweep(self._h))

    def phi(self, domain):
        p = self.problems[domain]
        out = np.zeros((p.ncorner, self.groups[domain]))
        check("umt_batch_get_phi", _lib.load().umt_batch_get_phi(self._h, domain, _dp(out)))
        return out

    def source_iteration(self, tol=1e-8, max_iters=1000):
        """(all converged, per-domain iterations, per-domain residuals, total iterations)."""
        n = len(self.problems)
        its = np.zeros(n, dtype=np.int32)
        res = np.zeros(n)
        tot = C.c_int(0)
        rc = _lib.load().umt_batch_source_iteration(self._h, tol, max_iters,
                                                    its.ctypes.data_as(_pi), _dp(res), C.byref(tot))
        if rc not in (0, 1):
            check("umt_batch_source_iteration", rc)
        return rc == 0, its, res, tot.value

    def set_graphs(self, on):
        check("umt_batch_set_graphs", _lib.load().umt_batch_set_graphs(self._h, int(bool(on))))

    def set_layout(self, groups_per_lane, angles_per_block):
        check("umt_batch_set_layout",
              _lib.load().umt_batch_set_layout(self._h, groups_per_lane, angles_per_block))

    def set_persistent(self, on, max_polls=0):
        """on: every level of an angle block in one persistent launch, with a
        grid barrier between levels. A workgroup polls a barrier at most
        max_polls times (0: about a second; < 0: not at all, which makes a
        sweep of two or more workgroups fail) before the sweep fails."""
        check("umt_batch_set_persistent",
              _lib.load().umt_batch_set_persistent(self._h, int(bool(on)), int(max_polls)))

    def persistent(self):
        """(on, workgroups of the persistent launch)."""
        on, blocks = C.c_int(0), C.c_int(0)
        check("umt_batch_get_persistent",
              _lib.load().umt_batch_get_persistent(self._h, C.byref(on), C.byref(blocks)))
        return bool(on.value), blocks.value

    def time_iterations(self, warmup=1, reps=3):
        ms = C.c_double(0.0)
        check("umt_batch_time_iterations",
              _lib.load().umt_batch_time_iterations(self._h, warmup, reps, C.byref(ms)))
        return ms.value

    def autotune(self, reps=2):
        """(best groups per lane, best angles per block, ms per iteration)."""
        k, n, ms = C.c_int(0), C.c_int(0), C.c_double(0.0)
        check("umt_batch_autotune",
              _lib.load().umt_batch_autotune(self._h, reps, C.byref(k), C.byref(n), C.byref(ms)))
        return k.value, n.value, ms.value
