"""ORACLE — test infrastructure only. CPU restatement of Teton's grey-transport
acceleration (GTA) as the product applies it: one one-group ("grey") problem
per energy group, in lockstep.

Follows
  rt/setGTAOpacity.F90:195-233   grey opacities (one group, no emission coupling)
  snac/GTASweep.F90:89-102       grey and stretched-TSA sweep sources
  rt/GreySweep.F90:60-96         grey S2 sweep followed by a stretched-TSA sweep
  rt/GTASolver.F90:185-402       BiCGStab on the grey corrections, sigma_s V
                                 weighted inner products (rt/scat_prod*.F90)
  rt/getCollisionRate.F90:64,87  grey source = sigma_s (phi_half - phi_prev)
  rt/LinearSolver.F90:55-111     sweep, grey solve, add the correction
  rt/rtmainsn.F90:129,220        epsGrey = 1e-5, then residual / 20
The S_N and S2 sweeps are the oracle's C restatement of SweepUCBxyz.
Only tests/ may import this module.
"""
import numpy as np

from . import oracle as O

WTISO = 1.0 / (4.0 * np.pi)
MIN_RATIO = 1.0e-10
TINY = 1.0e-300
PAR_TOL = 1.0e-12  # parallel-face tolerance of the S2 schedule and sweeps


def grey_opacities(sigt, sigs):
    """(grey sigma_t, grey sigma_s, eps, stretched sigma_t), arrays like sigt (Z, G)."""
    siga = sigt - sigs
    scat = sigs.copy()
    tot = sigt.copy()
    low = sigs / sigt <= MIN_RATIO
    scat[low] = 0.0
    tot[low] = siga[low]
    eps = 1.0 / np.sqrt(1.0 - scat / tot)
    return tot, scat, eps, tot / eps


class GreySolver:
    def __init__(self, mesh, geom, quad2, sched2, sigt, sigs, nthreads=1):
        self.m, self.g, self.q2, self.s2 = mesh, geom, quad2, sched2
        self.nthreads = nthreads
        self.sigt_g, sigs_g, eps, self.sigt_s = grey_opacities(sigt, sigs)
        self.sigs_c = np.repeat(sigs_g, 8, axis=0)
        self.eps_c = np.repeat(eps, 8, axis=0)
        self.V = geom["Volume"]
        self.VZ = geom["VolumeZone"]
        self.w = self.sigs_c * self.V[:, None]  # GreySigScatVol
        self.G = sigt.shape[1]
        self.sweeps = 0

    def _sweep(self, S, sig):
        # S2 directions lie in the tile's diagonal faces: those faces are
        # parallel in the sweep as in the S2 schedule.
        phi, _ = O.sweep_all(self.m, self.g, self.q2, self.s2, np.ascontiguousarray(S),
                             np.ascontiguousarray(sig), nthreads=self.nthreads,
                             par_tol=PAR_TOL)
        return phi

    def grey_sweep(self, P, gsrc=None):
        """M(P) (+ the source's response when gsrc is given)."""
        src = self.sigs_c * P + (gsrc if gsrc is not None else 0.0)
        P1 = self._sweep(WTISO * src, self.sigt_g)
        P2 = self._sweep(WTISO * self.sigs_c * self.eps_c * (P1 - P), self.sigt_s)
        self.sweeps += 2
        return P1 + P2

    def zavg(self, x):
        return (x * self.V[:, None]).reshape(-1, 8, self.G).sum(1) / self.VZ[:, None]

    def solve(self, gsrc, phi_half, eps_grey, max_sweeps=21, eps_point=1.0e-6):
        G = self.G
        F = np.zeros_like(gsrc)
        pz_old = np.zeros((self.VZ.shape[0], G))
        rad = self.zavg(phi_half)
        r = self.grey_sweep(np.zeros_like(gsrc), gsrc)
        d = r.copy()
        rr_old = (r * self.w).sum(0)
        active = np.ones(G, bool)
        n_grey = 1
        while True:
            active &= ~(np.abs(rr_old) < TINY)
            if not active.any():
                if n_grey <= 2:
                    F = r.copy()
                break
            n_grey += 2
            Ad = d - self.grey_sweep(d)
            dAd = (Ad * self.w).sum(0)
            active &= ~(np.abs(dAd) < TINY)
            alpha = np.where(active, rr_old / np.where(active, dAd, 1.0), 0.0)
            r = r - alpha * Ad
            As = r - self.grey_sweep(r)
            wn = (As * r * self.w).sum(0)
            wd = (As * As * self.w).sum(0)
            stop = active & ((np.abs(wd) < TINY) | (np.abs(wn) < TINY))
            omega = np.where(active & ~stop, wn / np.where(active & ~stop, wd, 1.0), 0.0)
            F = F + alpha * d + omega * r
            r = r - omega * As
            rr = (r * self.w).sum(0)
            ok = active & ~stop & (omega != 0.0)
            beta = np.where(ok, (rr * alpha) / np.where(ok, rr_old * omega, 1.0), 0.0)
            d = beta * d - beta * omega * Ad + r
            active &= ~stop
            pz = self.zavg(F)
            err = pz - pz_old
            phi_new = rad + pz
            if not (np.isfinite(err).all() and np.isfinite(phi_new).all()):
                raise FloatingPointError("grey solver produced a NaN")
            nz = phi_new != 0.0
            point = np.max(np.abs(err[nz] / phi_new[nz])) if nz.any() else 0.0
            e2 = (self.VZ[:, None] * err * err).sum()
            p2 = (self.VZ[:, None] * phi_new * phi_new).sum()
            l2 = np.sqrt(abs(e2 / p2)) if p2 != 0.0 else 0.0
            err_grey = max(point, l2)
            pz_old = pz
            if (err_grey < eps_point or n_grey >= max_sweeps) and err_grey < eps_grey:
                break
            if n_grey >= 100 * max_sweeps:
                raise RuntimeError("grey solver not converging")
            rr_old = rr
        return F, n_grey


def source_iteration_gta(mesh, geom, quad, sched, quad2, sched2, sigt, sigs, q, tol, max_iters,
                         phi0=None, nthreads=1):
    """GTA-accelerated steady source iteration. Returns (phi, iterations,
    residual, history, grey sweeps)."""
    G = sigt.shape[1]
    Cn = mesh["nCFaces"].shape[0]
    V, VZ = geom["Volume"], geom["VolumeZone"]
    grey = GreySolver(mesh, geom, quad2, sched2, sigt, sigs, nthreads)
    sigs_c = np.repeat(sigs, 8, axis=0)
    q_c = np.repeat(q, 8, axis=0)
    phi = np.zeros((Cn, G)) if phi0 is None else np.array(phi0, dtype=np.float64)

    def zavg(x):
        return (x * V[:, None]).reshape(-1, 8, G).sum(1) / VZ[:, None]

    old = zavg(phi)
    eps_grey = 1.0e-5
    hist = []
    res = 1.0
    it = 0
    while it < max_iters:
        prev = phi
        S = WTISO * (sigs_c * phi + q_c)
        half, _ = O.sweep_all(mesh, geom, quad, sched, S, sigt, nthreads=nthreads)
        gsrc = sigs_c * (half - prev)
        F, _ = grey.solve(gsrc, half, eps_grey)
        phi = half + F
        cur = zavg(phi)
        thr = 1.0e-6 * cur.max()
        mask = cur > thr
        res = float(np.max(np.abs((cur[mask] - old[mask]) / cur[mask])))
        old = cur
        hist.append(res)
        it += 1
        eps_grey = res / 20.0
        if res < tol:
            break
    return phi, it, res, np.array(hist), grey.sweeps
