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

    # GTA%enforceHardGTAIterMax: off for the old solver (GreyAcceleration_mod.F90:122),
    # on for the new one (:169); teton_setgtaoptions changes it (setGTAOptions.F90:15-34).
    hard_max_default = False

    def solve(self, gsrc, phi_half, eps_grey, max_sweeps=21, eps_point=1.0e-6, hard_max=None):
        if hard_max is None:
            hard_max = self.hard_max_default
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
            if hard_max and n_grey >= max_sweeps:  # GTASolver.F90:373-375
                break
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


# ---------------------------------------------------------------------------
# The reference's "new" GTA solver (Size%useNewGTASolver): per-zone corner
# transfer matrices and a direct solve of each zone's grey scattering problem
# with in-place LU factors, in place of the stretched-TSA grey sweeps.
#   snac/InitSweepGreyUCBxyz.F90:10-253   transfer matrices TT
#   snac/SweepGreyUCBxyz.F90:137-322      grey sweep with the incident part pInc
#   snac/GTASweep.F90:80-115              TSA source, PhiInc = sum_a w pInc
#   snac/UpdateScalarIntensity.F90:11-170 LU of I - TT wtiso sigma_s, solve
#   rt/GreySweep.F90:12-47                GreySweepNEW
#   rt/GTASolver.F90:158-198              matrices built once per solve, the
#                                         first (with-source) sweep decomposes

def _dot_tol(o, a, tol):
    d = o[0] * a[0] + o[1] * a[1] + o[2] * a[2]
    if tol > 0.0 and abs(d) <= tol * (abs(a[0]) + abs(a[1]) + abs(a[2])):
        return 0.0
    return d


def _sez_coefficients(aez, sigv, sig, area_opp):
    """(B1, B2, gtau): the EZ term of an outgoing face of corner c towards
    cez as sez = B1 S_c - B2 S_cez + gtau sigv psi_opp, S per steradian
    (InitSweepGreyUCBxyz.F90:166-214; Q = S / sigma)."""
    if area_opp > 0.0:
        sigv2 = sigv * sigv
        gnum = aez * aez * (1.82 * sigv2 + aez * (4.0 * sigv + 3.0 * aez))
        gtau = gnum / (gnum + 4.0 * sigv2 * sigv2 + aez * sigv * (6.0 * sigv2 + 2.0 * aez *
                                                                   (2.0 * sigv + aez)))
        B0 = 0.5 * aez * (1.0 - gtau)
        return (B0 - gtau * sigv) / sig, B0 / sig, gtau
    return 0.5 * aez / sig, 0.5 * aez / sig, None


def _zone_faces(mesh, geom, z, omega, par_tol):
    """Per corner of zone z: afp, aez (with the parallel tolerance), cEZ
    (0-based, local) and the downstream lists of the corner solve."""
    c0 = int(mesh["cOffSet"][z])
    nC = int(mesh["numCorner"][z])
    afp = [[_dot_tol(omega, geom["A_fp"][c0 + c, f], par_tol)
            for f in range(int(mesh["nCFaces"][c0 + c]))] for c in range(nC)]
    aez = [[_dot_tol(omega, geom["A_ez"][c0 + c, f], par_tol)
            for f in range(int(mesh["nCFaces"][c0 + c]))] for c in range(nC)]
    cez = [[int(mesh["cEZ"][c0 + c, f]) - 1 for f in range(int(mesh["nCFaces"][c0 + c]))]
           for c in range(nC)]
    need = [0] * nC
    ez_exit = [[] for _ in range(nC)]
    coef = [[] for _ in range(nC)]
    for c in range(nC):
        for f in range(len(aez[c])):
            e = cez[c][f]
            if e > c:
                if aez[c][f] > 0.0:
                    need[e] += 1
                    ez_exit[c].append(e)
                    coef[c].append(aez[c][f])
                elif aez[c][f] < 0.0:
                    need[c] += 1
                    ez_exit[e].append(c)
                    coef[e].append(-aez[c][f])
    return c0, nC, afp, aez, cez, need, ez_exit, coef


def _area_opp(afp_c, f):
    """Incoming area of the FP faces opposite EZ face f (the next face for a
    3-face corner, else the n - 2 following ones), and those faces."""
    n = len(afp_c)
    faces = [(f + 1) % n] if n == 3 else [(f + k) % n for k in range(1, n - 1)]
    area = 0.0
    inc = []
    for ifp in faces:
        if afp_c[ifp] < 0.0:
            area -= afp_c[ifp]
            inc.append(ifp)
    return area, inc


def transfer_matrices(mesh, geom, quad2, sigt_g, par_tol=PAR_TOL):
    """TT (Z, G, 8, 8): TT[z, g, s, c] = scalar flux at corner c of zone z
    from a unit per-steradian source at corner s of the same zone, summed
    over the grey directions with their weights, no incident flux
    (GTA%TT(s, c0 + c), InitSweepGreyUCBxyz.F90:10-253)."""
    Z = mesh["numCorner"].shape[0]
    G = sigt_g.shape[1]
    V = geom["Volume"]
    TT = np.zeros((Z, G, 8, 8))
    for z in range(Z):
        sig = sigt_g[z]
        for a in range(quad2["weight"].shape[0]):
            omega = quad2["omega"][a]
            c0, nC, afp, aez, cez, need, ez_exit, coef = _zone_faces(mesh, geom, z, omega, par_tol)
            Pvv = np.zeros((G, nC, nC))  # Pvv[g, column (source), row (flux)]
            for c in range(nC):
                Pvv[:, c, c] = V[c0 + c]
            denom = np.zeros((G, nC))
            for c in range(nC):
                sigv = V[c0 + c] * sig
                denom[:, c] = sigv
                for f in range(len(afp[c])):
                    if afp[c][f] > 0.0:
                        denom[:, c] += afp[c][f]
                for f in range(len(aez[c])):
                    a_ez = aez[c][f]
                    if a_ez > 0.0:
                        e = cez[c][f]
                        denom[:, c] += a_ez
                        area, _ = _area_opp(afp[c], f)
                        B1, B2, _ = _sez_coefficients(a_ez, sigv, sig, area)
                        Pvv[:, c, c] += B1
                        Pvv[:, e, c] -= B2
                        Pvv[:, c, e] -= B1
                        Pvv[:, e, e] += B2
            nd = list(need)
            for _ in range(nC):
                c = int(np.argmin(nd))
                Pvv[:, :, c] *= (1.0 / denom[:, c])[:, None]
                for e, k in zip(ez_exit[c], coef[c]):
                    nd[e] -= 1
                    Pvv[:, :, e] += k * Pvv[:, :, c]
                nd[c] = 99
            TT[z, :, :nC, :nC] += quad2["weight"][a] * Pvv
    return TT


def pinc_sweep(mesh, geom, quad2, sched2, sigt_g, tsa, par_tol=PAR_TOL):
    """The grey sweep of the new solver for every grey direction, vacuum
    boundaries: returns (phi, PhiInc), phi = sum_a w tPsi and PhiInc =
    sum_a w pInc, pInc being the part of tPsi carried by incident fluxes
    (SweepGreyUCBxyz.F90:137-322, GTASweep.F90:108-115). tsa is TsaSource
    (C, G) per steradian."""
    Z = mesh["numCorner"].shape[0]
    Cn = mesh["nCFaces"].shape[0]
    G = sigt_g.shape[1]
    V = geom["Volume"]
    phi = np.zeros((Cn, G))
    phi_inc = np.zeros((Cn, G))
    for a in range(quad2["weight"].shape[0]):
        omega = quad2["omega"][a]
        w = quad2["weight"][a]
        tpsi = np.zeros((Cn, G))  # incident boundary fluxes are zero (vacuum)
        pinc = np.zeros((Cn, G))
        # nextZ order: every zone after the zones it depends on
        order = [abs(int(z)) - 1 for z in sched2["nextZ"][a]]
        nextC = sched2["nextC"][a]
        for z in order:
            sig = sigt_g[z]
            c0, nC, afp, aez, cez, _, ez_exit, coef = _zone_faces(mesh, geom, z, omega, par_tol)
            Q = tsa[c0:c0 + nC] / sig
            src = V[c0:c0 + nC, None] * tsa[c0:c0 + nC]
            denom = np.zeros((nC, G))
            psifp = {}
            for c in range(nC):
                sigv = V[c0 + c] * sig
                denom[c] = sigv
                for f in range(len(afp[c])):
                    cfp = int(mesh["cFP"][c0 + c, f]) - 1
                    if afp[c][f] > 0.0:
                        denom[c] += afp[c][f]
                    elif afp[c][f] < 0.0:
                        psifp[(c, f)] = tpsi[cfp] if cfp < Cn else np.zeros(G)
                        src[c] -= afp[c][f] * psifp[(c, f)]
                        pinc[c0 + c] -= afp[c][f] * psifp[(c, f)]
                for f in range(len(aez[c])):
                    a_ez = aez[c][f]
                    if a_ez <= 0.0:
                        continue
                    e = cez[c][f]
                    denom[c] += a_ez
                    area, inc = _area_opp(afp[c], f)
                    if area > 0.0:
                        psi_opp = sum(-afp[c][i] * psifp[(c, i)] for i in inc) / area
                        B1, B2, gtau = _sez_coefficients(a_ez, sigv, sig, area)
                        sez = gtau * sigv * (psi_opp - Q[c]) + 0.5 * a_ez * (1.0 - gtau) * (Q[c] - Q[e])
                        src[c] += sez
                        src[e] -= sez
                        pinc[c0 + c] += gtau * sigv * psi_opp
                        pinc[c0 + e] -= gtau * sigv * psi_opp
                    else:
                        sez = 0.5 * a_ez * (Q[c] - Q[e])
                        src[c] += sez
                        src[e] -= sez
            for i in range(nC):
                c = int(nextC[c0 + i]) - 1
                tpsi[c0 + c] = src[c] / denom[c]
                pinc[c0 + c] = pinc[c0 + c] / denom[c]
                for e, k in zip(ez_exit[c], coef[c]):
                    src[e] += k * tpsi[c0 + c]
                    pinc[c0 + e] += k * pinc[c0 + c]
        phi += w * tpsi
        phi_inc += w * pinc
    return phi, phi_inc


def scalar_intensity_decompose(TT, sigs_g):
    """In-place LU factors (Doolittle, no pivoting) of each zone's
    A = I - TT wtiso sigma_s, A[z, g, c, s] = delta - wtiso sigma_s(s) TT[z, g, s, c]
    (UpdateScalarIntensity.F90:50-97). Returns LU (Z, G, 8, 8) holding the
    unit lower factor below and the upper factor on and above the diagonal."""
    A = -WTISO * sigs_g[:, :, None, None] * np.swapaxes(TT, 2, 3)
    n = A.shape[-1]
    A = A + np.eye(n)
    for i in range(n):
        t = np.zeros(A.shape[:2])
        for k in range(i):
            t = t + A[:, :, i, k] * A[:, :, k, i]
        A[:, :, i, i] = A[:, :, i, i] - t
        dinv = 1.0 / A[:, :, i, i]
        for j in range(i + 1, n):
            t = np.zeros(A.shape[:2])
            v = np.zeros(A.shape[:2])
            for k in range(i):
                t = t + A[:, :, i, k] * A[:, :, k, j]
                v = v + A[:, :, j, k] * A[:, :, k, i]
            A[:, :, i, j] = A[:, :, i, j] - t
            A[:, :, j, i] = dinv * (A[:, :, j, i] - v)
    return A


def scalar_intensity_solve(LU, rhs):
    """Forward and back substitution with the factors (ScalarIntensitySolve,
    UpdateScalarIntensity.F90:113-170); rhs (Z, G, 8)."""
    x = np.array(rhs, dtype=np.float64)
    n = x.shape[-1]
    for j in range(1, n):
        t = np.zeros(x.shape[:2])
        for i in range(j):
            t = t - LU[:, :, j, i] * x[:, :, i]
        x[:, :, j] = x[:, :, j] + t
    x[:, :, n - 1] = x[:, :, n - 1] / LU[:, :, n - 1, n - 1]
    for k in range(n - 2, -1, -1):
        t = np.zeros(x.shape[:2])
        for i in range(k + 1, n):
            t = t + x[:, :, i] * LU[:, :, k, i]
        x[:, :, k] = (x[:, :, k] - t) / LU[:, :, k, k]
    return x


def _zone_vec(x, G):
    """(C, G) corner array -> (Z, G, 8)."""
    return np.swapaxes(x.reshape(-1, 8, G), 1, 2)


def _corner_vec(x):
    """(Z, G, 8) -> (C, G)."""
    return np.ascontiguousarray(np.swapaxes(x, 1, 2)).reshape(-1, x.shape[1])


class NewGreySolver(GreySolver):
    """GTA with the reference's new solver: the same BiCGStab (solve()),
    whose operator M(P) is one grey sweep giving PhiInc followed by each
    zone's direct solve (I - TT wtiso sigma_s) P = PhiInc [+ TT wtiso gsrc]
    with LU factors made by the with-source call (GreySweepNEW)."""

    hard_max_default = True

    def __init__(self, *args, **kw):
        super().__init__(*args, **kw)
        self.sigs_g = self.sigs_c[::8]
        self.TT = None
        self.LU = None

    def build(self):
        self.TT = transfer_matrices(self.m, self.g, self.q2, self.sigt_g)
        self.LU = None

    def grey_sweep(self, P, gsrc=None):
        G = self.G
        tsa = WTISO * (self.sigs_c * P + (gsrc if gsrc is not None else 0.0))
        _, phi_inc = pinc_sweep(self.m, self.g, self.q2, self.s2, self.sigt_g, tsa)
        self.sweeps += 1
        rhs = _zone_vec(phi_inc, G)
        if gsrc is not None:
            if self.TT is None:
                self.build()
            # PhiInc + TT wtiso GreySource, then decompose (UpdateScalarIntensity.F90:53-62)
            rhs = rhs + np.einsum("zgsc,zgs->zgc", self.TT, WTISO * _zone_vec(gsrc, G))
            self.LU = scalar_intensity_decompose(self.TT, self.sigs_g)
        return _corner_vec(scalar_intensity_solve(self.LU, rhs))

    def solve(self, gsrc, phi_half, eps_grey, max_sweeps=21, eps_point=1.0e-6, hard_max=None):
        self.build()  # rt/GTASolver.F90:158-178: matrices built for every solve
        return super().solve(gsrc, phi_half, eps_grey, max_sweeps, eps_point, hard_max)
