"""The RZ oracle (oracle/oracle_rz.c, oracle/rz.py) pinned by what the
mathematics fixes, and the product's RZ host arrays against it.

The reference holds no test or fixture for this path and cannot be built
here. Pins:
  - quadrature: the QRS45 polar and RZ azimuthal tables of the fixture, sum
    of weights 2 pi, mu rising through each xi level from the starting
    direction (mu = -sqrt(1 - xi^2)) to the finishing one, the angular cell
    edges closing at the finishing direction, W1 - W2 = 1;
  - geometry: closure of every corner, corner areas summing to the shoelace
    zone area, volumes to the exact r-weighted zone volume, outward normals,
    and the divergence identity sum_f R_f Omega.A_f = mu Area_c;
  - schedule: hyperplanes are the longest-path layering of the zone graph,
    corner orders the smallest topological orders (brute force);
  - sweep: the constant isotropic solution is exact through every direction
    (angular redistribution included); every corner satisfies its balance
    equation with the coupling term written in the grey form of
    snac/SweepGreyUCBrz.F90:289-297 (an independent transcription); every
    zone conserves particles over the directions (the angular terms
    telescope to zero);
  - source iteration: global balance at convergence;
  - time-dependent sweep (tau psi_prev source): the constant solution
    (sigma + tau) psi = S + tau psi0 is exact, every zone conserves particles
    with an angle-dependent psi_prev, psi_prev of one direction reaches only
    that direction and the later ones of its xi level; the time step reduces
    to source iteration at tau = 0 and satisfies the global balance with the
    tau terms at convergence.
"""
import itertools
import json
import os

import numpy as np
import pytest

import umt_amd as U
from oracle import rz as R

FIXTURE = os.path.join(os.path.dirname(__file__), "golden", "qrs45_n1to4.json")


@pytest.fixture(scope="module")
def jittered():
    pts, quads = R.structured(5, 4, r0=0.0, r1=1.0, z0=0.0, z1=2.0, jitter=0.25, seed=3)
    m = R.quad_mesh(pts, quads)
    g = R.geometry(m)
    q = R.quadrature(2, 3)
    return pts, quads, m, g, q, R.schedule(m, g, q)


@pytest.mark.parametrize("npolar,naz", [(1, 1), (2, 3), (4, 4), (3, 2)])
def test_quadrature_levels(npolar, naz):
    q = R.quadrature(npolar, naz)
    tab = json.load(open(FIXTURE))["orders"]
    A = npolar * 2 * (2 * naz + 2)
    assert q["weight"].shape == (A,)
    assert abs(q["weight"].sum() - 2 * np.pi) < 1e-13
    n = 2 * naz + 2
    xis = np.array(tab[str(npolar)]["cosTheta"])[::-1]
    for lvl in range(2 * npolar):
        sl = slice(lvl * n, (lvl + 1) * n)
        mu, xi = q["omega"][sl, 0], q["omega"][sl, 1]
        assert np.all(xi == xi[0])
        assert abs(abs(xi[0]) - xis[lvl // 2]) == 0 and np.sign(xi[0]) == (-1 if lvl % 2 == 0 else 1)
        assert q["starting"][sl].tolist() == [1] + [0] * (n - 1)
        assert q["finishing"][sl].tolist() == [0] * (n - 1) + [1]
        assert mu[0] == -np.sqrt(1 - xi[0] ** 2) and mu[-1] == np.sqrt(1 - xi[0] ** 2)
        assert np.all(np.diff(mu) > 0)
        assert np.all(q["weight"][sl][1:-1] > 0) and q["weight"][sl][0] == q["weight"][sl][-1] == 0
        # the swept directions are the azimuthal table at sin(theta) scale, mirrored
        cphi = np.array(tab[str(naz)]["cosPhiRZ"])
        st = np.sqrt(1 - xis[lvl // 2] ** 2)
        np.testing.assert_allclose(mu[1:-1], np.concatenate([-st * cphi[::-1], st * cphi]), rtol=0,
                                   atol=1e-15)
        # level weights: pi * weightTheta (both mu signs), scaled to sum 2 pi overall
        wt = np.array(tab[str(npolar)]["weightTheta"])[::-1][lvl // 2]
        np.testing.assert_allclose(q["weight"][sl].sum(), np.pi * wt, rtol=1e-14)


@pytest.mark.parametrize("npolar,naz", [(1, 1), (2, 3), (4, 4)])
def test_angular_coefficients_close_each_level(npolar, naz):
    q = R.quadrature(npolar, naz)
    n = 2 * naz + 2
    for lvl in range(2 * npolar):
        sl = slice(lvl * n, (lvl + 1) * n)
        mu, xi, w = q["omega"][sl, 0], q["omega"][sl, 1], q["weight"][sl]
        fac, w1, w2 = q["angDerivFac"][sl], q["quadTauW1"][sl], q["quadTauW2"][sl]
        assert fac[0] == fac[-1] == 0 and w1[0] == 1 and w2[0] == 0
        # alpha_{m+1/2} = alpha_{m-1/2} - w_m mu_m ends at zero (sum w mu = 0)
        alpha = -np.cumsum(w[1:-1] * mu[1:-1])
        assert abs(alpha[-1]) < 1e-14
        # cell edges: phi_{m+1/2} = pi - pi * (cumulative weight) / (level weight)
        edges = np.sqrt(1 - xi[0] ** 2) * np.cos(np.pi - np.pi * np.cumsum(w[1:-1]) / w.sum())
        lo = np.concatenate([[mu[0]], edges[:-1]])
        tau = (mu[1:-1] - lo) / (edges - lo)
        assert np.all((tau > 0) & (tau <= 1))
        np.testing.assert_allclose(w1[1:-1], 1 / tau, rtol=1e-13)
        np.testing.assert_allclose(w1[1:-1] - w2[1:-1], 1.0, rtol=0, atol=1e-13)
        np.testing.assert_allclose(fac[1:-1], mu[1:-1] + alpha / (w[1:-1] * tau), rtol=1e-12,
                                   atol=1e-13)
        # the last edge is the finishing direction
        assert abs(edges[-1] - mu[-1]) < 1e-14


def test_geometry_identities(jittered):
    pts, quads, m, g, q, s = jittered
    Z = quads.shape[0]
    Afp, Aez = g["A_fp"], g["A_ez"]
    # closure of every corner's four faces
    assert np.abs(Afp.sum(1) + Aez.sum(1)).max() < 1e-15
    # corner areas and r-weighted volumes per zone (clockwise quads)
    P = pts[quads]  # (Z, 4, 2)
    r, z = P[..., 0], P[..., 1]
    rn, zn = np.roll(r, -1, 1), np.roll(z, -1, 1)
    cross = r * zn - rn * z
    # clockwise polygons: area = -(1/2) sum cross, int r dA = -(1/6) sum (r_i + r_{i+1}) cross
    area = -0.5 * cross.sum(1)
    vol = -(1.0 / 6.0) * np.sum((r + rn) * cross, 1)
    np.testing.assert_allclose(g["Area"].reshape(Z, 4).sum(1), area, rtol=1e-13)
    np.testing.assert_allclose(g["Volume"].reshape(Z, 4).sum(1), vol, rtol=1e-13)
    np.testing.assert_allclose(g["VolumeZone"], vol, rtol=1e-13)
    # outward FP normals: away from the zone centre
    centre = P.mean(1)
    for c in range(4 * Z):
        for f in range(2):
            assert Afp[c, f] @ (centre[c // 4] - m["px"][c]) < 0
    # divergence identity for random directions
    rng = np.random.default_rng(0)
    for _ in range(5):
        om = rng.normal(size=2)
        lhs = np.einsum("cf,cfk,k->c", g["RadiusFP"], Afp, om) + np.einsum(
            "cf,cfk,k->c", g["RadiusEZ"], Aez, om)
        np.testing.assert_allclose(lhs, om[0] * g["Area"], rtol=0, atol=1e-14)


def test_schedule_is_brute_force_layering(jittered):
    pts, quads, m, g, q, s = jittered
    Z = quads.shape[0]
    starts = np.concatenate([[0], np.cumsum(s["nHyperPlanes"])])
    for a in range(len(q["weight"])):
        if q["finishing"][a]:
            assert s["nHyperPlanes"][a] == 0
            continue
        om = q["omega"][a]
        up = [set() for _ in range(Z)]
        for c in range(4 * Z):
            cfp = m["cFP"][c, 0] - 1  # face 1 of each corner: every edge once, as snneed reads it
            if cfp < 4 * Z:
                d = g["A_fp"][c, 0] @ om
                if d < 0:
                    up[c // 4].add(cfp // 4)
                elif d > 0:
                    up[cfp // 4].add(c // 4)
        depth = [-1] * Z

        def dep(z):
            if depth[z] < 0:
                depth[z] = 1 + max((dep(u) for u in up[z]), default=-1)
            return depth[z]

        for z in range(Z):
            dep(z)
        zip_ = s["zonesInPlane"][starts[a]:starts[a + 1]]
        nz = s["nextZ"][a] - 1
        k = 0
        for h, cnt in enumerate(zip_):
            assert all(depth[z] == h for z in nz[k:k + cnt]), (a, h)
            k += cnt
        # corner order: smallest topological order of the within-zone graph
        for z in range(Z):
            before = []
            for c in range(4):
                for f in range(2):
                    n = m["cEZ"][4 * z + c, f] - 1
                    d = g["A_ez"][4 * z + c, f] @ om
                    if n > c and d != 0:
                        before.append((c, n) if d > 0 else (n, c))
            best = next(pm for pm in itertools.permutations(range(4))
                        if all(pm.index(u) < pm.index(v) for u, v in before))
            np.testing.assert_array_equal(s["nextC"][a][4 * z:4 * z + 4] - 1, best)


def _half_angle(q, psi):
    """PsiM entering each direction, rebuilt from the swept psi."""
    A = len(q["weight"])
    pm_in = np.zeros_like(psi)
    pm = None
    for a in range(A):
        if q["starting"][a]:
            pm = psi[a].copy()
            continue
        if q["finishing"][a]:
            continue
        pm_in[a] = pm
        pm = q["quadTauW1"][a] * psi[a] - q["quadTauW2"][a] * pm
    return pm_in


def test_constant_solution_is_exact(jittered):
    pts, quads, m, g, q, s = jittered
    Cn, nb, A = 4 * quads.shape[0], m["nbelem"], len(q["weight"])
    rng = np.random.default_rng(1)
    sig = rng.uniform(0.1, 30.0, (quads.shape[0], 2))
    S = np.repeat(sig, 4, 0) * 1.5
    phi, _, psi = R.sweep_all(m, g, q, s, S, sig, PsiB=np.full((A, nb, 2), 1.5), want_psi=True)
    assert np.abs(psi / 1.5 - 1).max() < 1e-13  # finishing directions included
    assert np.abs(phi / (1.5 * 2 * np.pi) - 1).max() < 1e-13


def _grey_corner_residual(m, g, q, S, sig, psi, pm_in, a):
    """psi_c (fac A + incoming R|a| + sigma V) - [V Q + incoming R|a| psi_up
    + EZ upwind + sez out - sez in + fac A PsiM] with sez in grey form."""
    Cn = m["cFP"].shape[0]
    om = q["omega"][a]
    fac = q["angDerivFac"][a]
    res = np.zeros_like(psi[a])
    sez_in = np.zeros_like(psi[a])
    lhs = np.zeros_like(psi[a])
    rhs = np.zeros_like(psi[a])
    for c in range(Cn):
        z = c // 4
        st = sig[z]
        A = g["Area"][c]
        afp = g["A_fp"][c] @ om
        aez = g["A_ez"][c] @ om
        den = fac * A + st * g["Volume"][c]
        r = g["Volume"][c] * S[c] + fac * A * pm_in[a][c]
        for f in range(2):
            up = m["cFP"][c, f] - 1
            if afp[f] < 0:
                den = den - g["RadiusFP"][c, f] * afp[f]
                if up < Cn:
                    r = r - g["RadiusFP"][c, f] * afp[f] * psi[a][up]
            n = 4 * z + m["cEZ"][c, f] - 1
            Rz = g["RadiusEZ"][c, f]
            if aez[f] < 0:
                den = den - Rz * aez[f]
                r = r - Rz * aez[f] * psi[a][n]
            elif aez[f] > 0:
                Qh, Qn = S[c] / st, S[n] / st
                if afp[f] < 0:
                    sA = st * A
                    gnum = aez[f] ** 2 * (1.82 * sA * sA + aez[f] * (4 * sA + 3 * aez[f]))
                    gtau = gnum / (gnum + 4 * sA ** 4 + aez[f] * sA * (6 * sA * sA + 2 * aez[f] * (2 * sA + aez[f])))
                    popp = psi[a][up] if up < Cn else 0.0
                    sez = Rz * (gtau * sA * (popp - Qh) + 0.5 * aez[f] * (1 - gtau) * (Qh - Qn))
                else:
                    sez = 0.5 * Rz * aez[f] * (Qh - Qn)
                r = r + sez
                sez_in[n] += sez
        lhs[c] = psi[a][c] * den
        rhs[c] = r
    rhs -= sez_in
    return np.abs(lhs - rhs).max() / np.abs(rhs).max()


def test_every_corner_balances_with_grey_form_coupling(jittered):
    pts, quads, m, g, q, s = jittered
    rng = np.random.default_rng(2)
    Z = quads.shape[0]
    sig = rng.uniform(0.5, 40.0, (Z, 1))
    S = rng.uniform(0.1, 1.0, (4 * Z, 1))
    _, _, psi = R.sweep_all(m, g, q, s, S, sig, want_psi=True)
    pm_in = _half_angle(q, psi)
    for a in range(len(q["weight"])):
        if q["finishing"][a]:
            continue
        assert _grey_corner_residual(m, g, q, S[:, 0], sig[:, 0], psi[..., 0], pm_in[..., 0], a) < 1e-13


def test_zones_conserve_particles(jittered):
    """sum_m w_m [net radius-weighted outflow + sigma V psi] = sum_m w_m V Q
    for every zone: the angular redistribution terms telescope to zero."""
    pts, quads, m, g, q, s = jittered
    rng = np.random.default_rng(3)
    Z = quads.shape[0]
    Cn = 4 * Z
    sig = rng.uniform(0.5, 40.0, (Z, 1))
    S = rng.uniform(0.1, 1.0, (Cn, 1))
    _, _, psi = R.sweep_all(m, g, q, s, S, sig, want_psi=True)
    lhs = np.zeros(Z)
    for a in range(len(q["weight"])):
        w = q["weight"][a]
        if w == 0:
            continue
        om = q["omega"][a]
        for c in range(Cn):
            z = c // 4
            lhs[z] += w * sig[z, 0] * g["Volume"][c] * psi[a, c, 0]
            for f in range(2):
                d = g["A_fp"][c, f] @ om
                up = m["cFP"][c, f] - 1
                flux = psi[a, c, 0] if d > 0 else (psi[a, up, 0] if up < Cn else 0.0)
                lhs[z] += w * g["RadiusFP"][c, f] * d * flux
    rhs = 2 * np.pi * (g["Volume"] * S[:, 0]).reshape(Z, 4).sum(1)
    np.testing.assert_allclose(lhs, rhs, rtol=1e-12)


def test_source_iteration_global_balance(jittered):
    pts, quads, m, g, q, s = jittered
    rng = np.random.default_rng(4)
    Z = quads.shape[0]
    sig = rng.uniform(1.0, 20.0, (Z, 1))
    sigs = 0.6 * sig
    qq = rng.uniform(0.1, 1.0, (Z, 1))
    phi, it, res, _ = R.source_iteration(m, g, q, s, sig, sigs, qq, 1e-13, 300)
    assert res < 1e-13
    S = (np.repeat(sigs, 4, 0) * phi + np.repeat(qq, 4, 0)) / (2 * np.pi)
    _, _, psi = R.sweep_all(m, g, q, s, S, sig, want_psi=True)
    leak = 0.0
    for a in range(len(q["weight"])):
        for c in range(4 * Z):
            for f in range(2):
                d = g["A_fp"][c, f] @ q["omega"][a]
                if m["cFP"][c, f] > 4 * Z and d > 0:
                    leak += q["weight"][a] * g["RadiusFP"][c, f] * d * psi[a, c, 0]
    absorb = ((np.repeat(sig - sigs, 4, 0) * phi)[:, 0] * g["Volume"]).sum()
    source = (np.repeat(qq, 4, 0)[:, 0] * g["Volume"]).sum()
    assert abs(absorb + leak - source) / source < 1e-10


def _zone_balance_lhs(m, g, q, psi, sig):
    """sum_m w_m [net radius-weighted FP outflow + sigma V psi] per zone."""
    Cn = m["cFP"].shape[0]
    Z = Cn // 4
    lhs = np.zeros(Z)
    for a in range(len(q["weight"])):
        w = q["weight"][a]
        if w == 0:
            continue
        om = q["omega"][a]
        for c in range(Cn):
            z = c // 4
            lhs[z] += w * sig[z, 0] * g["Volume"][c] * psi[a, c, 0]
            for f in range(2):
                d = g["A_fp"][c, f] @ om
                up = m["cFP"][c, f] - 1
                flux = psi[a, c, 0] if d > 0 else (psi[a, up, 0] if up < Cn else 0.0)
                lhs[z] += w * g["RadiusFP"][c, f] * d * flux
    return lhs


def test_time_dependent_constant_solution_is_exact(jittered):
    """(sigma + tau) psi = S + tau psi0 with psi0 isotropic and the same
    incoming boundary flux: psi is constant through every direction."""
    pts, quads, m, g, q, s = jittered
    Z, nb, A = quads.shape[0], m["nbelem"], len(q["weight"])
    rng = np.random.default_rng(11)
    sig = rng.uniform(0.1, 30.0, (Z, 2))
    tau, psi0, psic = 7.5, 0.4, 1.5
    S = np.repeat(sig + tau, 4, 0) * psic - tau * psi0
    prev = np.full((A, 4 * Z, 2), psi0)
    _, _, psi = R.sweep_all(m, g, q, s, S, sig + tau, PsiB=np.full((A, nb, 2), psic),
                            want_psi=True, tau=tau, psi_prev=prev)
    assert np.abs(psi / psic - 1).max() < 1e-13
    # without the tau psi0 source the same data is not that solution
    _, _, psi_no = R.sweep_all(m, g, q, s, S, sig + tau, PsiB=np.full((A, nb, 2), psic),
                               want_psi=True)
    assert np.abs(psi_no / psic - 1).max() > 0.1


def test_time_dependent_zones_conserve_particles(jittered):
    """With an angle-dependent psi_prev the tau source enters each direction's
    balance: sum_m w_m [outflow + (sigma + tau) V psi] = sum_m w_m V (S + tau
    psi_prev[m]) for every zone."""
    pts, quads, m, g, q, s = jittered
    rng = np.random.default_rng(12)
    Z, A = quads.shape[0], len(q["weight"])
    Cn = 4 * Z
    sig = rng.uniform(0.5, 40.0, (Z, 1))
    tau = 3.0
    S = rng.uniform(0.1, 1.0, (Cn, 1))
    prev = rng.uniform(0.0, 2.0, (A, Cn, 1))
    _, _, psi = R.sweep_all(m, g, q, s, S, sig + tau, want_psi=True, tau=tau, psi_prev=prev)
    lhs = _zone_balance_lhs(m, g, q, psi, sig + tau)
    src = S[None, :, 0] + tau * prev[:, :, 0]
    rhs = ((q["weight"][:, None] * src).sum(0) * g["Volume"]).reshape(Z, 4).sum(1)
    np.testing.assert_allclose(lhs, rhs, rtol=1e-12)


def test_tau_source_of_one_direction_reaches_only_its_successors(jittered):
    """psi_prev nonzero in one direction changes that direction and the later
    directions of its xi level (through the half-angle flux), nothing else."""
    pts, quads, m, g, q, s = jittered
    rng = np.random.default_rng(13)
    Z, A = quads.shape[0], len(q["weight"])
    sig = rng.uniform(0.5, 10.0, (Z, 1))
    S = rng.uniform(0.1, 1.0, (4 * Z, 1))
    starts = np.flatnonzero(q["starting"])
    lev = np.searchsorted(starts, np.arange(A), side="right") - 1
    target = starts[1] + 2  # inside the second xi level
    assert not q["starting"][target] and not q["finishing"][target]
    prev = np.zeros((A, 4 * Z, 1))
    prev[target] = rng.uniform(0.5, 1.5, (4 * Z, 1))
    _, _, base = R.sweep_all(m, g, q, s, S, sig + 2.0, want_psi=True, tau=2.0,
                             psi_prev=np.zeros_like(prev))
    _, _, psi = R.sweep_all(m, g, q, s, S, sig + 2.0, want_psi=True, tau=2.0, psi_prev=prev)
    changed = np.abs(psi - base).max(axis=(1, 2)) > 0
    expect = (lev == lev[target]) & (np.arange(A) >= target)
    np.testing.assert_array_equal(changed, expect)


def test_time_step_reduces_to_source_iteration_at_tau_zero(jittered):
    pts, quads, m, g, q, s = jittered
    rng = np.random.default_rng(14)
    Z, A = quads.shape[0], len(q["weight"])
    sig = rng.uniform(1.0, 20.0, (Z, 2))
    sigs, qq = 0.5 * sig, rng.uniform(0.1, 1.0, (Z, 2))
    phi, new, it, res, hist = R.time_step(m, g, q, s, sig, sigs, qq, 0.0,
                                          rng.uniform(0, 1, (A, 4 * Z, 2)), 0.0, 5)
    ref, it_ref, _, hist_ref = R.source_iteration(m, g, q, s, sig, sigs, qq, 0.0, 5)
    assert it == it_ref == 5
    np.testing.assert_allclose(hist, hist_ref, rtol=1e-10)
    # phi after the final (savePsi) sweep is one more iteration; new holds its psi
    S = (np.repeat(sigs, 4, 0) * ref + np.repeat(qq, 4, 0)) / (2 * np.pi)
    ref6, _, psi6 = R.sweep_all(m, g, q, s, S, sig, want_psi=True)
    np.testing.assert_allclose(phi, ref6, rtol=1e-13)
    np.testing.assert_allclose(new, psi6, rtol=1e-13, atol=1e-300)


def test_time_step_global_balance(jittered):
    """At convergence: leakage + sum V (sigma - sigs + tau) phi = sum V q +
    tau sum_m w_m sum V psi_prev[m]."""
    pts, quads, m, g, q, s = jittered
    rng = np.random.default_rng(15)
    Z, A = quads.shape[0], len(q["weight"])
    Cn = 4 * Z
    sig = rng.uniform(1.0, 20.0, (Z, 1))
    sigs, qq = 0.7 * sig, rng.uniform(0.1, 1.0, (Z, 1))
    tau = 5.0
    prev = rng.uniform(0.0, 1.0, (A, Cn, 1))
    phi, new, it, res, _ = R.time_step(m, g, q, s, sig, sigs, qq, tau, prev, 1e-14, 300)
    assert res < 1e-14
    leak = 0.0
    for a in range(A):
        for c in range(Cn):
            for f in range(2):
                d = g["A_fp"][c, f] @ q["omega"][a]
                if m["cFP"][c, f] > Cn and d > 0:
                    leak += q["weight"][a] * g["RadiusFP"][c, f] * d * new[a, c, 0]
    V = g["Volume"]
    absorb = ((np.repeat(sig - sigs + tau, 4, 0) * phi)[:, 0] * V).sum()
    source = (np.repeat(qq, 4, 0)[:, 0] * V).sum() + tau * (q["weight"] @ (prev[:, :, 0] @ V))
    assert abs(absorb + leak - source) / source < 1e-10
    # phi is the weighted sum of the saved psi
    np.testing.assert_allclose(np.tensordot(q["weight"], new, 1), phi, rtol=1e-13)


def test_product_host_arrays_equal_oracle(jittered):
    pts, quads, m, g, q, s = jittered
    p = U.RzProblem.from_quad_mesh(pts, quads, 2, 3)
    pm, pg, pq, ps = p.mesh(), p.geometry(), p.quadrature(), p.schedule()
    for k in ("numCorner", "cOffSet", "cFP", "cEZ", "CToFace", "zoneOpp", "faceOpp", "px"):
        np.testing.assert_array_equal(pm[k], m[k], err_msg=k)
    assert p.nbelem == m["nbelem"]
    for d, ref in ((pg, g), (pq, q), (ps, s)):
        for k, v in ref.items():
            np.testing.assert_array_equal(d[k], v, err_msg=k)


def test_product_orients_counterclockwise_zones():
    pts, quads = R.structured(3, 2, jitter=0.2, seed=7)
    ccw = quads[:, [0, 3, 2, 1]]
    a = U.RzProblem.from_quad_mesh(pts, quads, 1, 2)
    b = U.RzProblem.from_quad_mesh(pts, ccw, 1, 2)
    np.testing.assert_array_equal(a.points()[1], b.points()[1])
    np.testing.assert_array_equal(a.geometry()["A_fp"], b.geometry()["A_fp"])


def test_product_structured_mesh_matches_oracle_rebuild():
    p = U.RzProblem(6, 5, 2, 2, r0=0.5, r1=2.0, z0=-1.0, z1=1.0, jitter=0.3, seed=11)
    pts, quads = p.points()
    m = R.quad_mesh(pts, quads)
    g = R.geometry(m)
    for k in ("cFP", "cEZ", "px"):
        np.testing.assert_array_equal(p.mesh()[k], m[k])
    for k, v in g.items():
        np.testing.assert_array_equal(p.geometry()[k], v, err_msg=k)
    assert pts[:, 0].min() == 0.5 and pts[:, 1].max() == 1.0


def test_bad_quad_meshes_are_refused():
    pts, quads = R.structured(2, 2)
    bad = quads.copy()
    bad[1, 2] = len(pts)
    with pytest.raises(U.UmtError):
        U.RzProblem.from_quad_mesh(pts, bad, 1, 1)
    neg = pts.copy()
    neg[0, 0] = -0.1
    with pytest.raises(U.UmtError):
        U.RzProblem.from_quad_mesh(neg, quads, 1, 1)
    with pytest.raises(U.UmtError):
        U.RzProblem(2, 2, 5, 1)
