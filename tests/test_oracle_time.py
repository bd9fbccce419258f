"""The oracle's time-dependent step (oracle.time_step) against what the
transport equation fixes: tau = 0 is the steady iteration, every sweep
balances with the tau terms, and a very large tau leaves Psi unchanged."""
import numpy as np
import pytest

import umt_amd as U
from oracle import oracle as O


@pytest.fixture(scope="module")
def prob():
    p = U.TiledProblem(2, 2, 1, 2, 2)
    G = 2
    st, ss, qq = U.draw_cross_sections(31, p.nzones, G, (0.1, 10.0), (0.0, 0.9), (0.5, 1.0))
    return p, G, st, ss, qq


def test_tau_zero_is_the_steady_iteration(prob):
    p, G, st, ss, qq = prob
    m, g, q, s = p.mesh(), p.geometry(), p.quadrature(), p.schedule()
    psi0 = np.zeros((p.nangles, p.ncorner, G))
    phi, _, it, res, hist = O.time_step(m, g, q, s, st, ss, qq, 0.0, psi0, 1e-9, 300)
    ref, rit, rres, rhist = O.source_iteration(m, g, q, s, st, ss, qq, 1e-9, 300)
    assert it == rit
    np.testing.assert_allclose(hist, rhist, rtol=1e-6, atol=1e-15)
    # phi of time_step comes from one more (the save) sweep
    assert np.max(np.abs(phi - ref) / ref) < 1e-8


def leakage(p, m, g, q, psib):
    Cn = p.ncorner
    bc, bf = np.nonzero(m["cFP"] > Cn)
    bidx = m["cFP"][bc, bf] - Cn - 1
    leak = 0.0
    for a in range(p.nangles):
        afp = g["A_fp"][bc, bf] @ q["omega"][a]
        out = afp > 0
        leak = leak + q["weight"][a] * (afp[out, None] * psib[a, bidx[out]]).sum(0)
    return leak


def test_time_dependent_sweep_balances(prob):
    """Each sweep: leakage + (sigma_t + tau) V phi = (sigma_s phi_src + q) V
    + tau V phi_prev, per group (corner balance integrated over angle)."""
    p, G, st, ss, qq = prob
    m, g, q, s = p.mesh(), p.geometry(), p.quadrature(), p.schedule()
    tau = 7.0
    rng = np.random.default_rng(2)
    psi_prev = rng.uniform(0.1, 1.0, (p.nangles, p.ncorner, G))
    phi_src = rng.uniform(0.1, 1.0, (p.ncorner, G))
    phi, psi_new, _, _, _ = O.time_step(m, g, q, s, st, ss, qq, tau, psi_prev, 0.0, 0, phi0=phi_src)
    S = (np.repeat(ss, 8, axis=0) * phi_src + np.repeat(qq, 8, axis=0)) / (4 * np.pi)
    # the same sweep through the all-angle oracle gives the exiting fluxes
    Cn = p.ncorner
    V = g["Volume"][:, None]
    phi_prev = np.einsum("a,acg->cg", q["weight"], psi_prev)
    starts = np.concatenate([[0], np.cumsum(s["nHyperPlanes"])])
    nb = int(np.count_nonzero(m["cFP"] > Cn))
    psib = np.zeros((p.nangles, nb, G))
    for a in range(p.nangles):
        Psi1 = np.zeros((Cn + nb, G))
        PsiB = np.zeros((nb, G))
        O.sweep_angle(m, g, np.ascontiguousarray(S), np.ascontiguousarray(st + tau), q["omega"][a],
                      q["weight"][a], s["nextZ"][a], s["nextC"][a],
                      s["zonesInPlane"][starts[a]:starts[a + 1]], Psi1, PsiB, np.zeros((Cn, G)),
                      tau=tau, Psi=np.ascontiguousarray(psi_prev[a]))
        psib[a] = PsiB
    lhs = leakage(p, m, g, q, psib) + ((np.repeat(st, 8, axis=0) + tau) * V * phi).sum(0)
    rhs = (4 * np.pi * S * V).sum(0) + (tau * V * phi_prev).sum(0)
    np.testing.assert_allclose(lhs, rhs, rtol=1e-12)
    np.testing.assert_allclose(phi, np.einsum("a,acg->cg", q["weight"], psi_new), rtol=1e-13)


def test_large_tau_keeps_psi(prob):
    p, G, st, ss, qq = prob
    m, g, q, s = p.mesh(), p.geometry(), p.quadrature(), p.schedule()
    rng = np.random.default_rng(3)
    psi_prev = rng.uniform(0.5, 1.0, (p.nangles, p.ncorner, G))
    _, psi_new, _, _, _ = O.time_step(m, g, q, s, st, ss, qq, 1e9, psi_prev, 0.0, 2)
    assert np.max(np.abs(psi_new - psi_prev) / psi_prev) < 1e-6
