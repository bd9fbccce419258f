"""The oracle's GTA restatement (oracle/gta.py) against what the mathematics
fixes: one-group grey opacities, the grey-correction equation solved by the
BiCGStab, no correction without scattering, and the accelerated iteration
reaching plain source iteration's fixed point in far fewer sweeps."""
import numpy as np
import pytest

import umt_amd as U
from oracle import gta as OG
from oracle import oracle as O


@pytest.fixture(scope="module")
def prob():
    p = U.TiledProblem(2, 2, 1, 2, 2)
    m, g = p.mesh(), p.geometry()
    gq = p.gta()
    s2 = O.schedule(m, g["A_fp"], g["A_ez"], gq["omega"], par_tol=1e-12)
    q2 = dict(omega=gq["omega"], weight=gq["weight"])
    return p, q2, s2


def test_level_symmetric_s2_quadrature(prob):
    p, q2, _ = prob
    om, w = q2["omega"], q2["weight"]
    assert om.shape == (8, 3)
    np.testing.assert_allclose(np.abs(om), 0.577350269189625, rtol=0, atol=0)
    assert abs(w.sum() - 4 * np.pi) < 1e-13
    # same octant numbering as the product quadrature
    assert np.array_equal(np.sign(om), np.sign(p.quadrature()["omega"][:8]))


def test_grey_opacities_one_group():
    sigt = np.array([[2.0, 1.0, 5.0]])
    sigs = np.array([[1.5, 1e-12, 0.0]])
    tot, scat, eps, stretched = OG.grey_opacities(sigt, sigs)
    np.testing.assert_allclose(tot, [[2.0, 1.0 - 1e-12, 5.0]])
    np.testing.assert_allclose(scat, [[1.5, 0.0, 0.0]])
    np.testing.assert_allclose(eps, [[2.0, 1.0, 1.0]])   # 1/sqrt(1 - 0.75) = 2
    np.testing.assert_allclose(stretched, [[1.0, 1.0 - 1e-12, 5.0]])
    # stretched problem has no scattering: sigt/eps - eps sigma_a = 0
    np.testing.assert_allclose(tot[0, 0] / eps[0, 0] - eps[0, 0] * (tot[0, 0] - scat[0, 0]), 0.0,
                               atol=1e-15)


def test_bicgstab_solves_the_grey_correction_equation(prob):
    """F from the solver satisfies (I - M) F = M_src(gsrc) to the grey tolerance."""
    p, q2, s2 = prob
    G = 2
    st, ss, _ = U.draw_cross_sections(3, p.nzones, G, (1e-1, 1e2), (0.5, 0.99), (0, 0))
    solver = OG.GreySolver(p.mesh(), p.geometry(), q2, s2, st, ss)
    rng = np.random.default_rng(0)
    gsrc = rng.uniform(0.0, 1.0, size=(p.ncorner, G)) * np.repeat(ss, 8, axis=0)
    half = np.ones((p.ncorner, G))
    F, n = solver.solve(gsrc, half, eps_grey=1e-12, max_sweeps=201, eps_point=1e-12)
    b = solver.grey_sweep(np.zeros_like(gsrc), gsrc)
    resid = F - solver.grey_sweep(F) - b
    assert np.linalg.norm(resid) < 1e-8 * np.linalg.norm(b)


def test_no_scattering_means_no_correction(prob):
    p, q2, s2 = prob
    G = 1
    st = np.full((p.nzones, G), 2.0)
    ss = np.zeros((p.nzones, G))
    qq = np.ones((p.nzones, G))
    phi, it, res, _, _ = OG.source_iteration_gta(p.mesh(), p.geometry(), p.quadrature(),
                                                 p.schedule(), q2, s2, st, ss, qq, 1e-12, 10)
    ref, rit, _, _ = O.source_iteration(p.mesh(), p.geometry(), p.quadrature(), p.schedule(),
                                        st, ss, qq, 1e-12, 10)
    assert it == rit == 2
    np.testing.assert_array_equal(phi, ref)


def test_accelerated_iteration_reaches_the_plain_fixed_point(prob):
    p, q2, s2 = prob
    G = 2
    st, ss, qq = U.draw_cross_sections(7, p.nzones, G, (1e-2, 1e3), (0.9, 0.999), (0.0, 1.0))
    args = (p.mesh(), p.geometry(), p.quadrature(), p.schedule())
    phi, it, res, hist, sweeps = OG.source_iteration_gta(*args, q2, s2, st, ss, qq, 1e-11, 200,
                                                         nthreads=4)
    ref, rit, rres, _ = O.source_iteration(*args, st, ss, qq, 1e-13, 50000, nthreads=4)
    assert res < 1e-11 and rres < 1e-13
    assert it * 10 < rit
    np.testing.assert_allclose(phi, ref, rtol=1e-7)


# ---------------------------------------------------------------------------
# The transfer-matrix ("new") solver of the oracle, pinned independently of
# its own code: the matrices by unit-source sweeps of the oracle's C
# restatement of SweepUCBxyz, the incident-flux sweep by the same C sweep
# and by linearity, the LU factors by their residual, and the solver by the
# grey problem's solution that two other methods reach.


@pytest.fixture(scope="module")
def tile():
    p = U.TiledProblem(1, 1, 1, 2, 2)
    m, g = p.mesh(), p.geometry()
    gq = p.gta()
    s2 = O.schedule(m, g["A_fp"], g["A_ez"], gq["omega"], par_tol=1e-12)
    q2 = dict(omega=gq["omega"], weight=gq["weight"])
    G = 2
    st, ss, _ = U.draw_cross_sections(3, p.nzones, G, (1e-2, 1e3), (0.5, 0.99), (0, 0))
    return p, q2, s2, st, ss


def test_transfer_matrices_equal_unit_source_sweeps(tile):
    """T[z, :, s, :] is zone z's scalar flux from a unit source at its corner
    s alone (no other source anywhere, so no incident flux reaches z)."""
    p, q2, s2, st, ss = tile
    sig_g, _, _, _ = OG.grey_opacities(st, ss)
    TT = OG.transfer_matrices(p.mesh(), p.geometry(), q2, sig_g)
    G = st.shape[1]
    for z in range(p.nzones):
        for s in range(8):
            S = np.zeros((p.ncorner, G))
            S[8 * z + s] = 1.0
            phi, _ = O.sweep_all(p.mesh(), p.geometry(), q2, s2, S, np.ascontiguousarray(sig_g),
                                 par_tol=1e-12)
            ref = phi[8 * z:8 * z + 8].T
            assert np.max(np.abs(TT[z, :, s, :] - ref)) < 1e-12 * np.abs(ref).max(), (z, s)


def test_incident_flux_sweep_is_the_sweep_minus_the_local_response(tile):
    p, q2, s2, st, ss = tile
    sig_g, _, _, _ = OG.grey_opacities(st, ss)
    G = st.shape[1]
    rng = np.random.default_rng(1)
    tsa = rng.uniform(0.0, 1.0, (p.ncorner, G))
    phi, phi_inc = OG.pinc_sweep(p.mesh(), p.geometry(), q2, s2, sig_g, tsa)
    ref, _ = O.sweep_all(p.mesh(), p.geometry(), q2, s2, tsa, np.ascontiguousarray(sig_g),
                         par_tol=1e-12)
    np.testing.assert_allclose(phi, ref, rtol=1e-13)
    TT = OG.transfer_matrices(p.mesh(), p.geometry(), q2, sig_g)
    local = OG._corner_vec(np.einsum("zgsc,zgs->zgc", TT, OG._zone_vec(tsa, G)))
    assert np.max(np.abs(phi - local - phi_inc)) < 1e-14 * np.abs(phi).max()


def test_zone_lu_factors_solve_the_zone_systems(tile):
    p, q2, s2, st, ss = tile
    sig_g, scat_g, _, _ = OG.grey_opacities(st, ss)
    TT = OG.transfer_matrices(p.mesh(), p.geometry(), q2, sig_g)
    LU = OG.scalar_intensity_decompose(TT, scat_g)
    A = np.eye(8) - OG.WTISO * scat_g[:, :, None, None] * np.swapaxes(TT, 2, 3)
    L = np.tril(LU, -1) + np.eye(8)
    Uu = np.triu(LU)
    assert np.max(np.abs(L @ Uu - A)) < 1e-14
    rng = np.random.default_rng(2)
    b = rng.uniform(0.0, 1.0, (p.nzones, st.shape[1], 8))
    x = OG.scalar_intensity_solve(LU, b)
    np.testing.assert_allclose(x, np.linalg.solve(A, b[..., None])[..., 0], rtol=1e-12)


def test_new_solver_reaches_the_grey_solution(tile):
    """Kernel 2 and kernel 1 of the grey solve give the grey problem's
    solution, which plain grey source iteration reaches to 1e-14, and
    kernel 2 uses far fewer grey sweeps."""
    p, q2, s2, st, ss = tile
    G = st.shape[1]
    rng = np.random.default_rng(4)
    gsrc = rng.uniform(0.0, 1.0, (p.ncorner, G))
    half = rng.uniform(1.0, 2.0, (p.ncorner, G))
    old = OG.GreySolver(p.mesh(), p.geometry(), q2, s2, st, ss)
    new = OG.NewGreySolver(p.mesh(), p.geometry(), q2, s2, st, ss)
    Fo, _ = old.solve(gsrc, half, 1e-11)
    Fn, _ = new.solve(gsrc, half, 1e-11, hard_max=False)
    sig_g, scat_g, _, _ = OG.grey_opacities(st, ss)
    sc = np.repeat(scat_g, 8, axis=0)
    F = np.zeros_like(gsrc)
    for _ in range(20000):
        Fx, _ = O.sweep_all(p.mesh(), p.geometry(), q2, s2, OG.WTISO * (sc * F + gsrc),
                            np.ascontiguousarray(sig_g), par_tol=1e-12)
        done = np.max(np.abs(Fx - F)) < 1e-15 * np.abs(Fx).max()
        F = Fx
        if done:
            break
    assert np.max(np.abs(Fn - F)) < 1e-9 * np.abs(F).max()
    assert np.max(np.abs(Fo - F)) < 1e-9 * np.abs(F).max()
    assert 2 * new.sweeps < old.sweeps


def test_new_solver_stops_at_the_hard_grey_iteration_cap(tile):
    """The new solver enforces the grey iteration maximum as a hard cap
    (GTA%enforceHardGTAIterMax, GreyAcceleration_mod.F90:169; the exit at
    GTASolver.F90:373-375): with a tolerance it cannot reach it stops at
    nGreyIter = max_sweeps exactly. Without the cap the same solve runs on
    past max_sweeps, and the old solver has no cap by default."""
    p, q2, s2, st, ss = tile
    G = st.shape[1]
    rng = np.random.default_rng(5)
    gsrc = rng.uniform(0.0, 1.0, (p.ncorner, G))
    half = rng.uniform(1.0, 2.0, (p.ncorner, G))
    new = OG.NewGreySolver(p.mesh(), p.geometry(), q2, s2, st, ss)
    _, n_cap = new.solve(gsrc, half, 1e-300, max_sweeps=5, eps_point=1e-300)
    assert n_cap == 5
    _, n_more = new.solve(gsrc, half, 1e-13, max_sweeps=5, eps_point=1e-13, hard_max=False)
    assert n_more > 5
    old = OG.GreySolver(p.mesh(), p.geometry(), q2, s2, st, ss)
    _, n_old = old.solve(gsrc, half, 1e-13, max_sweeps=5, eps_point=1e-13)
    assert n_old > 5
