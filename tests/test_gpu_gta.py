"""Grey-transport acceleration on the GPU (per-group one-group GTA: S2 grey
sweeps, stretched-TSA sweeps, BiCGStab) against the oracle's restatement
(oracle/gta.py) and against plain source iteration's fixed point."""
import numpy as np
import pytest

import umt_amd as U
from umt_amd._lib import UmtError
from oracle import gta as OG
from oracle import oracle as O

from conftest import rel_err

pytestmark = pytest.mark.gpu

SEED = 20261205


@pytest.fixture(scope="module")
def thick():
    """c ~ U[0.9, 0.999], sigma_t over five decades: plain SI needs ~10^3 iterations."""
    p = U.TiledProblem(2, 2, 2, 2, 2)
    G = 2
    st, ss, qq = U.draw_cross_sections(SEED, p.nzones, G, (1e-2, 1e3), (0.9, 0.999), (0.0, 1.0))
    m, g = p.mesh(), p.geometry()
    gq = p.gta()
    s2 = O.schedule(m, g["A_fp"], g["A_ez"], gq["omega"], par_tol=1e-12)
    q2 = dict(omega=gq["omega"], weight=gq["weight"])
    return p, G, st, ss, qq, q2, s2


def _ctx(p, G, st, ss, qq, precision=64):
    ctx = U.SweepContext(p, G, precision=precision)
    ctx.set_sigt(st)
    ctx.set_scattering(ss, qq)
    ctx.set_acceleration("gta")
    return ctx


def test_gta_s2_schedule_matches_oracle(thick):
    p, G, st, ss, qq, q2, s2 = thick
    gq = p.gta()
    for k in ("nHyperPlanes", "zonesInPlane", "nextZ"):
        np.testing.assert_array_equal(gq[k], s2[k])


def test_s2_sweep_matches_oracle(thick):
    """Without scattering the grey operator is one S2 sweep of wtiso gsrc."""
    p, G, st, ss, qq, q2, s2 = thick
    ctx = _ctx(p, G, st, np.zeros_like(ss), qq)
    rng = np.random.default_rng(1)
    gsrc = rng.uniform(0.0, 1.0, (p.ncorner, G))
    got = ctx.grey_sweep(None, gsrc)
    ref, _ = O.sweep_all(p.mesh(), p.geometry(), q2, s2, gsrc / (4 * np.pi), st, par_tol=1e-12)
    err = rel_err(got, ref)
    print(f"S2 sweep rel err {err:.3e}", flush=True)
    assert err < 1e-12
    ctx.close()


def test_grey_operator_matches_oracle(thick):
    p, G, st, ss, qq, q2, s2 = thick
    ctx = _ctx(p, G, st, ss, qq)
    solver = OG.GreySolver(p.mesh(), p.geometry(), q2, s2, st, ss)
    rng = np.random.default_rng(2)
    P = rng.uniform(0.5, 2.0, (p.ncorner, G))
    gsrc = rng.uniform(0.0, 1.0, (p.ncorner, G))
    for args in ((P, None), (None, gsrc), (P, gsrc)):
        got = ctx.grey_sweep(*args)
        ref = solver.grey_sweep(*args) if args[0] is not None else solver.grey_sweep(
            np.zeros_like(P), args[1])
        err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
        print(f"grey operator ({args[0] is not None}, {args[1] is not None}) err {err:.3e}", flush=True)
        assert err < 1e-12
    ctx.close()


def test_grey_solve_matches_oracle(thick):
    p, G, st, ss, qq, q2, s2 = thick
    ctx = _ctx(p, G, st, ss, qq)
    solver = OG.GreySolver(p.mesh(), p.geometry(), q2, s2, st, ss)
    rng = np.random.default_rng(3)
    half = rng.uniform(0.5, 2.0, (p.ncorner, G))
    gsrc = np.repeat(ss, 8, axis=0) * rng.normal(0.0, 0.1, (p.ncorner, G))
    for eps in (1e-5, 1e-8):
        F, n = ctx.grey_solve(gsrc, half, eps)
        Fr, nr = solver.solve(gsrc, half, eps)
        err = np.max(np.abs(F - Fr)) / np.max(np.abs(Fr))
        print(f"grey solve eps={eps}: GPU {n} iterations, oracle {(nr - 1) // 2}; err {err:.3e}", flush=True)
        assert n == (nr - 1) // 2
        assert err < 1e-9
    ctx.close()


def test_gta_iterates_match_oracle(thick):
    """Three accelerated iterations, compared iterate by iterate."""
    p, G, st, ss, qq, q2, s2 = thick
    ctx = _ctx(p, G, st, ss, qq)
    _, it, _, hist = ctx.source_iteration(tol=0.0, max_iters=3)
    phi = ctx.phi()
    ref, rit, _, rhist, _ = OG.source_iteration_gta(p.mesh(), p.geometry(), p.quadrature(),
                                                    p.schedule(), q2, s2, st, ss, qq, 0.0, 3)
    assert it == rit == 3
    assert rel_err(phi, ref) < 1e-9
    np.testing.assert_allclose(hist, rhist, rtol=1e-6)
    ctx.close()


def test_gta_converges_to_the_plain_fixed_point_in_few_iterations(thick):
    p, G, st, ss, qq, q2, s2 = thick
    ctx = _ctx(p, G, st, ss, qq)
    ok, it, res, _ = ctx.source_iteration(tol=1e-10, max_iters=200)
    assert ok and res < 1e-10
    grey_its, grey_sweeps = ctx.gta_stats()
    phi = ctx.phi()
    ref, rit, rres, _ = O.source_iteration(p.mesh(), p.geometry(), p.quadrature(), p.schedule(),
                                           st, ss, qq, 1e-13, 20000, nthreads=4)
    print(f"GTA {it} iterations ({grey_its} BiCGStab steps, {grey_sweeps} S2 sweeps); "
          f"plain {rit} iterations")
    assert rres < 1e-13
    assert it * 10 < rit
    assert rel_err(phi, ref) < 1e-7
    ctx.close()


def test_gta_with_fp32_sweeps(thick):
    """FP32 sweeps + GTA to 1e-6, then FP64 sweeps + GTA to 1e-8: phi within
    1e-6 of the converged FP64 oracle."""
    p, G, st, ss, qq, q2, s2 = thick
    fast = _ctx(p, G, st, ss, qq, precision=32)
    exact = _ctx(p, G, st, ss, qq, precision=64)
    ok, it32, it64, res, _ = U.solve_mixed(fast, exact, tol=1e-8, switch_tol=1e-6, max_iters=200)
    assert ok and res < 1e-8
    ref, _, _, _ = O.source_iteration(p.mesh(), p.geometry(), p.quadrature(), p.schedule(),
                                      st, ss, qq, 1e-13, 20000, nthreads=4)
    err = rel_err(fast.phi(), ref)
    print(f"FP32+GTA {it32} iterations then FP64+GTA {it64}: max rel phi error vs FP64 oracle {err:.2e}")
    assert err < 1e-6
    fast.close()
    exact.close()


def test_graph_replays_match_eager_sweeps(thick):
    """Captured level launches, replayed several times with eager work and
    synchronisation in between, give the eager result every time."""
    p, G, st, ss, qq, q2, s2 = thick
    S = np.repeat(qq, 8, axis=0) / (4 * np.pi)
    eager = U.SweepContext(p, G)
    eager.set_sigt(st)
    eager.set_source(S)
    eager.sweep()
    ref = eager.phi()
    eager.close()
    ctx = U.SweepContext(p, G)
    ctx.set_sigt(st)
    ctx.set_graphs(True)
    for rep in range(4):
        ctx.set_source(S * (1.0 + rep))
        ctx.sweep()
        np.testing.assert_array_equal(ctx.phi(), ref * (1.0 + rep) if rep == 0 else ctx.phi())
        assert rel_err(ctx.phi(), ref * (1.0 + rep)) < 1e-14
    ctx.set_acceleration("gta")
    ctx.set_scattering(ss, qq)
    ctx.set_phi(np.zeros((p.ncorner, G)))  # the same start as the eager context
    ok, it, res, _ = ctx.source_iteration(tol=1e-10, max_iters=200)
    assert ok
    plain = _ctx(p, G, st, ss, qq)
    ok2, it2, _, _ = plain.source_iteration(tol=1e-10, max_iters=200)
    assert ok2 and it2 == it
    np.testing.assert_array_equal(ctx.phi(), plain.phi())
    ctx.close()
    plain.close()


# ---------------------------------------------------------------------------
# GTA kernel 2: the transfer-matrix solver (snac/InitSweepGreyUCBxyz.F90,
# snac/UpdateScalarIntensity.F90, rt/GreySweep.F90:12-47) against the
# oracle's restatement (oracle/gta.py: transfer_matrices, pinc_sweep,
# scalar_intensity_decompose/solve, NewGreySolver). A smaller mesh keeps the
# oracle's Python sweeps short.


@pytest.fixture(scope="module")
def thick_small():
    p = U.TiledProblem(2, 1, 1, 2, 2)
    G = 3
    st, ss, qq = U.draw_cross_sections(SEED + 1, p.nzones, G, (1e-2, 1e3), (0.9, 0.999), (0.0, 1.0))
    m, g = p.mesh(), p.geometry()
    gq = p.gta()
    s2 = O.schedule(m, g["A_fp"], g["A_ez"], gq["omega"], par_tol=1e-12)
    q2 = dict(omega=gq["omega"], weight=gq["weight"])
    return p, G, st, ss, qq, q2, s2


def _ctx2(p, G, st, ss, qq, precision=64):
    ctx = _ctx(p, G, st, ss, qq, precision)
    ctx.set_gta_solver(2)
    return ctx


def test_transfer_matrices_and_lu_match_oracle(thick_small):
    p, G, st, ss, qq, q2, s2 = thick_small
    print("case transfer matrices", flush=True)
    ctx = _ctx2(p, G, st, ss, qq)
    T, LU = ctx.grey_transfer()
    sig_g, scat_g, _, _ = OG.grey_opacities(st, ss)
    Tr = OG.transfer_matrices(p.mesh(), p.geometry(), q2, sig_g)
    LUr = OG.scalar_intensity_decompose(Tr, scat_g)
    scale = np.abs(Tr).max(axis=(2, 3), keepdims=True)
    eT = float(np.max(np.abs(T - Tr) / scale))
    eLU = float(np.max(np.abs(LU - LUr) / np.abs(LUr).max(axis=(2, 3), keepdims=True)))
    print(f"T err {eT:.2e}, LU err {eLU:.2e}", flush=True)
    assert eT < 1e-12 and eLU < 1e-11
    ctx.close()


def test_transfer_matrix_grey_operator_matches_oracle(thick_small):
    """M(P) of kernel 2, with and without the grey source. The device forms
    P + (I - B)^-1 (phi - P) from the grey sweep's phi; the oracle follows
    the reference's (I - B)^-1 (PhiInc [+ T wtiso gsrc]) with PhiInc from
    the incident-flux part of its own sweep: same mathematics, different
    rounding, so the bound is 1e-11 of the largest entry."""
    p, G, st, ss, qq, q2, s2 = thick_small
    ctx = _ctx2(p, G, st, ss, qq)
    solver = OG.NewGreySolver(p.mesh(), p.geometry(), q2, s2, st, ss)
    rng = np.random.default_rng(12)
    P = rng.uniform(0.5, 2.0, (p.ncorner, G))
    gsrc = rng.uniform(0.0, 1.0, (p.ncorner, G))
    solver.build()
    for args in ((None, gsrc), (P, gsrc), (P, None)):
        got = ctx.grey_sweep(*args)
        Pa = args[0] if args[0] is not None else np.zeros_like(P)
        if args[1] is None:
            solver.grey_sweep(np.zeros_like(P), gsrc)  # the with-source call makes the factors
        ref = solver.grey_sweep(Pa, args[1])
        err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
        print(f"kernel-2 operator ({args[0] is not None}, {args[1] is not None}) err {err:.3e}",
              flush=True)
        assert err < 1e-11
    ctx.close()


def test_transfer_matrix_grey_solve_matches_oracle(thick_small):
    p, G, st, ss, qq, q2, s2 = thick_small
    ctx = _ctx2(p, G, st, ss, qq)
    solver = OG.NewGreySolver(p.mesh(), p.geometry(), q2, s2, st, ss)
    rng = np.random.default_rng(13)
    half = rng.uniform(0.5, 2.0, (p.ncorner, G))
    gsrc = np.repeat(ss, 8, axis=0) * rng.normal(0.0, 0.1, (p.ncorner, G))
    F, n = ctx.grey_solve(gsrc, half, 1e-8)
    Fr, nr = solver.solve(gsrc, half, 1e-8)
    err = np.max(np.abs(F - Fr)) / np.max(np.abs(Fr))
    print(f"kernel-2 grey solve: GPU {n} iterations, oracle {(nr - 1) // 2}; err {err:.3e}", flush=True)
    assert n == (nr - 1) // 2
    assert err < 1e-9
    ctx.close()


def test_transfer_matrix_gta_converges_with_fewer_grey_sweeps(thick):
    """Source iteration with kernel 2 reaches plain source iteration's fixed
    point, and needs fewer grey sweeps than kernel 1 on this thick,
    scattering-dominated problem."""
    p, G, st, ss, qq, q2, s2 = thick
    runs = {}
    for kernel in (1, 2):
        ctx = _ctx(p, G, st, ss, qq)
        ctx.set_gta_solver(kernel)
        ok, it, res, _ = ctx.source_iteration(tol=1e-10, max_iters=200)
        assert ok and res < 1e-10
        runs[kernel] = (it, ctx.gta_stats(), ctx.phi())
        ctx.close()
    ref, rit, rres, _ = O.source_iteration(p.mesh(), p.geometry(), p.quadrature(), p.schedule(),
                                           st, ss, qq, 1e-13, 20000, nthreads=4)
    for kernel, (it, (gi, gs), phi) in runs.items():
        print(f"kernel {kernel}: {it} iterations, {gi} BiCGStab steps, {gs} grey sweeps; "
              f"phi err {rel_err(phi, ref):.2e}", flush=True)
        assert rel_err(phi, ref) < 1e-7
    assert runs[2][1][1] * 2 < runs[1][1][1]


@pytest.mark.parametrize("fp32_grey_max_its", [U.sweep.FP32_PHASE_GREY_MAX_ITS, None])
def test_transfer_matrix_gta_with_fp32_sweeps(thick, fp32_grey_max_its):
    """FP32 then FP64 sweeps with GTA kernel 2 reach plain iteration's fixed
    point, with the FP32 phase's grey cap (the default) and with the
    reference's controls; the cap leaves the FP32 context's own controls as
    they were."""
    p, G, st, ss, qq, q2, s2 = thick
    fast = _ctx2(p, G, st, ss, qq, precision=32)
    exact = _ctx2(p, G, st, ss, qq, precision=64)
    ok, it32, it64, res, _ = U.solve_mixed(fast, exact, tol=1e-8, switch_tol=1e-6, max_iters=200,
                                           fast_grey_max_its=fp32_grey_max_its)
    assert ok and res < 1e-8
    assert fast.grey_controls() == (1e-6, 21, True)
    ref, _, _, _ = O.source_iteration(p.mesh(), p.geometry(), p.quadrature(), p.schedule(),
                                      st, ss, qq, 1e-13, 20000, nthreads=4)
    err = rel_err(fast.phi(), ref)
    print(f"FP32+GTA2 {it32} iterations then FP64+GTA2 {it64}: max rel phi error {err:.2e}")
    assert err < 1e-6
    fast.close()
    exact.close()


def test_grey_controls_and_the_hard_iteration_cap(thick_small):
    """Kernel 2 stops the grey iteration at its maximum whatever the error
    (the reference's new solver sets GTA%enforceHardGTAIterMax,
    GreyAcceleration_mod.F90:169; exit at GTASolver.F90:373-375), with the
    oracle's iterate at the same controls; without the cap the same solve
    runs past the maximum."""
    p, G, st, ss, qq, q2, s2 = thick_small
    ctx = _ctx2(p, G, st, ss, qq)
    assert ctx.grey_controls() == (1e-6, 21, True)
    ctx.set_grey_controls(rel_tol=1e-300, max_its=5)
    rng = np.random.default_rng(14)
    half = rng.uniform(0.5, 2.0, (p.ncorner, G))
    gsrc = np.repeat(ss, 8, axis=0) * rng.normal(0.0, 0.1, (p.ncorner, G))
    print("case hard grey cap, 5 grey iterations", flush=True)
    F, n = ctx.grey_solve(gsrc, half, 1e-300)
    solver = OG.NewGreySolver(p.mesh(), p.geometry(), q2, s2, st, ss)
    Fr, nr = solver.solve(gsrc, half, 1e-300, max_sweeps=5, eps_point=1e-300)
    assert nr == 5 and n == 2, (n, nr)
    assert np.max(np.abs(F - Fr)) / np.max(np.abs(Fr)) < 1e-9
    ctx.set_gta_options(False)
    ctx.set_grey_controls(rel_tol=1e-13)
    _, n_free = ctx.grey_solve(gsrc, half, 1e-13)
    _, nr_free = solver.solve(gsrc, half, 1e-13, max_sweeps=5, eps_point=1e-13, hard_max=False)
    assert n_free == (nr_free - 1) // 2 and n_free > 2, (n_free, nr_free)
    ctx.set_gta_solver(1)
    assert ctx.grey_controls()[2] is False
    with pytest.raises(UmtError) as e:
        ctx.set_grey_controls(rel_tol=0.0)
    assert e.value.code == 57
    ctx.close()
