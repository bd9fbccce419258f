"""Meshes whose schedules have dependency cycles, on the GPU: the HIP sweep
with cycle lists, cyclic zones and split cyclic planes against the oracle
(FP64 within 1e-12), for every angle-block and groups-per-lane layout, with
graph replay, through the drop-in, in time-dependent cycles and in the
FP32 -> FP64 mixed solve."""
import ctypes as C

import numpy as np
import pytest

import umt_amd as U
from umt_amd import _lib
from oracle import oracle as O

from conftest import CYCLIC_MESHES, cyclic_problem, rel_err

pytestmark = pytest.mark.gpu

XS = dict(sigt=(0.5, 50.0), c=(0.0, 0.9), q=(0.0, 1.0))


@pytest.fixture(scope="module", params=CYCLIC_MESHES, ids=lambda p: f"n{p[0]}-a{p[1]}-t{p[3]}")
def cyc(request):
    p = cyclic_problem(*request.param)
    assert len(p.schedule()["cycleList"]) > 0
    return p


def _ctx(p, G, xs, precision=64):
    st, ss, qq = xs
    ctx = U.SweepContext(p, G, precision=precision)
    ctx.set_sigt(st)
    ctx.set_scattering(ss, qq)
    return ctx


def _oracle_si(p, xs, iters):
    st, ss, qq = xs
    phi, it, _, _ = O.source_iteration(p.mesh(), p.geometry(), p.quadrature(), p.schedule(), st,
                                       ss, qq, 0.0, iters)
    assert it == iters
    return phi


@pytest.mark.parametrize("G,K,n", [(3, 1, 1000), (4, 1, 3), (4, 2, 1), (4, 4, 1000), (8, 4, 5),
                                   (16, 8, 1000)])
def test_cyclic_iterations_match_oracle(cyc, G, K, n):
    """Four source iterations (the lagged cycle psi carried from sweep to
    sweep) give the oracle's phi for every layout."""
    p = cyc
    print(f"case nz={p.nzones} G={G} K={K} angles_per_block={n}", flush=True)
    xs = U.draw_cross_sections(17, p.nzones, G, **XS)
    ctx = _ctx(p, G, xs)
    ctx.set_groups_per_lane(K)
    ctx.set_angles_per_block(n)
    ctx.source_iteration(tol=0.0, max_iters=4)
    err = rel_err(ctx.phi(), _oracle_si(p, xs, 4))
    ctx.close()
    assert err < 1e-12, err


def test_cyclic_fp32_sweeps(cyc):
    p = cyc
    G = 4
    xs = U.draw_cross_sections(18, p.nzones, G, **XS)
    ctx = _ctx(p, G, xs, precision=32)
    ctx.source_iteration(tol=0.0, max_iters=3)
    err = rel_err(ctx.phi(), _oracle_si(p, xs, 3))
    print(f"FP32 cyclic sweeps: max rel err {err:.2e}", flush=True)
    ctx.close()
    assert err < 1e-4


def test_cyclic_source_iteration_graphs_and_fixed_point():
    """Graph replay gives the eager bits; both converge to the oracle's
    fixed point in the same number of iterations."""
    p = cyclic_problem(*CYCLIC_MESHES[2])
    G = 4
    xs = U.draw_cross_sections(19, p.nzones, G, **XS)
    runs = []
    for graphs in (False, True):
        ctx = _ctx(p, G, xs)
        ctx.set_graphs(graphs)
        ok, it, res, _ = ctx.source_iteration(tol=1e-10, max_iters=400)
        runs.append((ok, it, ctx.phi()))
        ctx.close()
    (ok0, it0, phi0), (ok1, it1, phi1) = runs
    print(f"cyclic SI: {it0} iterations", flush=True)
    assert ok0 and ok1 and it0 == it1
    np.testing.assert_array_equal(phi0, phi1)
    assert rel_err(phi0, _oracle_si(p, xs, it0)) < 1e-11


def test_dropin_with_cycle_lists_matches_oracle(cyc):
    """gpu_sweepucbxyz angle by angle with Teton's cyclePsi: phi and the
    updated lagged psi equal the oracle's over two sweeps."""
    p = cyc
    G = 3
    st, ss, qq = U.draw_cross_sections(20, p.nzones, G, **XS)
    m, g, q, s = p.mesh(), p.geometry(), p.quadrature(), p.schedule()
    S = np.ascontiguousarray(np.repeat(qq, 8, axis=0) / (4 * np.pi))
    total = len(s["cycleList"])
    cp_gpu, cp_ref = np.zeros((total, G)), np.zeros((total, G))
    for sweep in range(2):
        Phi = np.zeros((p.ncorner, G))
        for a in range(1, p.nangles + 1):
            Psi1 = np.zeros((p.ncorner + p.nbelem, G))
            PsiB = np.zeros((p.nbelem, G))
            assert U.sweep_ucbxyz(p, a, S, st, Psi1, PsiB, Phi, cycle_psi=cp_gpu) == 0
        ref, _ = O.sweep_all(m, g, q, s, S, st, cycle_psi=cp_ref)
        assert rel_err(Phi, ref) < 1e-12, sweep
        assert np.abs(cp_gpu - cp_ref).max() <= 1e-12 * np.abs(cp_ref).max()


def test_cyclic_zones_need_cycle_lists():
    """A context made from Teton arrays whose nextZ holds cyclic zones
    refuses to sweep until the cycle lists are set, then matches the oracle."""
    p = cyclic_problem(*CYCLIC_MESHES[2])
    s = p.schedule()
    assert s["nextZ"].min() < 0
    G = 2
    lib = _lib.load()
    m, g, q = p.mesh(), p.geometry(), p.quadrature()
    ip = lambda a: np.ascontiguousarray(a, np.int32).ctypes.data_as(C.POINTER(C.c_int))
    dp = lambda a: np.ascontiguousarray(a, np.float64).ctypes.data_as(C.POINTER(C.c_double))
    keep = [m, g, q, s]
    h = C.c_void_p()
    rc = lib.umt_ctx_create(p.nzones, p.ncorner, p.nbelem, G, p.nangles, 64, ip(m["numCorner"]),
                            ip(m["cOffSet"]), ip(m["nCFaces"]), ip(m["cFP"]), ip(m["cEZ"]),
                            dp(g["A_fp"]), dp(g["A_ez"]), dp(g["Volume"]), dp(g["VolumeZone"]),
                            dp(q["omega"]), dp(q["weight"]), ip(s["nHyperPlanes"]),
                            ip(s["zonesInPlane"]), ip(s["nextZ"]), C.byref(h))
    assert rc == 0, lib.umt_last_error()
    try:
        st, ss, qq = U.draw_cross_sections(21, p.nzones, G, **XS)
        assert lib.umt_ctx_set_sigt(h, dp(st)) == 0
        S = np.repeat(qq, 8, axis=0) / (4 * np.pi)
        assert lib.umt_ctx_set_source(h, dp(S)) == 0
        rc = lib.umt_ctx_sweep(h)
        assert rc == 15 and b"cycle" in lib.umt_last_error()
        assert lib.umt_ctx_set_cycles(h, ip(s["numCycles"]), ip(s["cycleList"])) == 0
        assert lib.umt_ctx_sweep(h) == 0
        phi = np.zeros((p.ncorner, G))
        assert lib.umt_ctx_get_phi(h, dp(phi)) == 0
        ref, _ = O.sweep_all(m, g, q, s, S, st)
        assert rel_err(phi, ref) < 1e-12
    finally:
        lib.umt_ctx_destroy(h)
    del keep


def test_time_step_with_cycles_matches_oracle():
    """A time-dependent cycle on a cyclic mesh: the lagged psi starts from
    the previous cycle's Psi (initCyclePsi) and the cycle ends with savePsi."""
    p = cyclic_problem(*CYCLIC_MESHES[1])
    G = 2
    st, ss, qq = U.draw_cross_sections(22, p.nzones, G, **XS)
    tau = 3.0
    rng = np.random.default_rng(4)
    psi_prev = rng.uniform(0.0, 0.2, (p.nangles, p.ncorner, G))
    ctx = _ctx(p, G, (st, ss, qq))
    ctx.set_tau(tau)
    ctx.set_psi_prev(psi_prev)
    ok, it, res, _ = ctx.source_iteration(tol=1e-10, max_iters=200)
    assert ok
    ctx.save_psi()
    phi_ref, psi_ref, it_ref, _, _ = O.time_step(p.mesh(), p.geometry(), p.quadrature(),
                                                 p.schedule(), st, ss, qq, tau, psi_prev, 0.0, it)
    assert it_ref == it
    assert rel_err(ctx.phi(), phi_ref) < 1e-11
    assert rel_err(ctx.psi_prev(), psi_ref) < 1e-10
    ctx.close()


def test_mixed_solve_carries_lagged_psi():
    """FP32 sweeps then FP64 sweeps on a cyclic mesh reach the FP64 fixed point."""
    p = cyclic_problem(*CYCLIC_MESHES[2])
    G = 4
    xs = U.draw_cross_sections(23, p.nzones, G, **XS)
    exact_only = _ctx(p, G, xs)
    ok, it64, _, _ = exact_only.source_iteration(tol=1e-10, max_iters=400)
    assert ok
    fast, exact = _ctx(p, G, xs, precision=32), _ctx(p, G, xs)
    ok, it1, it2, res, _ = U.solve_mixed(fast, exact, tol=1e-10, switch_tol=1e-6, max_iters=400)
    print(f"mixed: {it1} FP32 + {it2} FP64 iterations (FP64 only: {it64})", flush=True)
    assert ok
    assert rel_err(exact.phi(), exact_only.phi()) < 1e-8
    assert it2 < it64
    for c in (exact_only, fast, exact):
        c.close()


def test_batch_refuses_cyclic_domain():
    p = cyclic_problem(*CYCLIC_MESHES[0])
    with pytest.raises(U.UmtError):
        U.SweepBatch([U.TiledProblem(1, 1, 1, 1, 1), p], [2, 2])
