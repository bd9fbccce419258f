"""The zone layout of the all-angle sweep (one lane per group of a zone,
carrying its eight corners) on the GPU.

Every quantity it forms is the corner layout's expression in the corner
layout's order; the compiler's fused multiply-adds still fall differently
in places, so the two layouts agree to rounding (a few ulps: FP64 within
1e-13, FP32 within 1e-5) rather than bit for bit. They are compared on
tiled meshes for each slot width (16, 32, 64 groups per slot, one and
several slots per item) in FP64 and FP32, with angle blocks and graph
replay, on meshes whose schedules have cyclic zones and cycle lists, in
time-dependent cycles, with incident boundary fluxes and inside both GTA
solvers (whose S2 grey sweeps run the same kernel). FP64 is held to the
oracle at 1e-12, FP32 at 2e-5, as the corner layout is."""
import numpy as np
import pytest

import umt_amd as U
from oracle import oracle as O

from conftest import CYCLIC_MESHES, cyclic_problem, rel_err

pytestmark = pytest.mark.gpu

ZONE = U.ZONE_LAYOUT


def _xs(p, G, seed, sigt=(1e-2, 1e3), c=(0.0, 0.95), q=(0.0, 1.0)):
    return U.draw_cross_sections(seed, p.nzones, G, sigt, c, q)


def _same(a, b, what, tol):
    """Max |a - b| over each leading-axis block (an angle of psi, a corner
    of phi), relative to that block of b's largest magnitude (psi changes
    sign in thick zones)."""
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    assert a.shape == b.shape, what
    axes = tuple(range(1, a.ndim))
    scale = np.maximum(np.abs(b).max(axis=axes, keepdims=True), 1e-300)
    err = float(np.max(np.abs(a - b) / scale))
    print(f"{what}: zone vs corner layout {err:.2e}", flush=True)
    assert err < tol, (what, err)


def _tol(precision):
    return 1e-13 if precision == 64 else 1e-5


@pytest.mark.parametrize("tiles,quad,G,K", [((2, 3, 2), (3, 2), 16, 4), ((3, 2, 2), (2, 2), 32, 8),
                                            ((2, 2, 2), (2, 2), 64, 8), ((2, 2, 1), (2, 2), 48, 2),
                                            ((2, 2, 2), (1, 2), 128, 8)])
@pytest.mark.parametrize("precision", [64, 32])
def test_zone_sweep_equals_corner_sweep(tiles, quad, G, K, precision):
    """One sweep of all angles: phi and psi of every angle (the boundary
    tails included) against the corner layout's, phi against the oracle."""
    p = U.TiledProblem(*tiles, *quad)
    st, _, qq = _xs(p, G, 5)
    S = np.repeat(qq, 8, axis=0) / (4 * np.pi)
    print(f"case tiles={tiles} quad={quad} G={G} K={K} precision={precision}", flush=True)
    out = {}
    for layout in (K, ZONE):
        ctx = U.SweepContext(p, G, precision=precision)
        ctx.set_groups_per_lane(layout)
        assert ctx.groups_per_lane() == layout
        ctx.set_sigt(st)
        ctx.set_source(S)
        ctx.sweep()
        out[layout] = (ctx.phi(), ctx.psi())
        ctx.close()
    _same(out[ZONE][0], out[K][0], "phi", _tol(precision))
    _same(out[ZONE][1], out[K][1], "psi", _tol(precision))
    ref, _ = O.sweep_all(p.mesh(), p.geometry(), p.quadrature(), p.schedule(), S, st)
    assert rel_err(out[ZONE][0], ref) < (1e-12 if precision == 64 else 2e-5)


@pytest.mark.parametrize("graphs", [False, True])
def test_zone_source_iteration_with_angle_blocks(graphs):
    """Ten source iterations with angle blocks of 5 (and graph replay): the
    corner layout's phi, and the oracle's within 1e-12."""
    p = U.TiledProblem(3, 3, 2, 2, 2)
    G = 16
    st, ss, qq = _xs(p, G, 6)
    phis = {}
    for layout in (4, ZONE):
        ctx = U.SweepContext(p, G, precision=64)
        ctx.set_groups_per_lane(layout)
        ctx.set_angles_per_block(5)
        ctx.set_graphs(graphs)
        ctx.set_sigt(st)
        ctx.set_scattering(ss, qq)
        ctx.source_iteration(tol=0.0, max_iters=10)
        phis[layout] = ctx.phi()
        ctx.close()
    _same(phis[ZONE], phis[4], "phi", 1e-13)
    ref, it, _, _ = O.source_iteration(p.mesh(), p.geometry(), p.quadrature(), p.schedule(), st, ss,
                                       qq, 0.0, 10)
    assert it == 10
    assert rel_err(phis[ZONE], ref) < 1e-12


@pytest.mark.parametrize("mesh", CYCLIC_MESHES, ids=lambda m: f"n{m[0]}-a{m[1]}-t{m[3]}")
def test_zone_layout_on_cyclic_meshes(mesh):
    """Cyclic zones (solved once from lagged values) and cycle lists over
    four source iterations with angle blocks of 3: the corner layout's phi,
    and the oracle's within 1e-12."""
    p = cyclic_problem(*mesh)
    assert len(p.schedule()["cycleList"]) > 0
    G = 16
    st, ss, qq = U.draw_cross_sections(17, p.nzones, G, (0.5, 50.0), (0.0, 0.9), (0.0, 1.0))
    print(f"case cyclic mesh {mesh}: {p.nzones} zones", flush=True)
    phis = {}
    for layout in (8, ZONE):
        ctx = U.SweepContext(p, G, precision=64)
        ctx.set_groups_per_lane(layout)
        ctx.set_angles_per_block(3)
        ctx.set_sigt(st)
        ctx.set_scattering(ss, qq)
        ctx.source_iteration(tol=0.0, max_iters=4)
        phis[layout] = ctx.phi()
        ctx.close()
    _same(phis[ZONE], phis[8], "phi", 1e-13)
    ref, it, _, _ = O.source_iteration(p.mesh(), p.geometry(), p.quadrature(), p.schedule(), st, ss,
                                       qq, 0.0, 4)
    assert rel_err(phis[ZONE], ref) < 1e-12


def test_zone_layout_fp32_on_a_cyclic_mesh():
    """FP32 sweeps with cyclic zones and cycle lists: the corner layout's phi
    to rounding, the oracle's within the FP32 bound of the cyclic tests."""
    p = cyclic_problem(*CYCLIC_MESHES[2])
    G = 32
    st, ss, qq = U.draw_cross_sections(18, p.nzones, G, (0.5, 50.0), (0.0, 0.9), (0.0, 1.0))
    phis = {}
    for layout in (4, ZONE):
        ctx = U.SweepContext(p, G, precision=32)
        ctx.set_groups_per_lane(layout)
        ctx.set_sigt(st)
        ctx.set_scattering(ss, qq)
        ctx.source_iteration(tol=0.0, max_iters=3)
        phis[layout] = ctx.phi()
        ctx.close()
    _same(phis[ZONE], phis[4], "phi", 1e-5)
    ref, _, _, _ = O.source_iteration(p.mesh(), p.geometry(), p.quadrature(), p.schedule(), st, ss,
                                      qq, 0.0, 3)
    assert rel_err(phis[ZONE], ref) < 1e-4


@pytest.mark.parametrize("precision", [64, 32])
def test_zone_time_cycles_and_incident_flux(precision):
    """Two time-dependent cycles (tau Psi source, savePsi sweep) with
    incident boundary fluxes: phi, the saved Psi and the exiting boundary
    fluxes against the corner layout's."""
    p = U.TiledProblem(2, 2, 2, 2, 2)
    G = 32
    st, ss, qq = _xs(p, G, 7, sigt=(0.1, 10.0))
    rng = np.random.default_rng(8)
    psi0 = rng.uniform(0.1, 1.0, (p.nangles, p.ncorner, G))
    inc = rng.uniform(0.0, 0.5, (p.nangles, p.nbelem, G))
    out = {}
    for layout in (8, ZONE):
        ctx = U.SweepContext(p, G, precision=precision)
        ctx.set_groups_per_lane(layout)
        ctx.set_sigt(st)
        ctx.set_scattering(ss, qq)
        ctx.set_tau(2.5)
        ctx.set_psi_prev(psi0)
        ctx.set_incident_flux(inc)
        for _ in range(2):
            ctx.source_iteration(tol=0.0, max_iters=4)
            ctx.save_psi()
        out[layout] = (ctx.phi(), ctx.psi_prev(), ctx.boundary_flux())
        ctx.close()
    for i, what in enumerate(("phi", "saved Psi", "boundary psi")):
        _same(out[ZONE][i], out[8][i], what, _tol(precision))


@pytest.mark.parametrize("mode,precision", [("gta", 64), ("gta2", 64), ("gta", 32)])
def test_zone_layout_inside_gta(mode, precision):
    """GTA's S2 grey sweeps run the context's layout too: the accelerated
    solve converges as with the corner layout, in the same iterations give
    or take one, to the same answer within the convergence tolerance."""
    p = U.TiledProblem(2, 2, 2, 2, 2)
    G = 16
    st, ss, qq = _xs(p, G, 9, sigt=(0.1, 100.0), c=(0.9, 0.999))
    runs = {}
    for layout in (1, ZONE):
        ctx = U.SweepContext(p, G, precision=precision)
        ctx.set_groups_per_lane(layout)
        ctx.set_sigt(st)
        ctx.set_scattering(ss, qq)
        ctx.set_acceleration(mode)
        ok, it, res, _ = ctx.source_iteration(tol=1e-8 if precision == 64 else 1e-5, max_iters=200)
        runs[layout] = (ok, it, ctx.phi())
        ctx.close()
    assert runs[ZONE][0] and runs[1][0]
    assert abs(runs[ZONE][1] - runs[1][1]) <= 1, (runs[ZONE][1], runs[1][1])
    _same(runs[ZONE][2], runs[1][2], "phi", 1e-6 if precision == 64 else 1e-4)


def test_autotuner_times_the_zone_layout():
    p = U.TiledProblem(2, 2, 2, 2, 2)
    G = 32
    st, _, qq = _xs(p, G, 10)
    ctx = U.SweepContext(p, G, precision=64)
    ctx.set_sigt(st)
    ctx.set_source(np.repeat(qq, 8, axis=0) / (4 * np.pi))
    best, times = ctx.autotune(1)
    assert set(times) == {1, 2, 4, 8, ZONE} and best in times
    assert ctx.groups_per_lane() == best
    ctx.close()
    # Group counts that are not a multiple of 16 have no zone layout.
    ctx = U.SweepContext(p, 8, precision=64)
    with pytest.raises(U.UmtError):
        ctx.set_groups_per_lane(ZONE)
    _, times = ctx.autotune(1)
    assert set(times) == {1, 2, 4, 8}
    ctx.close()


@pytest.mark.parametrize("precision", [64, 32])
def test_grey_autotuner_times_the_zone_layout(precision):
    """The S2 grey sweeps' autotuner times the zone layout beside the corner
    layouts; GTA then converges with the S2 layout it kept, to the answer of
    a context whose S2 sweeps keep the default layout."""
    p = U.TiledProblem(2, 2, 2, 2, 2)
    G = 32
    st, ss, qq = _xs(p, G, 11, sigt=(0.1, 100.0), c=(0.9, 0.999))
    tol = 1e-8 if precision == 64 else 1e-5
    runs = {}
    for tune in (False, True):
        ctx = U.SweepContext(p, G, precision=precision)
        ctx.set_sigt(st)
        ctx.set_scattering(ss, qq)
        if tune:
            best, times = ctx.autotune_grey(1)
            print(f"grey autotune FP{precision}: best {best}, ms {times}", flush=True)
            assert set(times) == {1, 2, 4, 8, ZONE} and best in times
        ctx.set_acceleration("gta")
        ok, it, _, _ = ctx.source_iteration(tol=tol, max_iters=200)
        runs[tune] = (ok, it, ctx.phi())
        ctx.close()
    assert runs[True][0] and runs[False][0]
    assert abs(runs[True][1] - runs[False][1]) <= 1, (runs[True][1], runs[False][1])
    _same(runs[True][2], runs[False][2], "phi", 1e-6 if precision == 64 else 1e-4)
