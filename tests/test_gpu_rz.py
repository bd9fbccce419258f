"""The RZ sweep on the GPU against the oracle: every direction of every xi
level in one pass of level launches (the half-angle flux chained through
the levels) gives the oracle's psi, finishing-direction psi, exiting
boundary psi and phi (FP64 within 1e-12), for every groups-per-lane choice;
FP32 sweeps within 2e-5; source iteration step by step and converged."""
import numpy as np
import pytest

import umt_amd as U
from oracle import rz as R

from conftest import rel_err

pytestmark = pytest.mark.gpu

# (nr, nz, r0, r1, z0, z1, jitter, seed, npolar, nazimuthal)
CASES = [(5, 4, 0.0, 1.0, 0.0, 2.0, 0.25, 3, 2, 3), (7, 6, 0.5, 2.0, -1.0, 1.0, 0.3, 5, 3, 2),
         (4, 9, 0.0, 0.5, 0.0, 3.0, 0.0, 0, 1, 4)]


def _oracle(p):
    pts, quads = p.points()
    m = R.quad_mesh(pts, quads)
    g = R.geometry(m)
    q = R.quadrature(*p._quad_orders)
    return m, g, q, R.schedule(m, g, q)


@pytest.fixture(scope="module", params=CASES, ids=lambda c: f"{c[0]}x{c[1]}-r0{c[2]}-q{c[8]}{c[9]}")
def prob(request):
    nr, nz, r0, r1, z0, z1, jit, seed, npol, naz = request.param
    p = U.RzProblem(nr, nz, npol, naz, r0=r0, r1=r1, z0=z0, z1=z1, jitter=jit, seed=seed)
    p._quad_orders = (npol, naz)
    return p, _oracle(p)


def _xs(p, G, seed):
    rng = np.random.default_rng(seed)
    sig = np.exp(rng.uniform(np.log(0.05), np.log(200.0), (p.nzones, G)))
    S = rng.uniform(0.0, 1.0, (p.ncorner, G))
    return sig, S


@pytest.mark.parametrize("G,K", [(1, 1), (3, 1), (4, 2), (4, 4), (8, 4)])
def test_rz_sweep_matches_oracle(prob, G, K):
    p, (m, g, q, s) = prob
    print(f"case RZ {p.nzones} zones, {p.nangles} directions, G={G} K={K}", flush=True)
    sig, S = _xs(p, G, 10 + G)
    ctx = U.RzContext(p, G)
    ctx.set_groups_per_lane(K)
    ctx.set_sigt(sig)
    ctx.set_source(S)
    ctx.sweep()
    phi, psi = ctx.phi(), ctx.psi()
    ctx.close()
    ref_phi, ref_B, ref_psi = R.sweep_all(m, g, q, s, S, sig, want_psi=True)
    assert rel_err(phi, ref_phi) < 1e-12
    Cn = p.ncorner
    for a in range(p.nangles):
        scale = np.abs(ref_psi[a]).max()
        assert np.abs(psi[a, :Cn] - ref_psi[a]).max() <= 1e-12 * scale, a
        # exiting boundary elements of this direction
        out = np.abs(ref_B[a]).max(axis=1) > 0
        if out.any():
            assert np.abs(psi[a, Cn:][out] - ref_B[a][out]).max() <= 1e-12 * scale, a


def test_rz_fp32_sweep(prob):
    p, (m, g, q, s) = prob
    G = 4
    sig, S = _xs(p, G, 21)
    ctx = U.RzContext(p, G, precision=32)
    ctx.set_sigt(sig)
    ctx.set_source(S)
    ctx.sweep()
    ref, _, _ = R.sweep_all(m, g, q, s, S, sig)
    err = rel_err(ctx.phi(), ref)
    print(f"FP32 RZ sweep: max rel phi error {err:.2e}", flush=True)
    ctx.close()
    assert err < 2e-5


def test_rz_source_iteration_matches_oracle(prob):
    p, (m, g, q, s) = prob
    G = 2
    rng = np.random.default_rng(7)
    sig = np.exp(rng.uniform(np.log(0.1), np.log(50.0), (p.nzones, G)))
    sigs = sig * rng.uniform(0.0, 0.9, (p.nzones, G))
    qq = rng.uniform(0.0, 1.0, (p.nzones, G))
    ctx = U.RzContext(p, G)
    ctx.set_sigt(sig)
    ctx.set_scattering(sigs, qq)
    ok, it, res, hist = ctx.source_iteration(tol=0.0, max_iters=6)
    ref, it_ref, res_ref, hist_ref = R.source_iteration(m, g, q, s, sig, sigs, qq, 0.0, 6)
    assert it == it_ref == 6
    assert rel_err(ctx.phi(), ref) < 1e-12
    np.testing.assert_allclose(hist, hist_ref, rtol=1e-9)
    # and to convergence
    ctx.set_phi(np.zeros((p.ncorner, G)))
    ok, it, res, _ = ctx.source_iteration(tol=1e-10, max_iters=500)
    ref, it_ref, _, _ = R.source_iteration(m, g, q, s, sig, sigs, qq, 1e-10, 500)
    print(f"RZ SI: {it} iterations (oracle {it_ref})", flush=True)
    assert ok and abs(it - it_ref) <= 1
    assert rel_err(ctx.phi(), ref) < 1e-9
    ctx.close()


def test_rz_level_pipeline_is_shorter_than_serial_directions(prob):
    """The chained directions of a xi level run one level apart: a sweep takes
    far fewer launches than sweeping the directions one after another. Where
    mu changes sign inside a xi level the wavefront reverses along r and the
    next direction waits for the far side, so the depth is at most two plane
    depths plus one level per direction."""
    p, (m, g, q, s) = prob
    ctx = U.RzContext(p, 1)
    d = ctx.dims()
    ctx.close()
    serial = int(s["nHyperPlanes"].sum())
    assert d["items"] == p.nzones * int(np.sum(q["finishing"] == 0))
    assert d["levels"] < serial / 2
    per_level = 2 * p._quad_orders[1] + 1  # swept directions of one xi level
    assert d["levels"] <= 2 * int(s["nHyperPlanes"].max()) + per_level


def test_rz_graph_replay_equals_eager(prob):
    """Replayed sweeps (with the program's own eager work in between) give
    the eager sweep's bits, several times in a row."""
    p, (m, g, q, s) = prob
    G = 4
    rng = np.random.default_rng(31)
    sig = np.exp(rng.uniform(np.log(0.1), np.log(50.0), (p.nzones, G)))
    sigs = 0.5 * sig
    qq = rng.uniform(0.0, 1.0, (p.nzones, G))
    runs = []
    for graphs in (False, True):
        ctx = U.RzContext(p, G)
        ctx.set_graphs(graphs)
        ctx.set_sigt(sig)
        ctx.set_scattering(sigs, qq)
        phis = []
        for _ in range(3):
            ctx.source_iteration(tol=0.0, max_iters=2)
            phis.append(ctx.phi())
        runs.append(phis)
        ctx.close()
    for a, b in zip(*runs):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("precision", [64, 32])
def test_rz_two_cycles_match_oracle(prob, precision):
    """Time-dependent cycles (sigma_t + tau, the tau psi_prev source of each
    direction, the savePsi sweep) against the oracle's time step, with a
    psi_prev that differs between directions."""
    p, (m, g, q, s) = prob
    print(f"case RZ time step {p.nzones} zones, precision={precision}", flush=True)
    G, tau = 2, 3.0
    rng = np.random.default_rng(17)
    sig = np.exp(rng.uniform(np.log(0.1), np.log(20.0), (p.nzones, G)))
    sigs = sig * rng.uniform(0.0, 0.9, (p.nzones, G))
    qq = rng.uniform(0.0, 1.0, (p.nzones, G))
    psi0 = rng.uniform(0.0, 1.0, (p.nangles, p.ncorner, G))
    ctx = U.RzContext(p, G, precision=precision)
    with pytest.raises(RuntimeError):
        ctx.set_psi_prev(psi0)  # needs tau > 0 first
    ctx.set_sigt(sig)
    ctx.set_scattering(sigs, qq)
    ctx.set_tau(tau)
    ctx.set_psi_prev(psi0)
    tol = 1e-12 if precision == 64 else 2e-5
    psi_ref, phi_ref = psi0, None
    for cycle in range(2):
        ctx.source_iteration(tol=0.0, max_iters=5)
        ctx.save_psi()
        phi_ref, psi_ref, _, _, _ = R.time_step(m, g, q, s, sig, sigs, qq, tau, psi_ref, 0.0, 5,
                                                phi0=phi_ref)
        e_phi = rel_err(ctx.phi(), phi_ref)
        got = ctx.psi_prev()
        e_psi = np.max(np.abs(got - psi_ref)) / np.max(np.abs(psi_ref))
        print(f"cycle {cycle}: phi {e_phi:.2e}, psi {e_psi:.2e}", flush=True)
        assert e_phi < tol and e_psi < tol, cycle
    ctx.close()


def test_rz_autotune_keeps_a_supported_layout(prob):
    """The autotuner times every groups-per-lane value that divides G, keeps
    the fastest, and the sweep afterwards still gives the oracle's phi."""
    p, (m, g, q, s) = prob
    G = 4
    sig, S = _xs(p, G, 41)
    ctx = U.RzContext(p, G)
    ctx.set_sigt(sig)
    ctx.set_source(S)
    K, ms = ctx.autotune(reps=2)
    print(f"RZ autotune: K={K}, ms per sweep {ms}", flush=True)
    assert set(ms) == {1, 2, 4} and all(t > 0 for t in ms.values())
    assert ms[K] == min(ms.values())
    ctx.sweep()
    ref, _, _ = R.sweep_all(m, g, q, s, S, sig)
    assert rel_err(ctx.phi(), ref) < 1e-12
    ctx.close()


@pytest.mark.parametrize("nr,nz,npol,naz", [(16, 12, 2, 3), (9, 23, 3, 2), (64, 48, 4, 4)])
def test_rz_tile_levels_give_the_zone_level_bits(nr, nz, npol, naz):
    """On regular meshes the sweep runs in tile levels (a workgroup walking a
    tile's zone levels between barriers): far fewer launches, the same bits
    as one launch per zone level for every groups-per-lane choice, and the
    oracle's phi."""
    p = U.RzProblem(nr, nz, npol, naz, r0=0.0, r1=1.0, z0=0.0, z1=1.5)
    G = 8
    sig, S = _xs(p, G, 51)
    print(f"case RZ tiles {nr}x{nz} q{npol}{naz}", flush=True)
    out = {}
    for tiles in (True, False):
        ctx = U.RzContext(p, G)
        ctx.set_tiles(tiles)
        d = ctx.dims()
        assert d["tiled"] == int(tiles)
        ctx.set_sigt(sig)
        ctx.set_source(S)
        for K in (1, 2, 4):
            ctx.set_groups_per_lane(K)
            ctx.sweep()
            out[(tiles, K)] = (ctx.phi(), ctx.psi(), d["levels"])
        ctx.close()
    for K in (1, 2, 4):
        np.testing.assert_array_equal(out[(True, K)][0], out[(False, K)][0])
        np.testing.assert_array_equal(out[(True, K)][1], out[(False, K)][1])
    assert out[(True, 1)][2] < out[(False, 1)][2] / 2
    pts, quads = p.points()
    m = R.quad_mesh(pts, quads)
    g = R.geometry(m)
    q = R.quadrature(npol, naz)
    ref, _, _ = R.sweep_all(m, g, q, R.schedule(m, g, q), S, sig)
    assert rel_err(out[(True, 4)][0], ref) < 1e-12
