"""Time-dependent cycles on the GPU (tau Psi source, Sigt + tau, the savePsi
sweep) against the oracle's time step, in FP64, FP32, with angle blocks and
with GTA."""
import numpy as np
import pytest

import umt_amd as U
from oracle import oracle as O

from conftest import rel_err

pytestmark = pytest.mark.gpu

TAU = 3.0


@pytest.fixture(scope="module")
def prob():
    p = U.TiledProblem(2, 2, 1, 2, 2)
    G = 2
    st, ss, qq = U.draw_cross_sections(41, p.nzones, G, (0.1, 10.0), (0.0, 0.95), (0.5, 1.0))
    rng = np.random.default_rng(4)
    phi0 = rng.uniform(0.5, 2.0, (p.ncorner, G))
    psi0 = np.broadcast_to(phi0 / (4 * np.pi), (p.nangles, p.ncorner, G)).copy()
    return p, G, st, ss, qq, psi0


def ctx_for(p, G, st, ss, qq, psi0, precision=64, blocks=None, accel="none"):
    ctx = U.SweepContext(p, G, precision=precision)
    if blocks:
        ctx.set_angles_per_block(blocks)
    ctx.set_sigt(st)
    ctx.set_scattering(ss, qq)
    ctx.set_tau(TAU)
    ctx.set_psi_prev(psi0)
    ctx.set_acceleration(accel)
    return ctx


@pytest.mark.parametrize("precision,blocks", [(64, None), (64, 5), (32, None)])
def test_two_cycles_match_oracle(prob, precision, blocks):
    p, G, st, ss, qq, psi0 = prob
    print(f"case precision={precision} blocks={blocks}", flush=True)
    m, g, q, s = p.mesh(), p.geometry(), p.quadrature(), p.schedule()
    ctx = ctx_for(p, G, st, ss, qq, psi0, precision, blocks)
    psi_ref, phi_ref = psi0, None
    tol = 1e-12 if precision == 64 else 2e-5
    for cycle in range(2):
        ctx.source_iteration(tol=0.0, max_iters=6)
        ctx.save_psi()
        phi_ref, psi_ref, _, _, _ = O.time_step(m, g, q, s, st, ss, qq, TAU, psi_ref, 0.0, 6,
                                                phi0=phi_ref)
        assert rel_err(ctx.phi(), phi_ref) < tol, cycle
        got = ctx.psi_prev()
        err = np.max(np.abs(got - psi_ref)) / np.max(np.abs(psi_ref))
        assert err < tol, (cycle, err)
    ctx.close()


@pytest.mark.parametrize("precision", [64, 32])
def test_two_cycles_with_eight_groups_per_lane(precision):
    """The tau Psi source read K = 8 groups at a time (8 groups, one lane
    block per corner), against the oracle's time step."""
    p = U.TiledProblem(2, 1, 1, 2, 2)
    G = 8
    st, ss, qq = U.draw_cross_sections(43, p.nzones, G, (0.1, 10.0), (0.0, 0.95), (0.5, 1.0))
    rng = np.random.default_rng(5)
    psi0 = rng.uniform(0.0, 0.2, (p.nangles, p.ncorner, G))
    print(f"case K=8 precision={precision}", flush=True)
    m, g, q, s = p.mesh(), p.geometry(), p.quadrature(), p.schedule()
    ctx = ctx_for(p, G, st, ss, qq, psi0, precision)
    ctx.set_groups_per_lane(8)
    psi_ref, phi_ref = psi0, None
    tol = 1e-12 if precision == 64 else 2e-5
    for cycle in range(2):
        ctx.source_iteration(tol=0.0, max_iters=4)
        ctx.save_psi()
        phi_ref, psi_ref, _, _, _ = O.time_step(m, g, q, s, st, ss, qq, TAU, psi_ref, 0.0, 4,
                                                phi0=phi_ref)
        assert rel_err(ctx.phi(), phi_ref) < tol, cycle
        err = np.max(np.abs(ctx.psi_prev() - psi_ref)) / np.max(np.abs(psi_ref))
        assert err < tol, (cycle, err)
    ctx.close()


def test_gta_time_step_reaches_the_plain_answer(prob):
    p, G, st, ss, qq, psi0 = prob
    plain = ctx_for(p, G, st, ss, qq, psi0)
    ok0, it0, _, _ = plain.source_iteration(tol=1e-11, max_iters=500)
    gta = ctx_for(p, G, st, ss, qq, psi0, accel="gta")
    ok1, it1, _, _ = gta.source_iteration(tol=1e-11, max_iters=500)
    print(f"time step: plain {it0} iterations, GTA {it1}", flush=True)
    assert ok0 and ok1 and it1 <= it0
    assert rel_err(gta.phi(), plain.phi()) < 1e-9
    plain.close()
    gta.close()
