"""Failure handling of source iteration on the GPU: a non-finite scalar flux
must end the iteration with an error (status 34), never with "converged",
and more scattering than collision (sigma_s > sigma_t, a scattering ratio
above 1, where the iterates grow without bound) is refused (status 52).

The non-finite flux is forced from the host (set_phi with a NaN or an Inf
entry); every kernel still runs to its normal end. The convergence kernels
reduce the zone averages' maximum and the residual through atomicMax on
their bit patterns; these tests pin that a NaN or Inf anywhere reaches the
host's finiteness check instead of being dropped by the max."""
import numpy as np
import pytest

import umt_amd as U
from umt_amd._lib import UmtError

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def prob():
    return U.TiledProblem(2, 2, 2, 2, 2)


def _ctx(p, G, precision, c=(0.0, 0.9)):
    st, ss, qq = U.draw_cross_sections(7, p.nzones, G, (0.1, 10.0), c, (1.0, 1.0))
    ctx = U.SweepContext(p, G, precision=precision)
    ctx.set_sigt(st)
    ctx.set_scattering(ss, qq)
    return ctx, st, ss, qq


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("bad", [np.nan, np.inf, -np.inf])
@pytest.mark.parametrize("accel", ["none", "gta"])
def test_non_finite_phi_fails_source_iteration(prob, precision, bad, accel):
    print(f"case non-finite phi {bad} FP{precision} accel {accel}", flush=True)
    G = 4
    ctx, *_ = _ctx(prob, G, precision)
    if accel == "gta":
        ctx.autotune_grey(1)
    ctx.set_acceleration(accel)
    phi = np.ones((prob.ncorner, G))
    phi[prob.ncorner // 3, 2] = bad
    ctx.set_phi(phi)
    with pytest.raises(UmtError) as e:
        ctx.source_iteration(tol=1e-8, max_iters=20)
    ctx.close()
    # GTA meets the non-finite grey source first and reports its own status.
    assert e.value.code == (34 if accel == "none" else 39), str(e.value)


def test_finite_phi_still_converges(prob):
    """The same start without the bad entry converges (the check is not
    tripped by ordinary iterates)."""
    G = 4
    ctx, *_ = _ctx(prob, G, 64)
    ctx.set_phi(np.ones((prob.ncorner, G)))
    ok, it, res, _ = ctx.source_iteration(tol=1e-8, max_iters=500)
    ctx.close()
    assert ok and res < 1e-8, (it, res)


@pytest.mark.parametrize("precision", [32, 64])
def test_scattering_above_collision_is_refused(prob, precision):
    print(f"case sigma_s > sigma_t FP{precision}", flush=True)
    G = 2
    ctx, st, ss, qq = _ctx(prob, G, precision)
    ss = ss.copy()
    ss[5, 1] = 1.5 * st[5, 1]
    ctx.set_scattering(ss, qq)
    ctx.set_phi(np.zeros((prob.ncorner, G)))
    with pytest.raises(UmtError) as e:
        ctx.source_iteration(tol=1e-8, max_iters=50)
    ctx.close()
    assert e.value.code == 52, str(e.value)


def test_batch_refuses_scattering_above_collision(prob):
    G = 4
    st, ss, qq = U.draw_cross_sections(9, prob.nzones, G, (0.1, 10.0), (0.0, 0.9), (1.0, 1.0))
    ss = ss.copy()
    ss[0, 0] = 2.0 * st[0, 0]
    b = U.SweepBatch([prob], [G], precision=32)
    with pytest.raises(UmtError) as e:
        b.set_cross_sections(0, st, ss, qq)
    b.close()
    assert e.value.code == 52


def test_rz_non_finite_phi_and_scattering_above_collision():
    p = U.RzProblem(6, 5, 2, 2, r0=0.0, r1=1.0, z0=0.0, z1=1.0)
    G = 2
    rng = np.random.default_rng(3)
    st = np.exp(rng.uniform(np.log(0.1), np.log(10.0), (p.nzones, G)))
    ss = st * rng.uniform(0.0, 0.9, (p.nzones, G))
    qq = np.ones((p.nzones, G))
    ctx = U.RzContext(p, G)
    ctx.set_sigt(st)
    ctx.set_scattering(ss, qq)
    phi = np.ones((p.ncorner, G))
    phi[3, 1] = np.nan
    ctx.set_phi(phi)
    with pytest.raises(UmtError) as e:
        ctx.source_iteration(tol=1e-8, max_iters=20)
    assert e.value.code == 34, str(e.value)
    ss2 = ss.copy()
    ss2[2, 0] = 1.1 * st[2, 0]
    ctx.set_scattering(ss2, qq)
    ctx.set_phi(np.zeros((p.ncorner, G)))
    with pytest.raises(UmtError) as e:
        ctx.source_iteration(tol=1e-8, max_iters=20)
    ctx.close()
    assert e.value.code == 52, str(e.value)
