"""Batches of uneven domains on the GPU: each domain's answer equals a
single-domain context's (and the oracle's), for every groups-per-lane and
angle-block layout, with and without graph replay."""
import numpy as np
import pytest

import umt_amd as U
from oracle import oracle as O

from conftest import rel_err

pytestmark = pytest.mark.gpu

# (tiles, quadrature, groups, seed): uneven meshes, angle counts and group
# counts, including group counts that need padding.
DOMAINS = [((2, 2, 2), (2, 2), 3, 101), ((3, 2, 2), (3, 1), 5, 102), ((2, 3, 1), (1, 1), 8, 103),
           ((1, 2, 2), (2, 1), 2, 104)]
XS = dict(sigt=(1e-2, 1e3), c=(0.0, 0.95), q=(0.0, 1.0))


@pytest.fixture(scope="module")
def domains():
    out = []
    for tiles, quad, G, seed in DOMAINS:
        p = U.TiledProblem(*tiles, *quad)
        out.append((p, G, U.draw_cross_sections(seed, p.nzones, G, **XS)))
    return out


def make_batch(domains, precision=64):
    b = U.SweepBatch([d[0] for d in domains], [d[1] for d in domains], precision=precision)
    for i, (p, G, (st, ss, qq)) in enumerate(domains):
        b.set_cross_sections(i, st, ss, qq)
    return b


def single(p, G, xs, iters, precision=64):
    st, ss, qq = xs
    ctx = U.SweepContext(p, G, precision=precision)
    ctx.set_sigt(st)
    ctx.set_scattering(ss, qq)
    ctx.source_iteration(tol=0.0, max_iters=iters)
    phi = ctx.phi()
    ctx.close()
    return phi


@pytest.mark.parametrize("K,n", [(1, 1000), (2, 5), (4, 1000), (4, 3)])
def test_batched_steps_match_single_domain_contexts(domains, K, n):
    print(f"case K={K} angles_per_block={n}", flush=True)
    b = make_batch(domains)
    b.set_layout(K, n)
    d = b.dims()
    assert d["groups_per_lane"] == K and d["domains"] == len(domains)
    assert d["unknowns"] == sum(p.ncorner * p.nangles * G for p, G, _ in domains)
    b.sweep()
    b.sweep()
    for i, (p, G, xs) in enumerate(domains):
        ref = single(p, G, xs, 2)
        err = rel_err(b.phi(i), ref)
        assert err < 1e-13, (i, err)
    b.close()


def test_batched_sweep_matches_oracle(domains):
    b = make_batch(domains)
    assert b.dims()["angles_per_block"] == 32
    b.sweep()  # from phi = 0: S = q / 4 pi
    for i, (p, G, (st, ss, qq)) in enumerate(domains):
        S = np.repeat(qq, 8, axis=0) / (4 * np.pi)
        ref, _ = O.sweep_all(p.mesh(), p.geometry(), p.quadrature(), p.schedule(), S, st)
        assert rel_err(b.phi(i), ref) < 1e-12, i
    b.close()


def test_batched_source_iteration_and_graph_replay(domains):
    """Graph-replayed iterations give the eager bits; every domain converges
    to its single-domain fixed point."""
    eager = make_batch(domains)
    eager.set_graphs(False)
    ok0, its0, res0, tot0 = eager.source_iteration(tol=1e-8, max_iters=500)
    graph = make_batch(domains)
    graph.set_graphs(True)
    ok1, its1, res1, tot1 = graph.source_iteration(tol=1e-8, max_iters=500)
    print(f"batched SI: iterations per domain {list(its1)}, total {tot1}", flush=True)
    assert ok0 and ok1 and tot0 == tot1
    np.testing.assert_array_equal(its0, its1)
    for i, (p, G, (st, ss, qq)) in enumerate(domains):
        np.testing.assert_array_equal(graph.phi(i), eager.phi(i))
        ctx = U.SweepContext(p, G)
        ctx.set_sigt(st)
        ctx.set_scattering(ss, qq)
        ok, it, _, _ = ctx.source_iteration(tol=1e-8, max_iters=500)
        assert ok and it == its1[i]
        # the batch kept iterating every domain until the last converged
        ctx.set_phi(np.zeros((p.ncorner, G)))
        ctx.source_iteration(tol=0.0, max_iters=tot1)
        assert rel_err(graph.phi(i), ctx.phi()) < 1e-12
        ctx.close()
    eager.close()
    graph.close()


def test_batch_fp32_and_autotuner(domains):
    b = make_batch(domains, precision=32)
    K, n, ms = b.autotune(reps=1)
    print(f"autotuned K={K} angles_per_block={n} ({ms:.3f} ms per iteration)", flush=True)
    assert K in (1, 2, 4) and n >= 1 and ms > 0
    d = b.dims()
    assert (d["groups_per_lane"], d["angles_per_block"]) == (K, n)
    b.sweep()
    for i, (p, G, (st, ss, qq)) in enumerate(domains):
        S = np.repeat(qq, 8, axis=0) / (4 * np.pi)
        ref, _ = O.sweep_all(p.mesh(), p.geometry(), p.quadrature(), p.schedule(), S, st)
        assert rel_err(b.phi(i), ref) < 2e-5, i
    b.close()


# ---------------------------------------------------------------------------
# Batches drawn like the benchmark's 64 uneven domains: n^3 tiles, 2x2, 3x3
# or 4x4 product quadrature per octant (32, 72 or 128 angles), G ~ U{8..64}
# (most group counts need padding to a multiple of 4), cross sections
# sigma_t logU[1e-2, 1e3], c U[0, 0.95], q U[0, 1]. Groups are independent
# in a sweep, so the oracle runs each domain on three of its groups (first,
# middle, last: the last sits next to the padding) with every angle.


def _draw(n_domains, seed, n_range):
    rng = np.random.default_rng(seed)
    spec = [(int(rng.integers(n_range[0], n_range[1] + 1)), int(rng.choice([2, 3, 4])),
             int(rng.integers(8, 65))) for _ in range(n_domains)]
    out = []
    for i, (n, a, G) in enumerate(spec):
        p = U.TiledProblem(n, n, n, a, a)
        out.append((p, G, U.draw_cross_sections(seed + 1 + i, p.nzones, G, **XS)))
    return out


def _check_against_oracle(b, doms, which, tol):
    for i in which:
        p, G, (st, ss, qq) = doms[i]
        sel = sorted({0, G // 2, G - 1})
        S = np.repeat(qq, 8, axis=0) / (4 * np.pi)
        ref, _ = O.sweep_all(p.mesh(), p.geometry(), p.quadrature(), p.schedule(),
                             np.ascontiguousarray(S[:, sel]), np.ascontiguousarray(st[:, sel]),
                             nthreads=THREADS, fast=True)
        err = rel_err(b.phi(i)[:, sel], ref)
        assert err < tol, (i, p.nzones, p.nangles, G, err)


THREADS = 16


@pytest.mark.parametrize("precision,layout", [(64, (4, 32)), (64, (2, 128)), (32, "auto")])
def test_config3_distribution_matches_oracle(precision, layout):
    """16 domains with n in 4..8 tiles, every domain against the oracle."""
    doms = _draw(16, 20261215, (4, 8))
    print(f"case config-3 distribution FP{precision} layout {layout}: "
          f"{[(d[0].nzones, d[0].nangles, d[1]) for d in doms]}", flush=True)
    b = make_batch(doms, precision=precision)
    if layout == "auto":
        K, n, ms = b.autotune(reps=1)
        print(f"autotuned K={K} angles per block {n}", flush=True)
    else:
        b.set_layout(*layout)
    b.sweep()
    _check_against_oracle(b, doms, range(len(doms)), 1e-12 if precision == 64 else 2e-5)
    b.close()


def test_config3_full_batch_fp32():
    """The benchmark's 64-domain batch itself (n in 4..16 tiles, seed
    20261205: 1.98e6 zones, 4.2e10 unknowns per sweep), FP32 with the benchmark's first layout
    (4 groups per lane, 32 angles per block): one sweep from phi = 0, eight
    domains (the largest, the smallest, and six more spread over the list)
    against the oracle."""
    rng = np.random.default_rng(20261205)
    spec = [(int(rng.integers(4, 17)), int(rng.choice([2, 3, 4])), int(rng.integers(8, 65)))
            for _ in range(64)]
    doms = []
    for i, (n, a, G) in enumerate(spec):
        p = U.TiledProblem(n, n, n, a, a)
        doms.append((p, G, U.draw_cross_sections(20261205 + 1 + i, p.nzones, G, **XS)))
    print(f"case config-3 full batch: {sum(d[0].nzones for d in doms)} zones", flush=True)
    b = make_batch(doms, precision=32)
    b.set_layout(4, 32)
    b.sweep()
    sizes = [d[0].nzones * d[0].nangles for d in doms]
    which = sorted({int(np.argmax(sizes)), int(np.argmin(sizes)), 5, 17, 29, 41, 53, 63})
    _check_against_oracle(b, doms, which, 2e-5)
    b.close()


# ---------------------------------------------------------------------------
# The persistent sweep: every level of an angle block in one launch, with a
# grid barrier between levels. The per-task work is the per-level kernel's,
# so both must give the same bits.


@pytest.mark.parametrize("precision,layout", [(64, (4, 1000)), (64, (2, 5)), (32, (1, 7))])
def test_persistent_sweep_gives_the_per_level_bits(domains, precision, layout):
    print(f"case persistent FP{precision} layout {layout}", flush=True)
    out = {}
    for persistent in (False, True):
        b = make_batch(domains, precision=precision)
        b.set_layout(*layout)
        b.set_persistent(persistent)
        on, blocks = b.persistent()
        assert on == persistent and blocks >= 1
        d = b.dims()
        assert d["launches"] == (d["blocks"] if persistent else d["launches"])
        b.sweep()
        b.sweep()
        out[persistent] = [b.phi(i) for i in range(len(domains))]
        b.close()
    for i in range(len(domains)):
        np.testing.assert_array_equal(out[True][i], out[False][i])


def test_persistent_source_iteration_with_graph_replay_on_config3_domains():
    """Sixteen domains drawn like the benchmark's batch: graph-replayed
    persistent source iteration takes the per-level sweep's iterations to its
    bits, and each domain's phi matches the oracle's sweep afterwards."""
    doms = _draw(16, 20261217, (4, 6))
    runs = []
    for persistent in (False, True):
        b = make_batch(doms)
        b.set_layout(4, 1000)
        b.set_persistent(persistent)
        b.set_graphs(True)
        ok, its, res, tot = b.source_iteration(tol=1e-8, max_iters=300)
        runs.append((ok, list(its), tot, [b.phi(i) for i in range(len(doms))], b.persistent()))
        b.close()
    (ok0, its0, tot0, phi0, _), (ok1, its1, tot1, phi1, (on, blocks)) = runs
    print(f"persistent SI: {tot1} iterations on {blocks} workgroups", flush=True)
    assert on and blocks > 1
    assert ok0 and ok1 and its0 == its1 and tot0 == tot1
    for a, c in zip(phi0, phi1):
        np.testing.assert_array_equal(a, c)


def test_persistent_barrier_that_gives_up_is_an_error():
    """A grid barrier that waits for nobody (max_polls < 0) makes every
    workgroup but the last to arrive give up: the sweep fails with an error
    instead of returning wrong results, and after it the barrier is clear
    and a persistent sweep with the default wait matches per-level launches."""
    doms = _draw(8, 20261219, (5, 6))
    b = make_batch(doms)
    b.set_layout(4, 1000)
    b.set_persistent(True, max_polls=-1)
    assert b.persistent()[1] > 1
    with pytest.raises(U.UmtError):
        b.sweep()
    b.reset()
    b.set_persistent(True)
    b.sweep()
    got = [b.phi(i) for i in range(len(doms))]
    b.close()
    ref = make_batch(doms)
    ref.set_layout(4, 1000)
    ref.sweep()
    for i in range(len(doms)):
        np.testing.assert_array_equal(got[i], ref.phi(i))
    ref.close()


def test_batch_autotuner_tries_the_persistent_sweep(domains):
    b = make_batch(domains)
    K, n, ms = b.autotune(reps=1)
    on, blocks = b.persistent()
    print(f"autotuned K={K} n={n} persistent={on} ({blocks} workgroups), {ms:.3f} ms", flush=True)
    assert blocks >= 1 and ms > 0
    b.sweep()
    for i, (p, G, (st, ss, qq)) in enumerate(domains):
        S = np.repeat(qq, 8, axis=0) / (4 * np.pi)
        ref, _ = O.sweep_all(p.mesh(), p.geometry(), p.quadrature(), p.schedule(), S, st)
        assert rel_err(b.phi(i), ref) < 1e-12, i
    b.close()
