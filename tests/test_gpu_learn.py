"""Learned correction on the GPU (BF16 MFMA network kernels, features,
corrected source iteration) against the oracle's restatement, and what the
trained network gains over plain source iteration."""
import sys

import numpy as np
import pytest

import umt_amd as U
from oracle import learn as OL
from oracle import oracle as O

pytestmark = pytest.mark.gpu

CONFIG4 = dict(sigt=(0.1, 10.0), c=(0.95, 0.999), q=(1.0, 1.0))
SEED4 = 20261206


def say(msg):
    print(msg, file=sys.stderr, flush=True)


def random_params(seed, scale=0.3):
    rng = np.random.default_rng(seed)
    p = rng.uniform(-scale, scale, OL.NPARAMS).astype(np.float32)
    p[OL.W1:OL.B1].reshape(OL.NF, OL.NH)[OL.NUSED:] = 0.0
    return p


def random_samples(seed, n):
    rng = np.random.default_rng(seed)
    X = np.zeros((n, OL.NF), np.float32)
    X[:, :OL.NUSED] = rng.normal(0, 1, (n, OL.NUSED))
    a = rng.normal(0, 1, n).astype(np.float32)
    b = rng.normal(0, 1, n).astype(np.float32)
    return X, a, b


@pytest.fixture(scope="module")
def net():
    n = U.Learner(seed=5)
    yield n
    n.close()


def test_init_params_are_seeded_and_bounded(net):
    p = net.params()
    assert p.shape == (OL.NPARAMS,)
    w1 = p[OL.W1:OL.B1].reshape(OL.NF, OL.NH)
    assert not w1[OL.NUSED:].any()
    assert np.abs(w1).max() <= 1 / np.sqrt(OL.NUSED) + 1e-7
    assert np.abs(p[OL.W2:]).max() <= 1 / 8 + 1e-7
    other = U.Learner(seed=5)
    np.testing.assert_array_equal(other.params(), p)
    other.close()


@pytest.mark.parametrize("n", [1, 16, 1000, 4099])
def test_forward_matches_oracle(net, n):
    say(f"forward n={n}")
    p = random_params(1)
    net.set_params(p)
    X, _, _ = random_samples(2, n)
    r = net.forward(X)
    ref, _ = OL.forward(p.astype(np.float64), X, round=True)
    err = np.abs(r - ref)
    scale = max(1.0, np.abs(ref).max())
    # FP32 accumulation order and tanhf may flip a BF16 rounding of a hidden
    # activation (2^-8 relative); anything structural is orders larger.
    assert err.max() < 5e-3 * scale, err.max()
    assert np.median(err) < 1e-5 * scale
    exact, _ = OL.forward(p.astype(np.float64), X, round=False)
    assert np.abs(r - exact).max() > err.max() or err.max() == 0.0


@pytest.mark.parametrize("n", [32, 777, 8192])
def test_gradient_matches_oracle(net, n):
    say(f"gradient n={n}")
    p = random_params(3)
    net.set_params(p)
    X, a, b = random_samples(4, n)
    g, sse = net.gradient(X, a, b)
    ref, sse_ref = OL.gradient(p.astype(np.float64), X, a, b, round=True)
    assert abs(sse - sse_ref) < 1e-3 * sse_ref
    for lo, hi, name in ((OL.W1, OL.B1, "W1"), (OL.B1, OL.W2, "b1"), (OL.W2, OL.B2, "W2"),
                         (OL.B2, OL.W3, "b2"), (OL.W3, OL.B3, "w3"), (OL.B3, OL.NPARAMS, "b3")):
        d = np.linalg.norm(g[lo:hi] - ref[lo:hi])
        assert d <= 1e-2 * np.linalg.norm(ref[lo:hi]) + 1e-7, (name, d, np.linalg.norm(ref[lo:hi]))
    w1 = g[OL.W1:OL.B1].reshape(OL.NF, OL.NH)
    assert not w1[OL.NUSED:].any()  # empty feature slots get no gradient


def test_adam_step_matches_oracle(net):
    p = random_params(5)
    net.set_params(p)
    X, a, b = random_samples(6, 4096)
    net.step(X, a, b, lr=1e-3)
    got = net.params()
    g, _ = OL.gradient(p.astype(np.float64), X, a, b, round=True)
    ref, m, v = p.astype(np.float64), np.zeros(OL.NPARAMS), np.zeros(OL.NPARAMS)
    OL.adam_step(ref, m, v, g, 1, 1e-3)
    # The first Adam step moves each parameter by about lr sign(g): compare
    # where the gradient is clearly non-zero.
    sure = np.abs(g) > 1e-2 * np.abs(g).max()
    np.testing.assert_allclose(got[sure], ref[sure], rtol=0, atol=2e-5)
    assert np.abs(got - p).max() <= 1e-3 * 1.001


def test_training_fits_a_learnable_map(net):
    """b = a * f(X) with a smooth f: Adam on the GPU drives the loss down."""
    rng = np.random.default_rng(7)
    n = 20000
    X = np.zeros((n, OL.NF), np.float32)
    X[:, :OL.NUSED] = rng.uniform(-1, 1, (n, OL.NUSED))
    a = rng.normal(0, 1, n).astype(np.float32)
    f = 0.5 + 0.8 * np.tanh(2 * X[:, 0] - X[:, 6]) + 0.3 * X[:, 1] * X[:, 4]
    b = (a * f).astype(np.float32)
    net.set_params(random_params(8, scale=0.1))
    net.clear_samples()
    net.add_samples(X, a, b)
    assert net.num_samples() == n
    Xs, as_, bs = net.samples()
    np.testing.assert_array_equal(Xs, X)
    _, sse0 = net.gradient(X, a, b)
    loss = net.train(epochs=30, batch=2048, lr=3e-3, seed=1)
    _, sse1 = net.gradient(X, a, b)
    say(f"training loss {sse0 / n:.3e} -> {loss:.3e} (eval {sse1 / n:.3e})")
    assert sse1 / n < 0.05 * sse0 / n
    assert abs(loss - sse1 / n) < 0.5 * sse0 / n
    net.clear_samples()


@pytest.fixture(scope="module")
def cfg4():
    p = U.TiledProblem(4, 4, 4, 2, 2)
    return p, p.mesh(), p.geometry(), p.quadrature(), p.schedule()


def context(p, G, seed, precision=64):
    st, ss, qq = U.draw_cross_sections(seed, p.nzones, G, **CONFIG4)
    ctx = U.SweepContext(p, G, precision=precision)
    ctx.set_sigt(st)
    ctx.set_scattering(ss, qq)
    return ctx, (st, ss, qq)


def test_features_match_oracle(cfg4):
    p, m, *_ = cfg4
    G = 2
    ctx, (st, ss, qq) = context(p, G, 21)
    rng = np.random.default_rng(3)
    ph = rng.uniform(0.5, 2.0, (p.nzones, G))
    pp = ph - rng.normal(0, 0.05, (p.nzones, G))
    dp = rng.normal(0, 0.05, (p.nzones, G))
    dp[0, 0] = 0.0
    for d in (dp, None):
        X = ctx.learn_features(ph, pp, d)
        ref = OL.features(ph, pp, d, st, ss, qq, OL.neighbours(m["zoneOpp"]))
        np.testing.assert_allclose(X, ref, rtol=2e-6, atol=2e-6)
    ctx.close()


def test_zero_network_is_plain_source_iteration(cfg4):
    p = cfg4[0]
    ctx, _ = context(p, 2, 22)
    plain = ctx.source_iteration(tol=1e-8, max_iters=300)
    phi_plain = ctx.phi()
    zero = U.Learner(0)
    zero.set_params(np.zeros(OL.NPARAMS, np.float32))
    ctx.set_phi(np.zeros_like(phi_plain))
    ctx.set_learner(zero, every=2)
    ctx.set_acceleration("learned")
    got = ctx.source_iteration(tol=1e-8, max_iters=300)
    assert got[1] == plain[1]
    np.testing.assert_array_equal(got[3], plain[3])
    np.testing.assert_array_equal(ctx.phi(), phi_plain)
    assert ctx.learn_stats()[0] > 0
    ctx.close()
    zero.close()


def test_guard_falls_back_and_reaches_the_plain_fixed_point(cfg4):
    p = cfg4[0]
    ctx, _ = context(p, 2, 23)
    conv, _, _, _ = ctx.source_iteration(tol=1e-10, max_iters=500)
    assert conv
    ref = ctx.phi()
    bad = U.Learner(0)
    par = np.zeros(OL.NPARAMS, np.float32)
    par[OL.B3] = 8.0
    bad.set_params(par)
    ctx.set_phi(np.zeros_like(ref))
    ctx.set_learner(bad, every=2, guard=True)
    ctx.set_acceleration("learned")
    conv, it, res, _ = ctx.source_iteration(tol=1e-10, max_iters=500)
    assert conv and ctx.learn_stats()[1] > 0
    V, VZ = p.geometry()["Volume"], p.geometry()["VolumeZone"]
    a, b = OL.zone_average(ctx.phi(), V, VZ), OL.zone_average(ref, V, VZ)
    assert np.max(np.abs(a - b) / b) < 1e-8
    ctx.close()
    bad.close()


@pytest.fixture(scope="module")
def trained(cfg4):
    p = cfg4[0]
    say("training on config-4 iterates (seeds 1-4)")
    net, info = U.train_on_iterates(p, 2, seeds=[1, 2, 3, 4], rounds=4, epochs=40, **CONFIG4,
                                    log=say)
    say(f"trained in {info['seconds']:.2f} s, losses {info['losses']}")
    yield net, info
    net.close()


def test_learned_iteration_matches_oracle(cfg4, trained):
    """The GPU's corrected iteration follows the oracle's with the same weights."""
    p, m, g, q, s = cfg4
    net, _ = trained
    params = net.params().astype(np.float64)
    ctx, (st, ss, qq) = context(p, 2, 31)
    ctx.set_learner(net, every=2)
    ctx.set_acceleration("learned")
    conv, it, res, hist = ctx.source_iteration(tol=1e-8, max_iters=300)
    ref = OL.source_iteration(m, g, q, s, st, ss, qq, params, 1e-8, 300, nthreads=8)
    say(f"learned SI: GPU {it} iterations, oracle {ref['iterations']}")
    assert conv and ref["residual"] < 1e-8
    assert abs(it - ref["iterations"]) <= 1
    k = min(4, it, ref["iterations"])
    np.testing.assert_allclose(hist[:k], ref["history"][:k], rtol=1e-3)
    V, VZ = g["Volume"], g["VolumeZone"]
    a, b = OL.zone_average(ctx.phi(), V, VZ), OL.zone_average(ref["phi"], V, VZ)
    assert np.max(np.abs(a - b) / b) < 1e-7
    ctx.close()


def test_trained_network_cuts_iterations(cfg4, trained):
    """On held-out cross sections (config 4's seed) the trained correction
    reaches the 1e-8 residual in far fewer sweeps than plain source
    iteration, at the same fixed point."""
    p = cfg4[0]
    net, _ = trained
    ctx, _ = context(p, 2, SEED4, 64)
    conv0, it0, _, _ = ctx.source_iteration(tol=1e-8, max_iters=1000)
    plain = ctx.phi()
    ctx.set_phi(np.zeros_like(plain))
    ctx.set_learner(net, every=2)
    ctx.set_acceleration("learned")
    conv1, it1, _, _ = ctx.source_iteration(tol=1e-8, max_iters=1000)
    corr, fell = ctx.learn_stats()
    say(f"FP64: plain {it0} iterations, learned {it1} ({corr} corrections, fell back at {fell})")
    assert conv0 and conv1
    assert it1 <= 0.7 * it0
    V, VZ = p.geometry()["Volume"], p.geometry()["VolumeZone"]
    a, b = OL.zone_average(ctx.phi(), V, VZ), OL.zone_average(plain, V, VZ)
    assert np.max(np.abs(a - b) / b) < 1e-6
    # The bound is the training's doing: the network with its initial
    # (untrained) weights does not meet it.
    untrained = U.Learner(0)
    ctx.set_phi(np.zeros_like(plain))
    ctx.set_learner(untrained, every=2)
    conv2, it2, _, _ = ctx.source_iteration(tol=1e-8, max_iters=1000)
    say(f"FP64: untrained network {it2} iterations")
    assert conv2 and it2 > 0.7 * it0
    ctx.close()
    untrained.close()


def test_fp32_learned_then_fp64_matches_plain_fp64(cfg4, trained):
    """FP32 sweeps with the learned correction to 1e-6, FP64 sweeps with it to
    1e-8: fewer sweeps than plain FP64 and phi within 1e-6 of its fixed point."""
    p = cfg4[0]
    net, _ = trained
    ref_ctx, _ = context(p, 2, SEED4, 64)
    conv0, it0, _, _ = ref_ctx.source_iteration(tol=1e-8, max_iters=1000)
    plain = ref_ctx.phi()
    fast, _ = context(p, 2, SEED4, 32)
    exact, _ = context(p, 2, SEED4, 64)
    for c in (fast, exact):
        c.set_learner(net, every=2)
        c.set_acceleration("learned")
    ok, it32, it64, res, _ = U.solve_mixed(fast, exact, tol=1e-8, switch_tol=1e-6)
    say(f"plain FP64 {it0} iterations; FP32+learned {it32} + FP64+learned {it64}")
    assert conv0 and ok
    assert it32 + it64 < it0
    V, VZ = p.geometry()["Volume"], p.geometry()["VolumeZone"]
    a, b = OL.zone_average(fast.phi(), V, VZ), OL.zone_average(plain, V, VZ)
    assert np.max(np.abs(a - b) / b) < 1e-6
    for c in (ref_ctx, fast, exact):
        c.close()


def test_learned_correction_on_top_of_gta(cfg4):
    """Mode 3: the network trained on GTA-accelerated iterates corrects them
    further; GTA stays the fallback and the arbiter, the fixed point is plain
    iteration's, and it takes no more outer iterations than GTA alone."""
    p = cfg4[0]
    net, info = U.train_on_iterates(p, 2, seeds=[1, 2, 3], rounds=3, epochs=40, base="gta",
                                    **CONFIG4)
    ctx, _ = context(p, 2, SEED4, 64)
    ok0, it0, _, _ = ctx.source_iteration(tol=1e-10, max_iters=2000)
    plain = ctx.phi()
    ctx.set_phi(np.zeros_like(plain))
    ctx.set_acceleration("gta")
    ok1, it1, _, _ = ctx.source_iteration(tol=1e-8, max_iters=200)
    ctx.set_phi(np.zeros_like(plain))
    ctx.set_learner(net, every=2)
    ctx.set_acceleration("gta+learned")
    ok2, it2, _, _ = ctx.source_iteration(tol=1e-8, max_iters=200)
    corr, fell = ctx.learn_stats()
    say(f"plain {it0} (to 1e-10), GTA {it1}, GTA+learned {it2} ({corr} corrections, fell back at {fell})")
    assert ok0 and ok1 and ok2
    assert it2 <= it1
    V, VZ = p.geometry()["Volume"], p.geometry()["VolumeZone"]
    a, b = OL.zone_average(ctx.phi(), V, VZ), OL.zone_average(plain, V, VZ)
    assert np.max(np.abs(a - b) / b) < 1e-6
    ctx.close()
    net.close()


def test_non_finite_correction_falls_back_on_the_gpu(cfg4):
    """Finite weights whose FP32 output overflows: the correction is not
    applied, corrections stop, and plain iteration reaches its fixed point."""
    from test_oracle_learn import overflowing_params
    p = cfg4[0]
    ctx, _ = context(p, 2, 24)
    conv, _, _, _ = ctx.source_iteration(tol=1e-10, max_iters=500)
    ref = ctx.phi()
    net = U.Learner(0)
    net.set_params(overflowing_params().astype(np.float32))
    ctx.set_phi(np.zeros_like(ref))
    ctx.set_learner(net, every=2)
    ctx.set_acceleration("learned")
    conv, it, res, _ = ctx.source_iteration(tol=1e-10, max_iters=500)
    assert conv and ctx.learn_stats()[1] > 0
    V, VZ = p.geometry()["Volume"], p.geometry()["VolumeZone"]
    a, b = OL.zone_average(ctx.phi(), V, VZ), OL.zone_average(ref, V, VZ)
    assert np.max(np.abs(a - b) / b) < 1e-8
    ctx.close()
    net.close()


def test_network_gain_on_top_of_gta_kernel_2():
    """What the trained network adds to the reference's best accelerator, on
    cross sections drawn like the 32^3-tile workload's (sigma_t logU[1e-2,
    1e3], c U[0.9, 0.999], q U[0, 1]): a 6^3-tile problem, 32 angles, 8
    groups, held-out seed. Trained on GTA-kernel-2 iterates of 4^3-tile
    problems (seeds 1-3), the correction takes at most 3/4 of the outer
    iterations GTA kernel 2 takes alone, a network with its initial
    (untrained) weights does not, and every run reaches plain source
    iteration's fixed point (converged to 1e-10) within 1e-6."""
    xs = dict(sigt=(1e-2, 1e3), c=(0.9, 0.999), q=(0.0, 1.0))
    G = 8
    small = U.TiledProblem(4, 4, 4, 2, 2)
    net, info = U.train_on_iterates(small, G, seeds=[1, 2, 3], rounds=3, batch=8192, steps=1500,
                                    base="gta2", max_iters=500, **xs)
    say(f"trained in {info['seconds']:.1f} s on {info['samples']} samples, losses {info['losses']}")
    p = U.TiledProblem(6, 6, 6, 2, 2)
    st, ss, qq = U.draw_cross_sections(20261213, p.nzones, G, xs["sigt"], xs["c"], xs["q"])

    def run(accel, learner=None, tol=1e-8, max_iters=500):
        ctx = U.SweepContext(p, G)
        ctx.set_sigt(st)
        ctx.set_scattering(ss, qq)
        if learner is not None:
            ctx.set_learner(learner, every=2)
        ctx.set_acceleration(accel)
        ok, it, res, _ = ctx.source_iteration(tol=tol, max_iters=max_iters)
        phi = ctx.phi()
        ctx.close()
        return ok, it, phi

    ok0, it0, plain = run("none", tol=1e-10, max_iters=20000)
    ok1, it1, phi1 = run("gta2")
    ok2, it2, phi2 = run("gta2+learned", net)
    untrained = U.Learner(0)
    ok3, it3, phi3 = run("gta2+learned", untrained)
    say(f"plain {it0} (to 1e-10), GTA2 {it1}, GTA2 + trained {it2}, GTA2 + untrained {it3}")
    assert ok0 and ok1 and ok2 and ok3
    V, VZ = p.geometry()["Volume"], p.geometry()["VolumeZone"]
    ref = OL.zone_average(plain, V, VZ)
    for phi in (phi1, phi2, phi3):
        assert np.max(np.abs(OL.zone_average(phi, V, VZ) - ref) / ref) < 1e-6
    assert it2 <= 0.75 * it1, (it1, it2)
    assert it3 > 0.75 * it1, (it1, it3)
    untrained.close()
    net.close()
