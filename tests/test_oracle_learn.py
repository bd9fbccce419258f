"""The oracle's learned-correction restatement (oracle/learn.py) pinned
against torch (autograd, bfloat16 conversion, Adam), finite differences,
closed-form features and plain source iteration."""
import numpy as np
import pytest
import torch

import umt_amd as U
from oracle import learn as OL
from oracle import oracle as O


def random_params(seed, scale=0.3):
    rng = np.random.default_rng(seed)
    p = rng.uniform(-scale, scale, OL.NPARAMS)
    p[OL.W1:OL.B1].reshape(OL.NF, OL.NH)[OL.NUSED:] = 0.0
    return p.astype(np.float32).astype(np.float64)


def random_samples(seed, n):
    rng = np.random.default_rng(seed)
    X = np.zeros((n, OL.NF), np.float32)
    X[:, :OL.NUSED] = rng.normal(0, 1, (n, OL.NUSED))
    a = rng.normal(0, 1, n).astype(np.float32)
    b = rng.normal(0, 1, n).astype(np.float32)
    return X, a, b


def torch_net(p):
    w1, b1, w2, b2, w3, b3 = (torch.tensor(np.array(v, dtype=np.float64), requires_grad=True)
                              for v in OL.unpack(p))
    return [w1, b1, w2, b2, w3, b3]


def torch_loss(params, X, a, b):
    w1, b1, w2, b2, w3, b3 = params
    h1 = torch.tanh(torch.tensor(X, dtype=torch.float64) @ w1 + b1)
    h2 = torch.tanh(h1 @ w2 + b2)
    r = h2 @ w3 + b3
    e = r * torch.tensor(a, dtype=torch.float64) - torch.tensor(b, dtype=torch.float64)
    return (e * e).mean(), r


def test_bf16_matches_torch_round_to_nearest_even():
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.normal(0, 1, 20000), rng.normal(0, 1e-30, 100), rng.normal(0, 1e30, 100),
                        [0.0, -0.0, 1.0, 1.00390625, 1.01171875, 65504.0, 3.0e38, 1e-45]])
    x = x.astype(np.float32)
    ref = torch.tensor(x).to(torch.bfloat16).to(torch.float32).numpy().astype(np.float64)
    np.testing.assert_array_equal(OL.bf16(x), ref)
    # ties go to even: 1 + 2^-8 lies halfway between 1 and 1 + 2^-7
    assert OL.bf16(np.float32(1.00390625)) == 1.0
    assert OL.bf16(np.float32(1.01171875)) == 1.015625


def test_forward_and_gradient_match_torch_autograd():
    p = random_params(2)
    X, a, b = random_samples(3, 257)
    params = torch_net(p)
    loss, r = torch_loss(params, X, a, b)
    loss.backward()
    r_o, _ = OL.forward(p, X, round=False)
    np.testing.assert_allclose(r_o, r.detach().numpy(), rtol=1e-13, atol=1e-13)
    g, sse = OL.gradient(p, X, a, b, round=False)
    ref = np.concatenate([t.grad.numpy().ravel() for t in params])
    np.testing.assert_allclose(g, ref, rtol=1e-11, atol=1e-14)
    assert abs(sse / len(a) - loss.item()) < 1e-12 * max(1.0, loss.item())


def test_gradient_matches_finite_differences():
    p = random_params(4)
    X, a, b = random_samples(5, 64)
    g, _ = OL.gradient(p, X, a, b, round=False)
    rng = np.random.default_rng(6)
    idx = np.concatenate([rng.choice(np.arange(OL.W1, OL.W1 + OL.NUSED * OL.NH), 6, replace=False),
                          rng.choice(np.arange(OL.B1, OL.NPARAMS), 12, replace=False), [OL.B3]])
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        lp = OL.gradient(pp, X, a, b, round=False)[1] / len(a)
        lm = OL.gradient(pm, X, a, b, round=False)[1] / len(a)
        assert abs((lp - lm) / (2 * h) - g[i]) < 1e-7 * max(1.0, abs(g[i])), i


def test_bf16_rounding_stays_within_its_error_bound():
    """The rounded network differs from the exact one by BF16 relative
    error (2^-8) propagated through three layers, not more."""
    p = random_params(7)
    X, a, b = random_samples(8, 512)
    r0, _ = OL.forward(p, X, round=False)
    r1, _ = OL.forward(p, X, round=True)
    assert np.max(np.abs(r1 - r0)) < 0.05 * np.max(np.abs(r0))
    assert np.max(np.abs(r1 - r0)) > 0.0  # the rounding is actually applied
    g0, _ = OL.gradient(p, X, a, b, round=False)
    g1, _ = OL.gradient(p, X, a, b, round=True)
    assert np.linalg.norm(g1 - g0) < 0.05 * np.linalg.norm(g0)


def test_adam_step_matches_torch():
    p = random_params(9)
    X, a, b = random_samples(10, 128)
    params = torch_net(p)
    opt = torch.optim.Adam(params, lr=3e-3, betas=(0.9, 0.999), eps=1e-8)
    po, m, v = p.copy(), np.zeros_like(p), np.zeros_like(p)
    for t in range(1, 4):
        opt.zero_grad()
        loss, _ = torch_loss(params, X, a, b)
        loss.backward()
        opt.step()
        g, _ = OL.gradient(po, X, a, b, round=False)
        OL.adam_step(po, m, v, g, t, 3e-3)
    ref = np.concatenate([t.detach().numpy().ravel() for t in params])
    np.testing.assert_allclose(po, ref, rtol=1e-9, atol=1e-12)


@pytest.fixture(scope="module")
def tiny():
    p = U.TiledProblem(2, 2, 1, 2, 2)
    return p, p.mesh(), p.geometry(), p.quadrature(), p.schedule()


def test_features_closed_forms(tiny):
    p, m, *_ = tiny
    Z, G = p.nzones, 2
    nbr = OL.neighbours(m["zoneOpp"])
    sigt = np.full((Z, G), 2.0)
    sigs = np.full((Z, G), 1.5)
    q = np.full((Z, G), 0.5)            # q = sigma_a * ph: infinite-medium equilibrium
    ph = np.ones((Z, G))
    pp = np.full((Z, G), 0.75)
    dp = np.full((Z, G), 0.5)
    X = OL.features(ph, pp, dp, sigt, sigs, q, nbr)
    assert X.shape == (Z, G, 16) and X.dtype == np.float32
    np.testing.assert_allclose(X[..., 0], np.log10(2.0), rtol=1e-7)
    np.testing.assert_allclose(X[..., 1], 0.75, rtol=1e-7)
    np.testing.assert_allclose(X[..., 2], np.log10(0.25), rtol=1e-7)
    np.testing.assert_allclose(X[..., 3], 0.0, atol=1e-7)
    np.testing.assert_allclose(X[..., 4:6], 1.0, rtol=1e-7)   # uniform change: rel / max rel
    np.testing.assert_allclose(X[..., 6:8], 0.5, rtol=1e-7)   # d / dp = 0.25 / 0.5
    np.testing.assert_allclose(X[..., 8], 0.0, atol=1e-7)
    np.testing.assert_allclose(X[..., 9], np.log10(2.0), rtol=1e-7)
    assert not X[..., 10:].any()
    # ratio clamps at +-2, no previous change gives 0
    X2 = OL.features(ph, pp, np.full((Z, G), 0.01), sigt, sigs, q, nbr)
    np.testing.assert_allclose(X2[..., 6], 2.0)
    X3 = OL.features(ph, pp, None, sigt, sigs, q, nbr)
    assert not X3[..., 6:8].any()


def test_features_use_face_neighbours(tiny):
    p, m, *_ = tiny
    Z, G = p.nzones, 1
    nbr = OL.neighbours(m["zoneOpp"])
    sigt = np.ones((Z, G))
    sigt[0, 0] = 100.0
    X = OL.features(np.ones((Z, G)), np.zeros((Z, G)), None, sigt, 0.5 * sigt, np.ones((Z, G)), nbr)
    touching = [z for z in range(Z) if 0 in nbr[z]]
    k = np.count_nonzero(nbr >= 0, axis=1)
    for z in range(1, Z):
        want = (100.0 + (k[z] - 1)) / k[z] if z in touching else 1.0
        assert abs(10 ** X[z, 0, 9] - want) < 1e-5 * want


def test_targets_scale_by_rms_change():
    rng = np.random.default_rng(11)
    ph = rng.uniform(0.5, 2.0, (40, 3))
    d = rng.normal(0, 0.1, (40, 3))
    pstar = ph * 1.3
    a, b = OL.targets(ph, d, pstar)
    raw = d / ph
    s = np.sqrt(np.mean(raw * raw))
    np.testing.assert_allclose(a, raw / s, rtol=1e-6)
    np.testing.assert_allclose(b, 0.3 / s, rtol=1e-6)
    ph[0, 0] = 0.0
    a, b = OL.targets(ph, d, pstar)
    assert a[0, 0] == 0.0 and b[0, 0] == 0.0


def problem_xs(p, seed, G=2):
    return U.draw_cross_sections(seed, p.nzones, G, (0.1, 10.0), (0.95, 0.999), (1.0, 1.0))


def test_zero_network_is_plain_source_iteration(tiny):
    p, m, g, q, s = tiny
    st, ss, qq = problem_xs(p, 12)
    plain = O.source_iteration(m, g, q, s, st, ss, qq, 1e-8, 300)
    res = OL.source_iteration(m, g, q, s, st, ss, qq, np.zeros(OL.NPARAMS), 1e-8, 300)
    assert res["iterations"] == plain[1]
    assert res["corrections"] > 0
    np.testing.assert_array_equal(res["phi"], plain[0])
    np.testing.assert_allclose(res["history"], plain[3], rtol=0, atol=0)


def test_guard_falls_back_to_plain_iteration(tiny):
    """A network predicting r = 8 everywhere over-corrects; the guard stops it
    and plain iteration still reaches the plain fixed point."""
    p, m, g, q, s = tiny
    st, ss, qq = problem_xs(p, 13)
    par = np.zeros(OL.NPARAMS)
    par[OL.B3] = 8.0
    plain = O.source_iteration(m, g, q, s, st, ss, qq, 1e-10, 500)
    res = OL.source_iteration(m, g, q, s, st, ss, qq, par, 1e-10, 500)
    assert res["fell_back"] > 0 and res["corrections"] >= 1
    assert res["residual"] < 1e-10
    pz = OL.zone_average(plain[0], g["Volume"], g["VolumeZone"])
    rz = OL.zone_average(res["phi"], g["Volume"], g["VolumeZone"])
    assert np.max(np.abs(rz - pz) / pz) < 1e-8
    nog = OL.source_iteration(m, g, q, s, st, ss, qq, par, 1e-10, 60, guard=False)
    assert not nog["residual"] < 1e-3  # unguarded, the same correction diverges


def test_collected_states_match_the_iteration(tiny):
    p, m, g, q, s = tiny
    st, ss, qq = problem_xs(p, 14)
    res = OL.source_iteration(m, g, q, s, st, ss, qq, None, 1e-9, 300, collect=True)
    states = res["states"]
    # plain iteration: one state per sweep from the second on, except the last
    assert len(states) == res["iterations"] - 2
    for k, (X, ph, d, r) in enumerate(states):
        assert r == res["history"][k + 1]
        assert X.shape == ph.shape + (16,)


def overflowing_params():
    """Finite weights whose FP32 output overflows to infinity."""
    par = np.zeros(OL.NPARAMS)
    par[OL.B2:OL.W3] = 1.0          # h2 = tanh(1) everywhere
    par[OL.W3:OL.B3] = 3.0e38       # 64 terms of ~2.3e38 overflow FP32
    return par


def test_non_finite_correction_falls_back(tiny):
    p, m, g, q, s = tiny
    st, ss, qq = problem_xs(p, 15)
    plain = O.source_iteration(m, g, q, s, st, ss, qq, 1e-10, 500)
    res = OL.source_iteration(m, g, q, s, st, ss, qq, overflowing_params(), 1e-10, 500)
    assert res["fell_back"] > 0 and res["residual"] < 1e-10
    pz = OL.zone_average(plain[0], g["Volume"], g["VolumeZone"])
    rz = OL.zone_average(res["phi"], g["Volume"], g["VolumeZone"])
    assert np.max(np.abs(rz - pz) / pz) < 1e-8
