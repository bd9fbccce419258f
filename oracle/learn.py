"""ORACLE — test infrastructure only. CPU restatement (numpy, FP64) of the
learned correction of source iteration that libumt_amd.so runs on the GPU:
features, the 3-layer network (forward, gradient, Adam) with the GPU's BF16
rounding points reproduced, training targets, and the corrected source
iteration driven by the oracle's C sweeps. Only tests/ may import this module.

The correction has no counterpart in the reference; what it accelerates is
the steady source iteration of rt/rtmainsn.F90:167-235 with the residual of
rt/ConvergenceTest.F90:150-168 (zone averages, cutoff 1e-6 of the maximum).

How each part is pinned (tests/test_oracle_learn.py):
  bf16               torch's float32 -> bfloat16 conversion (round to nearest even)
  forward, gradient  torch autograd in float64 (round=False), central finite
                     differences, and the rounded path against the unrounded
                     one within the BF16 error bound
  adam_step          torch.optim.Adam with the same hyperparameters
  features           closed forms on uniform and equilibrium states
  source_iteration   zero network = plain source iteration exactly; the guard
                     falls back on a diverging correction and on a non-finite
                     one, and still reaches the plain fixed point
"""
import numpy as np

from . import oracle as O

NF, NH = 16, 64
NUSED = 10
W1, B1 = 0, NF * NH
W2 = B1 + NH
B2 = W2 + NH * NH
W3 = B2 + NH
B3 = W3 + NH
NPARAMS = B3 + 1
TINY = 1e-300
WTISO = 1.0 / (4.0 * np.pi)


def bf16(x):
    """float32 values rounded to bfloat16 (nearest even), returned as float64."""
    f = np.ascontiguousarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF
    r = np.where(nan, (u >> 16) | 0x40, r).astype(np.uint32) << 16
    return r.view(np.float32).astype(np.float64).reshape(np.shape(x))


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return (p[W1:B1].reshape(NF, NH), p[B1:W2], p[W2:B2].reshape(NH, NH), p[B2:W3], p[W3:B3],
            p[B3])


def forward(p, X, round=True):
    """r (n,) and the activations (h1, h2) for features X (n, 16)."""
    w1, b1, w2, b2, w3, b3 = unpack(p)
    X = np.asarray(X, dtype=np.float32).reshape(-1, NF)
    if round:
        Xb, w1b, w2b = bf16(X), bf16(w1.astype(np.float32)), bf16(w2.astype(np.float32))
        z1 = (Xb @ w1b).astype(np.float32).astype(np.float64) + b1
        h1 = np.tanh(z1.astype(np.float32).astype(np.float64))
        z2 = (bf16(h1.astype(np.float32)) @ w2b).astype(np.float32).astype(np.float64) + b2
        h2 = np.tanh(z2.astype(np.float32).astype(np.float64))
    else:
        h1 = np.tanh(X.astype(np.float64) @ w1 + b1)
        h2 = np.tanh(h1 @ w2 + b2)
    return h2 @ w3 + b3, (h1, h2)


def gradient(p, X, a, b, round=True):
    """Gradient of (1/n) sum (r a - b)^2 (flat, NPARAMS) and sum (r a - b)^2."""
    w1, b1, w2, b2, w3, b3 = unpack(p)
    X = np.asarray(X, dtype=np.float32).reshape(-1, NF)
    a = np.asarray(a, dtype=np.float32).astype(np.float64).ravel()
    b = np.asarray(b, dtype=np.float32).astype(np.float64).ravel()
    n = X.shape[0]
    r, (h1, h2) = forward(p, X, round)
    e = r * a - b
    g = 2.0 * a * e / n
    dz2 = g[:, None] * w3[None, :] * (1.0 - h2 * h2)
    rnd = bf16 if round else (lambda v: np.asarray(v, dtype=np.float64))
    w2o = bf16(w2.astype(np.float32)) if round else w2
    h1o = bf16(h1.astype(np.float32)) if round else h1
    dz2o = rnd(dz2.astype(np.float32)) if round else dz2
    dh1 = dz2o @ w2o.T
    dz1 = dh1 * (1.0 - h1 * h1)
    Xo = bf16(X) if round else X.astype(np.float64)
    dz1o = rnd(dz1.astype(np.float32)) if round else dz1
    out = np.zeros(NPARAMS)
    out[W1:B1] = (Xo.T @ dz1o).ravel()
    out[B1:W2] = dz1.sum(0)
    out[W2:B2] = (h1o.T @ dz2o).ravel()
    out[B2:W3] = dz2.sum(0)
    out[W3:B3] = (g[:, None] * h2).sum(0)
    out[B3] = g.sum()
    return out, float((e * e).sum())


def adam_step(p, m, v, grad, t, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """In place on p, m, v; t is the step number (1-based)."""
    m[:] = beta1 * m + (1.0 - beta1) * grad
    v[:] = beta2 * v + (1.0 - beta2) * grad * grad
    p -= lr * (m / (1.0 - beta1 ** t)) / (np.sqrt(v / (1.0 - beta2 ** t)) + eps)


def _nbmean(x, nbr):
    """Mean over each zone's face neighbours (nbr (Z, 6), 0-based, -1 none);
    a zone without neighbours keeps its own value."""
    acc = np.zeros_like(x)
    cnt = np.zeros((x.shape[0], 1))
    for f in range(nbr.shape[1]):
        j = nbr[:, f]
        ok = j >= 0
        acc[ok] += x[j[ok]]
        cnt[ok] += 1
    return np.where(cnt > 0, acc / np.maximum(cnt, 1), x)


def neighbours(zoneOpp):
    """0-based face neighbours from the 1-based zoneOpp (Z, 6)."""
    z = np.asarray(zoneOpp)
    return np.where(z > 0, z - 1, -1)


def features(ph, pp, dp, sigt, sigs, q, nbr):
    """Feature slots (Z, G, 16) float32 for zone averages (Z, G)."""
    d = ph - pp
    rel = d / np.maximum(np.abs(ph), TINY)
    if dp is None:
        ratio = np.zeros_like(d)
    else:
        den = np.where(np.abs(dp) > TINY, dp, TINY)
        ratio = np.clip(d / den, -2.0, 2.0)
    M = np.abs(rel).max(0) + TINY
    c = sigs / sigt
    sa = sigt - sigs
    X = np.zeros(ph.shape + (NF,))
    X[..., 0] = np.log10(sigt)
    X[..., 1] = c
    X[..., 2] = np.log10(np.maximum(1.0 - c, 1e-12))
    X[..., 3] = np.clip(np.log(np.maximum(ph * sa, TINY) / np.maximum(q, TINY)), -30, 30)
    X[..., 4] = rel / M
    X[..., 5] = _nbmean(rel, nbr) / M
    X[..., 6] = ratio
    X[..., 7] = _nbmean(ratio, nbr)
    X[..., 8] = np.clip(np.log(np.maximum(_nbmean(ph, nbr), TINY) / np.maximum(ph, TINY)), -30, 30)
    X[..., 9] = np.log10(np.maximum(_nbmean(sigt, nbr), TINY))
    return X.astype(np.float32)


def targets(ph, d, pstar):
    """(a, b) float32 of one state: d/ph and (pstar - ph)/ph, zero where ph <= 0,
    both scaled by 1/rms(a)."""
    pos = ph > 0
    safe = np.where(pos, ph, 1.0)
    a = np.where(pos, d / safe, 0.0)
    b = np.where(pos, (pstar - ph) / safe, 0.0)
    s = np.sum(a * a)
    scale = 1.0 / np.sqrt(s / a.size) if s > 0 else 0.0
    return (a * scale).astype(np.float32), (b * scale).astype(np.float32)


def zone_average(x, V, VZ):
    return (x * V[:, None]).reshape(-1, 8, x.shape[1]).sum(1) / VZ[:, None]


def residual(new, old, magnitude=False):
    """rt/ConvergenceTest.F90:150-168 on zone averages; magnitude: test |new|
    (what the corrected iteration uses, since its iterates may go negative)."""
    if not np.all(np.isfinite(new)):
        return float("inf")
    a = np.abs(new) if magnitude else new
    thr = 1e-6 * max(a.max(), 0.0)
    mask = a > thr
    if not mask.any():
        return 0.0
    return float(np.max(np.abs((new[mask] - old[mask]) / new[mask])))


def source_iteration(mesh, geom, quad, sched, sigt, sigs, q, params, tol, max_iters, every=2,
                     guard=True, correct=True, collect=False, phi0=None, nthreads=1):
    """Source iteration with the learned correction (params None or
    correct=False: plain). Returns dict(phi, iterations, residual, history,
    fell_back, corrections, states) with states = [(X, ph, d, res)] when collect."""
    V, VZ = geom["Volume"], geom["VolumeZone"]
    nbr = neighbours(mesh["zoneOpp"])
    G = sigt.shape[1]
    Cn = V.shape[0]
    sigs_c, q_c = np.repeat(sigs, 8, axis=0), np.repeat(q, 8, axis=0)
    phi = np.zeros((Cn, G)) if phi0 is None else np.array(phi0, dtype=np.float64)
    old = zone_average(phi, V, VZ)
    correcting = correct and params is not None
    track = correcting or collect
    since, have_dp, last_res = 0, False, -1.0
    pending_bad = False
    dp = None
    hist, states = [], []
    fell_back, corrections = 0, 0
    res, it = 1.0, 0
    while it < max_iters:
        prev = phi
        half, _ = O.sweep_all(mesh, geom, quad, sched, WTISO * (sigs_c * phi + q_c), sigt,
                              nthreads=nthreads)
        ph = zone_average(half, V, VZ)
        pp = old
        res = residual(ph, pp, magnitude=correct and params is not None)
        old = ph
        phi = half
        hist.append(res)
        it += 1
        if pending_bad and correcting:  # a non-finite correction: stop correcting
            correcting = False
            fell_back = it
        pending_bad = False
        if res < tol or not np.isfinite(res):
            break
        if track:
            if last_res >= 0.0:
                if guard and res > last_res and correcting:
                    correcting = False
                    fell_back = it
                last_res = -1.0
            since += 1
            d = ph - pp
            if since >= every and have_dp:
                X = features(ph, pp, dp, sigt, sigs, q, nbr)
                if collect:
                    states.append((X, ph.copy(), d.copy(), res))
                if correcting:
                    r, _ = forward(params, X.reshape(-1, NF))
                    with np.errstate(over="ignore", invalid="ignore"):
                        r = r.astype(np.float32).astype(np.float64).reshape(ph.shape)
                        new = half + np.repeat(r, 8, axis=0) * (half - prev)
                    ok = np.isfinite(new)
                    phi = np.where(ok, new, half)  # non-finite corrections are not applied
                    if not ok.all():
                        pending_bad = True
                    old = zone_average(phi, V, VZ)
                    since = 0
                    last_res = res
                    corrections += 1
            dp = d
            have_dp = True
    return dict(phi=phi, iterations=it, residual=res, history=np.array(hist),
                fell_back=fell_back, corrections=corrections, states=states)
