"""Learned correction of source iteration (host side of libumt_amd.so's
umt_learner_* and umt_ctx_learn_* entry points).

A per-(zone, group) network predicts r; every `every` sweeps the state
becomes phi_half + r (phi_half - phi_prev). Plain source iteration decides
convergence and is the fallback (guard). The network is trained on the
method's own iterates: solves with collection on record their states, and
after convergence each state's target is the correction that would have
reached the converged zone averages. Training alternates with solves that
use the current network, so later samples come from the states the
corrected iteration actually visits.
"""
import ctypes as C
import time

import numpy as np

from . import _lib
from ._lib import check
from .sweep import SweepContext, draw_cross_sections

NFEAT = 16
_pf = C.POINTER(C.c_float)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _fp(a):
    return a.ctypes.data_as(_pf)


class Learner:
    """The correction network and its training samples (device resident)."""

    def __init__(self, seed=0):
        lib = _lib.load()
        h = C.c_void_p()
        check("umt_learner_create", lib.umt_learner_create(seed, C.byref(h)))
        self._h = h
        self.nparams = int(lib.umt_learner_num_params())

    @property
    def handle(self):
        return self._h

    def close(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None and _lib.load is not None:
            _lib.load().umt_learner_destroy(h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        p = np.zeros(self.nparams, dtype=np.float32)
        check("umt_learner_get_params", _lib.load().umt_learner_get_params(self._h, _fp(p)))
        return p

    def set_params(self, p):
        a = _f32(p)
        assert a.shape == (self.nparams,)
        check("umt_learner_set_params", _lib.load().umt_learner_set_params(self._h, _fp(a)))

    def forward(self, X):
        X = _f32(X).reshape(-1, NFEAT)
        r = np.zeros(X.shape[0], dtype=np.float32)
        check("umt_learner_forward", _lib.load().umt_learner_forward(self._h, _fp(X), X.shape[0], _fp(r)))
        return r

    def gradient(self, X, a, b):
        """(gradient of mean (r a - b)^2, sum of squared errors)."""
        X = _f32(X).reshape(-1, NFEAT)
        a, b = _f32(a).ravel(), _f32(b).ravel()
        g = np.zeros(self.nparams, dtype=np.float32)
        sse = C.c_double(0.0)
        check("umt_learner_gradient", _lib.load().umt_learner_gradient(
            self._h, _fp(X), _fp(a), _fp(b), X.shape[0], _fp(g), C.byref(sse)))
        return g, sse.value

    def step(self, X, a, b, lr):
        X = _f32(X).reshape(-1, NFEAT)
        a, b = _f32(a).ravel(), _f32(b).ravel()
        check("umt_learner_step", _lib.load().umt_learner_step(
            self._h, _fp(X), _fp(a), _fp(b), X.shape[0], lr))

    def add_samples(self, X, a, b):
        X = _f32(X).reshape(-1, NFEAT)
        a, b = _f32(a).ravel(), _f32(b).ravel()
        check("umt_learner_add_samples", _lib.load().umt_learner_add_samples(
            self._h, _fp(X), _fp(a), _fp(b), X.shape[0]))

    def num_samples(self):
        n = C.c_int64(0)
        check("umt_learner_num_samples", _lib.load().umt_learner_num_samples(self._h, C.byref(n)))
        return n.value

    def samples(self):
        n = self.num_samples()
        X = np.zeros((n, NFEAT), dtype=np.float32)
        a = np.zeros(n, dtype=np.float32)
        b = np.zeros(n, dtype=np.float32)
        if n:
            check("umt_learner_get_samples", _lib.load().umt_learner_get_samples(
                self._h, _fp(X), _fp(a), _fp(b)))
        return X, a, b

    def clear_samples(self):
        check("umt_learner_clear_samples", _lib.load().umt_learner_clear_samples(self._h))

    def train(self, epochs=40, batch=4096, lr=3e-3, seed=0):
        """Adam over the held samples; returns the last epoch's mean loss."""
        loss = C.c_double(0.0)
        check("umt_learner_train", _lib.load().umt_learner_train(
            self._h, epochs, batch, lr, seed, C.byref(loss)))
        return loss.value


def train_on_iterates(problem, groups, seeds, sigt=(0.1, 10.0), c=(0.95, 0.999), q=(1.0, 1.0),
                      rounds=4, epochs=40, batch=4096, lr=3e-3, every=2, precision=64,
                      learner=None, net_seed=0, tol=1e-10, max_iters=2000, log=None,
                      collect_stride=1, steps=None, graphs=True, base="none"):
    """Train a Learner on source-iteration iterates of `problem` with cross
    sections drawn from `seeds`. Round 0 records plain iterations; each
    later round records the iterations corrected by the network trained so
    far. collect_stride keeps every n-th state; steps (if given) fixes the
    Adam steps per round instead of the epochs. base "gta" trains the
    correction of GTA-accelerated iterates (used as "gta+learned").
    Returns (learner, info) with per-round losses and timings."""
    net = learner if learner is not None else Learner(net_seed)
    ctx = SweepContext(problem, groups, precision=precision)
    ctx.set_graphs(graphs)
    info = {"losses": [], "samples": [], "seconds": 0.0}
    t0 = time.perf_counter()
    try:
        for rnd in range(rounds):
            for sd in seeds:
                st, ss, qq = draw_cross_sections(sd, problem.nzones, groups, sigt, c, q)
                ctx.set_sigt(st)
                ctx.set_scattering(ss, qq)
                ctx.set_phi(np.zeros((problem.ncorner, groups)))
                if rnd == 0:
                    ctx.set_learner(net, every=every, guard=False)
                    ctx.set_acceleration(base)
                else:
                    ctx.set_learner(net, every=every, guard=True)
                    ctx.set_acceleration("learned" if base == "none" else base + "+learned")
                ctx.learn_collect(True, collect_stride)
                conv, it, res, _ = ctx.source_iteration(tol=tol, max_iters=max_iters)
                ctx.learn_collect(False)
                if conv:
                    ctx.learn_commit(net)
                else:
                    ctx.learn_discard()
                if log:
                    log(f"round {rnd} seed {sd}: {it} iterations, converged={conv}")
            ep = epochs
            if steps is not None:
                ep = max(1, -(-steps * batch // max(net.num_samples(), 1)))
            loss = net.train(epochs=ep, batch=batch, lr=lr, seed=1000 + rnd)
            info["losses"].append(loss)
            info["samples"].append(net.num_samples())
            if log:
                log(f"round {rnd}: {net.num_samples()} samples, loss {loss:.3e}")
    finally:
        ctx.set_learner(None)
        ctx.close()
    info["seconds"] = time.perf_counter() - t0
    return net, info
