"""The largest single-GPU workload at its full size on the GPU: 32x32x32
tiles (786,432 zones, 6,291,456 corners), 4x4 product quadrature per octant
(128 angles), 32 groups, sigma_t logU[1e-2, 1e3], c U[0.9, 0.999],
q U[0, 1], seed 20261207, drawn as bench.py draws its headline workload
(2.6e10 unknowns per sweep).

Groups are independent in one sweep (the source is given), so the oracle's
C sweep runs on two of the 32 groups (the first and the last) with every
angle, corner and zone, and the GPU's phi of those groups must match it
elementwise: FP64 within 1e-12 with psi held in two angle blocks of 64
(107 GB), FP32 within 2e-5 with psi of all 128 angles (107 GB). Every group
is then covered by the FP32 sweep agreeing with the FP64 sweep within 2e-5
elementwise."""
import os

import numpy as np
import pytest

import umt_amd as U
from oracle import oracle as O

from conftest import rel_err

pytestmark = pytest.mark.gpu

SEED = 20261207
THREADS = max(1, min(16, os.cpu_count() or 1))
SAMPLE = [0, 31]


def test_config5_full_size_sweep_matches_oracle():
    p = U.TiledProblem(32, 32, 32, 4, 4)
    G = 32
    st, ss, qq = U.draw_cross_sections(SEED, p.nzones, G, (1e-2, 1e3), (0.9, 0.999), (0.0, 1.0))
    S = np.repeat(qq, 8, axis=0) / (4.0 * np.pi)
    print(f"case config 5 full size: {p.nzones} zones, {p.ncorner} corners, {p.nangles} angles, "
          f"{G} groups", flush=True)
    phis = {}
    for precision, apb in ((64, 64), (32, 0)):
        ctx = U.SweepContext(p, G, precision=precision)
        if apb:
            ctx.set_angles_per_block(apb)
        ctx.set_sigt(st)
        ctx.set_source(S)
        ctx.sweep()
        phis[precision] = ctx.phi()
        ctx.close()
        print(f"FP{precision} sweep done (angles per block {apb or p.nangles})", flush=True)
    Ss = np.ascontiguousarray(S[:, SAMPLE])
    sts = np.ascontiguousarray(st[:, SAMPLE])
    print(f"oracle on groups {SAMPLE}, {THREADS} threads", flush=True)
    ref, _ = O.sweep_all(p.mesh(), p.geometry(), p.quadrature(), p.schedule(), Ss, sts,
                         nthreads=THREADS, fast=True)
    e64 = rel_err(phis[64][:, SAMPLE], ref)
    e32 = rel_err(phis[32][:, SAMPLE], ref)
    e3264 = rel_err(phis[32], phis[64])
    print(f"max rel phi error: FP64 {e64:.2e}, FP32 {e32:.2e} (sampled groups); FP32 vs FP64 "
          f"over all groups {e3264:.2e}", flush=True)
    assert np.all(np.isfinite(phis[64])) and np.all(phis[64] >= 0)
    assert e64 < 1e-12, e64
    assert e32 < 2e-5, e32
    assert e3264 < 2e-5, e3264
