"""The largest single-GPU workload at its full size on the GPU: 32x32x32
tiles (786,432 zones, 6,291,456 corners), 4x4 product quadrature per octant
(128 angles), 32 groups, sigma_t logU[1e-2, 1e3], c U[0.9, 0.999],
q U[0, 1], seed 20261207, drawn as bench.py draws its headline workload
(2.6e10 unknowns per sweep).

Groups are independent in one sweep (the source is given), so the oracle's
C sweep runs on two of the 32 groups (the first and the last) with every
angle, corner and zone, and the GPU's phi of those groups must match it
elementwise: FP64 within 1e-12 with psi held in two angle blocks of 64
(107 GB), FP32 within 2e-5 with psi of all 128 angles (107 GB), for the zone
layout and the corner layout. Every group is then covered by the FP32 sweep
agreeing with the FP64 sweep within 2e-5 elementwise, and the two layouts'
FP64 sweeps agreeing within 1e-13."""
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
    # The zone layout (the autotuner's pick here) and the corner layout at 8
    # groups per lane.
    for layout in (U.ZONE_LAYOUT, 8):
        for precision, apb in ((64, 64), (32, 0)):
            ctx = U.SweepContext(p, G, precision=precision)
            ctx.set_groups_per_lane(layout)
            if apb:
                ctx.set_angles_per_block(apb)
            ctx.set_sigt(st)
            ctx.set_source(S)
            ctx.sweep()
            phis[(layout, precision)] = ctx.phi()
            ctx.close()
            print(f"layout {layout} FP{precision} sweep done (angles per block {apb or p.nangles})",
                  flush=True)
    Ss = np.ascontiguousarray(S[:, SAMPLE])
    sts = np.ascontiguousarray(st[:, SAMPLE])
    print(f"oracle on groups {SAMPLE}, {THREADS} threads", flush=True)
    ref, _ = O.sweep_all(p.mesh(), p.geometry(), p.quadrature(), p.schedule(), Ss, sts,
                         nthreads=THREADS, fast=True)
    for layout in (U.ZONE_LAYOUT, 8):
        p64, p32 = phis[(layout, 64)], phis[(layout, 32)]
        e64 = rel_err(p64[:, SAMPLE], ref)
        e32 = rel_err(p32[:, SAMPLE], ref)
        e3264 = rel_err(p32, p64)
        print(f"layout {layout}: max rel phi error FP64 {e64:.2e}, FP32 {e32:.2e} (sampled groups); "
              f"FP32 vs FP64 over all groups {e3264:.2e}", flush=True)
        assert np.all(np.isfinite(p64)) and np.all(p64 >= 0)
        assert e64 < 1e-12, (layout, e64)
        assert e32 < 2e-5, (layout, e32)
        assert e3264 < 2e-5, (layout, e3264)
    # The two layouts over every group: agreement to rounding.
    elay = rel_err(phis[(U.ZONE_LAYOUT, 64)], phis[(8, 64)])
    print(f"zone vs corner layout, FP64, all groups: {elay:.2e}", flush=True)
    assert elay < 1e-13, elay
