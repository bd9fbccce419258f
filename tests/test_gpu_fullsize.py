"""The sweeps at the benchmark's full sizes against the oracle's C sweep on
the host, with the inputs drawn as bench.py draws them:
  - xyz: 20x20x20 tiles (192000 zones), 4x4 product quadrature per octant
    (128 angles), 16 groups, 3.1e9 unknowns per sweep;
  - RZ: 512x512 quads, 4x4 RZ quadrature (72 swept directions), 16 groups,
    1.2e9 unknowns per sweep.
FP64 phi within 1e-12 and FP32 within 2e-5 of the oracle, elementwise over
every corner and group (xyz: for the corner and the zone layout). The oracle is fed the product's host arrays (mesh,
geometry, quadrature, schedule), which the oracle tests show bit-identical
to its own at smaller sizes."""
import os

import numpy as np
import pytest

import umt_amd as U
from oracle import oracle as O
from oracle import rz as R

from conftest import rel_err

pytestmark = pytest.mark.gpu

SEED = 20261204  # bench.py's default seed
THREADS = max(1, min(16, os.cpu_count() or 1))


def test_full_size_xyz_sweep_matches_oracle():
    p = U.TiledProblem(20, 20, 20, 4, 4)
    G = 16
    st, ss, qq = U.draw_cross_sections(SEED, p.nzones, G, (1e-2, 1e3), (0.0, 0.95), (0.0, 1.0))
    S = np.repeat(qq, 8, axis=0) / (4.0 * np.pi)
    print(f"case xyz full size: {p.nzones} zones, {p.nangles} angles, {G} groups; oracle on "
          f"{THREADS} threads", flush=True)
    ref, _ = O.sweep_all(p.mesh(), p.geometry(), p.quadrature(), p.schedule(), S, st,
                         nthreads=THREADS, fast=True)
    # The corner layout (4 groups per lane) and the zone layout, which the
    # autotuner picks at this size.
    for layout in (4, U.ZONE_LAYOUT):
        for precision, tol in ((64, 1e-12), (32, 2e-5)):
            ctx = U.SweepContext(p, G, precision=precision)
            ctx.set_groups_per_lane(layout)
            ctx.set_sigt(st)
            ctx.set_source(S)
            ctx.sweep()
            err = rel_err(ctx.phi(), ref)
            ctx.close()
            print(f"layout {layout} FP{precision}: max rel phi error {err:.2e}", flush=True)
            assert err < tol, (layout, precision, err)


def test_full_size_rz_sweep_matches_oracle():
    p = U.RzProblem(512, 512, 4, 4, r0=0.0, r1=1.0, z0=0.0, z1=1.0)
    G = 16
    rng = np.random.default_rng(SEED)
    sig = np.exp(rng.uniform(np.log(1e-2), np.log(1e3), (p.nzones, G)))
    S = rng.uniform(0.0, 1.0, (p.ncorner, G)) / (2 * np.pi)
    print(f"case RZ full size: {p.nzones} zones, {p.nangles} directions, {G} groups; oracle on "
          f"{min(THREADS, p.nlevels)} threads", flush=True)
    ref, _, _ = R.sweep_all(p.mesh(), p.geometry(), p.quadrature(), p.schedule(), S, sig,
                            nthreads=min(THREADS, p.nlevels), fast=True)
    for precision, tol in ((64, 1e-12), (32, 2e-5)):
        ctx = U.RzContext(p, G, precision=precision)
        ctx.set_sigt(sig)
        ctx.set_source(S)
        ctx.sweep()
        err = rel_err(ctx.phi(), ref)
        ctx.close()
        print(f"RZ FP{precision}: max rel phi error {err:.2e}", flush=True)
        assert err < tol, (precision, err)
