#!/usr/bin/env python3
"""Sweep throughput and time to convergence of the MI355X UCB xyz sweep.

Workload (one step = one full sweep of all angles and groups): the unit cube
as 20x20x20 tiles of UMT's 24-zone tile (192,000 hex zones, 1,536,000
corners), 4x4 product quadrature per octant (128 angles), 16 groups,
sigma_t ~ log-uniform[1e-2, 1e3] per (zone, group), c ~ U[0, 0.95],
q ~ U[0, 1], vacuum boundaries, seed 20261204. Unknowns per sweep =
corners x angles x groups = 3,145,728,000 (driver/test_driver.cc:1895).

Prints one JSON line. ``value`` is FP64 sweep unknowns/s with every input
resident in HBM. Also reported: the FP32 sweep (FP64 phi), the wall time of
source iteration to residual < 1e-8 for both, the sweep-level kernel against
the HBM roofline, and the oracle's CPU sweep timed on this host's cores.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

SEED = 20261204
HBM_PEAK_GBS = 8000.0


def bytes_per_unknown(precision, n_in, G):
    """Algorithmic HBM bytes per unknown of the reference dataflow (one angle
    at a time, no reuse across angles): source read, psi write, n_in upwind
    psi reads, phi read+write in FP64, sigma_t shared by 8 corners, and the
    per-corner geometry/connectivity bytes shared by the G groups."""
    if precision == 64:
        return 8 + 8 + 8 * n_in + 16 + 1 + 185.5 / G
    return 4 + 4 + 4 * n_in + 16 + 0.5 + 109.5 / G


def load_traffic(name):
    """HBM bytes per launch from a committed rocprofv3 --pmc summary, if any."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(p, G, st, S, seconds):
    """Oracle sweep (OpenMP over angles, one angle per thread) on a subset of
    angles, repeated for about `seconds` seconds."""
    from oracle import oracle as O

    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    threads = max(1, min(threads, 16, p.nangles))
    k = threads
    m, g, q, s = p.mesh(), p.geometry(), p.quadrature(), p.schedule()
    nplanes = int(np.sum(s["nHyperPlanes"][:k]))
    sub_q = dict(omega=np.ascontiguousarray(q["omega"][:k]),
                 weight=np.ascontiguousarray(q["weight"][:k]))
    sub_s = dict(nHyperPlanes=np.ascontiguousarray(s["nHyperPlanes"][:k]),
                 zonesInPlane=np.ascontiguousarray(s["zonesInPlane"][:nplanes]),
                 nextZ=np.ascontiguousarray(s["nextZ"][:k]),
                 nextC=np.ascontiguousarray(s["nextC"][:k]))
    O.load(fast=True)
    reps, t_total = 0, 0.0
    while t_total < seconds or reps == 0:
        t0 = time.perf_counter()
        O.sweep_all(m, g, sub_q, sub_s, S, st, nthreads=threads, fast=True)
        t_total += time.perf_counter() - t0
        reps += 1
    unknowns = p.ncorner * k * G * reps
    return {
        "value": unknowns / t_total,
        "unit": "unknowns/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"oracle C sweep (liboracle_fast.so, -O3 -march=x86-64-v3, OpenMP over "
                   f"angles) of {k} of the {p.nangles} angles x {G} groups x {p.ncorner} corners, "
                   f"{reps} repetitions in {t_total:.1f} s on {threads} threads of '{cpu_model()}'"),
    }


def measure(p, G, precision, st, S, warmup, steps, lib):
    import umt_amd as U

    ctx = U.SweepContext(p, G, precision=precision)
    ctx.set_sigt(st)
    ctx.set_source(S)
    best_k, tune_ms = ctx.autotune(2)
    ctx.sweep_n(warmup)
    lib.gpu_devicesynchronize()
    t0 = time.perf_counter()
    ctx.sweep_n(steps)
    lib.gpu_devicesynchronize()
    dt = time.perf_counter() - t0
    ms_sweep_ev, ms_levels, launches = ctx.time_sweeps(1, 3)
    n_in = ctx.incoming_faces()
    ctx.close()
    return dict(dt=dt, ms_levels=ms_levels, launches=launches, n_in=n_in,
                ms_sweep_events=ms_sweep_ev, groups_per_lane=best_k,
                autotune_ms_per_sweep={str(k): v for k, v in tune_ms.items()})


def converge(p, G, precision, st, ss, qq, tol, max_iters):
    import umt_amd as U

    ctx = U.SweepContext(p, G, precision=precision)
    ctx.set_sigt(st)
    ctx.set_scattering(ss, qq)
    best_k, _ = ctx.autotune(1)
    t0 = time.perf_counter()
    ok, it, res, _ = ctx.source_iteration(tol=tol, max_iters=max_iters)
    dt = time.perf_counter() - t0
    phi = ctx.phi()
    ctx.close()
    return dict(converged=ok, iterations=it, residual=res, seconds=dt,
                groups_per_lane=best_k), phi


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--tiles", type=int, nargs=3, default=[20, 20, 20])
    ap.add_argument("--quad", type=int, nargs=2, default=[4, 4], help="npolar nazimuthal")
    ap.add_argument("--groups", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-converge", action="store_true", help="skip time-to-convergence")
    ap.add_argument("--no-fp32", action="store_true", help="skip the FP32 sweep")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--max-iters", type=int, default=3000)
    ap.add_argument("--calibrate", action="store_true",
                    help="only run the HBM copy probe (for rocprofv3 --pmc calibration)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} is not supported: this build measures one GPU "
              "(one MI355X); multi-GPU sharding of the sweep is design-only.", file=sys.stderr)
        sys.exit(2)

    import umt_amd as U
    from umt_amd import _lib

    lib = _lib.load()
    if args.calibrate:
        n = 1 << 27
        ms = U.sweep.hbm_copy_probe(n, 5)
        print(json.dumps({"copy_probe_bytes_read": n * 8, "copy_probe_bytes_written": n * 8,
                          "ms_per_copy": ms, "GBps": 2 * n * 8 / ms / 1e6}))
        return

    nx, ny, nz = args.tiles
    npol, nazi = args.quad
    G = args.groups
    t0 = time.perf_counter()
    p = U.TiledProblem(nx, ny, nz, npol, nazi)
    setup_s = time.perf_counter() - t0
    st, ss, qq = U.draw_cross_sections(SEED, p.nzones, G, (1e-2, 1e3), (0.0, 0.95), (0.0, 1.0))
    S = np.repeat(qq, 8, axis=0) / (4.0 * np.pi)
    unknowns = p.unknowns(G)

    r64 = measure(p, G, 64, st, S, args.warmup, args.steps, lib)
    value = unknowns * args.steps / r64["dt"]
    B64 = bytes_per_unknown(64, r64["n_in"], G)
    ach64 = B64 * unknowns / (r64["ms_levels"] * 1e-3) / 1e9
    tr = load_traffic("ucb_sweep_level_fp64_20x20x20_traffic.json")
    traffic = tr.get("hbm_bytes_per_launch") if tr else None
    roofline = {
        "bound": "hbm", "achieved": ach64, "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": ach64 / HBM_PEAK_GBS, "traffic": traffic,
        "kernel": "ucb_sweep_level<double,K,false>",
        "launches_per_sweep": r64["launches"],
        "mean_launch_ms": r64["ms_levels"] / r64["launches"],
        "algorithmic_bytes_per_unknown": B64, "n_in": r64["n_in"],
        "unknowns_per_launch": unknowns / r64["launches"],
        "groups_per_lane": r64["groups_per_lane"],
        "autotune_ms_per_sweep": r64["autotune_ms_per_sweep"],
    }
    out = {
        "metric": "sweep unknowns/s (zone-corner x angle x group); wall time to converged phi at 1e-8",
        "value": value,
        "unit": "unknowns/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r64["dt"] * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded tiled hex mesh and cross sections)",
        "config": {
            "workload": (f"UCB xyz sweep, {nx}x{ny}x{nz} tiles ({p.nzones} zones, {p.ncorner} "
                         f"corners), {npol}x{nazi} product quadrature per octant ({p.nangles} "
                         f"angles), {G} groups, sigma_t logU[1e-2,1e3], c U[0,0.95], q U[0,1], "
                         f"vacuum"),
            "tiles": [nx, ny, nz], "zones": p.nzones, "corners": p.ncorner,
            "angles": p.nangles, "groups": G, "unknowns_per_sweep": unknowns,
            "parallelism": "1 GPU",
        },
        "roofline": roofline,
        "setup_s": setup_s,
    }

    if not args.no_fp32:
        r32 = measure(p, G, 32, st, S, args.warmup, args.steps, lib)
        B32 = bytes_per_unknown(32, r32["n_in"], G)
        ach32 = B32 * unknowns / (r32["ms_levels"] * 1e-3) / 1e9
        out["fp32_sweep"] = {
            "value": unknowns * args.steps / r32["dt"], "unit": "unknowns/s",
            "ms_per_step": r32["dt"] * 1e3 / args.steps,
            "roofline": {"bound": "hbm", "achieved": ach32, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": ach32 / HBM_PEAK_GBS,
                         "kernel": "ucb_sweep_level<float,K,false>",
                         "algorithmic_bytes_per_unknown": B32,
                         "groups_per_lane": r32["groups_per_lane"]},
        }

    if not args.no_converge:
        c64, phi64 = converge(p, G, 64, st, ss, qq, args.tol, args.max_iters)
        out["time_to_converge_fp64"] = c64
        if not args.no_fp32:
            c32, phi32 = converge(p, G, 32, st, ss, qq, args.tol, args.max_iters)
            rel = np.abs(phi32 - phi64) / np.maximum(np.abs(phi64), 1e-300)
            c32["max_rel_phi_error_vs_fp64"] = float(rel.max())
            c32["speedup_vs_fp64"] = c64["seconds"] / c32["seconds"]
            out["time_to_converge_fp32"] = c32

    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(p, G, st, S, args.cpu_seconds)
    else:
        out["cpu_baseline"] = None

    print(json.dumps(out))


if __name__ == "__main__":
    main()
