#!/usr/bin/env python3
"""Sweep throughput and time to convergence of the MI355X UCB xyz sweep.

Headline workload (one step = one full sweep of all angles and groups): the
unit cube as 32x32x32 tiles of UMT's 24-zone tile (786,432 hex zones,
6,291,456 corners), 4x4 product quadrature per octant (128 angles), 32
groups, sigma_t ~ log-uniform[1e-2, 1e3] per (zone, group), c ~ U[0.9,
0.999], q ~ U[0, 1], vacuum boundaries, seed 20261207: the largest
single-GPU workload (FP32 psi of every angle is 107 GB). Unknowns per sweep
= corners x angles x groups = 25,769,803,776 (driver/test_driver.cc:1895).

Prints one JSON line. ``value`` is FP64 sweep unknowns/s with every input
resident in HBM. Also reported for the same workload: the FP32 sweep (FP64
phi), the level kernel against the HBM roofline (byte model of this
design's dataflow, and the measured HBM traffic from committed rocprofv3
summaries), the wall time of source iteration to residual < 1e-8 (plain
FP64 run to convergence against FP32-then-FP64 with grey-transport
acceleration, with and without the learned correction), and the oracle's
CPU sweep timed on this host's cores. Secondary sections: the 20x20x20-tile
single-domain workload (sweeps, time to convergence of every method,
per-angle drop-in with host arrays), the learned correction's own
4x4x4-tile workload, 64 batched uneven domains (FP32 and FP64), and the 2D
cylindrical (RZ) sweep on a 512x512 quad mesh.
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

# name: (tiles, npolar x nazimuthal per octant, groups, c range, seed)
WORKLOADS = {
    "large": dict(tiles=(32, 32, 32), quad=(4, 4), groups=32, c=(0.9, 0.999), seed=20261207),
    "single": dict(tiles=(20, 20, 20), quad=(4, 4), groups=16, c=(0.0, 0.95), seed=20261204),
}
GRAPHS = [True]  # replay captured sweeps (off for profiler runs)
FP32_GREY = {}   # grey controls of the FP32 phase of mixed solves (empty: solve_mixed's)


def fp32_grey_max_its():
    import umt_amd as U
    return FP32_GREY.get("max_its", U.sweep.FP32_PHASE_GREY_MAX_ITS)


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


# ---------------------------------------------------------------------------
# Byte models


def design_bytes_per_unknown(precision, n_in, G, d):
    """Bytes per unknown the level kernel of this design must move between
    itself and memory: the psi write, the n_in upwind psi reads of interior
    faces, and per (level, zone) visit the zone's source S (G x 8 corners),
    sigma_t (G) and corner geometry (A_fp, A_ez, Volume, cFP: 164 B per
    corner in FP64, 88 B in FP32), shared by the d directions that meet the
    zone in the same launch; plus the (zone, angle) item (8 B per 8 corners
    x G groups). Re-reads inside a launch are not counted; psi written by an
    earlier launch and read from L2 or the MALL still counts."""
    r = 8 if precision == 64 else 4
    geo = 164.0 if precision == 64 else 88.0
    return r + r * n_in + (r + r / 8.0 + geo / G) / d + 1.0 / G


def reference_bytes_per_unknown(precision, n_in, G):
    """The reference dataflow's bytes per unknown (one angle at a time, no
    reuse across angles, phi read-modify-write per angle; SweepUCBxyz.F90),
    reported for comparison only: this design moves far less."""
    if precision == 64:
        return 8 + 8 + 8 * n_in + 16 + 1 + 185.5 / G
    return 4 + 4 + 4 * n_in + 16 + 0.5 + 109.5 / G


def load_profile(name):
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def level_kernel_name(precision, layout, G):
    """The level kernel a sweep layout runs, as rocprofv3 names it."""
    t = "double" if precision == 64 else "float"
    if layout == -1:  # zone layout: one lane per group, W lanes per zone
        W = 64 if G % 64 == 0 else (32 if G % 32 == 0 else 16)
        return f"umt::ucb_sweep_level_zone<{t}, {W}>"
    return f"umt::ucb_sweep_level<{t}, {layout}, false>"


def measured_traffic(precision, tiles, G, A, apb, layout):
    """HBM bytes per launch of the level kernel from the committed
    rocprofv3 --pmc summary for this workload, angle blocking and kernel,
    if any."""
    t = "x".join(str(v) for v in tiles)
    prof = load_profile(f"ucb_sweep_level_fp{precision}_{t}_traffic.json")
    if not prof:
        return None
    w = prof.get("workload_key", {})
    if (tuple(w.get("tiles", ())) != tuple(tiles) or w.get("groups") != G or w.get("angles") != A
            or w.get("angles_per_block") != apb
            or prof.get("kernel") != level_kernel_name(precision, layout, G)):
        return None
    return prof


# ---------------------------------------------------------------------------
# Host


def host_info():
    info = {"nproc": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0)),
            "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS")}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k in ("Model name", "Socket(s)", "Core(s) per socket", "Thread(s) per core", "CPU(s)",
                     "NUMA node(s)"):
                info["lscpu " + k] = v
    except (OSError, subprocess.SubprocessError):
        pass
    if "lscpu Model name" not in info:
        try:
            with open("/proc/cpuinfo") as f:
                for line in f:
                    if line.startswith("model name"):
                        info["cpuinfo model name"] = line.split(":", 1)[1].strip()
                        break
        except OSError:
            pass
    return info


def cpu_threads():
    """Every CPU this process may run on, or OMP_NUM_THREADS when the
    machine sets it lower (the GPU box sets it to this GPU's CPU share)."""
    n = len(os.sched_getaffinity(0))
    env = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return max(1, min(n, env) if env > 0 else n)


def cpu_baseline(p, G, st, S, seconds, host):
    """The oracle's C sweep (OpenMP over angles, one angle per thread as
    SetSweep runs one set per thread) on as many angles as threads,
    repeated until about `seconds` have passed (at least once)."""
    from oracle import oracle as O

    threads = cpu_threads()
    # Each thread holds its angle's Psi1 and its own phi (oracle.c): keep the
    # host memory of the sample under ~120 GB.
    per_thread = ((p.ncorner + p.nbelem) * G + p.ncorner * G) * 8
    mem_cap = max(1, int(120e9 // per_thread))
    k = max(1, min(threads, mem_cap, p.nangles))
    threads = k
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
    model = host.get("lscpu Model name") or host.get("cpuinfo model name", "unknown")
    return {
        "value": unknowns / t_total,
        "unit": "unknowns/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"oracle C sweep (liboracle_fast.so, -O3 -march=x86-64-v3, OpenMP over angles, "
                   f"one angle per thread) of {k} of the {p.nangles} angles x {G} groups x "
                   f"{p.ncorner} corners of the same workload, {reps} repetition(s) in "
                   f"{t_total:.1f} s on {threads} threads of '{model}' (nproc {host.get('nproc')}, "
                   f"{host.get('affinity_cpus')} CPUs in this process's affinity, "
                   f"OMP_NUM_THREADS={host.get('OMP_NUM_THREADS')}, memory cap {mem_cap} threads)"),
        "host": host,
    }


# ---------------------------------------------------------------------------
# Sweeps


def psi_bytes(p, G, precision, n_angles):
    return (p.ncorner + p.nbelem) * G * n_angles * (8 if precision == 64 else 4)


def context_extra_bytes(p, G, precision, accel="none"):
    """Device memory of a context besides its S_N psi: fields, GTA vectors
    and the S2 psi (8 directions); with GTA kernel 2 its transfer matrices
    (8 FP64 values per corner and group), with the learned correction its
    per-(zone, group) buffers (8 FP64 values and 16 FP32 features)."""
    rs = 8 if precision == 64 else 4
    extra = 16 * p.ncorner * G * 8 + (p.ncorner + p.nbelem) * G * 8 * rs + 2 * 10**9
    if "gta2" in accel:
        extra += 8 * p.ncorner * G * 8
    if "learned" in accel:
        extra += p.nzones * G * (8 * 8 + 16 * 4)
    return extra


def angles_per_block(p, G, precision, budget):
    """All angles in one block if their psi fits in `budget` bytes, else the
    widest block (A/2, A/4, ...) that does."""
    n = p.nangles
    while n > 1 and psi_bytes(p, G, precision, n) > budget:
        n = (n + 1) // 2
    return 0 if n == p.nangles else n


def measure_sweep(p, G, precision, st, S, warmup, steps, apb):
    import umt_amd as U
    from umt_amd import _lib

    lib = _lib.load()
    ctx = U.SweepContext(p, G, precision=precision)
    if apb:
        ctx.set_angles_per_block(apb)
    ctx.set_graphs(GRAPHS[0])
    ctx.set_sigt(st)
    ctx.set_source(S)
    best_k, tune_ms = ctx.autotune(2)
    ctx.sweep_n(warmup)
    lib.gpu_devicesynchronize()
    t0 = time.perf_counter()
    ctx.sweep_n(steps)
    lib.gpu_devicesynchronize()
    dt = time.perf_counter() - t0
    # HIP events on the context's stream around every level launch.
    ms_sweep_ev, ms_levels, launches = ctx.time_sweeps(1, 3)
    out = dict(dt=dt, ms_levels=ms_levels, launches=launches, n_in=ctx.incoming_faces(),
               ms_sweep_events=ms_sweep_ev, groups_per_lane=best_k,
               autotune_ms_per_sweep={str(k): v for k, v in tune_ms.items()},
               level_stats=ctx.level_stats())
    ctx.close()
    return out


def roofline_block(precision, r, unknowns, G, A, tiles, apb):
    ls = r["level_stats"]
    d = ls["items"] / ls["zone_visits"]
    B = design_bytes_per_unknown(precision, r["n_in"], G, d)
    launches = r["launches"]
    mean_launch_s = r["ms_levels"] * 1e-3 / launches
    per_launch = unknowns / launches
    achieved = B * per_launch / mean_launch_s / 1e9
    prof = measured_traffic(precision, tiles, G, A, apb, r["groups_per_lane"])
    traffic = prof["hbm_bytes_per_launch"] if prof else None
    rs = 8 if precision == 64 else 4
    # Whole sweep: the level kernels plus the phi reduction's psi re-read and
    # phi write, over the sweep's HIP-event time.
    sweep_B = B + rs + 8.0 / A
    sweep_ms = r["ms_sweep_events"]
    blk = {
        "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
        "kernel": level_kernel_name(precision, r["groups_per_lane"], G),
        "byte_model": ("psi write + n_in upwind psi reads + (S + sigma_t + corner geometry) once per "
                       "(level, zone) visit shared by d directions + (zone, angle) items"),
        "algorithmic_bytes_per_unknown": B, "n_in": r["n_in"], "directions_per_zone_visit": d,
        "unknowns_per_launch": per_launch, "launches_per_sweep": launches,
        "mean_launch_us": mean_launch_s * 1e6, "level_kernels_ms_per_sweep": r["ms_levels"],
        "groups_per_lane": r["groups_per_lane"], "angles_per_block": apb or A,
        "autotune_ms_per_sweep": r["autotune_ms_per_sweep"],
        "reference_dataflow_bytes_per_unknown": reference_bytes_per_unknown(precision, r["n_in"], G),
        "sweep": {"bytes_per_unknown": sweep_B, "ms_per_sweep_events": sweep_ms,
                  "frac": sweep_B * unknowns / (sweep_ms * 1e-3) / 1e9 / HBM_PEAK_GBS},
    }
    if prof:
        blk["measured_hbm_GBps"] = traffic / mean_launch_s / 1e9
        blk["measured_hbm_frac"] = blk["measured_hbm_GBps"] / HBM_PEAK_GBS
        blk["measured_bytes_per_unknown"] = traffic / per_launch
        blk["traffic_profile"] = prof.get("file", "")
    return blk


def sweep_section(p, G, st, S, args, tiles, label):
    """FP64 and FP32 sweep throughput with their roofline blocks."""
    import umt_amd as U

    unknowns = p.unknowns(G)
    free, total = U.sweep.device_memory()
    out = {}
    for prec in ((64,) if args.no_fp32 else (64, 32)):
        budget = free - context_extra_bytes(p, G, prec) - 4 * 10**9
        apb = angles_per_block(p, G, prec, budget)
        r = measure_sweep(p, G, prec, st, S, args.warmup, args.steps, apb)
        v = unknowns * args.steps / r["dt"]
        key = "fp64" if prec == 64 else "fp32"
        out[key] = {"value": v, "unit": "unknowns/s", "ms_per_step": r["dt"] * 1e3 / args.steps,
                    "dt": r["dt"],
                    "roofline": roofline_block(prec, r, unknowns, G, p.nangles, tiles, apb)}
        rf = out[key]["roofline"]
        log(f"{label} FP{prec} sweep: {v:.3e} unknowns/s, {r['dt'] * 1e3 / args.steps:.2f} ms/sweep, "
            f"level kernels {r['ms_levels']:.2f} ms ({rf['launches_per_sweep']} launches, "
            f"{rf['mean_launch_us']:.1f} us mean), design bytes {rf['algorithmic_bytes_per_unknown']:.2f} "
            f"B/unknown -> frac {rf['frac']:.3f}, measured frac {rf.get('measured_hbm_frac')}")
    out["device_memory_bytes"] = {"free_at_start": free, "total": total}
    return out


# ---------------------------------------------------------------------------
# Time to convergence


def solve(p, G, xs, accel, mixed, tol, max_iters, apb, net=None, warm_iters=2):
    """Source iteration to `tol` from phi = 0: FP64 sweeps (mixed False) or
    FP32 sweeps to 1e-6 then FP64 sweeps (mixed True), with acceleration
    accel ("none", "gta", "learned" or "gta+learned"), graph replay on.
    Every context first runs a warm-up solve of warm_iters iterations (which
    captures its graphs); the timed solve then starts again from phi = 0."""
    import umt_amd as U

    st, ss, qq = xs

    def context(precision):
        ctx = U.SweepContext(p, G, precision=precision)
        if apb.get(precision):
            ctx.set_angles_per_block(apb[precision])
        ctx.set_sigt(st)
        ctx.set_scattering(ss, qq)
        k, _ = ctx.autotune(1)
        if accel.startswith("gta"):
            k2, ms2 = ctx.autotune_grey(2)
            log(f"FP{precision} {accel}: S_N layout {k}, S2 layout {k2} "
                f"(ms per S2 sweep {', '.join(f'{kk}: {t:.2f}' for kk, t in ms2.items())})")
        if precision == 32 and accel.startswith("gta") and "rel_tol" in FP32_GREY:
            ctx.set_grey_controls(rel_tol=FP32_GREY["rel_tol"])
        ctx.set_graphs(GRAPHS[0])
        if accel.endswith("learned"):
            ctx.set_learner(net, every=2)
        ctx.set_acceleration(accel)
        return ctx

    exact = context(64)
    fast = context(32) if mixed else None
    zero = np.zeros((p.ncorner, G))
    for c in (exact, fast):
        if c is not None:
            c.set_phi(zero)
            c.source_iteration(tol=tol, max_iters=warm_iters)
            c.set_phi(zero)
            if hasattr(c, "reset_stats"):
                c.reset_stats()
    g0 = exact.gta_stats() if accel.startswith("gta") else (0, 0)
    t0 = time.perf_counter()
    if mixed:
        ok, it32, it64, res, hist = U.solve_mixed(fast, exact, tol=tol, switch_tol=1e-6,
                                                  max_iters=max_iters,
                                                  fast_grey_max_its=fp32_grey_max_its())
    else:
        ok, it64, res, hist = exact.source_iteration(tol=tol, max_iters=max_iters)
        it32 = 0
    dt = time.perf_counter() - t0
    out = dict(converged=ok, iterations_fp32=it32, iterations_fp64=it64, residual=res, seconds=dt)
    if accel.startswith("gta"):
        gi, gs = exact.gta_stats()
        out["grey_iterations_fp64"] = gi - g0[0]
        out["grey_sweeps_fp64"] = gs - g0[1]
        if fast is not None:
            out["grey_sweeps_fp32"] = fast.gta_stats()[1]
    phi = exact.phi()
    for c in (exact, fast):
        if c is not None:
            c.close()
    return out, phi


def rel_error(phi, ref):
    return float(np.max(np.abs(phi - ref) / np.maximum(np.abs(ref), 1e-300)))


ALL_METHODS = ["fp64_plain", "fp64_gta", "fp64_gta2", "fp32_then_fp64_gta", "fp32_then_fp64_gta2",
               "fp64_learned", "fp32_then_fp64_learned", "fp64_gta_learned",
               "fp32_then_fp64_gta_learned", "fp32_then_fp64_gta2_learned"]
# The large workload's default: plain FP64 (run to convergence), and
# FP32-then-FP64 with GTA kernel 2 with and without the learned correction.
LARGE_METHODS = ["fp64_plain", "fp32_then_fp64_gta2", "fp32_then_fp64_gta2_learned"]
# The 20^3-tile workload's default.
SINGLE_METHODS = ["fp64_plain", "fp64_gta2", "fp32_then_fp64_gta", "fp32_then_fp64_gta2",
                  "fp32_then_fp64_learned", "fp32_then_fp64_gta2_learned"]
# name: (acceleration, mixed, network). gta is GTA kernel 1 (grey and
# stretched-TSA sweeps), gta2 kernel 2 (transfer matrices and zone LU).
METHOD = {
    "fp64_plain": ("none", False, None), "fp64_gta": ("gta", False, None),
    "fp64_gta2": ("gta2", False, None),
    "fp32_then_fp64_gta": ("gta", True, None), "fp32_then_fp64_gta2": ("gta2", True, None),
    "fp64_learned": ("learned", False, "plain"),
    "fp32_then_fp64_learned": ("learned", True, "plain"),
    "fp64_gta_learned": ("gta+learned", False, "gta"),
    "fp32_then_fp64_gta_learned": ("gta+learned", True, "gta"),
    "fp32_then_fp64_gta2_learned": ("gta2+learned", True, "gta2"),
}


def time_to_converge(p, G, xs, tol, max_iters, methods, learn_tiles, c_range):
    """Plain FP64 source iteration (the reference method on the GPU), run to
    convergence, against grey-transport acceleration and the learned
    correction, FP64 and FP32->FP64. Errors are against the converged plain
    FP64 answer."""
    import umt_amd as U

    free, _ = U.sweep.device_memory()

    def blocks(accel):
        """Angles per block of the plain and the mixed solves with `accel`.
        FP64 psi of the mixed solves shares the device with the FP32 context:
        FP32 psi of every angle when that leaves room for half the angles'
        FP64 psi."""
        x32, x64 = context_extra_bytes(p, G, 32, accel), context_extra_bytes(p, G, 64, accel)
        budget32 = free - x32 - x64 - 8 * 10**9
        apb32 = angles_per_block(p, G, 32, budget32 - psi_bytes(p, G, 64, (p.nangles + 1) // 2))
        left = budget32 - psi_bytes(p, G, 32, apb32 or p.nangles)
        return ({64: angles_per_block(p, G, 64, free - x64 - 4 * 10**9)},
                {32: apb32, 64: angles_per_block(p, G, 64, left)})

    res = {"angles_per_block": {}, "timing": "each solve timed after a 2-iteration warm-up solve "
                                             "of the same contexts (graph capture), from phi = 0"}
    res["fp32_phase_grey_max_its"] = fp32_grey_max_its()
    if "rel_tol" in FP32_GREY:
        res["fp32_phase_grey_rel_tol"] = FP32_GREY["rel_tol"]

    def note_blocks(name, accel, mixed):
        plain, mix = blocks(accel)
        res["angles_per_block"][name] = (
            {"fp32": mix[32] or p.nangles, "fp64": mix[64] or p.nangles} if mixed
            else {"fp64": plain[64] or p.nangles})
        return mix if mixed else plain

    base, phi64 = None, None
    if "fp64_plain" in methods:
        base, phi64 = solve(p, G, xs, "none", False, tol, max_iters,
                            note_blocks("fp64_plain", "none", False))
        res["fp64_plain"] = base
        log(f"fp64_plain: {base}")
    nets = {}
    need = {METHOD[m][2] for m in methods if METHOD[m][2]}
    if learn_tiles:
        small = U.TiledProblem(learn_tiles, learn_tiles, learn_tiles, p.npolar, p.nazimuthal)
        for key in ("plain", "gta", "gta2"):
            if key not in need:
                continue
            on = "none" if key == "plain" else key
            t0 = time.perf_counter()
            net, info = U.train_on_iterates(small, G, seeds=[1, 2, 3], sigt=(1e-2, 1e3),
                                            c=c_range, q=(0.0, 1.0), rounds=3, batch=8192,
                                            steps=1500, collect_stride=4 if on == "none" else 1,
                                            max_iters=max_iters, base=on, graphs=GRAPHS[0])
            res[f"learned_training_{key}"] = {
                "seconds": time.perf_counter() - t0,
                "problem": f"{learn_tiles}^3 tiles, {p.nangles} angles, {G} groups, seeds 1-3 "
                           f"(cross sections drawn like the workload's), iterates of "
                           f"{'plain' if on == 'none' else on + '-accelerated'} source iteration",
                "losses": info["losses"], "samples": info["samples"]}
            log(f"training on {key} iterates: {res[f'learned_training_{key}']}")
            nets[key] = net
    for name in methods:
        if name == "fp64_plain":
            continue
        accel, mixed, netkey = METHOD[name]
        if netkey and netkey not in nets:
            continue
        r, phi = solve(p, G, xs, accel, mixed, tol, max_iters, note_blocks(name, accel, mixed),
                       net=nets.get(netkey))
        if base is not None:
            r["speedup_vs_fp64_plain"] = base["seconds"] / r["seconds"]
            if base["converged"]:
                r["max_rel_phi_error_vs_fp64_plain"] = rel_error(phi, phi64)
        res[name] = r
        log(f"{name}: {r}")
    # What the network gains over the same method without it.
    for a, b in (("fp32_then_fp64_gta_learned", "fp32_then_fp64_gta"),
                 ("fp32_then_fp64_gta2_learned", "fp32_then_fp64_gta2"),
                 ("fp64_gta_learned", "fp64_gta")):
        if a in res and b in res:
            res[f"learned_gain_{b}"] = res[b]["seconds"] / res[a]["seconds"]
    for net in nets.values():
        net.close()
    return res


def learned_config4(tol):
    """The learned correction on its own workload: 4x4x4 tiles, 2x2 product
    quadrature (32 angles), 2 groups, sigma_t logU[0.1, 10], c U[0.95, 0.999],
    q = 1, seed 20261206; trained on seeds 1-4 of the same distribution."""
    import umt_amd as U

    cfg = dict(sigt=(0.1, 10.0), c=(0.95, 0.999), q=(1.0, 1.0))
    p = U.TiledProblem(4, 4, 4, 2, 2)
    G = 2
    t0 = time.perf_counter()
    net, info = U.train_on_iterates(p, G, seeds=[1, 2, 3, 4], rounds=4, epochs=40,
                                    graphs=GRAPHS[0], **cfg)
    train_s = time.perf_counter() - t0
    xs = U.draw_cross_sections(20261206, p.nzones, G, **cfg)
    out = {"workload": "4x4x4 tiles (1536 zones), 32 angles, 2 groups, sigma_t logU[0.1,10], "
                       "c U[0.95,0.999], q=1, seed 20261206",
           "training_seconds": train_s, "training_losses": info["losses"]}
    base, phi64 = solve(p, G, xs, "none", False, tol, 2000, {})
    out["fp64_plain"] = base
    for name, accel, mixed in (("fp64_learned", "learned", False),
                               ("fp32_then_fp64_learned", "learned", True),
                               ("fp64_gta", "gta", False)):
        r, phi = solve(p, G, xs, accel, mixed, tol, 2000, {}, net=net)
        r["max_rel_phi_error_vs_fp64_plain"] = rel_error(phi, phi64)
        r["speedup_vs_fp64_plain"] = base["seconds"] / r["seconds"]
        out[name] = r
    log(f"learned config 4: {out}")
    net.close()
    return out


# ---------------------------------------------------------------------------
# Other sections


def dropin_rate(p, G, st, S, n):
    """The per-angle drop-in gpu_sweepucbxyz with host arrays: every call
    copies its inputs to the device and its outputs back, as the reference
    CUDA path does, so the rate is PCIe-bound. Not the benchmark's value."""
    import umt_amd as U

    C, nb = p.ncorner, p.nbelem
    Psi1 = np.zeros((C + nb, G))
    PsiB = np.zeros((nb, G))
    Phi = np.zeros((C, G))
    if U.sweep_ucbxyz(p, 1, S, st, Psi1, PsiB, Phi) != 0:  # first call allocates
        raise RuntimeError("drop-in sweep failed")
    t0 = time.perf_counter()
    for a in range(1, n + 1):
        if U.sweep_ucbxyz(p, a, S, st, Psi1, PsiB, Phi) != 0:
            raise RuntimeError(f"drop-in sweep failed at angle {a}")
    dt = time.perf_counter() - t0
    r = {"unknowns_per_s": C * G * n / dt, "ms_per_angle": dt * 1e3 / n, "angles_timed": n,
         "note": "host arrays in and out every call (PCIe-inclusive); not value"}
    log(f"drop-in (PCIe-inclusive): {r['unknowns_per_s']:.3e} unknowns/s, "
        f"{r['ms_per_angle']:.1f} ms per angle")
    return r


def batched_domains(n_domains, seed, precisions=(32, 64)):
    """Uneven domains in one batch: n_domains tiled meshes of n^3 tiles,
    n ~ U{4..16}, quadrature 2x2, 3x3 or 4x4 per octant (32, 72, 128 angles),
    G ~ U{8..64}, per-domain seeds, cross sections drawn like the single
    domain's. The first precision: the batched sweep (autotuned groups per
    lane and angles per block) against the same domains swept one context at
    a time; the others: the batched sweep with the first one's groups per
    lane and blocks of 32 angles. Each is timed with one launch per level
    and with one persistent launch per angle block."""
    import umt_amd as U

    rng = np.random.default_rng(seed)
    spec = [(int(rng.integers(4, 17)), int(rng.choice([2, 3, 4])), int(rng.integers(8, 65)))
            for _ in range(n_domains)]
    t0 = time.perf_counter()
    probs = []
    for i, (n, a, _) in enumerate(spec):
        probs.append(U.TiledProblem(n, n, n, a, a))
        if i % 16 == 15:
            log(f"batch: built {i + 1} of {n_domains} problems")
    build_s = time.perf_counter() - t0
    groups = [g for _, _, g in spec]
    xs = [U.draw_cross_sections(seed + 1 + i, p.nzones, g, (1e-2, 1e3), (0.0, 0.95), (0.0, 1.0))
          for i, (p, g) in enumerate(zip(probs, groups))]
    out = {
        "workload": (f"{n_domains} domains, n^3 tiles with n ~ U{{4..16}}, 32/72/128 angles, "
                     f"G ~ U{{8..64}}, sigma_t logU[1e-2,1e3], c U[0,0.95], q U[0,1], seed {seed}"),
        "domains": n_domains, "zones_total": int(sum(p.nzones for p in probs)),
        "problem_build_s": build_s,
    }
    K32 = None
    for prec in precisions:
        seq_ms, seq_launches = 0.0, 0
        # One context at a time is timed for the first precision only.
        for i, (p, g, (st, ss, qq)) in enumerate(zip(probs, groups, xs) if K32 is None else []):
            ctx = U.SweepContext(p, g, precision=prec)
            ctx.set_sigt(st)
            ctx.set_source(np.repeat(qq, 8, axis=0) / (4 * np.pi))
            k = 4 if g % 4 == 0 else (2 if g % 2 == 0 else 1)
            ctx.set_groups_per_lane(k)
            ctx.set_graphs(GRAPHS[0])
            ctx.sweep_n(1)
            t1 = time.perf_counter()
            ctx.sweep_n(3)
            seq_ms += (time.perf_counter() - t1) * 1e3 / 3
            seq_launches += ctx.levels
            ctx.close()
        b = U.SweepBatch(probs, groups, precision=prec)
        b.set_graphs(GRAPHS[0])
        for i, (st, ss, qq) in enumerate(xs):
            b.set_cross_sections(i, st, ss, qq)
        if K32 is None:
            b.set_layout(4, 32)
            K, n, ms_tuned = b.autotune(reps=2)
            K32 = K
        else:
            # The layout the autotuner chose for the first precision, with
            # angle blocks of 32 (the FP64 psi of every angle does not fit).
            K, n = K32, 32
            b.set_layout(K, n)
        # One launch per level against one persistent launch per angle block
        # (grid barrier between levels); "batched" is the faster.
        modes = {}
        for persistent in (False, True):
            b.set_persistent(persistent)
            ms_mode = b.time_iterations(1, 5)
            dm = b.dims()
            modes[persistent] = {"ms_per_iteration": ms_mode,
                                 "unknowns_per_s": dm["unknowns"] / (ms_mode * 1e-3),
                                 "launches_per_sweep": dm["launches"]}
        modes[True]["workgroups"] = b.persistent()[1]
        best = min(modes, key=lambda m: modes[m]["ms_per_iteration"])
        b.set_persistent(best)
        ms = modes[best]["ms_per_iteration"]
        d = b.dims()
        key = "fp64" if prec == 64 else "fp32"
        out["unknowns_per_sweep"] = d["unknowns"]
        out[key] = {
            "batched": {"ms_per_iteration": ms, "unknowns_per_s": d["unknowns"] / (ms * 1e-3),
                        "launches_per_sweep": d["launches"], "groups_per_lane": K,
                        "angles_per_block": n, "angle_blocks": d["blocks"], "persistent": best},
            "one_launch_per_level": modes[False],
            "persistent_launch_per_angle_block": modes[True],
        }
        if seq_ms > 0:
            out[key]["one_context_at_a_time"] = {"ms_per_sweep_sum": seq_ms,
                                                 "unknowns_per_s": d["unknowns"] / (seq_ms * 1e-3),
                                                 "level_launches_sum": seq_launches}
            out[key]["speedup_batched"] = seq_ms / ms
        log(f"batched domains FP{prec}: {out[key]}")
        b.close()
    return out


def rz_bytes_per_unknown(precision, n_in, G, start_frac):
    """Bytes per unknown of the RZ level kernel: source read, sigma_t
    shared by 4 corners, psi write, n_in upwind psi reads, half-angle flux
    read (not for starting directions) and write, and the per-corner
    geometry (A_fp, A_ez, RadiusFP, RadiusEZ, Area, Volume in the sweep
    precision, cFP as two int32) shared by the G groups."""
    r = 8 if precision == 64 else 4
    geo = 15 * r + 8
    return r + r / 4 + r + r * n_in + r * (1 - start_frac) + r + geo / G


def rz_sweep(nr, nz, npolar, nazimuthal, G, steps, warmup, cpu_seconds, tiles=False):
    """The 2D cylindrical sweep: FP64 and FP32 throughput, the level kernel
    against the HBM roofline, and the RZ oracle timed on this host's cores."""
    import umt_amd as U

    p = U.RzProblem(nr, nz, npolar, nazimuthal, r0=0.0, r1=1.0, z0=0.0, z1=1.0)
    q, g, m = p.quadrature(), p.geometry(), p.mesh()
    swept = q["finishing"] == 0
    unknowns = p.ncorner * int(swept.sum()) * G
    start_frac = float(q["starting"].sum()) / float(swept.sum())
    inner = m["cFP"] <= p.ncorner
    cnt = 0
    for a in np.where(swept)[0]:
        d = np.einsum("cfk,k->cf", g["A_fp"], q["omega"][a])
        cnt += int(np.count_nonzero((d < 0) & inner))
    n_in = cnt / (p.ncorner * int(swept.sum()))
    rng = np.random.default_rng(20261204)
    sig = np.exp(rng.uniform(np.log(1e-2), np.log(1e3), (p.nzones, G)))
    S = rng.uniform(0.0, 1.0, (p.ncorner, G)) / (2 * np.pi)
    out = {"config": {"workload": (f"UCB RZ sweep, {nr}x{nz} quad zones on [0,1]x[0,1] "
                                   f"({p.nzones} zones, {p.ncorner} corners), RZ product quadrature "
                                   f"{npolar}x{nazimuthal} ({p.nangles} directions in {p.nlevels} xi "
                                   f"levels, {int(swept.sum())} swept), {G} groups, sigma_t "
                                   f"logU[1e-2,1e3], vacuum"),
                      "unknowns_per_sweep": unknowns, "level_launches": p.sweep_levels},
           "unit": "unknowns/s (corners x swept directions x groups per second)"}
    for prec in (64, 32):
        ctx = U.RzContext(p, G, precision=prec)
        ctx.set_graphs(GRAPHS[0])
        if tiles:
            ctx.set_tiles(True)
        ctx.set_sigt(sig)
        ctx.set_source(S)
        K, ms_k = ctx.autotune(2)
        for _ in range(warmup):
            ctx.sweep()
        t0 = time.perf_counter()
        for _ in range(steps):
            ctx.sweep()
        dt = time.perf_counter() - t0
        ms_ev, ms_lv = ctx.time_sweeps(1, 3)
        ctx.close()
        B = rz_bytes_per_unknown(prec, n_in, G, start_frac)
        achieved = B * unknowns / (ms_lv * 1e-3) / 1e9
        key = "fp64" if prec == 64 else "fp32"
        out[key] = {"value": unknowns * steps / dt, "ms_per_step": dt * 1e3 / steps,
                    "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                                 "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                                 "kernel": f"rz_sweep_level<{'double' if prec == 64 else 'float'},K>",
                                 "level_kernels_ms_per_sweep": ms_lv,
                                 "algorithmic_bytes_per_unknown": B, "n_in": n_in,
                                 "groups_per_lane": K,
                                 "autotune_ms_per_sweep": {str(k): v for k, v in ms_k.items()}}}
        log(f"RZ FP{prec} sweep: {unknowns * steps / dt:.3e} unknowns/s, level kernels "
            f"{ms_lv:.2f} ms ({achieved:.0f} GB/s), groups per lane {K} (autotune {ms_k})")
    if cpu_seconds > 0:
        from oracle import rz as R

        Gc = 4
        threads = max(1, min(cpu_threads(), p.nlevels))
        s = p.schedule()
        Sc, sc = np.ascontiguousarray(S[:, :Gc]), np.ascontiguousarray(sig[:, :Gc])
        R.lib(fast=True)
        reps, t_total = 0, 0.0
        while t_total < cpu_seconds or reps == 0:
            t0 = time.perf_counter()
            R.sweep_all(m, g, q, s, Sc, sc, nthreads=threads, fast=True)
            t_total += time.perf_counter() - t0
            reps += 1
        out["cpu_baseline"] = {
            "value": p.ncorner * int(swept.sum()) * Gc * reps / t_total, "unit": "unknowns/s",
            "cores": threads, "kind": "port",
            "sample": (f"oracle C RZ sweep (liboracle_fast.so, OpenMP over xi levels, at most one "
                       f"thread per level) of the same mesh and directions with {Gc} groups, {reps} "
                       f"sweeps in {t_total:.1f} s on {threads} threads")}
    return out


# ---------------------------------------------------------------------------


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="large",
                    help="headline workload: large (32^3 tiles, 32 groups) or single (20^3, 16)")
    ap.add_argument("--no-secondary", action="store_true",
                    help="skip the 20^3-tile section when the headline is the large workload")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baselines")
    ap.add_argument("--no-converge", action="store_true", help="skip time-to-convergence")
    ap.add_argument("--no-fp32", action="store_true", help="skip the FP32 sweeps")
    ap.add_argument("--methods", default="",
                    help="comma-separated time-to-convergence methods of the headline workload "
                         f"(default: {', '.join(LARGE_METHODS)} for large; "
                         f"{', '.join(SINGLE_METHODS)} for single)")
    ap.add_argument("--dropin-angles", type=int, default=4,
                    help="angles timed through the per-angle drop-in with host arrays (0: skip)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--seed", type=int, default=None,
                    help="cross-section seed of the headline workload (default: the workload's own)")
    ap.add_argument("--fp32-grey-reltol", type=float, default=None,
                    help="grey tolerance of the FP32 phase of mixed solves (default: the reference's 1e-6)")
    ap.add_argument("--fp32-grey-maxits", type=int, default=None,
                    help="grey iteration maximum of the FP32 phase of mixed solves (default: solve_mixed's, 9)")
    ap.add_argument("--max-iters", type=int, default=3000)
    ap.add_argument("--batch-domains", type=int, default=64,
                    help="uneven domains in the batched-sweep section (0: skip it)")
    ap.add_argument("--learn-tiles", type=int, default=4,
                    help="train the learned correction on n^3-tile problems (0: skip it)")
    ap.add_argument("--no-graphs", action="store_true",
                    help="launch every sweep eagerly instead of replaying captured graphs")
    ap.add_argument("--rz", type=int, nargs=2, default=[512, 512],
                    help="RZ sweep section: nr nz quad zones (0 0: skip it)")
    ap.add_argument("--only-rz", action="store_true",
                    help="run only the RZ sweep section and print its JSON object")
    ap.add_argument("--rz-tiles", action="store_true",
                    help="RZ sweeps with the tile levels (fewer, larger launches) instead of zone levels")
    ap.add_argument("--calibrate", action="store_true",
                    help="only run the HBM copy probe (for rocprofv3 --pmc calibration)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} is not supported: this build measures one GPU "
              "(one MI355X); multi-GPU sharding of the sweep is design-only.", file=sys.stderr)
        sys.exit(2)

    import umt_amd as U
    from umt_amd import _lib

    _lib.load()
    GRAPHS[0] = not args.no_graphs
    if args.fp32_grey_reltol is not None:
        FP32_GREY["rel_tol"] = args.fp32_grey_reltol
    if args.fp32_grey_maxits is not None:
        FP32_GREY["max_its"] = args.fp32_grey_maxits
    if args.only_rz:
        print(json.dumps({"rz_sweep": rz_sweep(args.rz[0], args.rz[1], 4, 4, 16, args.steps,
                                               args.warmup, 0 if args.no_cpu else 8.0,
                                               args.rz_tiles)}))
        return
    if args.calibrate:
        n = 1 << 27
        ms = U.sweep.hbm_copy_probe(n, 5)
        print(json.dumps({"copy_probe_bytes_read": n * 8, "copy_probe_bytes_written": n * 8,
                          "ms_per_copy": ms, "GBps": 2 * n * 8 / ms / 1e6}))
        return

    sections_s = {}
    wl = WORKLOADS[args.workload]
    tiles, (npol, nazi), G, c_range, seed = wl["tiles"], wl["quad"], wl["groups"], wl["c"], wl["seed"]
    if args.seed is not None:
        seed = args.seed
    t0 = time.perf_counter()
    p = U.TiledProblem(*tiles, npol, nazi)
    st, ss, qq = U.draw_cross_sections(seed, p.nzones, G, (1e-2, 1e3), c_range, (0.0, 1.0))
    S = np.repeat(qq, 8, axis=0) / (4.0 * np.pi)
    setup_s = time.perf_counter() - t0
    unknowns = p.unknowns(G)
    log(f"headline problem built in {setup_s:.1f} s: {p.nzones} zones, {p.nangles} angles, {G} groups")

    t0 = time.perf_counter()
    sw = sweep_section(p, G, st, S, args, tiles, args.workload)
    sections_s["sweeps"] = time.perf_counter() - t0
    r64 = sw["fp64"]
    tstr = "x".join(str(v) for v in tiles)
    out = {
        "metric": "sweep unknowns/s (zone-corner x angle x group); wall time to converged phi at 1e-8",
        "value": r64["value"],
        "unit": "unknowns/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r64["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded tiled hex mesh and cross sections)",
        "config": {
            "workload": (f"UCB xyz sweep, {tstr} tiles ({p.nzones} zones, {p.ncorner} corners), "
                         f"{npol}x{nazi} product quadrature per octant ({p.nangles} angles), {G} "
                         f"groups, sigma_t logU[1e-2,1e3], c U[{c_range[0]},{c_range[1]}], q U[0,1], "
                         f"seed {seed}, vacuum"),
            "tiles": list(tiles), "zones": p.nzones, "corners": p.ncorner, "angles": p.nangles,
            "groups": G, "unknowns_per_sweep": unknowns, "parallelism": "1 GPU",
        },
        "roofline": r64["roofline"],
        "setup_s": setup_s,
    }
    if "fp32" in sw:
        out["fp32_sweep"] = sw["fp32"]
    out["device_memory_bytes"] = sw["device_memory_bytes"]

    if not args.no_converge:
        t0 = time.perf_counter()
        if args.methods:
            methods = [m.strip() for m in args.methods.split(",") if m.strip()]
        elif args.workload == "large":
            methods = LARGE_METHODS
        else:
            methods = ALL_METHODS
        out["time_to_converge"] = time_to_converge(p, G, (st, ss, qq), args.tol, args.max_iters,
                                                   methods, args.learn_tiles, c_range)
        sections_s["time_to_converge"] = time.perf_counter() - t0

    if not args.no_cpu:
        t0 = time.perf_counter()
        out["cpu_baseline"] = cpu_baseline(p, G, st, S, args.cpu_seconds, host_info())
        sections_s["cpu_baseline"] = time.perf_counter() - t0
    else:
        out["cpu_baseline"] = None
    del p, st, ss, qq, S

    if args.workload == "large" and not args.no_secondary:
        t0 = time.perf_counter()
        w2 = WORKLOADS["single"]
        p2 = U.TiledProblem(*w2["tiles"], *w2["quad"])
        G2 = w2["groups"]
        st2, ss2, qq2 = U.draw_cross_sections(w2["seed"], p2.nzones, G2, (1e-2, 1e3), w2["c"],
                                              (0.0, 1.0))
        S2 = np.repeat(qq2, 8, axis=0) / (4.0 * np.pi)
        sec = {"workload": (f"UCB xyz sweep, 20x20x20 tiles ({p2.nzones} zones, {p2.ncorner} "
                            f"corners), 4x4 product quadrature per octant ({p2.nangles} angles), "
                            f"{G2} groups, sigma_t logU[1e-2,1e3], c U[0,0.95], q U[0,1], seed "
                            f"{w2['seed']}, vacuum"),
               "unknowns_per_sweep": p2.unknowns(G2)}
        sec.update(sweep_section(p2, G2, st2, S2, args, w2["tiles"], "single"))
        if args.dropin_angles > 0:
            sec["dropin_pcie"] = dropin_rate(p2, G2, st2, S2, args.dropin_angles)
        if not args.no_converge:
            sec["time_to_converge"] = time_to_converge(p2, G2, (st2, ss2, qq2), args.tol,
                                                       args.max_iters, SINGLE_METHODS,
                                                       args.learn_tiles, w2["c"])
        out["single_domain_20x20x20"] = sec
        sections_s["single_domain_20x20x20"] = time.perf_counter() - t0
        del p2

    if not args.no_converge and args.learn_tiles:
        t0 = time.perf_counter()
        out["learned_config4"] = learned_config4(args.tol)
        sections_s["learned_config4"] = time.perf_counter() - t0

    if args.batch_domains:
        t0 = time.perf_counter()
        out["batched_domains"] = batched_domains(args.batch_domains, 20261205,
                                                 (32,) if args.no_fp32 else (32, 64))
        sections_s["batched_domains"] = time.perf_counter() - t0

    if args.rz[0] > 0:
        t0 = time.perf_counter()
        out["rz_sweep"] = rz_sweep(args.rz[0], args.rz[1], 4, 4, 16, args.steps, args.warmup,
                                   0 if args.no_cpu else 8.0)
        sections_s["rz_sweep"] = time.perf_counter() - t0

    out["section_seconds"] = sections_s
    print(json.dumps(out))


if __name__ == "__main__":
    main()
