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
source iteration to residual < 1e-8 (plain FP64 against grey-transport
acceleration and the learned correction, FP64 and FP32 then FP64), the
sweep-level kernel against the HBM roofline, and the oracle's CPU sweep
timed on this host's cores; and the 2D cylindrical (RZ) sweep on a regular
512x512 quad mesh with the RZ product quadrature 4x4 (FP64, FP32, roofline,
CPU oracle).
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
    if APB[precision]:
        ctx.set_angles_per_block(APB[precision])
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


def rz_bytes_per_unknown(precision, n_in, G, start_frac):
    """Algorithmic HBM bytes per unknown of the RZ level kernel: source read,
    sigma_t shared by 4 corners, psi write, n_in upwind psi reads, half-angle
    flux read (not for starting directions) and write, and the per-corner
    geometry (A_fp, A_ez, RadiusFP, RadiusEZ, Area, Volume in the sweep
    precision, cFP as two int32) shared by the G groups."""
    r = 8 if precision == 64 else 4
    geo = 15 * r + 8
    return r + r / 4 + r + r * n_in + r * (1 - start_frac) + r + geo / G


def rz_sweep(nr, nz, npolar, nazimuthal, G, steps, warmup, cpu_seconds, log):
    """The 2D cylindrical sweep: FP64 and FP32 throughput, the level kernel
    against the HBM roofline, and the RZ oracle timed on this host's cores."""
    import umt_amd as U

    p = U.RzProblem(nr, nz, npolar, nazimuthal, r0=0.0, r1=1.0, z0=0.0, z1=1.0)
    q, g, m = p.quadrature(), p.geometry(), p.mesh()
    swept = q["finishing"] == 0
    unknowns = p.ncorner * int(swept.sum()) * G
    start_frac = float(q["starting"].sum()) / float(swept.sum())
    # incoming interior FP faces per (corner, swept direction)
    inner = m["cFP"] <= p.ncorner
    cnt = 0
    for a in np.where(swept)[0]:
        d = np.einsum("cfk,k->cf", g["A_fp"], q["omega"][a])
        cnt += int(np.count_nonzero((d < 0) & inner))
    n_in = cnt / (p.ncorner * int(swept.sum()))
    rng = np.random.default_rng(SEED)
    sig = np.exp(rng.uniform(np.log(1e-2), np.log(1e3), (p.nzones, G)))
    S = rng.uniform(0.0, 1.0, (p.ncorner, G)) / (2 * np.pi)
    out = {"config": {"workload": (f"UCB RZ sweep, {nr}x{nz} quad zones on [0,1]x[0,1] "
                                   f"({p.nzones} zones, "
                                   f"{p.ncorner} corners), RZ product quadrature {npolar}x"
                                   f"{nazimuthal} ({p.nangles} directions in {p.nlevels} xi "
                                   f"levels, {int(swept.sum())} swept), {G} groups, sigma_t "
                                   f"logU[1e-2,1e3], vacuum"),
                      "unknowns_per_sweep": unknowns, "level_launches": p.sweep_levels},
           "unit": "unknowns/s (corners x swept directions x groups per second)"}
    for prec in (64, 32):
        ctx = U.RzContext(p, G, precision=prec)
        ctx.set_graphs(GRAPHS[0])
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
        kname = f"void umt::rz_sweep_level<{'double' if prec == 64 else 'float'}, {K}>(umt::RzLevelArgs<{'double' if prec == 64 else 'float'}>)"
        prof = load_traffic("rz_sweep_level_512x512_traffic.json")
        traffic = None
        if prof and (nr, nz, G) == (512, 512, 16) and kname in prof.get("kernels", {}):
            traffic = prof["kernels"][kname]["hbm_bytes_per_launch"]
        out[key] = {"value": unknowns * steps / dt, "ms_per_step": dt * 1e3 / steps,
                    "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                                 "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                                 "traffic": traffic,
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
        threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
        threads = max(1, min(threads, 16, p.nlevels))
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
            "sample": (f"oracle C RZ sweep (liboracle_fast.so, OpenMP over xi levels) of the "
                       f"same mesh and directions with {Gc} groups, {reps} sweeps in "
                       f"{t_total:.1f} s on {threads} threads of '{cpu_model()}'")}
    return out


APB = {64: 0, 32: 0}  # angles per block of the FP64 / FP32 contexts (0: all)
GRAPHS = [True]       # replay captured sweeps (off for profiler runs)


def solve(p, G, xs, accel, mixed, tol, max_iters, net=None, repeat=True, history=False):
    """Source iteration to `tol` from phi = 0: FP64 sweeps (mixed False) or
    FP32 sweeps to 1e-6 then FP64 sweeps (mixed True), with acceleration
    accel ("none", "gta" or "learned"), graph replay on. The timed solve is
    the second of two identical ones (the first captures the graphs)."""
    import umt_amd as U

    st, ss, qq = xs

    def context(precision):
        ctx = U.SweepContext(p, G, precision=precision)
        ctx.set_sigt(st)
        ctx.set_scattering(ss, qq)
        if APB[precision]:
            ctx.set_angles_per_block(APB[precision])
        ctx.autotune(1)
        if accel in ("gta", "gta+learned"):
            ctx.autotune_grey(2)
        ctx.set_graphs(GRAPHS[0])
        if accel in ("learned", "gta+learned"):
            ctx.set_learner(net, every=2)
        ctx.set_acceleration(accel)
        return ctx

    exact = context(64)
    fast = context(32) if mixed else None
    zero = np.zeros((p.ncorner, G))
    out = None
    for attempt in range(2 if repeat else 1):
        for c in (exact, fast):
            if c is not None:
                c.set_phi(zero)
        t0 = time.perf_counter()
        if mixed:
            ok, it32, it64, res, hist = U.solve_mixed(fast, exact, tol=tol, switch_tol=1e-6,
                                                      max_iters=max_iters)
        else:
            ok, it64, res, hist = exact.source_iteration(tol=tol, max_iters=max_iters)
            it32 = 0
        dt = time.perf_counter() - t0
        out = dict(converged=ok, iterations_fp32=it32, iterations_fp64=it64, residual=res,
                   seconds=dt)
        if history:
            out["history"] = [float(h) for h in hist]
    if accel in ("gta", "gta+learned"):
        gi, gs = exact.gta_stats()
        out["grey_iterations_fp64_total"] = gi
    phi = exact.phi()
    for c in (exact, fast):
        if c is not None:
            c.close()
    return out, phi


def rel_error(phi, ref):
    return float(np.max(np.abs(phi - ref) / np.maximum(np.abs(ref), 1e-300)))


def dropin_rate(p, G, st, S, n, log):
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


def time_to_converge(p, G, xs, tol, max_iters, learn_tiles, log, c_range=(0.0, 0.95),
                     gta_only=False):
    """Plain FP64 source iteration (the reference method on the GPU) against
    GTA and the learned correction, FP64 and FP32->FP64."""
    import umt_amd as U

    res = {}
    base, phi64 = solve(p, G, xs, "none", False, tol, max_iters, repeat=False, history=True)
    hist = base.pop("history")
    if not base["converged"] and len(hist) > 20:
        # Not converged within max_iters: extrapolate the iterations (and the
        # time) from the contraction over the last 20 iterations.
        rate = (hist[-1] / hist[-21]) ** (1.0 / 20)
        if 0.0 < rate < 1.0:
            more = np.log(tol / hist[-1]) / np.log(rate)
            base["estimated_iterations"] = float(len(hist) + more)
            base["estimated_seconds"] = base["seconds"] * base["estimated_iterations"] / len(hist)
            base["contraction_per_iteration"] = float(rate)
    res["fp64_plain"] = base
    log(f"fp64 plain: {base}")
    base_s = base.get("estimated_seconds", base["seconds"])
    cases = [("fp64_gta", "gta", False, None), ("fp32_then_fp64_gta", "gta", True, None)]
    nets = []
    if learn_tiles:
        small = U.TiledProblem(learn_tiles, learn_tiles, learn_tiles, p.npolar, p.nazimuthal)
        for key, on in ((("gta", "gta"),) if gta_only else (("plain", "none"), ("gta", "gta"))):
            t0 = time.perf_counter()
            net, info = U.train_on_iterates(small, G, seeds=[1, 2, 3], sigt=(1e-2, 1e3),
                                            c=c_range, q=(0.0, 1.0), rounds=3, batch=8192,
                                            steps=1500, collect_stride=4 if on == "none" else 1,
                                            max_iters=max_iters, base=on, graphs=GRAPHS[0])
            res[f"learned_training_{key}"] = {
                "seconds": time.perf_counter() - t0,
                "problem": f"{learn_tiles}^3 tiles, {p.nangles} angles, {G} groups, seeds 1-3 "
                           f"(cross sections drawn like the workload's), iterates of "
                           f"{'plain' if on == 'none' else 'GTA-accelerated'} source iteration",
                "losses": info["losses"], "samples": info["samples"]}
            log(f"training on {key} iterates: {res[f'learned_training_{key}']}")
            nets.append(net)
        if gta_only:
            nets.insert(0, None)
        if not gta_only:
            cases += [("fp64_learned", "learned", False, nets[0]),
                      ("fp32_then_fp64_learned", "learned", True, nets[0])]
        cases += [("fp64_gta_learned", "gta+learned", False, nets[1]),
                  ("fp32_then_fp64_gta_learned", "gta+learned", True, nets[1])]
    ref_phi, ref_name = (phi64, "fp64_plain") if base["converged"] else (None, "fp64_gta")
    for name, accel, mixed, net in cases:
        r, phi = solve(p, G, xs, accel, mixed, tol, max_iters, net=net)
        r["speedup_vs_fp64_plain"] = base_s / r["seconds"]
        if ref_phi is None and name == "fp64_gta" and r["converged"]:
            ref_phi = phi  # plain iteration did not converge: FP64 GTA is the reference answer
        if ref_phi is not None and name != ref_name:
            r[f"max_rel_phi_error_vs_{ref_name}"] = rel_error(phi, ref_phi)
        res[name] = r
        log(f"{name}: {r}")
    for net in nets:
        if net is not None:
            net.close()
    return res


def learned_config4(tol, log):
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
    base, phi64 = solve(p, G, xs, "none", False, tol, 2000)
    out["fp64_plain"] = base
    for name, accel, mixed in (("fp64_learned", "learned", False),
                               ("fp32_then_fp64_learned", "learned", True),
                               ("fp64_gta", "gta", False)):
        r, phi = solve(p, G, xs, accel, mixed, tol, 2000, net=net)
        r["max_rel_phi_error_vs_fp64_plain"] = rel_error(phi, phi64)
        r["speedup_vs_fp64_plain"] = base["seconds"] / r["seconds"]
        out[name] = r
    log(f"learned config 4: {out}")
    net.close()
    return out


def batched_domains(n_domains, seed, log):
    """Uneven domains in one batch: n_domains tiled meshes of n^3 tiles,
    n ~ U{4..16}, quadrature 2x2, 3x3 or 4x4 per octant (32, 72, 128 angles),
    G ~ U{8..64}, per-domain seeds, cross sections drawn like the workload's;
    FP32 sweeps. Reports the batched sweep (autotuned groups per lane and
    angles per block) against the same domains swept one context at a time."""
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
    log(f"batch: {n_domains} problems built in {build_s:.1f} s")
    # One context at a time (the same sweeps without batching).
    seq_ms, seq_launches = 0.0, 0
    for i, (p, g, (st, ss, qq)) in enumerate(zip(probs, groups, xs)):
        if i % 16 == 15:
            log(f"batch: swept {i + 1} of {n_domains} domains one at a time")
        ctx = U.SweepContext(p, g, precision=32)
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
    b = U.SweepBatch(probs, groups, precision=32)
    b.set_graphs(GRAPHS[0])
    for i, (st, ss, qq) in enumerate(xs):
        b.set_cross_sections(i, st, ss, qq)
    b.set_layout(4, 32)
    K, n, ms_tuned = b.autotune(reps=2)
    ms = b.time_iterations(1, 5)
    d = b.dims()
    out = {
        "workload": (f"{n_domains} domains, n^3 tiles with n ~ U{{4..16}}, 32/72/128 angles, "
                     f"G ~ U{{8..64}}, sigma_t logU[1e-2,1e3], c U[0,0.95], q U[0,1], FP32 sweeps, "
                     f"seed {seed}"),
        "domains": n_domains, "zones_total": int(sum(p.nzones for p in probs)),
        "unknowns_per_sweep": d["unknowns"],
        "batched": {"ms_per_iteration": ms, "unknowns_per_s": d["unknowns"] / (ms * 1e-3),
                    "launches_per_sweep": d["launches"], "groups_per_lane": K,
                    "angles_per_block": n, "angle_blocks": d["blocks"]},
        "one_context_at_a_time": {"ms_per_sweep_sum": seq_ms,
                                  "unknowns_per_s": d["unknowns"] / (seq_ms * 1e-3),
                                  "level_launches_sum": seq_launches},
        "speedup_batched": seq_ms / ms,
        "problem_build_s": build_s,
    }
    log(f"batched domains: {out}")
    b.close()
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--tiles", type=int, nargs=3, default=[20, 20, 20])
    ap.add_argument("--quad", type=int, nargs=2, default=[4, 4], help="npolar nazimuthal")
    ap.add_argument("--groups", type=int, default=16)
    ap.add_argument("--c", type=float, nargs=2, default=[0.0, 0.95],
                    help="scattering ratio range of the workload")
    ap.add_argument("--seed", type=int, default=SEED)
    ap.add_argument("--angles-per-block", type=int, nargs=2, default=[0, 0],
                    metavar=("FP64", "FP32"), help="sweep angle blocks (0: all angles at once)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-converge", action="store_true", help="skip time-to-convergence")
    ap.add_argument("--no-fp32", action="store_true", help="skip the FP32 sweep")
    ap.add_argument("--dropin-angles", type=int, default=4,
                    help="angles timed through the per-angle drop-in with host arrays (0: skip)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--max-iters", type=int, default=3000)
    ap.add_argument("--batch-domains", type=int, default=64,
                    help="uneven domains in the batched-sweep section (0: skip it)")
    ap.add_argument("--learn-tiles", type=int, default=4,
                    help="train the learned correction on n^3-tile problems (0: skip it)")
    ap.add_argument("--learned-with-gta-only", action="store_true",
                    help="run the learned correction only on top of GTA (for workloads where "
                         "iteration without GTA does not converge within --max-iters)")
    ap.add_argument("--no-graphs", action="store_true",
                    help="launch every sweep eagerly instead of replaying captured graphs")
    ap.add_argument("--rz", type=int, nargs=2, default=[512, 512],
                    help="RZ sweep section: nr nz quad zones (0 0: skip it)")
    ap.add_argument("--only-rz", action="store_true",
                    help="run only the RZ sweep section and print its JSON object")
    ap.add_argument("--calibrate", action="store_true",
                    help="only run the HBM copy probe (for rocprofv3 --pmc calibration)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} is not supported: this build measures one GPU "
              "(one MI355X); multi-GPU sharding of the sweep is design-only.", file=sys.stderr)
        sys.exit(2)

    import umt_amd as U
    from umt_amd import _lib

    def log(msg):
        print(f"[bench] {msg}", file=sys.stderr, flush=True)

    lib = _lib.load()
    if args.only_rz:
        GRAPHS[0] = not args.no_graphs
        print(json.dumps({"rz_sweep": rz_sweep(args.rz[0], args.rz[1], 4, 4, args.groups,
                                               args.steps, args.warmup,
                                               0 if args.no_cpu else 8.0, log)}))
        return

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
    c_range = tuple(args.c)
    APB[64], APB[32] = args.angles_per_block
    GRAPHS[0] = not args.no_graphs
    st, ss, qq = U.draw_cross_sections(args.seed, p.nzones, G, (1e-2, 1e3), c_range, (0.0, 1.0))
    S = np.repeat(qq, 8, axis=0) / (4.0 * np.pi)
    unknowns = p.unknowns(G)

    log(f"problem built in {setup_s:.1f} s: {p.nzones} zones, {p.nangles} angles, {G} groups")
    r64 = measure(p, G, 64, st, S, args.warmup, args.steps, lib)
    value = unknowns * args.steps / r64["dt"]
    log(f"FP64 sweep: {value:.3e} unknowns/s, {r64['dt'] * 1e3 / args.steps:.2f} ms per sweep")
    B64 = bytes_per_unknown(64, r64["n_in"], G)
    ach64 = B64 * unknowns / (r64["ms_levels"] * 1e-3) / 1e9
    # Measured HBM traffic is on file for the default workload only.
    profiled = (nx, ny, nz, G, p.nangles) == (20, 20, 20, 16, 128)
    tr = load_traffic("ucb_sweep_level_fp64_20x20x20_traffic.json") if profiled else None
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
                         f"angles), {G} groups, sigma_t logU[1e-2,1e3], c U[{c_range[0]},{c_range[1]}], "
                         f"q U[0,1], seed {args.seed}, "
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
        log(f"FP32 sweep: {unknowns * args.steps / r32['dt']:.3e} unknowns/s")
        B32 = bytes_per_unknown(32, r32["n_in"], G)
        ach32 = B32 * unknowns / (r32["ms_levels"] * 1e-3) / 1e9
        tr32 = load_traffic("ucb_sweep_level_fp32_20x20x20_traffic.json") if profiled else None
        traffic32 = tr32.get("hbm_bytes_per_launch") if tr32 else None
        ms32 = r32["ms_levels"] / r32["launches"]
        out["fp32_sweep"] = {
            "value": unknowns * args.steps / r32["dt"], "unit": "unknowns/s",
            "ms_per_step": r32["dt"] * 1e3 / args.steps,
            # frac can exceed 1: the algorithmic bytes re-read geometry and S
            # for every angle, and L2/MALL serve most of those re-reads.
            "roofline": {"bound": "hbm", "achieved": ach32, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": ach32 / HBM_PEAK_GBS, "traffic": traffic32,
                         "measured_hbm_frac": (traffic32 / (ms32 * 1e-3) / 1e9 / HBM_PEAK_GBS
                                               if traffic32 else None),
                         "kernel": "ucb_sweep_level<float,K,false>",
                         "mean_launch_ms": ms32,
                         "algorithmic_bytes_per_unknown": B32,
                         "groups_per_lane": r32["groups_per_lane"]},
        }

    if args.dropin_angles > 0:
        out["dropin_pcie"] = dropin_rate(p, G, st, S, args.dropin_angles, log)

    if not args.no_converge:
        out["time_to_converge"] = time_to_converge(
            p, G, (st, ss, qq), args.tol, args.max_iters, args.learn_tiles, log, c_range,
            args.learned_with_gta_only)

        if args.learn_tiles and not args.learned_with_gta_only:
            out["learned_config4"] = learned_config4(args.tol, log)

    if args.batch_domains:
        out["batched_domains"] = batched_domains(args.batch_domains, 20261205, log)

    if args.rz[0] > 0:
        out["rz_sweep"] = rz_sweep(args.rz[0], args.rz[1], 4, 4, G, args.steps, args.warmup,
                                   0 if args.no_cpu else 8.0, log)

    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(p, G, st, S, args.cpu_seconds)
    else:
        out["cpu_baseline"] = None

    print(json.dumps(out))


if __name__ == "__main__":
    main()
