"""The Fortran host (umt_amd/fortran, ISO_C_BINDING over the C-ABI) on the
GPU against the oracle: source iteration through the device-resident
context built from Fortran-held Teton-layout arrays, one angle through
the drop-in gpu_sweepucbxyz called with the reference's argument list, a
hex mesh handed over by the host (mesh ingestion), time-dependent cycles,
a batch of uneven domains, and the RZ context driven from Fortran (source
iteration and time-dependent cycles on a quad mesh the host hands over)."""
import os
import subprocess

import numpy as np
import pytest

import umt_amd as U
from oracle import oracle as O
from oracle import rz as R

from conftest import ROOT, hex_grid, rel_err

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(ROOT, "umt_amd", "fortran", "umt_si_driver")
RZ_DRIVER = os.path.join(ROOT, "umt_amd", "fortran", "umt_rz_driver")


def _run(args, out, tmp_path, driver=DRIVER, extra=None):
    path = str(tmp_path / out)
    cmd = [driver] + [str(a) for a in args] + [path] + ([str(extra)] if extra is not None else [])
    print("run:", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    return np.fromfile(path, dtype=np.float64), r.stdout


def test_fortran_source_iteration_matches_oracle(tmp_path):
    seed, G = 20261203, 2
    phi, out = _run(["si", 4, 4, 4, 2, 2, G, 64, 10, 0.0, seed, 0.1, 10.0, 0.0, 0.9, 1.0, 1.0, 1],
                    "phi_si.bin", tmp_path)
    assert "si iterations=    10" in out
    p = U.TiledProblem(4, 4, 4, 2, 2)
    st, ss, qq = U.draw_cross_sections(seed, p.nzones, G, (0.1, 10.0), (0.0, 0.9), (1.0, 1.0))
    ref, it, _, _ = O.source_iteration(p.mesh(), p.geometry(), p.quadrature(), p.schedule(),
                                       st, ss, qq, 0.0, 10)
    assert rel_err(phi.reshape(p.ncorner, G), ref) < 1e-12


def test_fortran_dropin_angle_matches_oracle(tmp_path):
    seed, G, angle = 7, 4, 11
    phi, _ = _run(["dropin", 2, 2, 2, 2, 2, G, 64, 0, 0.0, seed, 0.01, 100.0, 0.0, 0.9, 0.0, 1.0,
                   angle], "phi_dropin.bin", tmp_path)
    p = U.TiledProblem(2, 2, 2, 2, 2)
    m, g, q, s = p.mesh(), p.geometry(), p.quadrature(), p.schedule()
    st, ss, qq = U.draw_cross_sections(seed, p.nzones, G, (0.01, 100.0), (0.0, 0.9), (0.0, 1.0))
    S = np.repeat(ss + qq, 8, axis=0) / (4 * np.pi)
    a = angle - 1
    starts = np.concatenate([[0], np.cumsum(s["nHyperPlanes"])])
    Psi1 = np.zeros((p.ncorner + p.nbelem, G))
    PsiB = np.zeros((p.nbelem, G))
    ref = np.zeros((p.ncorner, G))
    O.sweep_angle(m, g, S, st, q["omega"][a], q["weight"][a], s["nextZ"][a], s["nextC"][a],
                  s["zonesInPlane"][starts[a]:starts[a + 1]], Psi1, PsiB, ref)
    assert rel_err(phi.reshape(p.ncorner, G), ref) < 1e-12


@pytest.mark.parametrize("mode", ["gta", "gta2"])
def test_fortran_gta_source_iteration_matches_python_host(tmp_path, mode):
    """GTA-accelerated source iteration (kernel 1, or kernel 2 through
    umt_ctx_set_gta_solver) driven from Fortran gives the Python host's
    answer on the same problem, and plain iteration's fixed point."""
    seed, G = 20261205, 2
    args = [mode, 2, 2, 2, 2, 2, G, 64, 200, 1e-10, seed, 1e-2, 1e3, 0.9, 0.999, 0.0, 1.0, 1]
    phi, out = _run(args, f"phi_{mode}.bin", tmp_path)
    assert "gta iterations=" in out
    p = U.TiledProblem(2, 2, 2, 2, 2)
    st, ss, qq = U.draw_cross_sections(seed, p.nzones, G, (1e-2, 1e3), (0.9, 0.999), (0.0, 1.0))
    ctx = U.SweepContext(p, G)
    ctx.set_sigt(st)
    ctx.set_scattering(ss, qq)
    ctx.set_acceleration(mode)
    ok, it, _, _ = ctx.source_iteration(tol=1e-10, max_iters=200)
    assert ok and f"gta iterations={it:6d}" in out
    assert rel_err(phi.reshape(p.ncorner, G), ctx.phi()) < 1e-12
    ref, _, rres, _ = O.source_iteration(p.mesh(), p.geometry(), p.quadrature(), p.schedule(),
                                         st, ss, qq, 1e-13, 20000, nthreads=4)
    assert rres < 1e-13
    assert rel_err(phi.reshape(p.ncorner, G), ref) < 1e-7
    ctx.close()


def _rz_mesh(tmp_path, npolar, naz):
    """A jittered quad mesh written for the Fortran host with counterclockwise
    zones (the library turns them clockwise), and the oracle's view of the
    problem the library builds from it."""
    pts, quads = R.structured(6, 5, r0=0.25, r1=1.5, z0=-1.0, z1=1.0, jitter=0.3, seed=9)
    ccw = np.ascontiguousarray(quads[:, ::-1], dtype=np.int32)
    path = tmp_path / "rz_mesh.bin"
    with open(path, "wb") as f:
        np.array([pts.shape[0], ccw.shape[0]], np.int32).tofile(f)
        np.ascontiguousarray(pts, np.float64).tofile(f)
        ccw.tofile(f)
    p = U.RzProblem.from_quad_mesh(pts, ccw, npolar, naz)
    pts2, quads2 = p.points()
    m = R.quad_mesh(pts2, quads2)
    g = R.geometry(m)
    q = R.quadrature(npolar, naz)
    return str(path), p, (m, g, q, R.schedule(m, g, q))


def test_fortran_rz_source_iteration_matches_oracle(tmp_path):
    seed, G = 20261208, 2
    mesh, p, (m, g, q, s) = _rz_mesh(tmp_path, 2, 3)
    phi, out = _run(["si", mesh, 2, 3, G, 64, 8, 0.0, seed, 0.1, 20.0, 0.0, 0.9, 0.5, 1.0, 0.0, 0],
                    "phi_rz_si.bin", tmp_path, RZ_DRIVER)
    assert "rz si iterations=     8" in out
    st, ss, qq = U.draw_cross_sections(seed, p.nzones, G, (0.1, 20.0), (0.0, 0.9), (0.5, 1.0))
    ref, it, _, _ = R.source_iteration(m, g, q, s, st, ss, qq, 0.0, 8)
    assert rel_err(phi.reshape(p.ncorner, G), ref) < 1e-12


def test_fortran_rz_time_cycles_match_oracle(tmp_path):
    seed, G, tau = 20261209, 2, 2.0
    mesh, p, (m, g, q, s) = _rz_mesh(tmp_path, 2, 2)
    out_all, out = _run(["time", mesh, 2, 2, G, 64, 5, 0.0, seed, 0.1, 20.0, 0.0, 0.9, 0.5, 1.0,
                         tau, 2], "phi_rz_time.bin", tmp_path, RZ_DRIVER)
    assert "rz cycle   2" in out
    A, Cn = p.nangles, p.ncorner
    phi = out_all[:Cn * G].reshape(Cn, G)
    psi = out_all[Cn * G:].reshape(A, Cn, G)
    st, ss, qq = U.draw_cross_sections(seed, p.nzones, G, (0.1, 20.0), (0.0, 0.9), (0.5, 1.0))
    # the host's starting Psi(g, c, a) = (1 + 0.25 mod(a, 3) + 0.1 g) / (2 pi), 1-based a and g
    a1 = np.arange(1, A + 1)[:, None, None]
    g1 = np.arange(1, G + 1)[None, None, :]
    psi_ref = np.broadcast_to((1 + 0.25 * (a1 % 3) + 0.1 * g1) / (2 * np.pi), (A, Cn, G))
    phi_ref = None
    for _ in range(2):
        phi_ref, psi_ref, _, _, _ = R.time_step(m, g, q, s, st, ss, qq, tau, psi_ref, 0.0, 5,
                                                phi0=phi_ref)
    assert rel_err(phi, phi_ref) < 1e-12
    assert np.abs(psi - psi_ref).max() <= 1e-12 * np.abs(psi_ref).max()


def test_fortran_mesh_ingestion_source_iteration_matches_oracle(tmp_path):
    """A distorted hex mesh handed over from Fortran as points and
    connectivity (umt_problem_from_hex_mesh) iterates like the oracle."""
    seed, G = 20261210, 2
    pts, hexes = hex_grid(3, 0.12, 5)
    p = U.TiledProblem.from_hex_mesh(pts, hexes, 2, 2)
    s = p.schedule()
    assert int(np.sum(s.get("numCycles", [0]))) == 0  # the array-built context takes no cycle lists
    mesh = tmp_path / "hex_mesh.bin"
    with open(mesh, "wb") as f:
        np.array([pts.shape[0], hexes.shape[0]], np.int32).tofile(f)
        np.ascontiguousarray(pts, np.float64).tofile(f)
        np.ascontiguousarray(hexes, np.int32).tofile(f)
    phi, out = _run(["mesh", 0, 0, 0, 2, 2, G, 64, 6, 0.0, seed, 0.1, 10.0, 0.0, 0.9, 0.5, 1.0, 1],
                    "phi_mesh.bin", tmp_path, extra=mesh)
    assert "si iterations=     6" in out
    st, ss, qq = U.draw_cross_sections(seed, p.nzones, G, (0.1, 10.0), (0.0, 0.9), (0.5, 1.0))
    ref, it, _, _ = O.source_iteration(p.mesh(), p.geometry(), p.quadrature(), s, st, ss, qq,
                                       0.0, 6)
    assert rel_err(phi.reshape(p.ncorner, G), ref) < 1e-12


def test_fortran_time_cycles_match_oracle(tmp_path):
    """Two time-dependent cycles (tau Psi source, savePsi) from a Psi the
    Fortran host holds, against the oracle's time step."""
    seed, G, tau, cycles = 20261211, 2, 3.0, 2
    phi_all, out = _run(["time", 2, 2, 1, 2, 2, G, 64, 5, 0.0, seed, 0.1, 10.0, 0.0, 0.9, 0.5, 1.0,
                         cycles], "phi_time.bin", tmp_path, extra=tau)
    assert "time cycle   2" in out
    p = U.TiledProblem(2, 2, 1, 2, 2)
    A, Cn = p.nangles, p.ncorner
    phi = phi_all[:Cn * G].reshape(Cn, G)
    psi = phi_all[Cn * G:].reshape(A, Cn, G)
    st, ss, qq = U.draw_cross_sections(seed, p.nzones, G, (0.1, 10.0), (0.0, 0.9), (0.5, 1.0))
    # the host's starting Psi(g, c, a) = (1 + 0.25 mod(a, 3) + 0.1 g) / (4 pi), 1-based a and g
    a1 = np.arange(1, A + 1)[:, None, None]
    g1 = np.arange(1, G + 1)[None, None, :]
    psi_ref = np.broadcast_to((1 + 0.25 * (a1 % 3) + 0.1 * g1) / (4 * np.pi), (A, Cn, G)).copy()
    phi_ref = None
    m, g, q, s = p.mesh(), p.geometry(), p.quadrature(), p.schedule()
    for _ in range(cycles):
        phi_ref, psi_ref, _, _, _ = O.time_step(m, g, q, s, st, ss, qq, tau, psi_ref, 0.0, 5,
                                                phi0=phi_ref)
    assert rel_err(phi, phi_ref) < 1e-12
    assert np.abs(psi - psi_ref).max() <= 1e-12 * np.abs(psi_ref).max()


def test_fortran_batch_matches_oracle(tmp_path):
    """Three uneven domains (different sizes and group counts) iterated in
    one batch from Fortran: each domain's phi is the oracle's."""
    seed, G, nd, iters = 20261212, 2, 3, 5
    phi_all, out = _run(["batch", 2, 2, 1, 2, 2, G, 64, iters, 0.0, seed, 0.1, 10.0, 0.0, 0.9, 0.5,
                         1.0, nd], "phi_batch.bin", tmp_path)
    assert f"iterations=     {iters}" in out
    off = 0
    for d in range(1, nd + 1):
        p = U.TiledProblem(2 + d - 1, 2, 1, 2, 2)
        Gd = G + d - 1
        st, ss, qq = U.draw_cross_sections(seed + d, p.nzones, Gd, (0.1, 10.0), (0.0, 0.9),
                                           (0.5, 1.0))
        ref, _, _, _ = O.source_iteration(p.mesh(), p.geometry(), p.quadrature(), p.schedule(),
                                          st, ss, qq, 0.0, iters)
        phi = phi_all[off:off + p.ncorner * Gd].reshape(p.ncorner, Gd)
        off += p.ncorner * Gd
        assert rel_err(phi, ref) < 1e-12, d
    assert off == phi_all.size


def _teton_psib(A, nb, G):
    """The driver's incident fluxes PsiB(g, b, a) = 0.01 (1 + mod(g + 3 b + 7 a, 11)),
    1-based, as (A, nb, G)."""
    a, b, g = np.meshgrid(np.arange(1, A + 1), np.arange(1, nb + 1), np.arange(1, G + 1),
                          indexing="ij")
    return 0.01 * (1 + (g + 3 * b + 7 * a) % 11)


@pytest.mark.parametrize("mesh_case", ["tiled", "cyclic"])
def test_fortran_teton_entry_point_matches_oracle(tmp_path, mesh_case):
    """SetSweep's angle loop through SweepUCBxyz(Set, setID, Groups, Angle,
    savePsi) of umt_teton_mod (the body teton/SweepUCBxyz.F90 compiles into
    a Teton build): Phi, the exiting boundary fluxes and the saved Psi
    against the oracle's sweep with the same incident fluxes; on a distorted
    mesh whose schedule has cycle lists as well."""
    seed, G = 20261212, 3
    print(f"case Teton entry point, {mesh_case} mesh", flush=True)
    if mesh_case == "tiled":
        p = U.TiledProblem(2, 2, 1, 2, 2)
        args = ["setsweep", 2, 2, 1]
        extra = None
    else:
        pts, hexes = hex_grid(4, 0.3, 1)
        p = U.TiledProblem.from_hex_mesh(pts, hexes, 2, 2)
        assert int(np.sum(p.schedule()["numCycles"])) > 0
        extra = tmp_path / "cyclic_mesh.bin"
        with open(extra, "wb") as f:
            np.array([pts.shape[0], hexes.shape[0]], np.int32).tofile(f)
            np.ascontiguousarray(pts, np.float64).tofile(f)
            np.ascontiguousarray(hexes, np.int32).tofile(f)
        args = ["setsweep", 0, 0, 0]
    out_all, out = _run(args + [2, 2, G, 64, 0, 0.0, seed, 0.1, 10.0, 0.0, 0.9, 0.0, 1.0, 1],
                        "phi_setsweep.bin", tmp_path, extra=extra)
    assert "setsweep angles=" in out
    A, Cn, nb = p.nangles, p.ncorner, p.nbelem
    phi = out_all[:Cn * G].reshape(Cn, G)
    psib = out_all[Cn * G:Cn * G + A * nb * G].reshape(A, nb, G)
    psi = out_all[Cn * G + A * nb * G:].reshape(A, Cn, G)
    st, ss, qq = U.draw_cross_sections(seed, p.nzones, G, (0.1, 10.0), (0.0, 0.9), (0.0, 1.0))
    S = np.repeat(ss + qq, 8, axis=0) / (4 * np.pi)
    s = p.schedule()
    cyc = np.zeros((max(len(s["cycleList"]), 1), G)) if len(s["cycleList"]) else None
    ref, ref_B = O.sweep_all(p.mesh(), p.geometry(), p.quadrature(), s, S, st,
                             PsiB=_teton_psib(A, nb, G), cycle_psi=cyc)
    assert rel_err(phi, ref) < 1e-12
    assert np.max(np.abs(psib - ref_B)) <= 1e-12 * np.abs(ref_B).max()
    w = p.quadrature()["weight"]
    np.testing.assert_allclose(np.tensordot(w, psi, axes=(0, 0)), phi, rtol=1e-12, atol=1e-300)
