"""The drop-in sweep (gpu_sweepucbxyz) on meshes of general polyhedral
zones, against the oracle's polyhedral build.

Meshes that are not the canonical hex (here hexahedra cut into six pyramids
about their centres: five-corner zones whose apex corners have four faces)
take the polyhedral kernel, which does the reference's arithmetic in its
order with IEEE divisions, so every angle's psi, exiting boundary psi, phi
and lagged cycle psi equal the oracle's serial sweep bit for bit: with and
without the time-dependent tau Psi source, on a mesh whose schedules have
cyclic zones and cycle lists, through both drop-in entry points. A hex mesh
handed over in the polyhedral layout (maxcf = 6) takes the same kernel and
agrees with the hex kernels to rounding."""
import numpy as np
import pytest

import umt_amd as U
from oracle import oracle as O
from polymesh import PolyProblem, hex_pyramid_mesh, widen

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def quad():
    return U.TiledProblem(1, 1, 1, 2, 2).quadrature()


def _inputs(p, G, seed):
    rng = np.random.default_rng(seed)
    st, _, _ = U.draw_cross_sections(seed, p.nzones, G, (0.05, 50.0), (0.0, 0.0), (0.0, 1.0))
    S = rng.uniform(0.0, 1.0, (p.ncorner, G))
    psib = rng.uniform(0.0, 0.5, (p.nangles, p.nbelem, G))
    return np.ascontiguousarray(st), S, psib


def _oracle_angle(p, a, S, st, Psi1, PsiB, Phi, tau=0.0, Psi=None, cycle_psi=None):
    s = p.schedule()
    start = int(np.sum(s["nHyperPlanes"][:a]))
    zip_ = s["zonesInPlane"][start:start + int(s["nHyperPlanes"][a])]
    ncyc = s["numCycles"]
    off, n = int(np.sum(ncyc[:a])), int(ncyc[a])
    q = p.quadrature()
    O.sweep_angle(p.mesh(), p.geometry(), S, st, q["omega"][a], q["weight"][a], s["nextZ"][a],
                  s["nextC"][a], zip_, Psi1, PsiB, Phi, tau=tau, Psi=Psi,
                  cycle_list=s["cycleList"][off:off + n] if n else None,
                  cycle_psi=cycle_psi[off:off + n] if n else None)


def _sweep_all_angles(p, G, seed, tau=0.0, noalloc=False, check=True):
    st, S, psib = _inputs(p, G, seed)
    rng = np.random.default_rng(seed + 1)
    Psi = rng.uniform(0.0, 1.0, (p.nangles, p.ncorner, G)) if tau else None
    s = p.schedule()
    total = max(int(s["numCycles"].sum()), 1)
    cyc_g = np.zeros((total, G))
    cyc_r = np.zeros((total, G))
    Phi_g = np.zeros((p.ncorner, G))
    Phi_r = np.zeros((p.ncorner, G))
    out = []
    for a in range(p.nangles):
        Psi1_g = np.zeros((p.ncorner + p.nbelem, G))
        Psi1_r = np.zeros((p.ncorner + p.nbelem, G))
        PsiB_g, PsiB_r = psib[a].copy(), psib[a].copy()
        Pa = np.ascontiguousarray(Psi[a]) if tau else None
        rc = U.sweep_ucbxyz(p, a + 1, S, st, Psi1_g, PsiB_g, Phi_g, tau=tau, Psi_angle=Pa,
                            cycle_psi=cyc_g, noalloc=noalloc)
        assert rc == 0, rc
        if check:
            _oracle_angle(p, a, S, st, Psi1_r, PsiB_r, Phi_r, tau=tau, Psi=Pa, cycle_psi=cyc_r)
            assert np.array_equal(Psi1_g[:p.ncorner], Psi1_r[:p.ncorner]), a
            assert np.array_equal(PsiB_g, PsiB_r), a
        out.append((Psi1_g[:p.ncorner].copy(), PsiB_g))
    if check:
        assert np.array_equal(Phi_g, Phi_r)
        assert np.array_equal(cyc_g, cyc_r)
    return out, Phi_g


@pytest.mark.parametrize("G,tau", [(1, 0.0), (5, 0.0), (16, 0.0), (4, 2.5)])
def test_pyramid_mesh_matches_the_oracle_bit_for_bit(quad, G, tau):
    p = PolyProblem(hex_pyramid_mesh(3, 3, 2), quad)
    print(f"case pyramid mesh {p.nzones} zones, {p.ncorner} corners, G={G}, tau={tau}", flush=True)
    _, phi = _sweep_all_angles(p, G, 11, tau=tau)
    # Finite and not trivially zero (the UCB scheme does not keep every corner
    # flux positive in optically thick zones, so no sign is asserted).
    assert np.all(np.isfinite(phi)) and phi.max() > 0


def test_cyclic_pyramid_mesh_matches_the_oracle_bit_for_bit(quad):
    p = PolyProblem(hex_pyramid_mesh(4, 4, 3, jitter=0.25, seed=1), quad)
    s = p.schedule()
    assert s["numCycles"].sum() > 0 and (s["nextZ"] < 0).any()
    print(f"case cyclic pyramid mesh: {int(s['numCycles'].sum())} cycle-list corners, "
          f"{int((s['nextZ'] < 0).sum())} cyclic zones", flush=True)
    _sweep_all_angles(p, 8, 12)


def test_noalloc_entry_point_gives_the_same_bits(quad):
    p = PolyProblem(hex_pyramid_mesh(3, 3, 2), quad)
    out_a, phi_a = _sweep_all_angles(p, 8, 13, check=False)
    out_b, phi_b = _sweep_all_angles(p, 8, 13, noalloc=True, check=False)
    assert np.array_equal(phi_a, phi_b)
    for (x1, y1), (x2, y2) in zip(out_a, out_b):
        assert np.array_equal(x1, x2) and np.array_equal(y1, y2)


def test_hex_mesh_in_the_polyhedral_layout(quad):
    tp = U.TiledProblem(2, 2, 1, 2, 2)
    wide = PolyProblem(widen(tp.mesh()), quad)
    hexp = PolyProblem(tp.mesh(), quad)
    G = 16
    print("case hex mesh through the polyhedral kernel", flush=True)
    out_w, phi_w = _sweep_all_angles(wide, G, 14)  # against the oracle, bit for bit
    out_h, phi_h = _sweep_all_angles(hexp, G, 14, check=False)  # the hex kernels
    assert np.abs(phi_w - phi_h).max() <= 1e-13 * np.abs(phi_h).max()
    for (x1, _), (x2, _) in zip(out_w, out_h):
        assert np.abs(x1 - x2).max() <= 1e-13 * np.abs(x2).max()
