"""The no-alloc drop-in gpu_sweepucbxyz_noalloc against gpu_sweepucbxyz
through the C-ABI: the same arguments give the same bits (the same kernels
run on the same inputs; only where the device buffers live differs), on two
problems in a row, the second smaller than the first (the scratch keeps the
first call's contents, so stale data would show), with and without the
time-dependent tau Psi source, and for every angle. A call whose buffers
exceed the scratch is refused with status 53 and leaves its outputs as they
were."""
import numpy as np
import pytest

import umt_amd as U
from umt_amd import _lib

pytestmark = pytest.mark.gpu


def _run(p, G, seed, noalloc, tau=0.0):
    rng = np.random.default_rng(seed)
    st, ss, qq = U.draw_cross_sections(seed, p.nzones, G, (0.1, 10.0), (0.0, 0.9), (0.0, 1.0))
    S = np.ascontiguousarray(rng.uniform(0.0, 1.0, (p.ncorner, G)) / (4 * np.pi))
    Psi1 = np.zeros((p.ncorner + p.nbelem, G))
    Phi = np.zeros((p.ncorner, G))
    PsiB = np.ascontiguousarray(rng.uniform(0.0, 0.1, (p.nangles, p.nbelem, G)))
    Psi = np.ascontiguousarray(rng.uniform(0.0, 0.1, (p.nangles, p.ncorner, G)))
    for a in range(1, p.nangles + 1):
        rc = U.sweep_ucbxyz(p, a, S, st, Psi1, PsiB[a - 1], Phi, tau=tau,
                            Psi_angle=Psi[a - 1], save_psi=True, noalloc=noalloc)
        assert rc == 0, (a, _lib.load().umt_last_error())
    return Psi1, PsiB, Phi, Psi


@pytest.mark.parametrize("tau", [0.0, 3.5])
def test_noalloc_matches_dropin_twice_second_smaller(tau):
    lib = _lib.load()
    assert lib.gpu_sweepucbxyz_noalloc_max_bytes() == 128 << 20
    for tiles, quad, G, seed in (((3, 3, 2), (2, 2), 6, 11), ((2, 1, 1), (1, 2), 3, 12)):
        p = U.TiledProblem(*tiles, *quad)
        print(f"case noalloc tiles {tiles} quad {quad} G={G} tau={tau}", flush=True)
        ref = _run(p, G, seed, False, tau)
        got = _run(p, G, seed, True, tau)
        for name, r, g in zip(("Psi1", "PsiB", "Phi", "Psi"), ref, got):
            np.testing.assert_array_equal(g, r, err_msg=name)
        assert np.all(np.isfinite(got[2])) and np.max(got[2]) > 0


def test_noalloc_refuses_calls_beyond_its_scratch():
    p = U.TiledProblem(8, 8, 8, 1, 1)
    G = 64  # about 217 MB of device buffers, more than the 128 MiB scratch
    st = np.ones((p.nzones, G))
    S = np.ones((p.ncorner, G))
    Psi1 = np.zeros((p.ncorner + p.nbelem, G))
    PsiB = np.zeros((p.nbelem, G))
    Phi = np.full((p.ncorner, G), 7.0)
    print(f"case noalloc refusal: {p.ncorner} corners, {G} groups", flush=True)
    rc = U.sweep_ucbxyz(p, 1, S, st, Psi1, PsiB, Phi, noalloc=True)
    assert rc == 53, rc
    assert b"scratch" in _lib.load().umt_last_error()
    assert np.all(Phi == 7.0) and np.all(Psi1 == 0.0)
    # the allocating drop-in still takes the same call
    assert U.sweep_ucbxyz(p, 1, S, st, Psi1, PsiB, Phi) == 0
