"""GPU tests of the direct vx-sxz mode (sxz_solver "direct": block
elimination over the coarse rows with inverse Schur complements factored on
the host, one workgroup per interface) and of the Jacobi tolerance setter.

The reference solves the vx-sxz system only by Jacobi sweeps
(update_ghost_sxz.cpp:229-236, 314-350), so the direct mode is pinned by
  * a sparse LU solve (scipy) of the oracle's own matrix and right-hand side
    formed from the same random state: the device solution to 1e-12
    relative (FP64 rounding of a diagonally dominant system);
  * full runs against the FP64 oracle's Jacobi at its 1e-12 tolerance:
    traces and energy to 1e-9 (the Jacobi's distance from the exact solve);
  * the composite energy, conserved with the direct ghosts as with the
    Jacobi ones.
"""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

import ecsg_amd
from helpers import composite, rel, sxz_sparse
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu


def _randomize(o, s, m, seed, scale=1.0):
    rng = np.random.default_rng(seed)
    for g in range(m.num_grids):
        for c in range(9):
            shape, _ = o.dims(g, c)
            a = rng.standard_normal(shape) * scale
            o.set(g, c, a)
            s.set(g, c, a)


def _coarse_ghosts(s, m, k=0):
    """sxz2(I, J, -1) over the interface's interior unknowns, [J-1][I-1]."""
    nx2, ny2 = m.grids[k + 1].n[0], m.grids[k + 1].n[1]
    return s.get(k + 1, 7)[0, 2:ny2 + 1, 2:nx2 + 1]


@pytest.mark.parametrize("ratio,coarse", [(2, (6, 5, 4)), (3, (20, 17, 4)), (3, (40, 33, 3))])
def test_direct_solve_equals_sparse_lu(ratio, coarse):
    m = composite((ratio,), coarse=coarse, nz_fine=(4,), nsteps=4, seed=40 + ratio)
    o = Oracle(m, eci_mode=1)
    s = ecsg_amd.Simulation(m, precision="f64", sxz_solver="direct")
    _randomize(o, s, m, seed=50 + ratio, scale=1e3)
    o.ghost_sxz(0)
    s.ghost_sxz(0)
    s.sync()
    A, B = o.sxz_system(0)
    x = spla.spsolve(sxz_sparse(A).tocsc(), B.ravel()).reshape(B.shape)
    got = _coarse_ghosts(s, m)
    assert np.abs(got - x).max() <= 1e-12 * np.abs(x).max()
    # the fine ghosts follow from the coarse ones as after the Jacobi solve
    for g in (0, 1):
        a, b = s.get(g, 7), o.get(g, 7)
        assert np.abs(a - b).max() <= 1e-9 * np.abs(b).max(), g
    assert s.jacobi_stats(0)[0] == 0
    s.close()
    o.close()


@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("ratios", [(2,), (3,), (2, 3)])
def test_direct_full_run_against_jacobi_oracle(ratios, graph):
    m = composite(ratios, nz_fine=(4,) * len(ratios), nsteps=80, seed=19)
    o = Oracle(m, eci_mode=1)
    o.run(m.nsteps)
    s = ecsg_amd.Simulation(m, precision="f64", sxz_solver="direct")
    s.step(m.nsteps, graph=graph)
    s.sync()
    assert np.abs(o.traces()).max() > 0
    assert rel(s.traces(), o.traces()) <= 1e-9
    assert np.allclose(s.energy(), o.energy(), rtol=1e-9, atol=0)
    for g in range(m.num_grids):
        for c in range(9):
            a, b = s.get(g, c), o.get(g, c)
            assert np.abs(a - b).max() <= 1e-8 * max(np.abs(b).max(), 1e-300), (g, c)
    s.close()
    o.close()


@pytest.mark.parametrize("precision,tol", [("f64", 1e-13), ("f32", 1e-5)])
@pytest.mark.parametrize("ratios", [(3,), (2, 3)])
def test_direct_energy_conserved_with_full_vy_mirror(ratios, precision, tol):
    m = composite(ratios, nz_fine=(4,) * len(ratios), nsteps=300, src_len=60, seed=13)
    s = ecsg_amd.Simulation(m, precision=precision, vy_mirror_all=True, sxz_solver="direct")
    s.step(m.nsteps)
    s.sync()
    e = s.energy()[61:]
    s.close()
    assert (e.max() - e.min()) / abs(e).max() <= tol


def test_direct_with_overlap_equals_serial():
    """The direct solve on the interface stream beside the stencils changes
    no value (51 x 51 unknowns, where the overlap is on by default)."""
    m = composite((3,), coarse=(50, 50, 3), nz_fine=(3,), nsteps=30, seed=12, cfl=0.8)
    out = []
    for ov in (False, True):
        s = ecsg_amd.Simulation(m, precision="f32", check_energy=False, sxz_solver="direct")
        s.set_overlap(ov)
        assert s.overlap_active == ov
        s.step(m.nsteps)
        s.sync()
        out.append(([s.get(g, c) for g in range(2) for c in range(9)], s.traces()))
        s.close()
    for a, b in zip(out[0][0], out[1][0]):
        assert np.array_equal(a, b)
    assert np.array_equal(out[0][1], out[1][1])


def test_jacobi_tolerance_setter():
    """FP32 with the double-precision Jacobi tolerance (1e-12, 1e-5): more
    sweeps than the single-precision default (1e-3, 1e-4), and interface
    ghosts that agree with the direct solve's to the FP32 rounding of the
    wavefield; bad values are refused; batch members must share it."""
    m = composite((3,), coarse=(20, 17, 4), nz_fine=(4,), nsteps=60, seed=23)
    runs = {}
    for name, kw in (("loose", {}), ("tight", {"jacobi_tol": (1e-12, 1e-5)}), ("direct", {"sxz_solver": "direct"})):
        s = ecsg_amd.Simulation(m, precision="f32", check_energy=False, **kw)
        s.step(m.nsteps)
        s.sync()
        runs[name] = (s.traces(), s.jacobi_stats(0)[0])
        s.close()
    assert runs["tight"][1] > runs["loose"][1] > 0
    assert runs["direct"][1] == 0
    assert rel(runs["tight"][0], runs["direct"][0]) <= 1e-5
    s = ecsg_amd.Simulation(m, precision="f32", check_energy=False)
    for bad in ((-1.0, 1e-4), (1e-3, float("nan")), (float("inf"), 1e-4)):
        with pytest.raises(ecsg_amd.EcsgError, match="tolerance"):
            s.set_jacobi_tol(*bad)
    s.close()
    a = ecsg_amd.Simulation(m, precision="f32", check_energy=False, sxz_solver="jacobi_lds")
    b = ecsg_amd.Simulation(m, precision="f32", check_energy=False, sxz_solver="jacobi_lds", jacobi_tol=(1e-12, 1e-5))
    with pytest.raises(ecsg_amd.EcsgError, match="share one Jacobi tolerance"):
        ecsg_amd.Batch([a, b])
    a.close()
    b.close()


def test_direct_mode_limits():
    """Coarse rows above 1024 unknowns are refused at creation."""
    m = composite((2,), coarse=(1026, 2, 2), nz_fine=(2,), nsteps=2, seed=3, iso=True)
    with pytest.raises(ecsg_amd.EcsgError, match="direct vx-sxz"):
        ecsg_amd.Simulation(m, precision="f32", check_energy=False, sxz_solver="direct")
