"""GPU tests of the boundary additions: parameters handed over as main.cpp
holds them after InitParField (ecsg_set_params_staged), the CFL time step
reduced on the device (ecsg_cfl_dt), and the guards of the cooperative
Jacobi solve (bounded grid-barrier wait, ecsg_set_barrier_timeout; no
overlap beside launches of more than one workgroup per CU)."""
import math

import numpy as np
import pytest

import ecsg_amd
from ecsg_amd.model import STIFF_POS, Grid, Model, cfl_dt, reshape_field
from helpers import composite

pytestmark = pytest.mark.gpu


def _run(m, precision, staged, nsteps, energy):
    s = ecsg_amd.Simulation(m, precision=precision, check_energy=energy, max_steps=nsteps, staged=staged)
    s.step(nsteps)
    s.sync()
    out = ([s.get(g, c) for g in range(m.num_grids) for c in range(19)], s.traces(),
           s.energy() if energy else None)
    s.close()
    return out


@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("iso", [True, False])
def test_staged_params_equal_primary(precision, iso):
    """The staged Fields (InitParField's output, isotropic pointers aliased)
    give the run of the primary-node inputs bit for bit: wavefields, staged
    parameters, traces and energy."""
    m = composite((2, 3), coarse=(4, 3, 4), nz_fine=(4, 4), nsteps=30, seed=21, iso=iso)
    f1, t1, e1 = _run(m, precision, False, 30, True)
    f2, t2, e2 = _run(m, precision, True, 30, True)
    for a, b in zip(f1, f2):
        assert np.array_equal(a, b)
    assert np.array_equal(t1, t2)
    assert np.array_equal(e1, e2)


def _stage_transpose(gs, n, pos):
    """Adjoint of ReshapeField: each staged value / 8 back to its 8 corners."""
    nx, ny, nz = n
    out = np.zeros((nz + 1, ny + 1, nx + 1))
    px, py, pz = pos
    ez, ey, ex = gs.shape
    for dz in (0, pz):
        for dy in (0, py):
            for dx in (0, px):
                out[dz:dz + ez, dy:dy + ey, dx:dx + ex] += gs / 8   # one term of the eight each
    return out


def test_staged_gradient_is_the_staged_chain_rule():
    """Gradients of a grid given staged Fields are w.r.t. those Fields: carried
    back through the staging (ReshapeField's transpose) they equal the
    gradients w.r.t. the primary-node inputs."""
    m = composite((3,), coarse=(4, 4, 3), nz_fine=(4,), nsteps=24, seed=5)
    rng = np.random.default_rng(3)
    obs = rng.standard_normal((len(m.rec_loc), 24))
    grads = []
    for staged in (False, True):
        s = ecsg_amd.Simulation(m, precision="f64", check_energy=False, max_steps=24, staged=staged)
        s.gradient(24, obs)
        grads.append([[s.grad(g, 10 + k) for k in range(9)] for g in range(2)])
        s.close()
    for g in range(2):
        for k in range(9):
            back = _stage_transpose(grads[1][g][k], m.grids[g].n, STIFF_POS[k])
            ref = grads[0][g][k]
            assert np.abs(back - ref).max() <= 1e-12 * np.abs(ref).max(), (g, k)


def test_cfl_dt_on_device():
    """ecsg_cfl_dt against the flows' bound h_min / (vp sqrt 3): the closed
    form on a homogeneous medium, and the host formula bit for bit on
    heterogeneous isotropic and orthorhombic composites."""
    n, h, rho, vp = (6, 5, 4), 12.5, 2100.0, 3300.0
    shape = (n[2] + 1, n[1] + 1, n[0] + 1)
    g = Grid(n, (h, 10.0, 15.0), np.full(shape, 1.0 / rho), [np.full(shape, rho * vp * vp), np.full(shape, 1e9)])
    m = Model(grids=[g], dt=1.0, nsteps=1, iso=True)
    want = 0.9 * 10.0 / (vp * math.sqrt(3))
    assert abs(ecsg_amd.cfl_dt(m, 0.9) - want) <= 2e-16 * want
    for iso in (True, False):
        m = composite((2, 3), coarse=(4, 3, 4), nz_fine=(4, 4), nsteps=4, seed=9, iso=iso)
        assert ecsg_amd.cfl_dt(m, 0.9) == cfl_dt(m.grids, iso, frac=0.9)
        assert ecsg_amd.cfl_dt(m, 0.5) == cfl_dt(m.grids, iso, frac=0.5)


def _coop_model(nsteps):
    # 51 x 51 coarse iterate: above the single-workgroup limit, tiled cooperative solve
    return composite((3,), coarse=(50, 50, 3), nz_fine=(3,), nsteps=nsteps, seed=14, cfl=0.8)


def test_barrier_timeout_is_reported_and_recovered():
    """A grid barrier that times out (forced: limit 0) ends the solve early and
    is reported by the next synchronisation; after ecsg_reset with the default
    limit the handle runs as a fresh one, bit for bit."""
    m = _coop_model(12)
    s = ecsg_amd.Simulation(m, precision="f32", check_energy=False, max_steps=12)
    s.set_barrier_timeout(0)
    s.step(4, graph=False)
    with pytest.raises(ecsg_amd.EcsgError, match="timed out"):
        s.sync()
    s.set_barrier_timeout(10000)
    s.reset()
    s.step(12, graph=False)
    s.sync()
    got = [s.get(g, c) for g in range(2) for c in range(9)], s.traces()
    s.close()
    f = ecsg_amd.Simulation(m, precision="f32", check_energy=False, max_steps=12)
    f.step(12, graph=False)
    f.sync()
    want = [f.get(g, c) for g in range(2) for c in range(9)], f.traces()
    f.close()
    for a, b in zip(got[0], want[0]):
        assert np.array_equal(a, b)
    assert np.array_equal(got[1], want[1])


def _cus():
    from helpers import cu_count
    return cu_count()


def test_no_overlap_beside_a_launch_of_more_tiles_than_cus():
    """An interface whose tiled Jacobi launch has more 32 x 32 tiles than the
    device has compute units is either refused at creation (the kernel fits
    one workgroup per CU: more could not all be resident) or solved with
    nothing beside it: the overlap stays off even when asked for, and the
    steps complete and converge."""
    t = int(math.isqrt(_cus())) + 1                 # t * t tiles > CUs
    nc = 32 * t + 1                                  # coarse cells per side
    m = composite((3,), coarse=(nc, nc, 2), nz_fine=(2,), nsteps=6, seed=15, iso=True, cfl=0.8)
    try:
        s = ecsg_amd.Simulation(m, precision="f32", check_energy=False, max_steps=6)
    except ecsg_amd.EcsgError as e:
        assert "too large for one resident Jacobi launch" in str(e)
        s = None
    if s is None:
        return
    s.set_barrier_timeout(5000)
    s.set_overlap(True)
    assert not s.overlap_active
    s.step(6, graph=False)
    s.sync()
    tot, last, capped = s.jacobi_stats(0)
    assert tot >= 6 and not capped
    assert all(np.isfinite(s.get(g, c)).all() for g in range(2) for c in range(9))
    s.close()


def test_overlap_beside_a_launch_of_few_tiles():
    small = ecsg_amd.Simulation(_coop_model(4), precision="f32", check_energy=False, max_steps=4)
    small.set_overlap(True)
    assert small.overlap_active
    small.close()
