"""GPU parity of the radius 2..6 single-grid stencils (update_vel.cpp and
update_stress.cpp `ir` loops, their wider ghost mirrors, FD coefficients
init_del_fdcoeff.cpp:38-85) against the oracle.

FP64: every array, halos included, equals the oracle's bit for bit after each
update and over whole runs (energy to 1e-12: fixed-order reductions in a
different grouping).  FP32 tracks the FP64 oracle to 1e-4 (receiver traces).
"""
import numpy as np
import pytest

import ecsg_amd
from ecsg_amd.model import FIELD_NAMES
from helpers import composite, rel, single
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu

RADII = [2, 3, 4, 5, 6]


def _model(R, top_bc=1, src_type=1, nsteps=30, seed=0, n=(14, 13, 12)):
    # the time step shrinks with the stencil's sum of |coefficients| (<= 1.37)
    m = single(n=n, nsteps=nsteps, seed=seed, top_bc=top_bc, src_type=src_type, cfl=0.6)
    m.radius = R
    return m


@pytest.mark.parametrize("top_bc", [0, 1])
@pytest.mark.parametrize("R", RADII)
def test_radius_updates_bitwise_fp64(R, top_bc):
    m = _model(R, top_bc=top_bc, seed=R)
    o = Oracle(m, eci_mode=1, check_energy=False)
    s = ecsg_amd.Simulation(m, precision="f64", check_energy=False)
    rng = np.random.default_rng(100 + R)
    for c in range(9):
        shape, r = o.dims(0, c)
        assert r == R and s.dims(0, c) == shape
        a = rng.standard_normal(shape)
        o.set(0, c, a)
        s.set(0, c, a)
    for phase, fo, fs in (("vel", o.update_vel, s.update_vel), ("stress", o.update_stress, s.update_stress),
                          ("vel2", o.update_vel, s.update_vel)):
        fo(0)
        fs(0)
        for c in range(9):
            a, b = s.get(0, c), o.get(0, c)
            assert np.array_equal(a, b), "%s: %s differs (max %g)" % (phase, FIELD_NAMES[c], np.abs(a - b).max())
    s.close()
    o.close()


@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("R", RADII)
def test_radius_full_run_fp64(R, graph):
    m = _model(R, top_bc=R % 2, src_type=R % 2, nsteps=40, seed=10 + R)
    o = Oracle(m, eci_mode=1, check_energy=True)
    s = ecsg_amd.Simulation(m, precision="f64", check_energy=True)
    o.run(m.nsteps)
    s.step(m.nsteps, graph=graph)
    s.sync()
    assert np.array_equal(s.traces(), o.traces())
    assert np.abs(s.traces()).max() > 0
    assert np.allclose(s.energy(), o.energy(), rtol=1e-12, atol=0)
    for c in range(9):
        assert np.array_equal(s.get(0, c), o.get(0, c)), FIELD_NAMES[c]
    s.close()
    o.close()


@pytest.mark.parametrize("R", [2, 4, 6])
def test_radius_fp32_tracks_fp64_oracle(R):
    m = _model(R, nsteps=80, seed=20 + R, n=(24, 22, 20))
    o = Oracle(m, eci_mode=1, check_energy=False)
    o.run(m.nsteps)
    s = ecsg_amd.Simulation(m, precision="f32", check_energy=False)
    s.step(m.nsteps)
    s.sync()
    assert rel(s.traces(), o.traces()) <= 1e-4
    s.close()
    o.close()


def test_radius_rejected_where_not_defined():
    m = composite((2,), nsteps=5, seed=3)
    m.radius = 2
    with pytest.raises(ecsg_amd.EcsgError):
        ecsg_amd.Simulation(m, precision="f64")
    m1 = _model(3, nsteps=5)
    s = ecsg_amd.Simulation(m1, precision="f32", check_energy=False)
    with pytest.raises(ecsg_amd.EcsgError):
        s.gradient(5, np.zeros((len(m1.rec_loc), 5)))
    with pytest.raises(ecsg_amd.EcsgError):
        s.set_fused(True)
    with pytest.raises(ecsg_amd.EcsgError):
        s.autotune(1)
    with pytest.raises(ecsg_amd.EcsgError):
        ecsg_amd.Batch([s])
    s.close()
