"""GPU parity of the II interpolation (imode 0, update_ghost_{vz,sxz,syz}.cpp
mode 0) against the oracle, and the reference's II golden traces.

FP64: the ghost formulas are evaluated in the reference's operation order, so
every array equals the oracle's bit for bit, ghost by ghost and over whole
runs.  FP32 runs track the FP64 oracle to 1e-4 relative L2 (receiver traces).
"""
import os

import numpy as np
import pytest

import ecsg_amd
from ecsg_amd.model import FIELD_NAMES, reference_flow
from helpers import composite, rel
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def _ii(m):
    m.imode = 0
    return m


@pytest.mark.parametrize("xend", [False, True])
def test_ii_ghosts_exact_fp64(xend):
    m = _ii(composite((2,), coarse=(6, 5, 4), nz_fine=(4,), iso=True, nsteps=4, seed=61))
    o = Oracle(m, eci_mode=0, check_energy=False, ii_xend=xend)
    s = ecsg_amd.Simulation(m, precision="f64", check_energy=False, ii_vz_xend=xend)
    rng = np.random.default_rng(62)
    for g in range(2):
        for c in range(9):
            a = rng.standard_normal(o.dims(g, c)[0])
            o.set(g, c, a)
            s.set(g, c, a)
    for name, fo, fs in (("vz", o.ghost_vz, s.ghost_vz), ("sxz", o.ghost_sxz, s.ghost_sxz), ("syz", o.ghost_syz, s.ghost_syz)):
        fo(0)
        fs(0)
        for g in range(2):
            for c in range(9):
                a, b = s.get(g, c), o.get(g, c)
                assert np.array_equal(a, b), "%s ghosts: grid %d %s (max %g)" % (name, g, FIELD_NAMES[c], np.abs(a - b).max())
    s.close()
    o.close()


@pytest.mark.parametrize("case", ["r2", "r2r2"])
@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("xend", [False, True])
def test_ii_full_run_fp64(case, graph, xend):
    if case == "r2":
        m = _ii(composite((2,), nsteps=60, seed=63))
    else:
        m = _ii(composite((2, 2), coarse=(4, 4, 4), nz_fine=(4, 4), nsteps=60, seed=64))
    o = Oracle(m, eci_mode=0, check_energy=True, ii_xend=xend)
    s = ecsg_amd.Simulation(m, precision="f64", check_energy=True, ii_vz_xend=xend)
    o.run(m.nsteps)
    s.step(m.nsteps, graph=graph)
    s.sync()
    assert np.array_equal(s.traces(), o.traces())
    assert np.allclose(s.energy(), o.energy(), rtol=1e-12, atol=0)
    for g in range(m.num_grids):
        for c in range(9):
            assert np.array_equal(s.get(g, c), o.get(g, c)), "grid %d %s" % (g, FIELD_NAMES[c])
    s.close()
    o.close()


def test_ii_phase_by_phase_equals_steps():
    m = _ii(composite((2,), nsteps=30, seed=65))
    a = ecsg_amd.Simulation(m, precision="f64", check_energy=True)
    b = ecsg_amd.Simulation(m, precision="f64", check_energy=True)
    a.step(m.nsteps)
    for _ in range(m.nsteps):
        for g in range(2):
            b.update_vel(g)
        b.record_receivers()
        b.add_source()
        b.ghost_vz(0)
        b.energy_now()
        for g in range(2):
            b.update_stress(g)
        b.ghost_sxz(0)
        b.ghost_syz(0)
        b.end_step()
    a.sync()
    b.sync()
    assert np.array_equal(a.traces(), b.traces())
    assert np.allclose(a.energy(), b.energy(), rtol=1e-13)
    a.close()
    b.close()


def test_ii_fp32_tracks_fp64_oracle():
    m = _ii(composite((2,), nsteps=120, seed=66))
    o = Oracle(m, eci_mode=0, check_energy=False)
    o.run(m.nsteps)
    s = ecsg_amd.Simulation(m, precision="f32", check_energy=False)
    s.step(m.nsteps)
    s.sync()
    assert rel(s.traces(), o.traces()) <= 1e-4
    s.close()
    o.close()


def test_ii_golden_reference_traces():
    """The reference's II flows (tests/test1-cmp-eci-ii/ii, par1.py and
    par3.py) against their golden traces; the ECI goldens of the same flows
    are far away."""
    gold = np.load(os.path.join(HERE, "golden", "reference_traces.npz"))
    for test in (1, 3):
        m = reference_flow(test, eci=False)
        s = ecsg_amd.Simulation(m, precision="f64", check_energy=False)
        s.step(m.nsteps)
        s.sync()
        tr = s.traces()
        err = rel(tr, gold["rec_ii_test%d" % test])
        err_eci = rel(tr, gold["rec_eci_test%d" % test])
        print("golden II test%d: relL2 vs II %.3e, vs ECI %.3e" % (test, err, err_eci))
        assert err <= 3e-7
        assert err_eci >= 1000 * err
        s.close()


def test_ii_rejected_where_not_defined():
    m3 = _ii(composite((3,), nsteps=5, seed=67))
    with pytest.raises(ecsg_amd.EcsgError):
        ecsg_amd.Simulation(m3, precision="f64")
    m = _ii(composite((2,), nsteps=5, seed=68))
    s = ecsg_amd.Simulation(m, precision="f32", check_energy=False)
    with pytest.raises(ecsg_amd.EcsgError):
        s.gradient(5, np.zeros((len(m.rec_loc), 5)))
    with pytest.raises(ecsg_amd.EcsgError):
        s.set_fused(True)
    with pytest.raises(ecsg_amd.EcsgError):
        ecsg_amd.Batch([s])
    s.close()
