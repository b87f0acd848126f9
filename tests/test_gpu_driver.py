"""GPU tests of the command-line driver ecsg_run: the reference's command
line and RSF files in, rec_out / energy_trace / res_trace / movies out, equal
to what the library computes for the same model, and the reference's own
composite flow against its golden traces."""
import os
import subprocess
import sys

import numpy as np
import pytest

import ecsg_amd
from ecsg_amd.model import reference_flow
from ecsg_amd.rsf import driver_path, read_rsf, write_flow_inputs
from helpers import composite, rel

pytestmark = pytest.mark.gpu


def _run(args, cwd, timeout=300):
    exe = driver_path()
    if not os.path.exists(exe):
        pytest.fail("ecsg_run has not been built (python -m ecsg_amd.build)")
    print("ecsg_run", " ".join(a for a in args if "/" not in a), flush=True)
    out = subprocess.run([exe] + list(args), capture_output=True, text=True, cwd=cwd, timeout=timeout)
    sys.stdout.write(out.stdout[-2000:])
    assert out.returncode == 0, out.stderr[-4000:]
    return out


def _interior(a):
    return a[1:-1, 1:-1, 1:-1]


def _model():
    m = composite((3,), coarse=(6, 5, 4), nz_fine=(5,), iso=True, nsteps=30, seed=41)
    m.rec_loc = np.array([[0, 2, 3, 1], [1, 2, 1, 0], [1, 5, 4, 3], [0, 9, 8, 5]], np.int32)
    return m


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_driver_equals_library(tmp_path, precision):
    """Graph-replayed steps between movie frames: rec_out, energy_trace and
    every vx/vz movie frame equal the library's own run of the same model."""
    m = _model()
    args = write_flow_inputs(m, str(tmp_path))
    _run(args + ["check_energy=y", "movie_vx_flag=y", "movie_vz_flag=y", "precision=" + precision], str(tmp_path))
    s = ecsg_amd.Simulation(m, precision=precision, check_energy=True)
    frames = {0: [], 2: []}
    for _ in range(m.nsteps - 1):  # one frame after each of the first nsteps-1 steps
        s.step(1, graph=False)
        s.sync()
        for c in frames:
            frames[c].append([_interior(s.get(g, c)).astype(np.float32) for g in range(m.num_grids)])
    s.step(1, graph=False)
    s.sync()
    kv, tr = read_rsf(str(tmp_path / "rec_out.rsf"))
    assert int(kv["n1"]) == m.nsteps and int(kv["n2"]) == len(m.rec_loc)
    assert np.array_equal(tr, s.traces().astype(np.float32))
    assert np.abs(tr).max() > 0
    kv, e = read_rsf(str(tmp_path / "energy_trace.rsf"))
    assert np.array_equal(e.ravel(), s.energy().astype(np.float32))
    for c, name in ((0, "vx"), (2, "vz")):
        for g in range(m.num_grids):
            kv, mv = read_rsf(str(tmp_path / ("movie_%s_%d.rsf" % (name, g + 1))))
            assert int(kv["n4"]) == m.nsteps - 1
            want = np.stack([f[g] for f in frames[c]])
            assert mv.shape == want.shape
            assert np.array_equal(mv, want), (name, g)
    s.close()


def test_driver_verbose_phase_by_phase(tmp_path):
    """verb=y runs the loop phase by phase (main.cpp:165-261) with the
    residual trace: traces equal the graph run, the flux residuals vanish
    to rounding and the AvgAbs columns are those of the coarse fields."""
    m = _model()
    args = write_flow_inputs(m, str(tmp_path))
    _run(args + ["check_energy=y", "verb=y"], str(tmp_path))
    s = ecsg_amd.Simulation(m, precision="f64", check_energy=True)
    s.step(m.nsteps)
    s.sync()
    kv, tr = read_rsf(str(tmp_path / "rec_out.rsf"))
    assert rel(tr.astype(np.float64), s.traces()) <= 1e-6  # float32 output
    kv, e = read_rsf(str(tmp_path / "energy_trace.rsf"))
    assert np.allclose(e.ravel(), s.energy(), rtol=1e-6)
    kv, res = read_rsf(str(tmp_path / "res_trace.rsf"))
    assert (int(kv["n1"]), int(kv["n2"]), int(kv["n3"])) == (m.nsteps, 9, 1)
    res = res.reshape(9, m.nsteps).astype(np.float64)
    for j in (0, 3, 6):  # res, then AvgAbs of the coarse velocity and stress
        scale = (res[j + 1] * res[j + 2]).max() * m.grids[1].nodes
        assert scale > 0
        assert np.abs(res[j]).max() <= 1e-9 * scale, j
    s.close()


def test_avg_abs_matches_numpy():
    m = _model()
    s = ecsg_amd.Simulation(m, precision="f64", check_energy=False)
    rng = np.random.default_rng(5)
    for g in range(m.num_grids):
        for c in range(9):
            len_ = s.get(g, c).shape
            a = rng.standard_normal(len_)
            s.set(g, c, a)
            want = np.abs(_interior(a)).mean()
            assert abs(s.avg_abs(g, c) - want) <= 1e-13 * want
    s.close()


@pytest.mark.timeout(600)
def test_driver_reference_flow_golden(tmp_path):
    """The reference's composite flow (tests/test1-cmp-eci-ii/eci, par1.py:
    2 grids at ratio 2, 1172 steps) with its own command line
    (eci/SConstruct:177-192: check_energy, verb; the vx movies, 1.7 GB of
    frames, are left off here and covered on a small model above): rec_out
    against the flow's golden ECI traces."""
    here = os.path.dirname(os.path.abspath(__file__))
    gold = np.load(os.path.join(here, "golden", "reference_traces.npz"))
    m = reference_flow(1)
    args = write_flow_inputs(m, str(tmp_path))
    _run(args + ["num_threads=16", "check_energy=y", "verb=y"], str(tmp_path), timeout=600)
    kv, tr = read_rsf(str(tmp_path / "rec_out.rsf"))
    err = rel(tr.astype(np.float64), gold["rec_eci_test1"])
    print("driver golden test1: relL2 vs ECI %.3e" % err)
    assert err <= 3e-7
    kv, e = read_rsf(str(tmp_path / "energy_trace.rsf"))
    e = e.ravel()
    assert abs(e[-1] - e[701]) / e[701] <= 1e-6
    kv, res = read_rsf(str(tmp_path / "res_trace.rsf"))
    res = res.reshape(9, m.nsteps).astype(np.float64)
    for j in (0, 3, 6):
        scale = (res[j + 1] * res[j + 2]).max() * m.grids[1].nodes
        assert np.abs(res[j]).max() <= 1e-9 * scale, j


@pytest.mark.timeout(600)
def test_driver_reference_flow_ii_golden(tmp_path):
    """The II flow (tests/test1-cmp-eci-ii/ii/SConstruct:173-184: the same
    command line with imode=0) through the driver: rec_out against the golden
    II traces."""
    here = os.path.dirname(os.path.abspath(__file__))
    gold = np.load(os.path.join(here, "golden", "reference_traces.npz"))
    m = reference_flow(1, eci=False)
    args = write_flow_inputs(m, str(tmp_path))
    assert "imode=0" in args
    _run(args + ["num_threads=16", "check_energy=y"], str(tmp_path), timeout=600)
    kv, tr = read_rsf(str(tmp_path / "rec_out.rsf"))
    err = rel(tr.astype(np.float64), gold["rec_ii_test1"])
    print("driver golden II test1: relL2 vs II %.3e" % err)
    assert err <= 3e-7
