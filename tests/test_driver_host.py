"""The command-line driver ecsg_run on the CPU: RSF input/output, the
reference's command line, location mapping (dry_run=y stops before any
device work) and its error behaviour."""
import os
import re
import subprocess

import numpy as np
import pytest

from ecsg_amd.model import reference_flow
from ecsg_amd.rsf import driver_path, read_rsf, write_flow_inputs, write_rsf
from helpers import composite


@pytest.fixture(scope="module")
def exe():
    p = driver_path()
    if not os.path.exists(p):
        pytest.fail("ecsg_run has not been built (python -m ecsg_amd.build)")
    return p


def _run(exe, args, cwd):
    return subprocess.run([exe] + list(args), capture_output=True, text=True, cwd=cwd, timeout=120)


def _mapped(stdout):
    locs = [tuple(int(v) for v in t) for t in re.findall(r"\(grid#(\d+), ix=(-?\d+),\s+iy=(-?\d+), iz=(-?\d+)\)", stdout)]
    steps = int(re.search(r"== (\d+) time steps", stdout).group(1))
    return locs[0], locs[1:], steps


def test_rsf_round_trip(tmp_path):
    a = np.arange(24, dtype=np.float32).reshape(2, 3, 4)
    p = str(tmp_path / "a.rsf")
    write_rsf(p, a, d1=0.5, label1="X")
    kv, b = read_rsf(p)
    assert kv["n1"] == "4" and kv["n2"] == "3" and kv["n3"] == "2" and kv["label1"] == "X"
    assert np.array_equal(a, b)


def test_dry_run_maps_the_reference_flow(exe, tmp_path):
    """tests/test1-cmp-eci-ii (par1.py, par3.py): the driver reads the flow's
    inputs and maps source and receivers to the same grid points as the
    flow's own mapping (init_src.cpp:33-45, init_del_rec_trace.cpp:31-42)."""
    for test in (1, 3):
        m = reference_flow(test)
        args = write_flow_inputs(m, str(tmp_path / ("t%d" % test)))
        out = _run(exe, args + ["check_energy=y", "num_threads=16", "dry_run=y"], str(tmp_path))
        assert out.returncode == 0, out.stderr
        src, rec, steps = _mapped(out.stdout)
        assert src == (m.src_grid,) + tuple(m.src_loc)
        assert [tuple(r) for r in m.rec_loc.tolist()] == rec
        assert steps == m.nsteps == 1172


def test_dry_run_generic_positions(exe, tmp_path):
    """Positions written from grid indices map back to them, including a
    receiver on a coarse grid's top plane."""
    m = composite((3,), coarse=(6, 5, 4), nz_fine=(5,), iso=True, nsteps=37, seed=3)
    m.rec_loc = np.array([[0, 2, 3, 1], [1, 2, 1, 0], [1, 5, 4, 3], [0, 17, 14, 5]], np.int32)
    out = _run(exe, write_flow_inputs(m, str(tmp_path)) + ["dry_run=y"], str(tmp_path))
    assert out.returncode == 0, out.stderr
    src, rec, steps = _mapped(out.stdout)
    assert src == (m.src_grid,) + tuple(m.src_loc)
    assert rec == [tuple(r) for r in m.rec_loc.tolist()]
    assert steps == 37


@pytest.mark.parametrize("drop,extra,message", [
    ("dt=", [], "dt must be given"),
    ("radius=", [], "radius must be given"),
    ("", ["radius=7"], "radius 7 is not supported"),
    ("movie_nt=", [], "movie_nt must be given"),
    ("", ["check_energy=maybe"], "not a boolean"),
    ("", ["num_grids=3"], "buoy_3"),
])
def test_driver_errors(exe, tmp_path, drop, extra, message):
    m = composite((2,), coarse=(4, 4, 4), nz_fine=(4,), iso=True, nsteps=10, seed=1)
    args = [a for a in write_flow_inputs(m, str(tmp_path)) if not (drop and a.startswith(drop))]
    args = [a for a in args if not any(a.split("=")[0] == e.split("=")[0] for e in extra)] + extra
    out = _run(exe, args + ["dry_run=y"], str(tmp_path))
    assert out.returncode == 1
    assert message in out.stderr, out.stderr


def test_driver_rejects_mismatched_and_misplaced_inputs(exe, tmp_path):
    m = composite((2,), coarse=(4, 4, 4), nz_fine=(4,), iso=True, nsteps=10, seed=2)
    args = write_flow_inputs(m, str(tmp_path))
    # s13 of grid 1 with a wrong size
    bad = str(tmp_path / "bad.rsf")
    write_rsf(bad, np.ones((5, 9, 8), np.float32))
    out = _run(exe, [a if not a.startswith("s13_1=") else "s13_1=" + bad for a in args] + ["dry_run=y"], str(tmp_path))
    assert out.returncode == 1 and "s13_1" in out.stderr
    # a receiver below the last grid
    rec = str(tmp_path / "rec_far.rsf")
    write_rsf(rec, np.array([[0.0, 0.0, 1e6]], np.float32), n1=3, n2=1)
    out = _run(exe, [a if not a.startswith("rec_in=") else "rec_in=" + rec for a in args] + ["dry_run=y"], str(tmp_path))
    assert out.returncode == 1 and "below the last grid" in out.stderr
