"""GPU tests of the diagnostics: the summation-by-parts residuals of
test_sbp.cpp (ecsg_sbp_residuals) against the numpy restatement, on random
fields and on stepped states, and through the driver (sbp=y)."""
import os
import subprocess

import numpy as np
import pytest

import ecsg_amd
from ecsg_amd.rsf import driver_path, read_rsf, write_flow_inputs
from helpers import composite, sbp_residuals, single

pytestmark = pytest.mark.gpu


def _fields(s, g):
    return {c: s.get(g, c) for c in range(9)}


@pytest.mark.parametrize("R", [1, 2])
def test_sbp_device_equals_numpy_on_random_fields(R):
    m = single(n=(14, 13, 12), nsteps=5, seed=7, cfl=0.6)
    m.radius = R
    s = ecsg_amd.Simulation(m, precision="f64", check_energy=False)
    rng = np.random.default_rng(8 + R)
    for c in range(9):
        s.set(0, c, rng.standard_normal(s.dims(0, c)))
    got = s.sbp_residuals(0)
    want, mag = sbp_residuals(_fields(s, 0), m.grids[0].n, r=R)
    assert np.all(np.abs(got - want) <= 1e-12 * mag), (got, want)
    assert np.all(np.abs(want) >= 1e-6 * mag)  # random fields: the identities do not hold
    s.close()


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_sbp_vanish_on_stepped_states(precision):
    m = single(n=(20, 18, 16), nsteps=40, seed=9, top_bc=1, src_type=1)
    s = ecsg_amd.Simulation(m, precision=precision, check_energy=False)
    s.step(m.nsteps)
    s.sync()
    r = s.sbp_residuals(0)
    _, mag = sbp_residuals(_fields(s, 0), m.grids[0].n)
    tol = 1e-13 if precision == "f64" else 1e-5
    assert np.all(np.abs(r) <= tol * mag), r / mag
    s.close()
    # the fine grid of a composite model carries the interface flux in z
    mc = composite((3,), nsteps=30, seed=4)
    sc = ecsg_amd.Simulation(mc, precision="f64", check_energy=False)
    sc.step(mc.nsteps)
    sc.sync()
    r = sc.sbp_residuals(0)
    want, mag = sbp_residuals(_fields(sc, 0), mc.grids[0].n)
    assert np.all(np.abs(r - want) <= 1e-12 * mag)
    assert np.all(np.abs(r[[0, 3, 4]]) >= 1e-4 * mag[[0, 3, 4]])
    sc.close()


def test_driver_sbp_trace(tmp_path):
    m = single(n=(14, 13, 12), nsteps=25, seed=11, top_bc=0, src_type=0, iso=True)
    args = write_flow_inputs(m, str(tmp_path)) + ["sbp=y", "sbp_trace=" + str(tmp_path / "sbp.rsf")]
    out = subprocess.run([driver_path()] + args, capture_output=True, text=True, cwd=str(tmp_path), timeout=120)
    assert out.returncode == 0, out.stderr[-3000:]
    assert "SBP_res1..7" in out.stderr
    kv, t = read_rsf(str(tmp_path / "sbp.rsf"))
    assert (int(kv["n1"]), int(kv["n2"])) == (m.nsteps, 7)
    s = ecsg_amd.Simulation(m, precision="f64", check_energy=False)
    s.step(m.nsteps)
    s.sync()
    assert np.allclose(t[:, -1], s.sbp_residuals(0).astype(np.float32), rtol=1e-5, atol=0)
    kv, tr = read_rsf(str(tmp_path / "rec_out.rsf"))
    assert np.array_equal(tr, s.traces().astype(np.float32))
    s.close()
