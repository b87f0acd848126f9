"""The benchmark workload at its full size -- 1207x1207x81 (h 5 m) +
403x403x101 (15 m) + 135x135x121 (45 m) nodes, ratio 3 at both interfaces,
136.6 M nodes, layered orthorhombic medium, explosive Ricker at 50 m depth,
1200 surface receivers -- checked through properties that do not need the
oracle to run at this size (the oracle pins the same code paths bit for bit
on small models):

* the FP32 run tracks the FP64 run (receiver traces and whole wavefields of
  the two upper grids, after the wave has crossed the first interface) far
  inside the 1e-4 relative-L2 bar;
* with vy mirrored on every plane, the composite energy (FP64) is constant
  once the source has stopped;
* the interface overlap changes no value of the FP32 step;
* the two-grid ratio-3 configuration (4.3 M nodes, 1500 steps) and the
  uniform 81^3 configuration (600 steps, energy every step) against the FP64
  oracle itself.
"""
import numpy as np
import pytest

import ecsg_amd
from ecsg_amd import model as M
from helpers import rel
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu

STEPS = 200


@pytest.fixture(scope="module")
def full_model():
    return M.three_grid_workload(nsteps=300)


def _run(m, precision, nsteps, overlap=None):
    s = ecsg_amd.Simulation(m, precision=precision, check_energy=False, max_steps=nsteps + 2)
    if overlap is not None:
        s.set_overlap(overlap)
    s.step(nsteps)
    s.sync()
    out = (s.traces(), s.get(0, 2), s.get(1, 0), [s.jacobi_stats(k) for k in range(2)])
    s.close()
    return out


def test_full_size_fp32_tracks_fp64(full_model):
    r64 = _run(full_model, "f64", STEPS)
    r32 = _run(full_model, "f32", STEPS)
    assert np.abs(r64[0]).max() > 0
    assert np.linalg.norm(r64[2]) > 0          # the wave is in the 403x403x101 grid
    assert r64[3][0][0] > STEPS                 # Jacobi sweeps on the first interface
    for k, what in ((0, "traces"), (1, "grid 0 vz"), (2, "grid 1 vx")):
        err = rel(r32[k], r64[k])
        assert err <= 1e-5, (what, err)         # measured: 8.5e-7, 1.6e-6, 1.1e-6


def test_full_size_energy_conserved(full_model):
    m = full_model
    src = m.src_trace
    m.src_trace = src[:80].copy()               # the Ricker (peak at step 35) ends at step 80
    try:
        s = ecsg_amd.Simulation(m, precision="f64", check_energy=True, vy_mirror_all=True, max_steps=262)
        s.step(260)
        s.sync()
        e = s.energy()[81:260]
        s.close()
    finally:
        m.src_trace = src
    assert abs(e).max() > 0
    assert (e.max() - e.min()) / abs(e).max() <= 1e-13   # measured: 5.7e-16


def test_full_size_overlap_bitwise(full_model):
    a = _run(full_model, "f32", 60, overlap=False)
    b = _run(full_model, "f32", 60, overlap=True)
    for k in range(3):
        assert np.array_equal(a[k], b[k]), k
    assert a[3] == b[3]


def test_two_grid_config_fp32_against_fp64_oracle():
    """The two-grid ratio-3 configuration at full size (241x241x61 h 5 m over
    81x81x121 h 15 m, smoothed +-5 % orthorhombic perturbation, Ricker 20 Hz
    at 30 m, 1500 steps): FP32 device traces against the FP64 oracle (CPU,
    OpenMP) within the 1e-4 relative-L2 bar, FP64 device traces within the
    1e-6 bar.  The FP64 traces are not bit for bit here: at step 104 the
    first interface's Jacobi solve stops one sweep later on the device (81
    against 80 sweeps) -- its norm and the AvgAbs tolerance are sums taken in
    another order, and this step's norm lies within their rounding of the
    tolerance; the extra sweep moves the ghosts by less than the tolerance."""
    m = M.two_grid_workload()
    o = Oracle(m, prec="f64", eci_mode=1, check_energy=False)
    o.run(m.nsteps)
    ref = o.traces()
    o.close()
    assert np.abs(ref).max() > 0
    s = ecsg_amd.Simulation(m, precision="f32", check_energy=False)
    s.step(m.nsteps)
    s.sync()
    err = rel(s.traces(), ref)
    s.close()
    assert err <= 1e-4, err                     # measured: 1.4e-6
    s = ecsg_amd.Simulation(m, precision="f64", check_energy=False)
    s.step(m.nsteps)
    s.sync()
    err64 = rel(s.traces(), ref)
    s.close()
    print("two-grid 1500 steps: FP32 %.3e, FP64 %.3e relative L2 against the FP64 oracle" % (err, err64))
    assert err64 <= 1e-6, err64                 # measured: 8.7e-16


def test_uniform_config_fp64_against_oracle():
    """The reference's CPU-runnable configuration: one uniform 81x81x81 grid,
    h 10 m, homogeneous orthorhombic medium at isotropic values, explosive
    Ricker 15 Hz at (40, 40, 20), free top, 81 receivers at z = 2, energy
    every step, 600 steps, FP64: traces and final wavefields bit for bit
    against the oracle, energy to 1e-12 (its reductions are summed in another
    order)."""
    m = M.uniform_model()
    o = Oracle(m, prec="f64", check_energy=True)
    o.run(m.nsteps)
    s = ecsg_amd.Simulation(m, precision="f64", check_energy=True)
    s.step(m.nsteps)
    s.sync()
    assert np.abs(o.traces()).max() > 0
    assert np.array_equal(s.traces(), o.traces())
    assert np.allclose(s.energy(), o.energy(), rtol=1e-12, atol=0)
    for c in range(9):
        assert np.array_equal(s.get(0, c), o.get(0, c)), c
    s.close()
    o.close()
