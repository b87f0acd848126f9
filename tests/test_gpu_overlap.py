"""GPU tests of the interface overlap (ecsg_set_overlap): the interface ghost
updates on a second stream beside the stencil planes away from the
interfaces must give the serial step's results bit for bit -- every field
entry, halos and ghosts included, and the receiver traces -- in FP64 and
FP32, eagerly and graph-replayed, for two- and three-grid models (grids of
two z cells included, whose every plane but one touches an interface), with
the source on the planes next to an interface, from random states, and with
interfaces too large for the single-workgroup Jacobi solve (cooperative
solve, eager steps).  The gradient's forward steps use the same schedule."""
import numpy as np
import pytest

import ecsg_amd
from helpers import composite

pytestmark = pytest.mark.gpu


def _model(case, nsteps):
    if case == "r3":
        return composite((3,), coarse=(6, 5, 4), nz_fine=(5,), nsteps=nsteps, seed=4)
    if case == "r2r3":
        return composite((2, 3), coarse=(4, 3, 4), nz_fine=(4, 4), nsteps=nsteps, seed=6)
    if case == "thin":                 # middle and top grids of two z cells
        return composite((3, 2), coarse=(4, 4, 3), nz_fine=(2, 2), nsteps=nsteps, seed=11, src_type=1)
    if case == "src_fine_plane":       # explosive source on the fine grid's interface plane
        m = composite((3,), coarse=(6, 5, 4), nz_fine=(5,), nsteps=nsteps, seed=7, src_type=1)
        m.src_loc = (9, 7, 5)
        return m
    if case == "src_coarse_plane":     # explosive source on the coarse grid's first plane
        m = composite((3,), coarse=(6, 5, 4), nz_fine=(5,), nsteps=nsteps, seed=8, src_type=1)
        m.src_grid, m.src_loc = 1, (3, 2, 0)
        return m
    if case == "deep":                 # source and receivers two planes or more from the interface
        m = composite((3,), coarse=(6, 5, 4), nz_fine=(6,), nsteps=nsteps, seed=13, src_type=1)
        m.src_loc = (9, 7, 3)
        m.rec_loc = np.array([[0, 4, 5, 1], [0, 10, 3, 4], [1, 2, 2, 2], [1, 5, 4, 4]], np.int32)
        return m
    if case == "large_iface":          # 101 x 101 coarse interface: cooperative Jacobi, no graph
        return composite((3,), coarse=(100, 100, 3), nz_fine=(3,), nsteps=nsteps, seed=12, cfl=0.8)
    raise ValueError(case)


def _run(m, precision, overlap, graph, nsteps, random_start=False):
    s = ecsg_amd.Simulation(m, precision=precision, check_energy=False, max_steps=nsteps + 2)
    s.set_overlap(overlap)
    assert s.overlap_active == overlap
    if random_start:
        rng = np.random.default_rng(5)
        for g in range(m.num_grids):
            for c in range(9):
                s.set(g, c, rng.standard_normal(s.dims(g, c)) * (1.0 if c < 3 else 1e5))
    s.step(3, graph=graph)
    s.step(nsteps - 3, graph=graph)
    s.sync()
    out = ([s.get(g, c) for g in range(m.num_grids) for c in range(9)], s.traces())
    s.close()
    return out


def _same(got, ref, what):
    assert np.array_equal(got[1], ref[1]), "traces (%s)" % what
    for k, (a, b) in enumerate(zip(got[0], ref[0])):
        assert np.array_equal(a, b), "grid %d field %d (%s)" % (k // 9, k % 9, what)


@pytest.mark.parametrize("case", ["r3", "r2r3", "thin", "src_fine_plane", "src_coarse_plane", "deep"])
@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_overlap_equals_serial_step(case, precision):
    nsteps = 24
    m = _model(case, nsteps)
    ref = _run(m, precision, False, False, nsteps)
    assert np.abs(ref[1]).max() > 0
    for graph in (False, True):
        _same(_run(m, precision, True, graph, nsteps), ref, "graph=%s" % graph)


@pytest.mark.parametrize("case", ["r2r3", "thin"])
def test_overlap_from_random_state(case):
    nsteps = 9
    m = _model(case, nsteps)
    m.src_trace = np.zeros(nsteps + 2, np.float32)
    ref = _run(m, "f64", False, False, nsteps, random_start=True)
    _same(_run(m, "f64", True, False, nsteps, random_start=True), ref, "random start")


@pytest.mark.parametrize("mode", [1, 2])
@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_overlap_large_interface(precision, mode):
    nsteps = 30
    m = _model("large_iface", nsteps)
    if mode == 2:  # receivers off the coarse grid's plane z = 1
        m.rec_loc = m.rec_loc[m.rec_loc[:, 0] == 0]
    assert _mode(m) == mode
    ref = _run(m, precision, False, False, nsteps)
    assert np.abs(ref[1]).max() > 0
    _same(_run(m, precision, True, True, nsteps), ref, "large interface")


def test_overlap_modes():
    """Mode 2 only when the source and every receiver are on stress-interior
    planes (two or more planes from each interface)."""
    assert _mode(_model("deep", 5)) == 2
    assert _mode(_model("src_fine_plane", 5)) == 1
    assert _mode(_model("src_coarse_plane", 5)) == 1
    m = _model("deep", 5)
    m.rec_loc = np.array([[0, 4, 5, 1], [1, 2, 2, 1]], np.int32)  # coarse z = 1 reads the velocity edge plane
    assert _mode(m) == 1
    m.rec_loc = np.array([[0, 4, 5, 5]], np.int32)  # fine z = nz-1
    assert _mode(m) == 1
    assert _mode(_model("large_iface", 5)) == 1


def _mode(m):
    s = ecsg_amd.Simulation(m, precision="f32", check_energy=False)
    s.set_overlap(True)
    k = s.overlap_mode
    s.close()
    return k


def test_overlap_not_active_where_not_applicable():
    m = composite((3,), coarse=(6, 5, 4), nz_fine=(5,), nsteps=5, seed=4)
    s = ecsg_amd.Simulation(m, precision="f64", check_energy=True)
    s.set_overlap(True)
    assert not s.overlap_active  # the energy check runs the serial step
    s.close()
    m1 = composite((), coarse=(8, 7, 6), nz_fine=(), nsteps=5, seed=1)
    s = ecsg_amd.Simulation(m1, precision="f64", check_energy=False)
    s.set_overlap(True)
    assert not s.overlap_active  # one grid: no interfaces
    s.close()


def test_overlap_default_follows_the_solver():
    """Default: on with the tiled multi-workgroup Jacobi solve (vx-sxz
    iterates above 2048 entries), off with the single-workgroup one of
    small interfaces or of sxz_solver "jacobi_lds" (state only, no steps)."""
    small = composite((3,), coarse=(6, 5, 4), nz_fine=(5,), nsteps=2, seed=4)      # 7 x 6 iterate
    s = ecsg_amd.Simulation(small, precision="f32", check_energy=False)
    assert not s.overlap_active
    s.close()
    big = composite((3,), coarse=(50, 50, 3), nz_fine=(3,), nsteps=2, seed=12)     # 51 x 51 iterate
    s = ecsg_amd.Simulation(big, precision="f32", check_energy=False)
    assert s.overlap_active
    s.set_overlap(False)
    assert not s.overlap_active
    s.set_overlap(None)
    assert s.overlap_active
    s.close()
    s = ecsg_amd.Simulation(big, precision="f32", check_energy=False, sxz_solver="jacobi_lds")
    assert not s.overlap_active
    s.close()


def test_gradient_with_overlap_equals_serial():
    nsteps = 16
    m = composite((3,), coarse=(6, 5, 4), nz_fine=(5,), nsteps=nsteps, seed=4)
    res = []
    for overlap in (False, True):
        s = ecsg_amd.Simulation(m, precision="f64", check_energy=False, max_steps=nsteps)
        s.set_overlap(overlap)
        s.set_checkpointing(5, "native")
        obs = np.zeros((len(m.rec_loc), nsteps))
        J = s.gradient(nsteps, obs)
        grads = [s.grad(g, sel) for g in range(m.num_grids) for sel in range(9, 19)]
        res.append((J, grads))
        s.close()
    assert res[0][0] == res[1][0] and res[0][0] > 0
    for a, b in zip(res[0][1], res[1][1]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("overlap", [False, True])
def test_step_timed_equals_step(overlap):
    """ecsg_step_n_timed advances the state as ecsg_step_n and times the
    grid-0 stencil launches."""
    nsteps = 12
    m = _model("r2r3", nsteps)
    outs = []
    for timed in (False, True):
        s = ecsg_amd.Simulation(m, precision="f32", check_energy=False, max_steps=nsteps + 2)
        s.set_overlap(overlap)
        if timed:
            v, st = s.step_timed(nsteps)
            assert v > 0 and st > 0
        else:
            s.step(nsteps)
        s.sync()
        outs.append(([s.get(g, c) for g in range(m.num_grids) for c in range(9)], s.traces()))
        s.close()
    _same(outs[1], outs[0], "timed")


def test_two_handles_interleaved():
    """Two handles whose interfaces use the cooperative Jacobi solve (101 x
    101 coarse interfaces), stepped in interleaved calls with no
    synchronisation between them, each on its own streams (the library
    issues cooperative launches of all handles in one order): every field
    entry and trace equals the same handle stepped alone."""
    nsteps = 16
    ms = [composite((3,), coarse=(100, 100, 3), nz_fine=(3,), nsteps=nsteps, seed=sd, cfl=0.8) for sd in (12, 14)]
    alone = []
    for m in ms:
        s = ecsg_amd.Simulation(m, precision="f32", check_energy=False, max_steps=nsteps + 2)
        assert s.overlap_active
        s.step(nsteps, graph=False)
        s.sync()
        alone.append(([s.get(g, c) for g in range(m.num_grids) for c in range(9)], s.traces()))
        s.close()
    sims = [ecsg_amd.Simulation(m, precision="f32", check_energy=False, max_steps=nsteps + 2) for m in ms]
    for _ in range(nsteps):
        for s in sims:
            s.step(1, graph=False)
    for s, m, ref in zip(sims, ms, alone):
        s.sync()
        _same(([s.get(g, c) for g in range(m.num_grids) for c in range(9)], s.traces()), ref, "interleaved")
        s.close()
