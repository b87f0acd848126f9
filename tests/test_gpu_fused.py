"""GPU tests of the fused velocity+stress pass (ecsg_set_fused): it must
reproduce the two separate in-place passes bit for bit -- every field entry,
halos and interface ghosts included, and the receiver traces -- in FP64 and
FP32, with and without graph replay, for odd and even step counts, with the
source anywhere (including on the interface planes whose normal stresses the
fused pass defers until after the vz-szz ghost solve)."""
import numpy as np
import pytest

import ecsg_amd
from helpers import composite

pytestmark = pytest.mark.gpu


def _model(case, nsteps):
    if case == "single_free":
        return composite((), coarse=(21, 17, 9), nz_fine=(), nsteps=nsteps, seed=1)
    if case == "single_rigid":
        return composite((), coarse=(21, 17, 9), nz_fine=(), top_bc=0, nsteps=nsteps, seed=2, src_type=1)
    if case == "r2":
        return composite((2,), coarse=(7, 6, 4), nz_fine=(5,), nsteps=nsteps, seed=3, iso=True)
    if case == "r3":
        return composite((3,), coarse=(6, 5, 4), nz_fine=(5,), nsteps=nsteps, seed=4)
    if case == "r3_rigid":
        return composite((3,), coarse=(6, 5, 4), nz_fine=(5,), top_bc=0, nsteps=nsteps, seed=5)
    if case == "r2r3":
        return composite((2, 3), coarse=(4, 3, 4), nz_fine=(4, 4), nsteps=nsteps, seed=6)
    if case == "src_fine_plane":      # explosive source on the fine grid's interface plane
        m = composite((3,), coarse=(6, 5, 4), nz_fine=(5,), nsteps=nsteps, seed=7, src_type=1)
        m.src_loc = (9, 7, 5)
        return m
    if case == "src_coarse_plane":    # explosive source on the coarse grid's first plane
        m = composite((3,), coarse=(6, 5, 4), nz_fine=(5,), nsteps=nsteps, seed=8, src_type=1)
        m.src_grid, m.src_loc = 1, (3, 2, 0)
        return m
    if case == "src_corner":          # vx force on a boundary corner node
        m = composite((3,), coarse=(6, 5, 4), nz_fine=(5,), nsteps=nsteps, seed=9)
        m.src_loc = (0, 0, 0)
        return m
    # grids large enough for interior tiles (k_fused_lean: columns 64.., rows
    # 4 x rows per wavefront.., planes 1 .. nz-3); the vx force sits where a
    # lean tile's halo column (x = 128), halo row (y = 32) and z chunk start
    # (z = 9 with 8-plane chunks) meet
    if case == "big_single_vx":
        m = composite((), coarse=(200, 70, 12), nz_fine=(), nsteps=nsteps, seed=10)
        m.src_loc = (128, 32, 9)
        return m
    if case == "big_r3_explosive":   # explosive source inside the fine grid's interior tiles
        m = composite((3,), coarse=(70, 30, 5), nz_fine=(12,), nsteps=nsteps, seed=11, src_type=1)
        m.src_loc = (100, 40, 6)
        return m
    if case == "big_r2r3_rigid":
        return composite((2, 3), coarse=(24, 12, 4), nz_fine=(11, 6), top_bc=0, nsteps=nsteps, seed=12)
    raise ValueError(case)


def _run(m, precision, fused, graph, nsteps, random_start=False):
    s = ecsg_amd.Simulation(m, precision=precision, check_energy=False, max_steps=nsteps + 2)
    s.set_fused(fused)
    if random_start:
        rng = np.random.default_rng(5)
        for g in range(m.num_grids):
            for c in range(9):
                s.set(g, c, rng.standard_normal(s.dims(g, c)) * (1.0 if c < 3 else 1e5))
    # an odd first chunk, then the rest: exercises both field-set parities
    s.step(3, graph=graph)
    s.step(nsteps - 3, graph=graph)
    s.sync()
    out = ([s.get(g, c) for g in range(m.num_grids) for c in range(9)], s.traces())
    s.close()
    return out


CASES = ["single_free", "single_rigid", "r2", "r3", "r3_rigid", "r2r3", "src_fine_plane", "src_coarse_plane",
         "src_corner", "big_single_vx", "big_r3_explosive", "big_r2r3_rigid"]


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_fused_pass_equals_separate_passes(case, precision):
    nsteps = 24
    m = _model(case, nsteps)
    ref = _run(m, precision, False, False, nsteps)
    for graph in (False, True):
        got = _run(m, precision, True, graph, nsteps)
        assert np.array_equal(got[1], ref[1]), "traces (graph=%s)" % graph
        for k, (a, b) in enumerate(zip(got[0], ref[0])):
            assert np.array_equal(a, b), "grid %d field %d (graph=%s)" % (k // 9, k % 9, graph)
    assert np.abs(ref[1]).max() > 0


@pytest.mark.parametrize("case", ["r3", "r2r3", "single_rigid", "big_single_vx", "big_r2r3_rigid"])
def test_fused_pass_from_random_state(case):
    """Random values in every entry (halos too) before stepping."""
    nsteps = 9
    m = _model(case, nsteps)
    m.src_trace = np.zeros(nsteps + 2, np.float32)
    ref = _run(m, "f64", False, False, nsteps, random_start=True)
    got = _run(m, "f64", True, False, nsteps, random_start=True)
    for k, (a, b) in enumerate(zip(got[0], ref[0])):
        assert np.array_equal(a, b), "grid %d field %d" % (k // 9, k % 9)
