"""GPU tests of batched stepping (ecsg_batch_*) and of the tiling autotuner.

A batch runs the same kernels' arithmetic as a single model (the stencil
updates are pointwise, the interface solves per interface), so a batch of
models of different sizes, ratios and shots must reproduce each model run on
its own bit for bit, in FP32 and FP64, with and without graph replay, with
the interface work beside the stencils (the default for batches with
interfaces: graphs of 8 steps and of one) or after them
(ECSG_BATCH_OVERLAP=0).
"""
import numpy as np
import pytest

import ecsg_amd
from helpers import composite

pytestmark = pytest.mark.gpu

FIELDS = range(9)


def _members(nsteps, seed=0):
    """Four models: ratio 3, ratio 2 (isotropic), a single grid and a
    ratio-2-then-3 stack, each with its own source position and wavelet."""
    ms = [
        composite((3,), coarse=(6, 5, 4), nz_fine=(5,), nsteps=nsteps, seed=seed + 1),
        composite((2,), coarse=(7, 6, 4), nz_fine=(6,), iso=True, nsteps=nsteps, seed=seed + 2),
        composite((), coarse=(20, 17, 9), nz_fine=(), nsteps=nsteps, seed=seed + 3, top_bc=0),
        composite((2, 3), coarse=(4, 3, 4), nz_fine=(4, 4), nsteps=nsteps, seed=seed + 4, src_type=1),
    ]
    return ms


def _state(sim, m):
    return [sim.get(g, c) for g in range(m.num_grids) for c in FIELDS], sim.traces()


@pytest.mark.parametrize("overlap", ["1", "0"])
@pytest.mark.parametrize("precision", ["f32", "f64"])
@pytest.mark.parametrize("graph", [False, True])
def test_batch_reproduces_each_model_alone(precision, graph, overlap, monkeypatch):
    monkeypatch.setenv("ECSG_BATCH_OVERLAP", overlap)
    nsteps = 43
    ms = _members(nsteps)
    solo = []
    for m in ms:
        s = ecsg_amd.Simulation(m, precision=precision, check_energy=False, max_steps=nsteps)
        s.step(nsteps, graph=False)
        s.sync()
        solo.append(_state(s, m))
        s.close()
    sims = [ecsg_amd.Simulation(m, precision=precision, check_energy=False, max_steps=nsteps) for m in ms]
    b = ecsg_amd.Batch(sims)
    b.step(nsteps, graph=graph)
    b.sync()
    for k, (s, m) in enumerate(zip(sims, ms)):
        fields, tr = _state(s, m)
        assert np.array_equal(tr, solo[k][1]), "member %d traces" % k
        for a, c in zip(fields, solo[k][0]):
            assert np.array_equal(a, c), "member %d fields" % k
        assert np.abs(tr).max() > 0
    b.close()


@pytest.mark.parametrize("overlap", ["1", "0"])
def test_persistent_stencil_launches_match_members(overlap, monkeypatch):
    # ECSG_BATCH_PERSIST=1: one workgroup per CU, each looping over several
    # tiles of the ~800-tile batch (two ratio-3 and two ratio-2 models of
    # 120 x 120 x 20 fine cells); the same results as each member alone
    monkeypatch.setenv("ECSG_BATCH_OVERLAP", overlap)
    monkeypatch.setenv("ECSG_BATCH_PERSIST", "1")
    nsteps = 24
    ms = [composite((3,), coarse=(40, 40, 10), nz_fine=(20,), nsteps=nsteps, seed=61),
          composite((3,), coarse=(40, 39, 12), nz_fine=(18,), nsteps=nsteps, seed=62),
          composite((2,), coarse=(60, 60, 10), nz_fine=(20,), nsteps=nsteps, seed=63),
          composite((2,), coarse=(58, 61, 8), nz_fine=(16,), iso=True, nsteps=nsteps, seed=64)]
    solo = []
    for m in ms:
        s = ecsg_amd.Simulation(m, precision="f32", check_energy=False, max_steps=nsteps)
        s.step(nsteps, graph=False)
        s.sync()
        solo.append(_state(s, m))
        s.close()
    sims = [ecsg_amd.Simulation(m, precision="f32", check_energy=False, max_steps=nsteps) for m in ms]
    b = ecsg_amd.Batch(sims)
    b.step(nsteps)
    b.sync()
    for k, (s, m) in enumerate(zip(sims, ms)):
        fields, tr = _state(s, m)
        assert np.array_equal(tr, solo[k][1]), "member %d traces" % k
        for a, c in zip(fields, solo[k][0]):
            assert np.array_equal(a, c), "member %d fields" % k
    b.close()


def test_batch_closed_with_steps_in_flight():
    # closing a batch right after enqueueing its steps (no synchronize):
    # the steps complete before its graphs and tables go, and the members
    # then read the stepped state
    nsteps = 40
    ms = _members(nsteps, seed=70)
    solo = []
    for m in ms:
        s = ecsg_amd.Simulation(m, precision="f32", check_energy=False, max_steps=nsteps)
        s.step(nsteps, graph=False)
        s.sync()
        solo.append(s.traces())
        s.close()
    sims = [ecsg_amd.Simulation(m, precision="f32", check_energy=False, max_steps=nsteps) for m in ms]
    b = ecsg_amd.Batch(sims)
    b.step(nsteps)
    b.close()
    for k, s in enumerate(sims):
        assert np.array_equal(s.traces(), solo[k]), "member %d" % k
        s.close()


def test_batch_graph_replays_interleaved_with_reads():
    """Graph replays in chunks with reads of the members in between: every
    chunk equals the eager batch."""
    nsteps = 30
    ms = _members(nsteps, seed=10)
    runs = {}
    for graph in (False, True):
        sims = [ecsg_amd.Simulation(m, precision="f32", check_energy=False, max_steps=nsteps) for m in ms]
        b = ecsg_amd.Batch(sims)
        snaps = []
        for chunk in (7, 11, 12):
            b.step(chunk, graph=graph)
            snaps.append([s.get(0, 0).copy() for s in sims])  # ordered after the batch's work
        b.sync()
        runs[graph] = (snaps, [s.traces() for s in sims])
        b.close()
    for a, c in zip(runs[True][0], runs[False][0]):
        for x, y in zip(a, c):
            assert np.array_equal(x, y)
    for x, y in zip(runs[True][1], runs[False][1]):
        assert np.array_equal(x, y)


def test_single_model_graph_with_lds_jacobi_matches_eager():
    """Composite single models whose interfaces fit the LDS solver are graph
    captured; chunks of replays with reads in between equal the eager run."""
    m = composite((3,), coarse=(6, 5, 4), nz_fine=(5,), nsteps=30, seed=21)
    out = {}
    for graph in (False, True):
        s = ecsg_amd.Simulation(m, precision="f32", check_energy=False, max_steps=30)
        snaps = []
        for chunk in (5, 10, 15):
            s.step(chunk, graph=graph)
            s.sync()
            snaps.append(s.get(1, 7).copy())
        out[graph] = (snaps, s.traces())
        s.close()
    for x, y in zip(out[True][0], out[False][0]):
        assert np.array_equal(x, y)
    assert np.array_equal(out[True][1], out[False][1])


def test_autotune_preserves_state_and_results():
    m = composite((3,), coarse=(6, 5, 4), nz_fine=(5,), nsteps=30, seed=31)
    ref = ecsg_amd.Simulation(m, precision="f32", check_energy=False, max_steps=30)
    ref.step(30, graph=False)
    ref.sync()
    s = ecsg_amd.Simulation(m, precision="f32", check_energy=False, max_steps=30)
    s.step(10, graph=False)
    s.sync()
    before = [s.get(g, c).copy() for g in range(2) for c in FIELDS]
    choice = s.autotune(reps=2)
    assert len(choice) == 2 and all(b in (1, 2, 4, 8) and z in (1, 2, 4, 8, 16) for b, z in choice)
    after = [s.get(g, c) for g in range(2) for c in FIELDS]
    for a, c in zip(after, before):
        assert np.array_equal(a, c)
    s.step(20, graph=False)
    s.sync()
    assert np.array_equal(s.traces(), ref.traces())
    # batch autotune
    ms = _members(20, seed=40)
    sims = [ecsg_amd.Simulation(mm, precision="f32", check_energy=False, max_steps=20) for mm in ms]
    b = ecsg_amd.Batch(sims)
    by, zc = b.autotune(reps=2)
    assert by in (1, 2, 4, 8) and zc in (2, 4, 8, 16)
    b.step(20)
    b.sync()
    for mm, sm in zip(ms, sims):
        alone = ecsg_amd.Simulation(mm, precision="f32", check_energy=False, max_steps=20)
        alone.step(20, graph=False)
        alone.sync()
        assert np.array_equal(alone.traces(), sm.traces())
        alone.close()
    b.close()


def test_tune_steps_keeps_results():
    # two candidates x (4 timed + 8 capture steps), then 6 more steps: the
    # members' traces equal those of each member stepped alone for 30 steps
    ms = _members(30, seed=44)
    sims = [ecsg_amd.Simulation(mm, precision="f32", check_energy=False, max_steps=30) for mm in ms]
    b = ecsg_amd.Batch(sims)
    by, zc = b.tune_steps([(8, 8), (4, 2)], steps=4)
    assert (by, zc) in ((8, 8), (4, 2))
    b.step(6)
    b.sync()
    for mm, sm in zip(ms, sims):
        alone = ecsg_amd.Simulation(mm, precision="f32", check_energy=False, max_steps=30)
        alone.step(30, graph=False)
        alone.sync()
        assert np.array_equal(alone.traces(), sm.traces())
        alone.close()
    b.close()
    with pytest.raises(ecsg_amd.EcsgError):
        ecsg_amd.Batch(sims).tune_steps([])


def test_batch_rejects_energy_check_members():
    m = composite((3,), coarse=(6, 5, 4), nz_fine=(5,), nsteps=5, seed=51)
    s = ecsg_amd.Simulation(m, precision="f32", check_energy=True, max_steps=5)
    with pytest.raises(ecsg_amd.EcsgError):
        ecsg_amd.Batch([s])
