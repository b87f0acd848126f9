"""GPU tests of batched gradients (ecsg_batch_gradient) and of the
batched-shots workload (16 uneven two-grid models x 4 shots).

The batched transposed step runs the single-model adjoint kernels' bodies
over flat work lists and the batched forward step the single-model step's
arithmetic, so every member's misfit and gradient must equal its own
ecsg_gradient bit for bit (FP32 and FP64, with recomputed segments, with
the forward steps' interface solves beside the stencils or after them); a
model's gradient -- its shots' accumulators added in member order on the
device, in the storage precision -- equals the sum of the shots' own
gradients to that precision's rounding (1e-5 FP32, 1e-12 FP64, relative).
Members use the single-workgroup Jacobi solve (sxz_solver 2), as batches do.
"""
import numpy as np
import pytest

import ecsg_amd
from ecsg_amd.model import batch_workload
from helpers import composite

pytestmark = pytest.mark.gpu


def _members(nsteps, seed=0):
    return [
        composite((3,), coarse=(6, 5, 4), nz_fine=(5,), nsteps=nsteps, seed=seed + 1),
        composite((2,), coarse=(7, 6, 4), nz_fine=(6,), iso=True, nsteps=nsteps, seed=seed + 2),
        composite((), coarse=(20, 17, 9), nz_fine=(), nsteps=nsteps, seed=seed + 3, top_bc=0),
        composite((2, 3), coarse=(4, 3, 4), nz_fine=(4, 4), nsteps=nsteps, seed=seed + 4, src_type=1),
    ]


def _observed(m, nsteps, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((len(m.rec_loc), nsteps)) * 1e-3


def _grads(sim, m):
    return [sim.grad(g, sel) for g in range(m.num_grids) for sel in range(9, 19)]


def _solo(m, precision, nsteps, every, obs):
    s = ecsg_amd.Simulation(m, precision=precision, check_energy=False, max_steps=nsteps, sxz_solver="jacobi_lds")
    s.set_checkpointing(every)
    J = s.gradient(nsteps, obs)
    out = J, _grads(s, m)
    s.close()
    return out


@pytest.mark.parametrize("overlap", ["1", "0"])
@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("every", [7, 0])
def test_batch_gradient_equals_each_member_alone(precision, every, overlap, monkeypatch):
    # forward and recomputed steps overlapped (the default for members with
    # interfaces) or serial (ECSG_BATCH_OVERLAP=0)
    monkeypatch.setenv("ECSG_BATCH_OVERLAP", overlap)
    nsteps = 30
    ms = _members(nsteps)
    obs = [_observed(m, nsteps, 10 + k) for k, m in enumerate(ms)]
    sims = [ecsg_amd.Simulation(m, precision=precision, check_energy=False, max_steps=nsteps, sxz_solver="jacobi_lds")
            for m in ms]
    for s in sims:
        s.set_checkpointing(every)
    b = ecsg_amd.Batch(sims)
    J = b.gradient(nsteps, obs)
    K = b.checkpoint_interval
    assert 1 <= K <= nsteps and (every == 0 or K == every)
    got = [_grads(s, m) for s, m in zip(sims, ms)]
    b.close()
    for s in sims:
        s.close()
    for k, m in enumerate(ms):
        Js, gs = _solo(m, precision, nsteps, K, obs[k])
        assert J[k] == Js, k
        for a, c in zip(got[k], gs):
            assert np.array_equal(a, c), "member %d" % k
        assert max(np.abs(a).max() for a in gs) > 0


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_batch_gradient_sums_each_models_shots(precision):
    """Three shots of one model and one of another: the first shot's handle
    holds the model's gradient, the sum of the shots' own gradients (to the
    rounding of the storage precision's additions); the other members keep
    their own bit for bit."""
    nsteps = 24
    shots = []
    for k, loc in enumerate([(9, 7, 2), (4, 10, 3), (14, 3, 1)]):
        m = composite((3,), coarse=(6, 5, 4), nz_fine=(5,), nsteps=nsteps, seed=3)
        m.src_loc = loc
        shots.append(m)
    other = composite((2,), coarse=(7, 6, 4), nz_fine=(6,), iso=True, nsteps=nsteps, seed=8)
    ms = shots + [other]
    obs = [_observed(m, nsteps, 30 + k) for k, m in enumerate(ms)]
    solo = [_solo(m, precision, nsteps, 6, o) for m, o in zip(ms, obs)]
    sims = [ecsg_amd.Simulation(m, precision=precision, check_energy=False, max_steps=nsteps, sxz_solver="jacobi_lds")
            for m in ms]
    for s in sims:
        s.set_checkpointing(6)
    b = ecsg_amd.Batch(sims)
    J = b.gradient(nsteps, obs, group=[0, 0, 0, 1])
    got = [_grads(s, m) for s, m in zip(sims, ms)]
    b.close()
    for s in sims:
        s.close()
    # accumulators added on the device in the storage precision, then carried
    # to the inputs in double: the per-shot gradients' sum to that rounding
    for q in range(len(got[0])):
        want = solo[0][1][q] + solo[1][1][q] + solo[2][1][q]
        scale = np.abs(want).max()
        assert np.abs(got[0][q] - want).max() <= (1e-5 if precision == "f32" else 1e-12) * scale, q
    for k in (1, 2, 3):
        for a, c in zip(got[k], solo[k][1]):
            assert np.array_equal(a, c)
    assert [J[k] for k in range(4)] == [solo[k][0] for k in range(4)]


def test_batched_shots_workload_forward_and_gradient():
    """The batched-shots workload (16 two-grid models of uneven sizes, 4 shots
    each, FP32): 300 graph-replayed batched steps equal a subset of members
    stepped alone bit for bit; a batched gradient over 60 steps (checkpoints
    every 20) gives, for the members of two models, each shot's own gradient
    bit for bit and the models' shot sums."""
    nsteps = 300
    ms = batch_workload(nsteps=nsteps)
    assert len(ms) == 64
    check = [0, 13, 34, 63]
    solo = {}
    for k in check:
        s = ecsg_amd.Simulation(ms[k], precision="f32", check_energy=False, max_steps=nsteps, sxz_solver="jacobi_lds")
        s.step(nsteps)
        s.sync()
        solo[k] = ([s.get(g, c) for g in range(2) for c in range(9)], s.traces())
        s.close()
    sims = [ecsg_amd.Simulation(m, precision="f32", check_energy=False, max_steps=nsteps, sxz_solver="jacobi_lds")
            for m in ms]
    b = ecsg_amd.Batch(sims)
    b.autotune(reps=1)
    b.step(nsteps, graph=True)
    b.sync()
    for k in check:
        f, tr = solo[k]
        assert np.array_equal(sims[k].traces(), tr), k
        for a, c in zip([sims[k].get(g, c) for g in range(2) for c in range(9)], f):
            assert np.array_equal(a, c), k
    # gradients: zero observed data, checkpoints every 20 of 60 steps
    ng = 60
    for s in sims:
        s.reset()
        s.set_checkpointing(20)
    obs = [np.zeros((len(m.rec_loc), ng)) for m in ms]
    groups = [k // 4 for k in range(64)]
    J = b.gradient(ng, obs, group=groups)
    assert b.checkpoint_interval == 20
    lead = {0: sims[0].grad(0, 15), 8: sims[8].grad(0, 15)}   # stiff[5] (code c12) of models 0 and 2, fine grid
    own = {k: sims[k].grad(1, 9) for k in (1, 2, 3, 9, 10, 11)}  # buoyancy, coarse grid
    b.close()
    for s in sims:
        s.close()
    per = {}
    for k in range(0, 12):
        s = ecsg_amd.Simulation(ms[k], precision="f32", check_energy=False, max_steps=ng, sxz_solver="jacobi_lds")
        s.set_checkpointing(20)
        Jk = s.gradient(ng, obs[k])
        assert Jk == J[k], k
        per[k] = (s.grad(0, 15), s.grad(1, 9))
        s.close()
    for k in (1, 2, 3, 9, 10, 11):
        assert np.array_equal(own[k], per[k][1]), k
    for L in (0, 8):
        want = per[L][0] + per[L + 1][0] + per[L + 2][0] + per[L + 3][0]
        assert np.abs(lead[L] - want).max() <= 1e-5 * np.abs(want).max(), L
        assert np.abs(want).max() > 0
