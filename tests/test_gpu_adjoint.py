"""GPU tests of the adjoint-state gradient (ecsg_gradient and the transposed
step), through the C ABI.

The adjoint has no counterpart in the reference, so it is pinned by what the
mathematics fixes rather than by the CPU restatement:
  * dot-product test of one transposed step: <(S - I) u, w> = <u, (S^T - I) w>
    for random u, w over every wavefield entry (halos and interface ghosts
    included), with the source switched off so that the step S is linear;
  * central finite differences of the misfit along seeded random directions
    of the set_params inputs (buoyancy and the nine stiffnesses), in FP64;
  * FP32 adjoint against the FP64 adjoint;
  * reduced-precision checkpoints against native ones;
  * a torch.autograd.Function wrapping forward + adjoint (test harness only)
    checked against the finite difference of the forward.
"""
import copy

import numpy as np
import pytest

import ecsg_amd
from ecsg_amd.model import Grid
from helpers import composite, rel

try:  # imported at collection: the first import on a fresh machine is slow
    import torch
except ImportError:  # pragma: no cover
    torch = None

pytestmark = pytest.mark.gpu

# Central differences of the misfit along random relative perturbations of
# every node's parameters: with eps = 1e-4 the O(eps^2) truncation is ~1e-6
# of the derivative, so the FP64 adjoint must agree to F64_FD_TOL; the
# coarser eps = 1e-3 check over 8 directions is held to 1e-3.
F64_FD_TOL = 2e-5
F32_VS_F64_TOL = 1e-2


def _scales(m):
    """Velocity and stress magnitudes for which one step changes each by a
    comparable relative amount (|dt/h * b * s| ~ |v|)."""
    G = m.grids[0]
    k = m.dt / G.h[0]
    return 1.0, 1.0 / (k * float(np.mean(G.buoy)))


def _dot_test(m, precision="f64", seed=0):
    s = ecsg_amd.Simulation(m, precision=precision, check_energy=False, max_steps=4)
    rng = np.random.default_rng(seed)
    sv, ss = _scales(m)
    u, w = {}, {}
    for g in range(m.num_grids):
        for c in range(9):
            shape = s.dims(g, c)
            su = sv if c < 3 else ss
            u[g, c] = rng.standard_normal(shape) * su
            w[g, c] = rng.standard_normal(shape) / su
            s.set(g, c, u[g, c])
            s.set_adjoint(g, c, w[g, c])
    s.step(1, graph=False)
    s.sync()
    s.adjoint_step()
    lhs = rhs = 0.0
    mag = 0.0
    for g in range(m.num_grids):
        for c in range(9):
            du = s.get(g, c) - u[g, c]
            dw = s.get_adjoint(g, c) - w[g, c]
            lhs += float(np.sum(du * w[g, c]))
            rhs += float(np.sum(u[g, c] * dw))
            mag += float(np.sum(np.abs(du * w[g, c])))
    s.close()
    return lhs, rhs, mag


@pytest.mark.parametrize("case", ["single_free", "single_rigid", "r2", "r3", "r2r3", "r3_rigid"])
def test_transposed_step_dot_product(case):
    if case.startswith("single"):
        m = composite((), coarse=(9, 7, 6), nz_fine=(), top_bc=1 if case == "single_free" else 0, nsteps=4, seed=3)
    elif case == "r2":
        m = composite((2,), coarse=(5, 4, 4), nz_fine=(5,), nsteps=4, seed=4)
    elif case == "r3":
        m = composite((3,), coarse=(5, 4, 4), nz_fine=(5,), nsteps=4, seed=5)
    elif case == "r3_rigid":
        m = composite((3,), coarse=(5, 4, 4), nz_fine=(5,), top_bc=0, nsteps=4, seed=6)
    else:
        m = composite((2, 3), coarse=(4, 3, 4), nz_fine=(4, 4), nsteps=4, seed=7)
    m.src_trace = np.zeros(4, np.float32)
    lhs, rhs, mag = _dot_test(m)
    print("dot test %s: <Su-u,w> = %.15e  <u,S'w-w> = %.15e  (sum |terms| %.3e)" % (case, lhs, rhs, mag))
    assert abs(lhs - rhs) <= 1e-9 * mag


def _with_params(m, factors):
    """Copy of m whose set_params inputs are multiplied by factors[g][a]
    (a = 0 buoyancy, 1.. stiffness inputs)."""
    m2 = copy.copy(m)
    grids = []
    for g, G in enumerate(m.grids):
        f = factors[g]
        grids.append(Grid(G.n, G.h, G.buoy * f[0], [a * f[1 + i] for i, a in enumerate(G.stiff)], G.origin_z))
    m2.grids = grids
    return m2


def _directions(m, rng, planes=None):
    """Random relative perturbations of every input array; planes: restrict
    them to the interface planes (fine iz = nz, coarse iz = 0)."""
    out = []
    for g, G in enumerate(m.grids):
        shape = G.buoy.shape
        arrs = [rng.uniform(-1, 1, size=shape) for _ in range(1 + len(G.stiff))]
        if planes:
            mask = np.zeros(shape)
            if g < m.num_grids - 1:
                mask[-1] = 1
            if g > 0:
                mask[0] = 1
            arrs = [a * mask for a in arrs]
        out.append(arrs)
    return out


def _misfit(m, nsteps, obs, precision="f64"):
    """Forward run only: 1/2 sum (traces - obs)^2."""
    s = ecsg_amd.Simulation(m, precision=precision, check_energy=False, max_steps=nsteps)
    s.step(nsteps, graph=False)
    s.sync()
    tr = s.traces()
    s.close()
    return 0.5 * float(np.sum((tr - obs) ** 2))


def _grad_dot(s, m, d):
    tot = 0.0
    for g, G in enumerate(m.grids):
        tot += float(np.sum(s.grad(g, 9) * G.buoy * d[g][0]))
        sels = [10, 14] if m.iso else list(range(10, 19))
        for i, sel in enumerate(sels):
            tot += float(np.sum(s.grad(g, sel) * G.stiff[i] * d[g][1 + i]))
    return tot


def _observed(m, nsteps, seed):
    """Observed traces: the same model with a smooth bump in the fine grid's
    shear moduli and density."""
    rng = np.random.default_rng(seed)
    G = m.grids[0]
    nz, ny, nx = G.buoy.shape
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    bump = np.exp(-(((x - nx / 2) / (nx / 5)) ** 2 + ((y - ny / 2) / (ny / 5)) ** 2 + ((z - nz / 2) / (nz / 3)) ** 2))
    m2 = copy.copy(m)
    grids = list(m.grids)
    st = list(G.stiff)
    if m.iso:
        st = [st[0], st[1] + 0.0]
        st[1] = st[1] + 0.08 * (st[0] - st[1]) * bump
    else:
        for i in (6, 7, 8):
            st[i] = st[i] * (1 - 0.08 * bump)
    grids[0] = Grid(G.n, G.h, G.buoy / (1 + 0.03 * bump), st, G.origin_z)
    m2.grids = grids
    s = ecsg_amd.Simulation(m2, precision="f64", check_energy=False, max_steps=nsteps)
    s.step(nsteps, graph=False)
    s.sync()
    tr = s.traces()
    s.close()
    return tr + 1e-3 * np.abs(tr).max() * rng.standard_normal(tr.shape)


def _fd_check(m, nsteps, ndir, planes=False, seed=0, eps=1e-4, every=7):
    obs = _observed(m, nsteps, seed)
    s = ecsg_amd.Simulation(m, precision="f64", check_energy=False, max_steps=nsteps)
    s.set_checkpointing(every, "native")
    J0 = s.gradient(nsteps, obs)
    rng = np.random.default_rng(100 + seed)
    errs = []
    for _ in range(ndir):
        d = _directions(m, rng, planes=planes)
        adj = _grad_dot(s, m, d)
        fp = [[1 + eps * a for a in dg] for dg in d]
        fm = [[1 - eps * a for a in dg] for dg in d]
        Jp = _misfit(_with_params(m, fp), nsteps, obs)
        Jm = _misfit(_with_params(m, fm), nsteps, obs)
        fd = (Jp - Jm) / (2 * eps)
        errs.append(abs(fd - adj) / max(abs(fd), 1e-300))
        print("J0 %.6e  fd %.10e  adjoint %.10e  rel %.2e" % (J0, fd, adj, errs[-1]))
    s.close()
    return max(errs)


@pytest.mark.parametrize("case", ["single_free", "single_rigid", "r2_iso", "r3", "r2r3"])
def test_gradient_matches_central_differences(case):
    if case == "single_free":
        m = composite((), coarse=(10, 8, 7), nz_fine=(), nsteps=50, seed=11)
    elif case == "single_rigid":
        m = composite((), coarse=(10, 8, 7), nz_fine=(), top_bc=0, nsteps=50, seed=12)
    elif case == "r2_iso":
        m = composite((2,), coarse=(5, 4, 4), nz_fine=(5,), iso=True, nsteps=50, seed=13)
    elif case == "r3":
        m = composite((3,), coarse=(5, 4, 4), nz_fine=(5,), nsteps=50, seed=14)
    else:
        m = composite((2, 3), coarse=(4, 3, 4), nz_fine=(4, 4), nsteps=50, seed=15)
    err = _fd_check(m, 50, ndir=3, seed=1)
    assert err <= F64_FD_TOL


def test_gradient_summed_over_shots():
    """Shots of one model share its media: the gradient of the misfit summed
    over the shots is the sum of the shots' gradients (each shot its own
    handle, as the batched configuration computes them), checked against
    central differences of the summed misfit along seeded directions."""
    nsteps, eps = 40, 1e-4
    base = composite((3,), coarse=(6, 5, 4), nz_fine=(5,), nsteps=nsteps, seed=23)
    shots = []
    for loc in ((4, 5, 2), (13, 9, 3), (8, 3, 1)):
        m = copy.copy(base)
        m.src_loc = loc
        shots.append(m)
    obs = [_observed(m, nsteps, seed=30 + k) for k, m in enumerate(shots)]
    rng = np.random.default_rng(41)
    dirs = [_directions(base, rng) for _ in range(3)]
    adj = np.zeros(len(dirs))
    for m, o in zip(shots, obs):
        s = ecsg_amd.Simulation(m, precision="f64", check_energy=False, max_steps=nsteps)
        s.set_checkpointing(9, "native")
        assert s.gradient(nsteps, o) > 0
        adj += [_grad_dot(s, m, d) for d in dirs]
        s.close()
    for d, a in zip(dirs, adj):
        J = {}
        for sign in (1, -1):
            f = [[1 + sign * eps * x for x in dg] for dg in d]
            J[sign] = sum(_misfit(_with_params(m, f), nsteps, o) for m, o in zip(shots, obs))
        fd = (J[1] - J[-1]) / (2 * eps)
        assert abs(fd - a) <= F64_FD_TOL * abs(fd), (fd, a)


def test_gradient_eight_directions_eps_1e3():
    """Two-grid ratio-3 model, central differences with eps = 1e-3 (relative)
    along 8 seeded directions, held to 1e-3."""
    m = composite((3,), coarse=(6, 5, 5), nz_fine=(6,), nsteps=60, seed=17)
    err = _fd_check(m, 60, ndir=8, seed=9, eps=1e-3, every=20)
    assert err <= 1e-3


@pytest.mark.parametrize("ratios", [(2,), (3,)])
def test_interface_parameter_gradient(ratios):
    """Directions confined to the interface planes, where the ghost solves'
    parameter derivatives carry a large share of the gradient."""
    m = composite(ratios, coarse=(5, 4, 4), nz_fine=(5,), nsteps=50, seed=21)
    err = _fd_check(m, 50, ndir=2, planes=True, seed=2)
    assert err <= F64_FD_TOL


def test_checkpoint_interval_does_not_change_the_gradient():
    """Native checkpoints reproduce the forward exactly: any interval gives
    the same gradient bit for bit."""
    m = composite((3,), coarse=(5, 4, 4), nz_fine=(5,), nsteps=40, seed=31)
    obs = _observed(m, 40, 3)
    res = []
    for every in (40, 7, 1):
        s = ecsg_amd.Simulation(m, precision="f64", check_energy=False, max_steps=40)
        s.set_checkpointing(every, "native")
        J = s.gradient(40, obs)
        res.append((J, [s.grad(g, sel) for g in range(2) for sel in range(9, 19)]))
        s.close()
    for J, gs in res[1:]:
        assert J == res[0][0]
        for a, b in zip(gs, res[0][1]):
            assert np.array_equal(a, b)


def _recomputed(N, K1, K2):
    """Forward steps the backward pass recomputes: one level, every segment
    but the last; two levels, every segment once and every sub-segment but
    each segment's last again (first segment / sub-segment: the remainder)."""
    nseg = -(-N // K1)
    lens = [N - (nseg - 1) * K1] + [K1] * (nseg - 1)
    if K2 >= K1:
        return N - min(K1, N)
    return sum(L + (L - K2 if L > K2 else 0) for L in lens)


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_two_level_checkpoints_equal_one_level(precision):
    """A ring shorter than the checkpoint interval (two levels: sub-segment
    checkpoints, each sub-segment recomputed into the ring) gives the
    one-level gradient and misfit bit for bit with native checkpoints, and
    reports its interval, ring and recomputed steps; with FP16 checkpoints it
    stays within the FP16 checkpoints' own error of the native gradient."""
    N = 37
    m = composite((3,), coarse=(5, 4, 4), nz_fine=(5,), nsteps=N, seed=33)
    obs = _observed(m, N, 3)
    res = {}
    for every, ring, fmt in ((10, 0, "native"), (10, 3, "native"), (37, 4, "native"), (7, 1, "native"),
                             (10, 0, "fp16"), (10, 3, "fp16")):
        s = ecsg_amd.Simulation(m, precision=precision, check_energy=False, max_steps=N)
        s.set_checkpointing(every, fmt)
        s.set_ring_steps(ring)
        J = s.gradient(N, obs)
        k1, k2, rec = s.checkpoint_levels
        assert (k1, k2) == (every, ring if ring else every), (k1, k2)
        assert rec == _recomputed(N, k1, k2), (every, ring, rec)
        # the state after the call is the one the last recomputation left
        res[(every, ring, fmt)] = (J, np.concatenate([s.grad(g, sel).ravel() for g in range(2) for sel in range(9, 19)]))
        s.close()
    J0, g0 = res[(10, 0, "native")]
    assert J0 > 0
    for key in ((10, 3, "native"), (37, 4, "native"), (7, 1, "native")):
        J, g = res[key]
        assert J == J0, key
        assert np.array_equal(g, g0), key
    e1 = rel(res[(10, 0, "fp16")][1], g0)
    e2 = rel(res[(10, 3, "fp16")][1], g0)
    print("%s fp16 checkpoints: one level %.3e, two levels %.3e" % (precision, e1, e2))
    assert e2 <= max(5 * e1, 1e-6), (e1, e2)


def test_interval_from_memory_with_a_ring_limit():
    """Interval 0: the longest one-level interval the memory (and the ring
    limit) allows -- here the whole run, or the ring limit (two levels are
    chosen only when no one-level interval fits in device memory); results
    equal bit for bit."""
    N = 30
    m = composite((3,), coarse=(5, 4, 4), nz_fine=(5,), nsteps=N, seed=35)
    obs = _observed(m, N, 3)
    out = []
    for ring in (0, 4):
        s = ecsg_amd.Simulation(m, precision="f64", check_energy=False, max_steps=N)
        s.set_checkpointing(0, "native")
        s.set_ring_steps(ring)
        J = s.gradient(N, obs)
        k1, k2, rec = s.checkpoint_levels
        if ring == 0:
            assert (k1, k2, rec) == (N, N, 0)        # everything fits: one segment
        else:
            assert (k1, k2, rec) == (4, 4, N - 4)
        out.append((J, np.concatenate([s.grad(g, sel).ravel() for g in range(2) for sel in range(9, 19)])))
        s.close()
    assert out[0][0] == out[1][0]
    assert np.array_equal(out[0][1], out[1][1])


@pytest.mark.parametrize("fmt,tol", [("fp16", 2e-2), ("bf16", 1e-1)])
def test_reduced_precision_checkpoints(fmt, tol):
    m = composite((3,), coarse=(5, 4, 4), nz_fine=(5,), nsteps=60, seed=41)
    obs = _observed(m, 60, 4)
    out = {}
    for f in ("native", fmt):
        s = ecsg_amd.Simulation(m, precision="f32", check_energy=False, max_steps=60)
        s.set_checkpointing(10, f)
        J = s.gradient(60, obs)
        out[f] = (J, np.concatenate([s.grad(g, sel).ravel() for g in range(2) for sel in range(9, 19)]))
        s.close()
    e = rel(out[fmt][1], out["native"][1])
    print("%s checkpoints: gradient rel L2 vs native %.3e" % (fmt, e))
    assert out[fmt][0] == out["native"][0]
    assert e <= tol


def test_fp32_adjoint_against_fp64():
    m = composite((3,), coarse=(5, 4, 4), nz_fine=(5,), nsteps=60, seed=51)
    obs = _observed(m, 60, 5)
    out = {}
    for p in ("f64", "f32"):
        s = ecsg_amd.Simulation(m, precision=p, check_energy=False, max_steps=60)
        s.set_checkpointing(20, "native")
        out[p] = (s.gradient(60, obs), [s.grad(g, sel) for g in range(2) for sel in range(9, 19)])
        s.close()
    ej = abs(out["f32"][0] - out["f64"][0]) / out["f64"][0]
    eg = max(rel(a, b) for a, b in zip(out["f32"][1], out["f64"][1]))
    print("FP32 vs FP64: misfit rel %.3e, worst gradient array rel L2 %.3e" % (ej, eg))
    assert ej <= F32_VS_F64_TOL
    assert eg <= F32_VS_F64_TOL


def test_torch_autograd_function_directional_derivative():
    """forward + adjoint wrapped as a torch.autograd.Function of per-array
    scale factors; its gradient checked against central differences."""
    if torch is None:
        pytest.skip("torch not importable")
    m = composite((3,), coarse=(5, 4, 4), nz_fine=(5,), nsteps=40, seed=61)
    obs = _observed(m, 40, 6)
    na = 10

    class Misfit(torch.autograd.Function):
        @staticmethod
        def forward(ctx, scales):
            f = scales.detach().cpu().numpy().reshape(m.num_grids, na)
            mm = _with_params(m, f)
            s = ecsg_amd.Simulation(mm, precision="f64", check_energy=False, max_steps=40)
            J = s.gradient(40, obs)
            g = np.zeros((m.num_grids, na))
            for gi, G in enumerate(mm.grids):
                g[gi, 0] = np.sum(s.grad(gi, 9) * G.buoy) / f[gi, 0]
                for a in range(9):
                    g[gi, 1 + a] = np.sum(s.grad(gi, 10 + a) * G.stiff[a]) / f[gi, 1 + a]
            s.close()
            ctx.save_for_backward(torch.from_numpy(g.reshape(-1)))
            return torch.tensor(J, dtype=torch.float64)

        @staticmethod
        def backward(ctx, gout):
            (g,) = ctx.saved_tensors
            return gout * g

    x = torch.ones(m.num_grids * na, dtype=torch.float64, requires_grad=True)
    J = Misfit.apply(x)
    J.backward()
    d = torch.from_numpy(np.random.default_rng(7).uniform(-1, 1, x.shape[0]))
    eps = 1e-4
    with torch.no_grad():
        Jp = Misfit.apply(x + eps * d)
        Jm = Misfit.apply(x - eps * d)
    fd = float((Jp - Jm) / (2 * eps))
    ad = float(torch.dot(x.grad, d))
    print("autograd directional derivative %.10e, central difference %.10e" % (ad, fd))
    assert abs(ad - fd) <= 1e-4 * abs(fd)
