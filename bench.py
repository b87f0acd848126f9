#!/usr/bin/env python3
"""Benchmark of the ECSG composite time step on one MI355X.

Workload (default): the three-grid composite -- 1207x1207x81 nodes at
h = 5 m, 403x403x101 at 15 m, 135x135x121 at 45 m, ratio 3 at both
interfaces, 136.6 M nodes, layered orthorhombic medium (9 Cij + rho per node),
FP32 wavefields and parameters, 2000 steps.  A step is one full composite
time step of main.cpp:165-261 (velocity stencils, receivers, source,
vz interface solves, stress stencils, sxz Jacobi and syz interface solves),
with all inputs resident in HBM.

Output: one JSON line.  value = grid-point updates per second (sum over grids
of nodes x steps / wall time).  roofline: the dominant kernel (stress stencil
of the 1207x1207x81 grid; with the interface overlap, its launch over the
planes away from the interface) timed with HIP events on its launch stream
during the timed steps, at 96 algorithmic bytes per node (24 FP32 array
accesses).
cpu_baseline: the CPU restatement (oracle/, OpenMP) on the same workload for a
bounded number of steps.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload three_grid]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Gpt-updates/s fwd & fwd+adjoint, 3-grid FP32 vs plain FP64 GPU; % HBM peak"
HBM_PEAK = 8.0e12          # MI355X HBM3E, bytes/s
BYTES_STEP_F32 = 148       # algorithmic bytes per node and step (37 array accesses x 4 B)
BYTES_STRESS_F32 = 96      # UpdateStress: 24 accesses x 4 B
BYTES_VEL_F32 = 52         # UpdateVel: 13 accesses x 4 B


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_model(name: str, steps: int):
    from ecsg_amd import model as M
    if name == "three_grid":
        return M.three_grid_workload(nsteps=steps)
    if name == "two_grid":
        return M.two_grid_workload(nsteps=steps)
    if name == "uniform":
        return M.uniform_model(nsteps=steps)
    if name == "reference_test1":
        return M.reference_flow(1)
    raise SystemExit("unknown workload %s" % name)


def gpu_run(m, precision: str, warmup: int, preroll: int, steps: int, profile: bool, prof_steps: int = 50,
            solver: str = "jacobi"):
    """warmup + preroll untimed steps from rest (the pre-roll carries the
    wavefield across the interfaces, so that the timed steps include the
    interface solves' steady-state sweep counts), then `steps` timed steps."""
    import ecsg_amd
    t0 = time.time()
    sim = ecsg_amd.Simulation(m, precision=precision, check_energy=False,
                              max_steps=warmup + preroll + steps + prof_steps + 2, sxz_solver=solver)
    log("[bench] %s state ready in %.1fs" % (precision, time.time() - t0))
    sim.step(warmup)
    sim.sync()
    log("[bench] %s warmup %d steps done" % (precision, warmup))
    pre = None
    if preroll > 0:
        s0 = [sim.jacobi_stats(k)[0] for k in range(m.num_grids - 1)]
        t = time.perf_counter()
        sim.step(preroll)
        sim.sync()
        ep = time.perf_counter() - t
        s1 = [sim.jacobi_stats(k)[0] for k in range(m.num_grids - 1)]
        pre = {"steps": preroll, "ms_per_step": round(ep / preroll * 1e3, 4),
               "jacobi_sweeps_per_step": [round((b - a) / preroll, 2) for a, b in zip(s0, s1)]}
        log("[bench] %s pre-roll %d steps: %s" % (precision, preroll, pre))
    # the timed steps, with HIP events around the grid-0 stencil launches
    s0 = [sim.jacobi_stats(k)[0] for k in range(m.num_grids - 1)]
    t = time.perf_counter()
    live = sim.step_timed(steps)
    el = time.perf_counter() - t
    s1 = [sim.jacobi_stats(k)[0] for k in range(m.num_grids - 1)]
    stats = [round((b - a) / steps, 2) for a, b in zip(s0, s1)]
    overlap = sim.overlap_mode
    if profile and prof_steps > 0:
        # every kernel kind: further steps with each launch bracketed by events
        ms, cnt = sim.profile(prof_steps)
        sim.sync()
        log("[bench] %s profiled steps, kernel avg ms: vel0 %.4f stress0 %.4f, interface %.4f (x%d), point %.4f"
            % (precision, ms[0], ms[1], ms[4], cnt[4], ms[5]))
    tr = sim.traces()
    sim.close()
    return el, live, stats, tr, overlap, pre


def adjoint_run(m, steps: int, every: int, fmt: str, precision: str = "f32"):
    """Forward with checkpoints + adjoint (gradient of 1/2 sum traces^2, i.e.
    zero observed data) over `steps` steps; the first call builds the
    interface term lists and allocates, the second is timed."""
    import ecsg_amd
    sim = ecsg_amd.Simulation(m, precision=precision, check_energy=False, max_steps=steps)
    sim.set_checkpointing(every, fmt)
    obs = np.zeros((len(m.rec_loc), steps))
    t0 = time.time()
    sim.gradient(steps, obs)
    sim.sync()
    log("[bench] adjoint set-up and first gradient (%d steps) in %.1fs" % (steps, time.time() - t0))
    s0 = sim.adjoint_sweeps()
    t = time.perf_counter()
    J = sim.gradient(steps, obs)
    sim.sync()
    el = time.perf_counter() - t
    sweeps = sim.adjoint_sweeps()
    levels = sim.checkpoint_levels   # (interval, ring, recomputed steps)
    sim.close()
    return el, J, sweeps, levels


def batch_run(steps: int, solo_steps: int, warmup: int, tiling=None):
    """Batched shots over uneven two-grid models (16 models x 4 shots):
    one launch per phase for all 64 members, graph-captured, tiling chosen
    from timed steps (or the fixed `tiling` = (rows, planes)); beside it the same members stepped one at a time (each with its
    own graph-captured step on its own stream, so that several are in flight
    at once; their vx-sxz systems are solved by one workgroup each,
    sxz_solver "jacobi_lds", which lets them share the GPU)."""
    import ecsg_amd
    from ecsg_amd import model as M
    t0 = time.time()
    cands = [(8, 8), (8, 4), (4, 8), (4, 4)]
    tune = (len(cands) + 1) * (16 + 8)
    ms = M.batch_workload(nsteps=tune + warmup + steps + solo_steps + 10)
    nodes = sum(m.total_nodes for m in ms)
    sims = [ecsg_amd.Simulation(m, precision="f32", check_energy=False,
                                max_steps=tune + warmup + steps + solo_steps + 10,
                                sxz_solver="jacobi_lds") for m in ms]
    log("[bench] batch: %d members, %.1f M nodes, ready in %.1fs" % (len(ms), nodes / 1e6, time.time() - t0))
    b = ecsg_amd.Batch(sims)
    # the stencil-only autotuner's pick and four tilings, each timed over 16
    # real steps (these steps are part of the warm-up)
    by, zc = b.autotune(reps=10)
    tuned = (by, zc)
    if tiling:
        by, zc = tiling
        b.set_tiling(by, zc)
    else:
        by, zc = b.tune_steps([tuned] + [c for c in cands if c != tuned], steps=16)
    b.step(warmup)
    b.sync()
    t = time.perf_counter()
    b.step(steps, graph=True)
    b.sync()
    el = time.perf_counter() - t
    b.close()
    # the same members one at a time
    t = time.perf_counter()
    for sm in sims:
        sm.step(solo_steps, graph=True)
    for sm in sims:
        sm.sync()
    el1 = time.perf_counter() - t
    for sm in sims:
        sm.close()
    return {"value": round(nodes * steps / el / 1e9, 3), "unit": "Gpt/s", "members": len(ms), "models": 16,
            "shots_per_model": 4, "nodes": nodes, "steps": steps, "ms_per_step": round(el / steps * 1e3, 4),
            "tiling": {"rows": by, "z_planes": zc, "stencil_only_autotune": list(tuned)}, "graph": True,
            "interface_overlap": os.environ.get("ECSG_BATCH_OVERLAP", "1").strip() not in ("0", ""),
            "one_at_a_time_value": round(nodes * solo_steps / el1 / 1e9, 3), "one_at_a_time_steps": solo_steps,
            "workload": "16 ratio-3 two-grid models, fine (nx,ny) in [120,240] (multiples of 3), nz in [30,60] cells, "
                        "coarse (nx/3, ny/3, 2nz); 4 Ricker shots each (10-25 Hz); FP32"}


def batch_gradient_run(steps: int, every: int, tiling=None):
    """The batched-shots workload's gradients with one launch per phase
    (ecsg_batch_gradient): 64 members (16 models x 4 shots), forward with
    checkpoints + recomputation + transposed steps + 9 Cij and buoyancy
    gradients, FP32, zero observed data (misfit 1/2 sum traces^2), each
    model's gradient summed over its shots on the device.  Checkpoint
    interval `every` (0: the largest that fits every member's ring in device
    memory).  A first call of the same length does the set-up (rings,
    checkpoints, interface term lists), as the first gradient of an iterative
    inversion would; the second is timed.  The stencils' tiling is the
    fixed `tiling` or, as in the batched forward leg, the fastest of the
    autotuner's pick and four tall tilings over 16 timed forward steps (the
    gradient starts from rest whatever these steps did)."""
    import ecsg_amd
    from ecsg_amd import model as M
    cands = [(8, 8), (8, 4), (4, 8), (4, 4)]
    tune = (len(cands) + 1) * (16 + 8)   # forward steps the tiling choice may take
    ms = M.batch_workload(nsteps=max(steps, tune) + 2)
    nodes = sum(m.total_nodes for m in ms)
    sims = [ecsg_amd.Simulation(m, precision="f32", check_energy=False, max_steps=max(steps, tune),
                                sxz_solver="jacobi_lds") for m in ms]
    for s in sims:
        s.set_checkpointing(every, "native")
    b = ecsg_amd.Batch(sims)
    tuned = b.autotune(reps=10)
    if tiling:
        b.set_tiling(*tiling)
    else:
        tiling = b.tune_steps([tuned] + [c for c in cands if c != tuned], steps=16)
    obs = [np.zeros((len(m.rec_loc), steps)) for m in ms]
    groups = [k // 4 for k in range(len(ms))]
    t0 = time.perf_counter()
    b.gradient(steps, obs, group=groups)
    setup = time.perf_counter() - t0
    t1 = time.perf_counter()
    Js = b.gradient(steps, obs, group=groups)
    el = time.perf_counter() - t1
    K = b.checkpoint_interval
    b.close()
    for s in sims:
        s.close()
    per_model = [float(sum(Js[k:k + 4])) for k in range(0, len(Js), 4)]
    return {"value": round(nodes * steps / el / 1e9, 3), "unit": "Gpt/s", "members": len(ms), "steps": steps,
            "checkpoint_every": K, "recomputed_steps": max(0, steps - K), "seconds": round(el, 3),
            "first_call_seconds": round(setup, 3), "ms_per_step": round(el / steps * 1e3, 3),
            "tiling": {"rows": tiling[0], "z_planes": tiling[1]},
            "misfit_per_model": [float("%.6e" % j) for j in per_model],
            "what": "ecsg_batch_gradient: forward with checkpoints + recomputation + transposed step + 9 Cij and "
                    "buoyancy gradients of all 64 members, one launch per phase; gradients summed per model on the "
                    "device"}


def stream_bound(nx: int, ny: int, planes: int, by: int = 4, zc: int = 1, reps: int = 10, tilings=(),
                 vel_planes: int = 0):
    """The streaming bound of the stress sweep's traversal: ecsg_stream_probe
    (18 FP32 arrays read, 6 written back in place, no arithmetic beyond one
    multiply-add) over the same nodes -- nx+1 by ny+1 nodes on `planes`
    planes, the same row pitch and plane stride -- in the stencil tiling
    (64 x by rows x zc planes, the sweeps' banded XCD order) and as one flat pass over the
    arrays' planes (halo rows and padding included); mean of `reps` launches
    each, HIP events on the launch stream.  `tilings`: further (by, zc)
    pairs, reported as "tiles_<by>x<zc>_ms".  vel_planes > 0: also the
    velocity sweep's traffic (10 arrays read, 3 written back, 52 B per node)
    over that many planes in the same tiling, "vel_tiles_ms"."""
    import torch
    import ecsg_amd
    L = ecsg_amd.load_library()
    nz = max(planes, vel_planes) - 1
    F = int(L.ecsg_stream_probe_floats(nx, ny, nz))
    buf = torch.zeros(F, dtype=torch.float32, device="cuda")
    st = torch.cuda.current_stream()
    P = (32 + nx + 2 + 31) // 32 * 32
    S2 = P * (ny + 3)
    out = {}
    runs = [("tiles", by, zc, 0, planes), ("flat", by, 0, 0, planes)]
    runs += [("tiles_%dx%d" % t, t[0], t[1], 0, planes) for t in tilings]
    if vel_planes > 0:
        runs.append(("vel_tiles", by, zc, 1, vel_planes))
    for name, byv, zcv, pat, npl in runs:
        bpn = 96 if pat == 0 else 52
        nbytes = bpn * npl * S2 if zcv == 0 else bpn * (nx + 1) * (ny + 1) * npl

        def launch():
            rc = L.ecsg_stream_probe(buf.data_ptr(), nx, ny, npl - 1, byv, zcv, pat, st.cuda_stream)
            if rc != 0:
                raise RuntimeError("ecsg_stream_probe failed (%d)" % rc)
        for _ in range(2):
            launch()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        for _ in range(reps):
            launch()
        e1.record(st)
        e1.synchronize()
        ms = e0.elapsed_time(e1) / reps
        out[name + "_ms"] = round(ms, 5)
        out[name + "_GBps"] = round(nbytes / (ms * 1e-3) / 1e9, 1)
    del buf
    torch.cuda.empty_cache()
    # the copy ceiling: the runtime's device-to-device copy of 4 GiB (bytes
    # read + bytes written per second)
    n = (4 << 30) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(2):
        dst.copy_(src)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(reps):
        dst.copy_(src)
    e1.record(st)
    e1.synchronize()
    ms = e0.elapsed_time(e1) / reps
    out["copy_GBps"] = round(2 * 4 * n / (ms * 1e-3) / 1e9, 1)
    del src, dst
    torch.cuda.empty_cache()
    return out


def cpu_baseline(m, steps: int, steps_1t: int, c0: bool):
    """The oracle (CPU restatement of the reference's loops, OpenMP) on the
    host cores, threads bound close (OMP_PROC_BIND=close, OMP_PLACES=cores):
    the same model for `steps` steps with every thread the process has
    (OMP_NUM_THREADS), then `steps_1t` steps on one thread; with c0, the
    81^3-node uniform configuration in full (600 steps, FP64, energy every
    step) on every thread.  Wall-clock times (perf_counter)."""
    os.environ["OMP_PROC_BIND"] = "close"      # read when the OpenMP runtime loads with the oracle
    os.environ["OMP_PLACES"] = "cores"
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle_lib import Oracle, lib
    L = lib("f32")
    threads = int(L.orc_num_threads())
    t0 = time.time()
    o = Oracle(m, prec="f32", check_energy=False, max_steps=steps + steps_1t + 2)
    log("[bench] cpu oracle ready in %.1fs (%d threads)" % (time.time() - t0, threads))
    o.run(1)  # first touch
    t = time.perf_counter()
    o.run(steps)
    el = time.perf_counter() - t
    out = {"threads": threads, "steps": steps, "seconds": el}
    if steps_1t > 0:
        L.orc_set_num_threads(1)
        t = time.perf_counter()
        o.run(steps_1t)
        out["one_thread"] = {"steps": steps_1t, "seconds": time.perf_counter() - t}
        L.orc_set_num_threads(threads)
    o.close()
    if c0:
        from ecsg_amd import model as M
        mc = M.uniform_model(nsteps=600)
        L64 = lib("f64")
        L64.orc_set_num_threads(threads)
        oc = Oracle(mc, prec="f64", check_energy=True, max_steps=600)
        t = time.perf_counter()
        oc.run(600)
        ec = time.perf_counter() - t
        e = oc.energy()
        oc.close()
        out["c0"] = {"steps": 600, "nodes": mc.total_nodes, "seconds": ec, "energy_last": float(e[-1])}
    return out


def profiled_traffic(workload: str):
    """HBM bytes per launch of the dominant kernel, measured with rocprofv3
    FETCH_SIZE / WRITE_SIZE (scaled by a copy of known size) and committed
    under profiles/; None when no profile of this workload exists."""
    path = os.path.join(ROOT, "profiles", "stencil_kernels_%s.md" % workload)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return int(line.split(":")[1])
    return None


def copy_calibration():
    """1 GiB device copies of known size, so that FETCH_SIZE / WRITE_SIZE of
    the stencil kernels can be scaled: torch's copy (plain accesses), the
    library's FP32 copy with the stencils' non-temporal loads and stores
    (hints 3) and without them (hints 0), disjoint and in place (each value
    read and written back, as the stencils update their fields); three
    launches each."""
    import torch
    import ecsg_amd
    L = ecsg_amd.load_library()
    n = (1 << 30) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    for in_place in (False, True):
        for hints in (0, 3):
            for _ in range(3):
                rc = L.ecsg_copy_probe(dst.data_ptr(), dst.data_ptr() if in_place else src.data_ptr(), n, hints, None)
                if rc != 0:
                    raise RuntimeError("ecsg_copy_probe failed (%d)" % rc)
    torch.cuda.synchronize()
    assert bool(torch.equal(dst, src))
    print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "launches": 3,
                      "kernels": ["__amd_rocclr_copyBuffer", "k_copy_probe<0, false>", "k_copy_probe<3, false>",
                                  "k_copy_probe<0, true>", "k_copy_probe<3, true>"]}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--preroll", type=int, default=1000,
                    help="untimed steps after the warm-up, before the timed steps (the wavefield crosses the "
                         "interfaces; default: half of the 2000-step run)")
    ap.add_argument("--workload", default="three_grid")
    ap.add_argument("--fp64-steps", type=int, default=500)
    ap.add_argument("--cpu-steps", type=int, default=120)
    ap.add_argument("--cpu-steps-1t", type=int, default=8, help="steps of the CPU baseline's one-thread leg")
    ap.add_argument("--no-cpu-c0", action="store_true", help="skip the CPU baseline's full 81^3 uniform run")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--prof-steps", type=int, default=0,
                    help="extra steps timing every kernel kind with events (log only)")
    ap.add_argument("--solver", default="jacobi", choices=["jacobi", "chebyshev"],
                    help="vx-sxz interface solver (jacobi = the reference's sweeps)")
    ap.add_argument("--adjoint-steps", type=int, default=400,
                    help="steps of the timed forward+adjoint gradient (0: skip)")
    ap.add_argument("--ckpt-every", type=int, default=0,
                    help="checkpoint interval of the adjoint leg (0: the largest that fits in device memory)")
    ap.add_argument("--ckpt-format", default="native", choices=["native", "fp16", "bf16"])
    ap.add_argument("--adjoint-mem-steps", type=int, default=40,
                    help="steps of a second gradient run that keeps every state in memory (0: skip)")
    ap.add_argument("--batch-steps", type=int, default=400, help="steps of the batched-shots leg (0: skip)")
    ap.add_argument("--batch-solo-steps", type=int, default=50)
    ap.add_argument("--batch-tiling", default=None,
                    help="fixed tiling ROWSxPLANES of the batched leg (default: chosen from timed steps)")
    ap.add_argument("--batch-grad-steps", type=int, default=200,
                    help="steps of each batch member's gradient (0: skip)")
    ap.add_argument("--batch-grad-every", type=int, default=0,
                    help="checkpoint interval of the batched gradient (0: the largest that fits in device memory)")
    ap.add_argument("--no-stream-bound", action="store_true",
                    help="skip the streaming probe of the stress sweep's traversal")
    ap.add_argument("--stream-bound-grid", type=int, nargs=3, metavar=("NX", "NY", "PLANES"),
                    help="only run the streaming probe over this grid and print its JSON line")
    ap.add_argument("--stream-vel-planes", type=int, default=0,
                    help="with --stream-bound-grid: also the velocity sweep's traffic over this many planes")
    ap.add_argument("--stream-tilings", default="",
                    help="with --stream-bound-grid: further tilings to time, e.g. 4x1,4x2,1x8 (rows x planes)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device copy (scale factor for HBM byte counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.stream_bound_grid:
        tilings = [tuple(int(v) for v in t.split("x")) for t in args.stream_tilings.split(",") if t]
        print(json.dumps(stream_bound(*args.stream_bound_grid, tilings=tilings, vel_planes=args.stream_vel_planes)),
              flush=True)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU only; the multi-GPU sharding is a design "
              "(see the sharding section of the design notes), no N>1 figure is produced", file=sys.stderr)
        sys.exit(2)

    affinity = len(os.sched_getaffinity(0))   # before the CPU baseline's OpenMP runtime binds this thread
    total_steps = args.warmup + args.preroll + max(args.steps, args.fp64_steps, args.adjoint_steps) + 60
    t0 = time.time()
    m = build_model(args.workload, total_steps)
    nodes = m.total_nodes
    log("[bench] model %s: %d grids, %.1f M nodes, dt=%.4g, built in %.1fs"
        % (m.name, m.num_grids, nodes / 1e6, m.dt, time.time() - t0))

    el, live, stats, tr32, overlap, pre32 = gpu_run(m, "f32", args.warmup, args.preroll, args.steps,
                                                    profile=args.prof_steps > 0, prof_steps=args.prof_steps,
                                                    solver=args.solver)
    gpts = nodes * args.steps / el / 1e9
    ms_step = el / args.steps * 1e3
    log("[bench] f32: %.3f s for %d steps -> %.2f Gpt/s, %.3f ms/step" % (el, args.steps, gpts, ms_step))
    g0 = m.grids[0]
    # the timed launches of grid 0: all its planes, or, with the interface
    # overlap, its interior planes (velocities 0..nz-1; stresses 0..nz-1 in
    # mode 1, 0..nz-2 in mode 2)
    planes0 = {0: g0.n[2] + 1, 1: g0.n[2], 2: g0.n[2] - 1}[overlap]
    planes0v = g0.n[2] + 1 if overlap == 0 else g0.n[2]
    n0 = (g0.n[0] + 1) * (g0.n[1] + 1) * planes0
    n0v = (g0.n[0] + 1) * (g0.n[1] + 1) * planes0v
    t_vel, t_stress = live[0] * 1e-3, live[1] * 1e-3
    achieved = BYTES_STRESS_F32 * n0 / t_stress if t_stress > 0 else 0.0
    vel_bw = BYTES_VEL_F32 * n0v / t_vel if t_vel > 0 else 0.0
    log("[bench] timed steps, grid-0 launches avg ms: vel %.4f stress %.4f" % live)
    sb = None
    if not args.no_stream_bound:
        # in a child process: torch allocates the probe's buffer there, and
        # torch's HIP start-up needs a process in which HIP is not running yet
        import subprocess
        res = subprocess.run([sys.executable, "-u", os.path.abspath(__file__), "--stream-bound-grid",
                              str(g0.n[0]), str(g0.n[1]), str(planes0), "--stream-vel-planes", str(planes0v)],
                             capture_output=True, text=True, timeout=600)
        if res.returncode != 0:
            raise RuntimeError("streaming probe failed:\n" + res.stdout + res.stderr)
        sb = json.loads(res.stdout.strip().splitlines()[-1])
        sb["kernel_vs_tiles"] = round(sb["tiles_ms"] / live[1], 4) if live[1] > 0 else None
        sb["vel_kernel_vs_tiles"] = round(sb["vel_tiles_ms"] / live[0], 4) if live[0] > 0 else None
        log("[bench] streaming bound of the stress sweep's traversal: %s" % sb)

    fp64 = None
    if not args.no_fp64 and args.fp64_steps > 0:
        el64, live64, st64, tr64, _, pre64 = gpu_run(m, "f64", args.warmup, args.preroll, args.fp64_steps,
                                                     profile=args.prof_steps > 0, prof_steps=args.prof_steps,
                                                     solver=args.solver)
        g64 = nodes * args.fp64_steps / el64 / 1e9
        # the FP32 run's receiver traces against the FP64 run's, over the steps both made
        nc = min(tr32.shape[1], tr64.shape[1])
        d = float(np.linalg.norm(tr32[:, :nc] - tr64[:, :nc]))
        ref = float(np.linalg.norm(tr64[:, :nc]))
        fp32_err = d / ref if ref > 0 else d
        fp64 = {"value": round(g64, 3), "unit": "Gpt/s", "steps": args.fp64_steps,
                "ms_per_step": round(el64 / args.fp64_steps * 1e3, 4),
                "hbm_frac_step": round(2 * BYTES_STEP_F32 * g64 * 1e9 / HBM_PEAK, 4),
                "stress0_ms": round(live64[1], 5), "vel0_ms": round(live64[0], 5),
                "jacobi_sweeps_per_timed_step": st64,
                "fp32_trace_rel_l2": float("%.3e" % fp32_err), "traces_compared_steps": nc}
        log("[bench] f64: %.2f Gpt/s; FP32 traces vs FP64 over %d steps: %.3e relative L2" % (g64, nc, fp32_err))

    adjoint = None
    if args.adjoint_steps > 0:
        ea, Ja, sweeps, (K, ring, rec) = adjoint_run(m, args.adjoint_steps, args.ckpt_every, args.ckpt_format)
        ga = nodes * args.adjoint_steps / ea / 1e9
        ms_a = ea / args.adjoint_steps * 1e3
        adjoint = {"value": round(ga, 3), "unit": "Gpt/s", "steps": args.adjoint_steps,
                   "what": "forward with checkpoints + recomputation (one level: all segments but the last; two "
                           "levels when the ring is shorter than the interval) + adjoint step "
                           "+ gradient of the 9 Cij and buoyancy, FP32, misfit 1/2 sum vx_rec^2",
                   "checkpoint_every": K, "ring_steps": ring, "checkpoint_format": args.ckpt_format,
                   "recomputed_steps": rec,
                   "ms_per_step": round(ms_a, 4),
                   "cost_vs_forward_f32": round(ms_a / ms_step, 3),
                   "cost_vs_forward_f64": round(ms_a / fp64["ms_per_step"], 3) if fp64 else None,
                   "adjoint_jacobi_sweeps_per_step": round(sweeps / args.adjoint_steps / max(1, m.num_grids - 1), 2)}
        log("[bench] fwd+adjoint: %.2f Gpt/s, %.3f ms/step (%.2fx the FP32 forward)" % (ga, ms_a, ms_a / ms_step))
        if args.adjoint_mem_steps > 0:
            # a run short enough for every state to stay in device memory
            em, _, _, (Km, _, recm) = adjoint_run(m, args.adjoint_mem_steps, 0, args.ckpt_format)
            ms_m = em / args.adjoint_mem_steps * 1e3
            adjoint["in_memory"] = {"steps": args.adjoint_mem_steps, "checkpoint_every": Km,
                                    "recomputed_steps": recm,
                                    "ms_per_step": round(ms_m, 4), "cost_vs_forward_f32": round(ms_m / ms_step, 3),
                                    "cost_vs_forward_f64": round(ms_m / fp64["ms_per_step"], 3) if fp64 else None}
            log("[bench] fwd+adjoint in memory (%d steps): %.3f ms/step (%.2fx the FP32 forward)"
                % (args.adjoint_mem_steps, ms_m, ms_m / ms_step))

    batch = None
    if args.batch_steps > 0:
        bt = tuple(int(x) for x in args.batch_tiling.split("x")) if args.batch_tiling else None
        batch = batch_run(args.batch_steps, args.batch_solo_steps, args.warmup, bt)
        log("[bench] batch: %s" % batch)
        if args.batch_grad_steps > 0:
            batch["gradient"] = batch_gradient_run(args.batch_grad_steps, args.batch_grad_every, bt)
            batch["gradient"]["cost_vs_batched_forward"] = round(batch["value"] / batch["gradient"]["value"], 3)
            log("[bench] batch gradients: %s" % batch["gradient"])

    cpu = None
    if not args.no_cpu:
        cb = cpu_baseline(m, args.cpu_steps, args.cpu_steps_1t, not args.no_cpu_c0)
        ce, threads = cb["seconds"], cb["threads"]
        try:
            import psutil
            phys = psutil.cpu_count(logical=False)
        except Exception:
            phys = None
        cpu = {"value": round(nodes * args.cpu_steps / ce / 1e9, 5), "unit": "Gpt/s", "cores": threads,
               "kind": "port",
               "sample": "the same %s model, FP32 restatement of the reference's loops (oracle/ecsg_oracle.c, "
                         "OpenMP, OMP_PROC_BIND=close), %d steps after one warm-up step, %.1f s wall on %d threads"
                         % (m.name, args.cpu_steps, ce, threads),
               "host_physical_cores": phys, "affinity_cpus": affinity}
        if "one_thread" in cb:
            o1 = cb["one_thread"]
            cpu["one_thread"] = {"value": round(nodes * o1["steps"] / o1["seconds"] / 1e9, 5), "unit": "Gpt/s",
                                 "steps": o1["steps"], "seconds": round(o1["seconds"], 2)}
        if "c0" in cb:
            c = cb["c0"]
            cpu["c0_full"] = {"value": round(c["nodes"] * c["steps"] / c["seconds"] / 1e9, 5), "unit": "Gpt/s",
                              "steps": c["steps"], "seconds": round(c["seconds"], 2), "threads": threads,
                              "what": "81^3-node uniform configuration in full, FP64, energy every step"}
        log("[bench] cpu: %s" % cpu)

    out = {
        "metric": METRIC,
        "value": round(gpts, 3),
        "unit": "Gpt/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": m.name, "grids": [list(g.n) for g in m.grids], "steps_h": [g.h[0] for g in m.grids],
                   "nodes": nodes, "ratio": m.ratios(), "receivers": int(len(m.rec_loc)),
                   "energy_check": False, "sxz_solver": args.solver, "parallelism": "single GPU"},
        "hbm_frac_step": round(BYTES_STEP_F32 * gpts * 1e9 / HBM_PEAK, 4),
        "roofline": {"bound": "hbm", "kernel": "k_stress<float> on planes 0..%d of the %dx%dx%d-node grid"
                     % ((planes0 - 1,) + tuple(x + 1 for x in g0.n)),
                     "achieved": round(achieved / 1e9, 1), "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK, 4), "traffic": profiled_traffic(args.workload),
                     "bytes_per_launch": BYTES_STRESS_F32 * n0, "avg_launch_ms": round(live[1], 5),
                     "vel_kernel_GBps": round(vel_bw / 1e9, 1), "interface_overlap": overlap,
                     "stream_bound": sb},
        "fp64": fp64,
        "adjoint": adjoint,
        "batch": batch,
        "preroll": pre32,
        "jacobi_sweeps_per_timed_step": stats,
        "fp32_trace_rel_l2_vs_fp64": fp64["fp32_trace_rel_l2"] if fp64 else None,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
