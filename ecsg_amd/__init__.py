"""ecsg_amd -- MI355X-native ECSG time-stepping path.

The compute path is libecsg_hip.so (hand-written gfx950 HIP kernels behind
the C ABI of include/ecsg_hip.h).  This module binds that ABI with ctypes and
mirrors the reference's driver vocabulary: a Simulation owns the device
state of all composite grids (main.cpp:66-155) and advances it with the
phase order of main.cpp:165-261.  There is no CPU fallback: if the library
is missing, importing works but creating a Simulation raises.
"""
from __future__ import annotations

import atexit
import ctypes as C
import os
import time
import weakref
from typing import Optional

import numpy as np

from . import model  # noqa: F401

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libecsg_hip.so")

F32, F64 = 0, 1
MAXG = 8

# symbols declared in include/ecsg_hip.h
ABI_SYMBOLS = (
    "ecsg_create", "ecsg_destroy", "ecsg_last_error", "ecsg_version", "ecsg_set_params", "ecsg_set_source",
    "ecsg_set_receivers", "ecsg_update_vel", "ecsg_record_receivers", "ecsg_add_source", "ecsg_update_ghost_vz",
    "ecsg_compute_energy", "ecsg_update_stress", "ecsg_update_ghost_sxz", "ecsg_update_ghost_syz", "ecsg_end_step",
    "ecsg_compute_res", "ecsg_step_n", "ecsg_steps_done", "ecsg_synchronize", "ecsg_field_dims", "ecsg_get_field",
    "ecsg_set_field", "ecsg_get_traces", "ecsg_get_energy", "ecsg_jacobi_stats", "ecsg_profile_kernels",
    "ecsg_reset", "ecsg_set_checkpointing", "ecsg_gradient", "ecsg_get_gradient", "ecsg_adjoint_step",
    "ecsg_get_adjoint_field", "ecsg_set_adjoint_field", "ecsg_adjoint_stats", "ecsg_autotune", "ecsg_set_fused",
    "ecsg_batch_create", "ecsg_batch_destroy", "ecsg_batch_last_error", "ecsg_batch_step_n", "ecsg_batch_synchronize",
    "ecsg_batch_set_tiling", "ecsg_batch_autotune", "ecsg_field_avgabs", "ecsg_sbp_residuals",
    "ecsg_checkpoint_interval", "ecsg_copy_probe", "ecsg_stream_probe_floats", "ecsg_stream_probe",
    "ecsg_set_overlap", "ecsg_overlap_active", "ecsg_step_n_timed", "ecsg_set_barrier_timeout", "ecsg_set_jacobi_tol",
    "ecsg_set_params_staged", "ecsg_cfl_dt", "ecsg_cfl_dt_noalloc", "ecsg_batch_gradient",
    "ecsg_batch_checkpoint_interval", "ecsg_set_ring_steps", "ecsg_checkpoint_levels",
)

CKPT_FORMATS = {"native": 0, "fp16": 1, "bf16": 2}


class EcsgConfig(C.Structure):
    _fields_ = [
        ("num_grids", C.c_int),
        ("n", (C.c_int * 3) * MAXG),
        ("h", (C.c_double * 3) * MAXG),
        ("dt", C.c_double),
        ("radius", C.c_int),
        ("top_bc", C.c_int),
        ("imode", C.c_int),
        ("check_energy", C.c_int),
        ("precision", C.c_int),
        ("max_steps", C.c_int),
        ("vy_mirror_all", C.c_int),
        ("sxz_solver", C.c_int),
        ("ii_vz_xend", C.c_int),
    ]


class EcsgError(RuntimeError):
    pass


_LIB = None
# open handles, closed at interpreter exit before the HIP runtime is torn down
_OPEN = weakref.WeakSet()


@atexit.register
def _close_open_handles():
    for kind in (Batch, Simulation):
        for obj in [o for o in list(_OPEN) if isinstance(o, kind)]:
            try:
                obj.close()
            except Exception:
                pass


def load_library(path: Optional[str] = None):
    """Load libecsg_hip.so; raises if it has not been built."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise EcsgError("libecsg_hip.so not built (%s); run python -m ecsg_amd.build" % p)
    L = C.CDLL(p)
    vp, i, d = C.c_void_p, C.c_int, C.c_double
    P = C.POINTER
    L.ecsg_create.argtypes = [P(EcsgConfig), P(vp)]
    L.ecsg_destroy.argtypes = [vp]
    L.ecsg_last_error.argtypes = [vp]
    L.ecsg_last_error.restype = C.c_char_p
    L.ecsg_version.restype = C.c_char_p
    L.ecsg_set_params.argtypes = [vp, i, P(d), P(P(d)), i]
    L.ecsg_set_source.argtypes = [vp, i, i, i, i, i, P(C.c_float), i]
    L.ecsg_set_receivers.argtypes = [vp, i, P(i)]
    for f in ("ecsg_update_vel", "ecsg_update_ghost_vz", "ecsg_update_stress", "ecsg_update_ghost_sxz",
              "ecsg_update_ghost_syz"):
        getattr(L, f).argtypes = [vp, i]
    for f in ("ecsg_record_receivers", "ecsg_add_source", "ecsg_end_step", "ecsg_synchronize", "ecsg_steps_done"):
        getattr(L, f).argtypes = [vp]
    L.ecsg_compute_energy.argtypes = [vp, P(d)]
    L.ecsg_compute_res.argtypes = [vp, i, P(d)]
    L.ecsg_field_avgabs.argtypes = [vp, i, i, P(d)]
    L.ecsg_sbp_residuals.argtypes = [vp, i, P(d)]
    L.ecsg_step_n.argtypes = [vp, i, i]
    L.ecsg_copy_probe.argtypes = [vp, vp, C.c_uint64, i, vp]
    L.ecsg_stream_probe_floats.argtypes = [i, i, i]
    L.ecsg_stream_probe_floats.restype = C.c_uint64
    L.ecsg_stream_probe.argtypes = [vp, i, i, i, i, i, i, vp]
    L.ecsg_set_overlap.argtypes = [vp, i]
    L.ecsg_overlap_active.argtypes = [vp]
    L.ecsg_step_n_timed.argtypes = [vp, i, P(d)]
    L.ecsg_set_barrier_timeout.argtypes = [vp, i]
    L.ecsg_set_jacobi_tol.argtypes = [vp, C.c_double, C.c_double]
    L.ecsg_set_params_staged.argtypes = [vp, i, P(d), P(P(d))]
    L.ecsg_cfl_dt.argtypes = [P(EcsgConfig), P(P(d)), P((P(d) * 9)), i, d, P(d)]
    L.ecsg_cfl_dt_noalloc.argtypes = [P(EcsgConfig), P(vp), P(vp * 9), i, d, vp, vp]
    L.ecsg_field_dims.argtypes = [vp, i, i, P(i)]
    L.ecsg_get_field.argtypes = [vp, i, i, P(d)]
    L.ecsg_set_field.argtypes = [vp, i, i, P(d)]
    L.ecsg_get_traces.argtypes = [vp, P(d)]
    L.ecsg_get_energy.argtypes = [vp, P(d)]
    L.ecsg_jacobi_stats.argtypes = [vp, i, P(C.c_longlong), P(i), P(i)]
    L.ecsg_profile_kernels.argtypes = [vp, i, P(d), P(i)]
    L.ecsg_reset.argtypes = [vp]
    L.ecsg_set_checkpointing.argtypes = [vp, i, i]
    L.ecsg_checkpoint_interval.argtypes = [vp]
    L.ecsg_set_ring_steps.argtypes = [vp, i]
    L.ecsg_checkpoint_levels.argtypes = [vp, P(i), P(i), P(C.c_longlong)]
    L.ecsg_gradient.argtypes = [vp, i, P(d), P(d)]
    L.ecsg_get_gradient.argtypes = [vp, i, i, P(d)]
    L.ecsg_adjoint_step.argtypes = [vp]
    L.ecsg_get_adjoint_field.argtypes = [vp, i, i, P(d)]
    L.ecsg_set_adjoint_field.argtypes = [vp, i, i, P(d)]
    L.ecsg_adjoint_stats.argtypes = [vp, P(C.c_longlong)]
    L.ecsg_autotune.argtypes = [vp, i, P(i), P(i)]
    L.ecsg_set_fused.argtypes = [vp, i]
    L.ecsg_batch_create.argtypes = [P(vp), i, P(vp)]
    L.ecsg_batch_destroy.argtypes = [vp]
    L.ecsg_batch_last_error.argtypes = [vp]
    L.ecsg_batch_last_error.restype = C.c_char_p
    L.ecsg_batch_step_n.argtypes = [vp, i, i]
    L.ecsg_batch_synchronize.argtypes = [vp]
    L.ecsg_batch_set_tiling.argtypes = [vp, i, i]
    L.ecsg_batch_autotune.argtypes = [vp, i, P(i), P(i)]
    L.ecsg_batch_gradient.argtypes = [vp, i, P(P(d)), P(d), P(i)]
    L.ecsg_batch_checkpoint_interval.argtypes = [vp]
    if path is None:
        _LIB = L
    return L


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def cfl_dt(m, frac: float = 0.9) -> float:
    """ecsg_cfl_dt: frac x min over grids of h_min / (vp sqrt 3), vp^2 the
    largest max(C11,C22,C33) x b, reduced on the device in FP64."""
    L = load_library()
    cfg = EcsgConfig()
    cfg.num_grids = m.num_grids
    keep = []
    buoys = (C.POINTER(C.c_double) * m.num_grids)()
    stiffs = ((C.POINTER(C.c_double) * 9) * m.num_grids)()
    for g, G in enumerate(m.grids):
        for k in range(3):
            cfg.n[g][k] = int(G.n[k])
            cfg.h[g][k] = float(G.h[k])
        b = np.ascontiguousarray(G.buoy, np.float64)
        keep.append(b)
        buoys[g] = _dp(b)
        for k, a in enumerate(list(G.stiff)[:9]):
            a = np.ascontiguousarray(a, np.float64)
            keep.append(a)
            stiffs[g][k] = _dp(a)
    out = C.c_double()
    rc = L.ecsg_cfl_dt(C.byref(cfg), buoys, stiffs, int(m.iso), float(frac), C.byref(out))
    if rc != 0:
        raise EcsgError("ecsg_cfl_dt failed (%d)" % rc)
    return out.value


class Simulation:
    """Device state of one composite model (all grids), advanced on one stream.

    precision: "f32" (wavefields and parameters in float; interface solves,
    energy and residuals in double) or "f64" (everything double).
    """

    def __init__(self, m, precision: str = "f64", check_energy: bool = True, max_steps: Optional[int] = None,
                 vy_mirror_all: bool = False, sxz_solver: str = "jacobi", ii_vz_xend: bool = False,
                 staged: bool = False, jacobi_tol=None):
        self.L = load_library()
        self.model = m
        cfg = EcsgConfig()
        cfg.num_grids = m.num_grids
        for g, G in enumerate(m.grids):
            for k in range(3):
                cfg.n[g][k] = int(G.n[k])
                cfg.h[g][k] = float(G.h[k])
        cfg.dt = float(m.dt)
        cfg.radius = m.radius
        cfg.top_bc = m.top_bc
        cfg.imode = m.imode
        cfg.check_energy = int(check_energy)
        cfg.precision = {"f32": F32, "f64": F64}[precision]
        cfg.max_steps = int(m.nsteps if max_steps is None else max_steps)
        cfg.vy_mirror_all = int(vy_mirror_all)
        cfg.sxz_solver = {"jacobi": 0, "chebyshev": 1, "jacobi_lds": 2, "direct": 3}[sxz_solver]
        cfg.ii_vz_xend = int(ii_vz_xend)
        self.max_steps = cfg.max_steps
        self.precision = precision
        self._staged = set()
        self.h = C.c_void_p()
        rc = self.L.ecsg_create(C.byref(cfg), C.byref(self.h))
        if rc != 0:
            msg = self.L.ecsg_last_error(self.h).decode() if self.h else ""
            self.L.ecsg_destroy(self.h)
            self.h = None
            raise EcsgError("ecsg_create failed (%d): %s" % (rc, msg))
        _OPEN.add(self)
        if jacobi_tol is not None:
            self.set_jacobi_tol(*jacobi_tol)
        for g in range(m.num_grids):
            if staged:
                self.set_params_staged(g)
            else:
                self._set_params(g)
        if m.src_trace is not None and len(m.src_trace):
            tr = np.ascontiguousarray(m.src_trace, np.float32)
            self._chk(self.L.ecsg_set_source(self.h, m.src_grid, *map(int, m.src_loc), m.src_type,
                                             tr.ctypes.data_as(C.POINTER(C.c_float)), len(tr)))
        rec = np.ascontiguousarray(m.rec_loc, np.int32).reshape(-1, 4)
        self.nrec = len(rec)
        self._chk(self.L.ecsg_set_receivers(self.h, len(rec), rec.ctypes.data_as(C.POINTER(C.c_int))))

    def _chk(self, rc):
        if rc != 0:
            raise EcsgError("ecsg error %d: %s" % (rc, self.L.ecsg_last_error(self.h).decode()))

    def _set_params(self, g):
        G = self.model.grids[g]
        st = [np.ascontiguousarray(a, np.float64) for a in self.model.stiff9(g)]
        arr = (C.POINTER(C.c_double) * 9)(*[_dp(a) for a in st])
        b = np.ascontiguousarray(G.buoy, np.float64)
        self._chk(self.L.ecsg_set_params(self.h, g, _dp(b), arr, int(self.model.iso)))

    def set_params_staged(self, g, stiff=None, buoy=None):
        """ecsg_set_params_staged: the nine staged Fields (default: the model's
        parameters staged as InitParField does, isotropic entries aliased)."""
        from .model import staged_params
        st = staged_params(self.model, g) if stiff is None else stiff
        keep = {}
        ptrs = []
        for a in st:
            if id(a) not in keep:
                keep[id(a)] = np.ascontiguousarray(a, np.float64)
            ptrs.append(_dp(keep[id(a)]))
        arr = (C.POINTER(C.c_double) * 9)(*ptrs)
        b = np.ascontiguousarray(self.model.grids[g].buoy if buoy is None else buoy, np.float64)
        self._chk(self.L.ecsg_set_params_staged(self.h, g, _dp(b), arr))
        self._staged.add(g)

    def set_barrier_timeout(self, ms: int):
        """Longest wait at one grid barrier of the cooperative Jacobi solve."""
        self._chk(self.L.ecsg_set_barrier_timeout(self.h, int(ms)))

    def set_jacobi_tol(self, tol_cap: float, tol_fac: float):
        """Exit tolerance min(tol_cap, tol_fac * AvgAbs(B)) of the forward
        Jacobi sweeps (defaults: 1e-3, 1e-4 in f32; 1e-12, 1e-5 in f64)."""
        self._chk(self.L.ecsg_set_jacobi_tol(self.h, float(tol_cap), float(tol_fac)))

    def close(self):
        if getattr(self, "h", None):
            self.L.ecsg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --- time stepping -----------------------------------------------------
    def step(self, n: int = 1, graph: bool = True):
        self._chk(self.L.ecsg_step_n(self.h, int(n), int(graph)))

    def sync(self):
        self._chk(self.L.ecsg_synchronize(self.h))

    @property
    def steps_done(self) -> int:
        return self.L.ecsg_steps_done(self.h)

    # per-phase entry points (main.cpp:179-251)
    def update_vel(self, g): self._chk(self.L.ecsg_update_vel(self.h, g))
    def record_receivers(self): self._chk(self.L.ecsg_record_receivers(self.h))
    def add_source(self): self._chk(self.L.ecsg_add_source(self.h))
    def ghost_vz(self, k): self._chk(self.L.ecsg_update_ghost_vz(self.h, k))
    def update_stress(self, g): self._chk(self.L.ecsg_update_stress(self.h, g))
    def ghost_sxz(self, k): self._chk(self.L.ecsg_update_ghost_sxz(self.h, k))
    def ghost_syz(self, k): self._chk(self.L.ecsg_update_ghost_syz(self.h, k))
    def end_step(self): self._chk(self.L.ecsg_end_step(self.h))

    def energy_now(self) -> float:
        e = C.c_double()
        self._chk(self.L.ecsg_compute_energy(self.h, C.byref(e)))
        return e.value

    def residuals(self, which: int) -> np.ndarray:
        out = np.zeros(max(1, self.model.num_grids - 1))
        self._chk(self.L.ecsg_compute_res(self.h, which, _dp(out)))
        return out

    def sbp_residuals(self, g: int = 0) -> np.ndarray:
        """TestSBP (test_sbp.cpp:5-199): the seven summation-by-parts residuals of grid g."""
        out = np.zeros(7)
        self._chk(self.L.ecsg_sbp_residuals(self.h, g, _dp(out)))
        return out

    def avg_abs(self, g: int, sel: int) -> float:
        """Field::AvgAbs (field.cpp:215-228) of wavefield sel on grid g."""
        v = C.c_double()
        self._chk(self.L.ecsg_field_avgabs(self.h, g, sel, C.byref(v)))
        return v.value

    # --- data ----------------------------------------------------------------
    def dims(self, g, sel):
        l3 = (C.c_int * 3)()
        self._chk(self.L.ecsg_field_dims(self.h, g, sel, l3))
        return (l3[2], l3[1], l3[0])

    def get(self, g, sel) -> np.ndarray:
        out = np.empty(self.dims(g, sel), np.float64)
        self._chk(self.L.ecsg_get_field(self.h, g, sel, _dp(out)))
        return out

    def set(self, g, sel, a):
        a = np.ascontiguousarray(a, np.float64).reshape(self.dims(g, sel))
        self._chk(self.L.ecsg_set_field(self.h, g, sel, _dp(a)))

    def traces(self) -> np.ndarray:
        n = min(self.steps_done, self.max_steps)
        out = np.empty((self.nrec, n), np.float64)
        self._chk(self.L.ecsg_get_traces(self.h, _dp(out)))
        return out

    def energy(self) -> np.ndarray:
        n = min(self.steps_done, self.max_steps)
        out = np.empty(n, np.float64)
        self._chk(self.L.ecsg_get_energy(self.h, _dp(out)))
        return out

    def jacobi_stats(self, k: int = 0):
        t, l, c = C.c_longlong(), C.c_int(), C.c_int()
        self._chk(self.L.ecsg_jacobi_stats(self.h, k, C.byref(t), C.byref(l), C.byref(c)))
        return t.value, l.value, c.value

    def step_timed(self, nsteps: int):
        """nsteps steps (no graph) timing the grid-0 velocity and stress
        stencil launches with HIP events; returns their average ms and
        blocks until done."""
        ms = (C.c_double * 2)()
        self._chk(self.L.ecsg_step_n_timed(self.h, int(nsteps), ms))
        return ms[0], ms[1]

    def profile(self, nsteps: int):
        ms = (C.c_double * 6)()
        cnt = (C.c_int * 6)()
        self._chk(self.L.ecsg_profile_kernels(self.h, int(nsteps), ms, cnt))
        return list(ms), list(cnt)

    # --- adjoint ------------------------------------------------------------
    def reset(self):
        """Zero wavefields, interface warm starts, traces and the step counter."""
        self._chk(self.L.ecsg_reset(self.h))

    def set_checkpointing(self, every: int = 50, fmt: str = "native"):
        """every = 0: the largest interval that fits in free device memory."""
        self._chk(self.L.ecsg_set_checkpointing(self.h, int(every), CKPT_FORMATS[fmt]))

    @property
    def checkpoint_interval(self) -> int:
        """The interval the last gradient call used."""
        return int(self.L.ecsg_checkpoint_interval(self.h))

    def set_ring_steps(self, steps: int = 0):
        """Longest ring of saved states a gradient may hold (0: as long as
        the interval, or as fits in device memory); a shorter ring than the
        interval means two checkpoint levels."""
        self._chk(self.L.ecsg_set_ring_steps(self.h, int(steps)))

    @property
    def checkpoint_levels(self):
        """(interval, ring length, recomputed forward steps) of the last
        gradient call."""
        k1, k2, rec = C.c_int(), C.c_int(), C.c_longlong()
        self._chk(self.L.ecsg_checkpoint_levels(self.h, C.byref(k1), C.byref(k2), C.byref(rec)))
        return k1.value, k2.value, rec.value

    def gradient(self, nsteps: int, observed: np.ndarray) -> float:
        """Misfit 1/2 sum (traces - observed)^2 over nsteps steps from rest and
        its gradient (read with grad()).  observed: nrec x nsteps."""
        obs = np.ascontiguousarray(observed, np.float64).reshape(self.nrec, nsteps) if self.nrec else np.zeros(1)
        J = C.c_double()
        self._chk(self.L.ecsg_gradient(self.h, int(nsteps), _dp(obs), C.byref(J)))
        return J.value

    def grad(self, g: int, sel: int) -> np.ndarray:
        """dJ/d(input) at primary nodes, (nz+1, ny+1, nx+1): sel 9 = buoyancy,
        10 + m = stiff[m] of set_params; for a grid given staged Fields, the
        gradient w.r.t. staged Field m in its own shape."""
        G = self.model.grids[g]
        if sel >= 10 and g in self._staged:
            out = np.empty(self.dims(g, sel), np.float64)
        else:
            out = np.empty((G.n[2] + 1, G.n[1] + 1, G.n[0] + 1), np.float64)
        self._chk(self.L.ecsg_get_gradient(self.h, g, sel, _dp(out)))
        return out

    def adjoint_step(self):
        self._chk(self.L.ecsg_adjoint_step(self.h))

    def get_adjoint(self, g, sel) -> np.ndarray:
        out = np.empty(self.dims(g, sel), np.float64)
        self._chk(self.L.ecsg_get_adjoint_field(self.h, g, sel, _dp(out)))
        return out

    def set_adjoint(self, g, sel, a):
        a = np.ascontiguousarray(a, np.float64).reshape(self.dims(g, sel))
        self._chk(self.L.ecsg_set_adjoint_field(self.h, g, sel, _dp(a)))

    def adjoint_sweeps(self) -> int:
        n = C.c_longlong()
        self._chk(self.L.ecsg_adjoint_stats(self.h, C.byref(n)))
        return n.value

    def set_fused(self, on: bool = True):
        """Fused velocity+stress pass in step() (identical results)."""
        self._chk(self.L.ecsg_set_fused(self.h, int(bool(on))))

    def set_overlap(self, on=True):
        """Interface ghost updates on a second stream beside the stencil
        planes away from the interfaces (identical results).  True / False,
        or None for the default: on when the steps use the multi-workgroup
        Jacobi solve (vx-sxz iterates above 2048 entries)."""
        self._chk(self.L.ecsg_set_overlap(self.h, -1 if on is None else int(bool(on))))

    @property
    def overlap_active(self) -> bool:
        return self.overlap_mode > 0

    @property
    def overlap_mode(self) -> int:
        """0: serial step; 1: interface overlap; 2: overlap with grid 0's
        stress interior also beside the sxz/syz ghost solves."""
        return int(self.L.ecsg_overlap_active(self.h))

    def autotune(self, reps: int = 5):
        """Per-grid stencil tiling chosen by timing (state preserved):
        returns [(rows per workgroup, z planes per thread), ...]."""
        ng = self.model.num_grids
        by, zc = (C.c_int * ng)(), (C.c_int * ng)()
        self._chk(self.L.ecsg_autotune(self.h, int(reps), by, zc))
        return [(by[g], zc[g]) for g in range(ng)]


class Batch:
    """Many Simulations (shots, models of any sizes) advanced together with one
    launch per phase of the time step (ecsg_batch_*).  The members keep their
    state and accessors and must outlive the batch."""

    def __init__(self, sims):
        self.L = load_library()
        self.sims = list(sims)
        arr = (C.c_void_p * len(self.sims))(*[s.h.value for s in self.sims])
        self.h = C.c_void_p()
        rc = self.L.ecsg_batch_create(arr, len(self.sims), C.byref(self.h))
        if rc != 0:
            msg = self.L.ecsg_batch_last_error(self.h).decode() if self.h else ""
            self.L.ecsg_batch_destroy(self.h)
            self.h = None
            raise EcsgError("ecsg_batch_create failed (%d): %s" % (rc, msg))
        _OPEN.add(self)

    def _chk(self, rc):
        if rc != 0:
            raise EcsgError("ecsg batch error %d: %s" % (rc, self.L.ecsg_batch_last_error(self.h).decode()))

    def step(self, n: int = 1, graph: bool = True):
        self._chk(self.L.ecsg_batch_step_n(self.h, int(n), int(graph)))

    def sync(self):
        self._chk(self.L.ecsg_batch_synchronize(self.h))

    def set_tiling(self, by: int, zc: int):
        self._chk(self.L.ecsg_batch_set_tiling(self.h, int(by), int(zc)))

    def autotune(self, reps: int = 5):
        by, zc = C.c_int(), C.c_int()
        self._chk(self.L.ecsg_batch_autotune(self.h, int(reps), C.byref(by), C.byref(zc)))
        return by.value, zc.value

    def tune_steps(self, candidates, steps: int = 16):
        """Times `steps` real batched steps with each (rows, z_planes) tiling
        of `candidates` and keeps the fastest; returns it.  Unlike autotune,
        which times the stencil launches alone on a saved copy of the state,
        this sees the interface solves that run beside them -- and advances
        every member by len(candidates) x (steps + 8) time steps (8 per
        candidate for the graph capture, untimed)."""
        best = None
        for by, zc in candidates:
            self.set_tiling(by, zc)
            self.step(8)
            self.sync()
            t = time.perf_counter()
            self.step(steps)
            self.sync()
            el = time.perf_counter() - t
            if best is None or el < best[0]:
                best = (el, int(by), int(zc))
        if best is None:
            raise EcsgError("tune_steps needs at least one candidate tiling")
        self.set_tiling(best[1], best[2])
        return best[1], best[2]

    def gradient(self, nsteps: int, observed, group=None) -> np.ndarray:
        """ecsg_batch_gradient: every member's misfit and gradient (read with
        each member's grad()) with one launch per phase; observed[m] is member
        m's nrec x nsteps traces.  group[m]: member m's model; the first member
        of each group then holds the model's summed gradient.  Returns the
        misfits."""
        obs = []
        for sm, o in zip(self.sims, observed):
            obs.append(np.ascontiguousarray(o, np.float64).reshape(sm.nrec, nsteps) if sm.nrec else np.zeros(1))
        ptrs = (C.POINTER(C.c_double) * len(obs))(*[_dp(o) for o in obs])
        mis = np.zeros(len(self.sims))
        grp = None
        if group is not None:
            grp = np.ascontiguousarray(group, np.int32)
        self._chk(self.L.ecsg_batch_gradient(self.h, int(nsteps), ptrs, _dp(mis),
                                             grp.ctypes.data_as(C.POINTER(C.c_int)) if grp is not None else None))
        return mis

    @property
    def checkpoint_interval(self) -> int:
        return int(self.L.ecsg_batch_checkpoint_interval(self.h))

    def close(self):
        if getattr(self, "h", None):
            self.L.ecsg_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
