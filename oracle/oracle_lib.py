"""ctypes binding of the CPU restatement (liboracle_f64.so / liboracle_f32.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product path never loads it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def build(quiet: bool = True) -> None:
    """Compile the restatement with its Makefile (gcc, OpenMP)."""
    out = subprocess.run(["make", "-C", HERE, "-B"], capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + out.stdout + out.stderr)
    if not quiet:
        print(out.stdout)


def _load(prec: str):
    path = os.path.join(HERE, "liboracle_%s.so" % prec)
    if not os.path.exists(path):
        build()
    lib = C.CDLL(path)
    vp, i, d = C.c_void_p, C.c_int, C.c_double
    P = C.POINTER
    lib.orc_create.restype = vp
    lib.orc_create.argtypes = [i, P(C.c_int), P(C.c_double), i, i, i, i, d, i, i]
    lib.orc_destroy.argtypes = [vp]
    lib.orc_set_params.argtypes = [vp, i, P(C.c_double), P(P(C.c_double)), i]
    lib.orc_set_source.argtypes = [vp, i, i, i, i, i, P(C.c_float), i]
    lib.orc_set_receivers.argtypes = [vp, i, P(C.c_int)]
    for f in ("orc_update_vel", "orc_update_stress", "orc_ghost_vz", "orc_ghost_sxz", "orc_ghost_syz", "orc_run",
              "orc_set_steps_done"):
        getattr(lib, f).argtypes = [vp, i]
    for f in ("orc_copy_vel", "orc_record", "orc_add_source", "orc_steps_done"):
        getattr(lib, f).argtypes = [vp]
    lib.orc_energy_now.argtypes = [vp]
    lib.orc_energy_now.restype = d
    lib.orc_residuals.argtypes = [vp, i, P(C.c_double)]
    lib.orc_field_dims.argtypes = [vp, i, i, P(C.c_int)]
    lib.orc_get_field.argtypes = [vp, i, i, P(C.c_double)]
    lib.orc_set_field.argtypes = [vp, i, i, P(C.c_double)]
    lib.orc_get_traces.argtypes = [vp, P(C.c_double)]
    lib.orc_get_energy.argtypes = [vp, P(C.c_double)]
    lib.orc_jacobi_stats.argtypes = [vp, i, P(C.c_long), P(C.c_int), P(C.c_int)]
    lib.orc_get_jacobi_state.argtypes = [vp, i, P(C.c_double)]
    lib.orc_get_sxz_system.argtypes = [vp, i, P(C.c_double), P(C.c_double)]
    lib.orc_fd_coeff.argtypes = [i, P(C.c_double)]
    lib.orc_set_vy_mirror_all.argtypes = [vp, i]
    lib.orc_set_ii_xend.argtypes = [vp, i]
    lib.orc_set_num_threads.argtypes = [i]
    return lib


_LIBS = {}


def lib(prec: str = "f64"):
    if prec not in _LIBS:
        _LIBS[prec] = _load(prec)
    return _LIBS[prec]


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Oracle:
    """One simulation in the CPU restatement.

    eci_mode 0 = the reference's ratio-2 interface formulas; 1 = ratio-generic.
    """

    def __init__(self, model, prec: str = "f64", eci_mode: int = None, check_energy: bool = True,
                 max_steps: int = None, vy_mirror_all: bool = False, ii_xend: bool = False):
        self.L = lib(prec)
        self.model = model
        ratios = model.ratios()
        if eci_mode is None:
            eci_mode = 0 if all(r == 2 for r in ratios) else 1
        self.eci_mode = eci_mode
        ng = model.num_grids
        n = np.array([x for G in model.grids for x in G.n], np.int32)
        h = np.array([x for G in model.grids for x in G.h], np.float64)
        self.max_steps = model.nsteps if max_steps is None else max_steps
        self.h = self.L.orc_create(ng, n.ctypes.data_as(C.POINTER(C.c_int)), _dp(h), model.radius, model.top_bc,
                                   model.imode, eci_mode, model.dt, int(check_energy), self.max_steps)
        if not self.h:
            raise RuntimeError("orc_create failed (sizes/ratio)")
        self.check_energy = check_energy
        self.L.orc_set_vy_mirror_all(self.h, int(vy_mirror_all))
        self.L.orc_set_ii_xend(self.h, int(ii_xend))
        for g in range(ng):
            self.set_params(g)
        if model.src_trace is not None and len(model.src_trace):
            tr = np.ascontiguousarray(model.src_trace, np.float32)
            self.L.orc_set_source(self.h, model.src_grid, *map(int, model.src_loc), model.src_type,
                                  tr.ctypes.data_as(C.POINTER(C.c_float)), len(tr))
        rec = np.ascontiguousarray(model.rec_loc, np.int32)
        self.L.orc_set_receivers(self.h, len(rec), rec.ctypes.data_as(C.POINTER(C.c_int)))

    def set_params(self, g):
        G = self.model.grids[g]
        st = [np.ascontiguousarray(a, np.float64) for a in self.model.stiff9(g)]
        self._keep = st
        arr = (C.POINTER(C.c_double) * 9)(*[_dp(a) for a in st])
        b = np.ascontiguousarray(G.buoy, np.float64)
        rc = self.L.orc_set_params(self.h, g, _dp(b), arr, int(self.model.iso))
        if rc:
            raise RuntimeError("orc_set_params failed")

    def close(self):
        if self.h:
            self.L.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # phases
    def run(self, nsteps: int):
        if self.L.orc_run(self.h, nsteps):
            raise RuntimeError("orc_run: step buffer exhausted")

    def update_vel(self, g): self.L.orc_update_vel(self.h, g)
    def update_stress(self, g): self.L.orc_update_stress(self.h, g)
    def ghost_vz(self, k): self.L.orc_ghost_vz(self.h, k)
    def ghost_sxz(self, k): self.L.orc_ghost_sxz(self.h, k)
    def ghost_syz(self, k): self.L.orc_ghost_syz(self.h, k)
    def copy_vel(self): self.L.orc_copy_vel(self.h)
    def energy_now(self): return self.L.orc_energy_now(self.h)

    @property
    def steps_done(self):
        return self.L.orc_steps_done(self.h)

    def dims(self, g, c):
        l3 = np.zeros(3, np.int32)
        r = self.L.orc_field_dims(self.h, g, c, l3.ctypes.data_as(C.POINTER(C.c_int)))
        if r < 0:
            raise ValueError("no such field")
        return (int(l3[2]), int(l3[1]), int(l3[0])), r

    def get(self, g, c):
        shape, _ = self.dims(g, c)
        out = np.empty(shape, np.float64)
        self.L.orc_get_field(self.h, g, c, _dp(out))
        return out

    def set(self, g, c, a):
        shape, _ = self.dims(g, c)
        a = np.ascontiguousarray(a, np.float64).reshape(shape)
        self.L.orc_set_field(self.h, g, c, _dp(a))

    def traces(self):
        n = self.steps_done
        out = np.empty((len(self.model.rec_loc), n), np.float64)
        self.L.orc_get_traces(self.h, _dp(out))
        return out

    def energy(self):
        out = np.empty(self.steps_done, np.float64)
        self.L.orc_get_energy(self.h, _dp(out))
        return out

    def residuals(self, which):
        out = np.zeros(max(1, self.model.num_grids - 1), np.float64)
        self.L.orc_residuals(self.h, which, _dp(out))
        return out

    def jacobi_stats(self, k=0):
        t, l, c = C.c_long(), C.c_int(), C.c_int()
        self.L.orc_jacobi_stats(self.h, k, C.byref(t), C.byref(l), C.byref(c))
        return t.value, l.value, c.value

    def sxz_system(self, k=0):
        G = self.model.grids[k + 1]
        nn = (G.n[0] - 1) * (G.n[1] - 1)
        A = np.empty(9 * nn)
        B = np.empty(nn)
        self.L.orc_get_sxz_system(self.h, k, _dp(A), _dp(B))
        return A.reshape(9, G.n[1] - 1, G.n[0] - 1), B.reshape(G.n[1] - 1, G.n[0] - 1)

    def jacobi_state(self, k=0):
        G = self.model.grids[k + 1]
        out = np.empty((G.n[1] + 1) * (G.n[0] + 1))
        self.L.orc_get_jacobi_state(self.h, k, _dp(out))
        return out.reshape(G.n[1] + 1, G.n[0] + 1)


def fd_coeff(radius: int):
    out = np.zeros(radius)
    if lib("f64").orc_fd_coeff(radius, _dp(out)):
        raise ValueError("radius")
    return out
