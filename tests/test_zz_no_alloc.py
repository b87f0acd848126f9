"""The allocation-free CFL entry point (ecsg_cfl_dt_noalloc) against the
allocating one (ecsg_cfl_dt), through the C ABI, on two calls in a row --
the second on a smaller input -- with the inputs already in device memory.
Its square roots are the device's: the results agree to one ulp."""
import ctypes as C

import numpy as np
import pytest

import ecsg_amd
from helpers import composite

pytestmark = pytest.mark.gpu


class _Hip:
    """The test's own device buffers (hipMalloc through the HIP runtime the
    library links; the entry point under test allocates nothing)."""

    def __init__(self):
        self.L = None
        for name in ("libamdhip64.so.7", "libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
            try:
                self.L = C.CDLL(name)
                break
            except OSError:
                continue
        assert self.L is not None, "HIP runtime not found"
        self.bufs = []

    def upload(self, a):
        a = np.ascontiguousarray(a, np.float64)
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), C.c_size_t(a.nbytes)) == 0
        self.bufs.append(p)
        assert self.L.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1) == 0  # host to device
        return p

    def download(self, p, n):
        out = np.empty(n, np.float64)
        assert self.L.hipDeviceSynchronize() == 0
        assert self.L.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, C.c_size_t(out.nbytes), 2) == 0  # device to host
        return out

    def free(self):
        for p in self.bufs:
            self.L.hipFree(p)
        self.bufs = []


def _device_call(m, frac):
    L = ecsg_amd.load_library()
    H = _Hip()
    cfg = ecsg_amd.EcsgConfig()
    cfg.num_grids = m.num_grids
    buoys = (C.c_void_p * m.num_grids)()
    stiffs = ((C.c_void_p * 9) * m.num_grids)()
    for g, G in enumerate(m.grids):
        for k in range(3):
            cfg.n[g][k] = int(G.n[k])
            cfg.h[g][k] = float(G.h[k])
        buoys[g] = H.upload(G.buoy)
        for k, a in enumerate(list(G.stiff)[:9]):
            stiffs[g][k] = H.upload(a)
    out = H.upload(np.full(1, -1.0))
    rc = L.ecsg_cfl_dt_noalloc(C.byref(cfg), buoys, stiffs, int(m.iso), float(frac), out, None)
    assert rc == 0, rc
    dt = float(H.download(out, 1)[0])
    H.free()
    return dt


@pytest.mark.parametrize("iso", [False, True])
def test_noalloc_cfl_equals_allocating(iso):
    big = composite((2, 3), coarse=(6, 5, 4), nz_fine=(4, 4), nsteps=4, seed=31, iso=iso)
    small = composite((3,), coarse=(3, 3, 2), nz_fine=(2,), nsteps=4, seed=32, iso=iso)
    for m, frac in ((big, 0.9), (small, 0.7)):
        want = ecsg_amd.cfl_dt(m, frac)
        got = _device_call(m, frac)
        assert abs(got - want) <= np.spacing(want), (got, want)


def test_noalloc_rejects_too_many_grids():
    L = ecsg_amd.load_library()
    cfg = ecsg_amd.EcsgConfig()
    cfg.num_grids = 9
    z = (C.c_void_p * 9)()
    s = ((C.c_void_p * 9) * 9)()
    assert L.ecsg_cfl_dt_noalloc(C.byref(cfg), z, s, 1, 0.9, C.c_void_p(1), None) == -1
