"""Host logic (CPU): model geometry, location mapping, wavelets, and the
C-ABI library's exported symbols (loaded, not called: no GPU here)."""
import ctypes
import math
import os
import re

import numpy as np
import pytest

import ecsg_amd
from ecsg_amd.model import Grid, _cround, map_location, reference_flow, ricker, ricker_sfmath

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_reference_flow_geometry():
    """tests/test1-cmp-eci-ii/eci/SConstruct:44-57,113,165,180 with par1.py."""
    m = reference_flow(1)
    (g1, g2) = m.grids
    assert g1.n == (384, 384, 72) and g2.n == (192, 192, 59)
    assert g1.h == (4.16667,) * 3 and g2.h == (8.33333,) * 3
    assert m.dt == 0.589378
    assert m.nsteps == 1172
    assert m.src_loc == (192, 192, 14) and m.src_grid == 0
    assert m.rec_loc.tolist() == [[0, 210, 192, 14], [0, 230, 192, 14], [0, 248, 192, 14], [0, 268, 192, 14]]
    assert m.total_nodes == 385 * 385 * 73 + 193 * 193 * 60
    # 5 bottom fine slices carry the lower layer (n_extra = 5)
    assert g1.buoy[-5:].min() == g1.buoy[-5:].max() == np.float32(0.377358)
    assert g1.buoy[:-5].max() == np.float32(0.4)
    m3 = reference_flow(3)
    assert m3.src_loc == (192, 192, 71)


def test_location_mapping_strict_stacking():
    grids = [Grid((4, 4, 10), (1.0, 1.0, 1.0), None, None), Grid((2, 2, 5), (2.0, 2.0, 2.0), None, None)]
    # exactly on the interface stays in the upper grid (strict >, init_src.cpp:21)
    assert map_location(grids, 1.0, 1.0, 10.0) == (0, 1, 1, 10)
    assert map_location(grids, 2.0, 2.0, 12.0) == (1, 1, 1, 1)
    assert _cround(2.5) == 3 and _cround(-2.5) == -3 and _cround(1.4999) == 1
    with pytest.raises(ValueError):
        map_location(grids, 0, 0, 100.0)


def test_ricker_peak_and_sfmath_rounding():
    dt = 1e-3
    w = ricker(200, dt, 15.0)
    assert abs(float(w[int(round(1 / 15.0 / dt))]) - 1.0) < 1e-2
    assert w.dtype == np.float32
    s = ricker_sfmath(700, 0.589378399797743 / 1000, 15.0, 1 / 15.0)
    assert s.dtype == np.float32 and s.shape == (700,)
    # printed constants: t0 = 0.066667 s, sampling 0.000589 s -> peak near sample 113
    assert int(np.argmax(s)) == 113


def _declared_functions():
    hdr = open(os.path.join(ROOT, "include", "ecsg_hip.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    return sorted(set(re.findall(r"\b(ecsg_[a-z_0-9]+)\s*\(", hdr)))


def test_header_and_binding_agree():
    assert set(_declared_functions()) == set(ecsg_amd.ABI_SYMBOLS)


def test_library_exports_every_declared_symbol():
    path = ecsg_amd.LIB_PATH
    if not os.path.exists(path):
        pytest.fail("libecsg_hip.so has not been built (python -m ecsg_amd.build)")
    lib = ctypes.CDLL(path)
    for name in _declared_functions():
        assert hasattr(lib, name), name
    lib.ecsg_version.restype = ctypes.c_char_p
    assert b"gfx950" in lib.ecsg_version()


def test_stream_probe_buffer_size():
    """ecsg_stream_probe_floats (host arithmetic only): 18 arrays of the
    radius-1 FP32 device layout, ix = 0 on a 128-byte line, a halo of 1,
    rows and arrays padded to 32 floats."""
    L = ecsg_amd.load_library()
    # 1 cell per dimension: rows of 64 floats (32 + 3 rounded up), 4 rows, 4 planes
    assert L.ecsg_stream_probe_floats(1, 1, 1) == 18 * 64 * 4 * 4
    # 1206 x 1206 x 80 cells: rows of 1248 floats, 1209 rows, 83 planes
    assert L.ecsg_stream_probe_floats(1206, 1206, 80) == 18 * 1248 * 1209 * 83
    assert L.ecsg_stream_probe_floats(0, 4, 4) == 0


def test_config_struct_layout_matches_header():
    # ecsg_config: int, int[8][3], (pad), double[8][3], double, 9 ints, (pad to 8)
    import ctypes as C
    assert C.sizeof(ecsg_amd.EcsgConfig) == 4 + 96 + 4 + 192 + 8 + 9 * 4 + 4
    assert ecsg_amd.EcsgConfig.ii_vz_xend.offset == 4 + 96 + 4 + 192 + 8 + 8 * 4


def test_bench_one_gpu_only():
    """bench.py runs on one GPU; asked for more it says so and exits with
    status 2 before touching any device."""
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    out = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True,
                         text=True, timeout=120)
    assert out.returncode == 2
    assert "one GPU" in out.stderr
