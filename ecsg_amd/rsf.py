"""RSF files for the command-line driver ``ecsg_run``.

Madagascar's RSF format: a text header of ``key=value`` lines naming the
binary data with ``in=``; here always 4-byte native floats, axis 1 fastest.
``write_flow_inputs`` writes what the reference's flow scripts hand to its
``main.exe`` (tests/test1-cmp-eci-ii/eci/SConstruct:177-192): ``buoy_<i>``,
``s11_<i>``, ``s13_<i>`` (or the nine ``c<ij>_<i>`` with ``aniso=y``),
``src`` (wavelet with the position in ``o2 o3 o4``, init_src.cpp:16-31) and
``rec_in`` (x y z per receiver, init_del_rec_trace.cpp:16-31), and returns the
matching command line.
"""
from __future__ import annotations

import math
import os
import shlex
from typing import Dict, List, Optional, Sequence

import numpy as np

from .model import Model

STIFF_FILES = ("c11", "c22", "c33", "c23", "c13", "c12", "c44", "c55", "c66")


def write_rsf(path: str, data: np.ndarray, **header) -> None:
    """Header at ``path``, native float32 data at ``path + '@'``.  ``n1..``
    default to the array's shape, last axis first."""
    a = np.ascontiguousarray(np.asarray(data, dtype=np.float32))
    dims = {"n%d" % (k + 1): v for k, v in enumerate(reversed(a.shape))}
    dims.update(header)
    dpath = path + "@"
    a.tofile(dpath)
    with open(path, "w") as f:
        f.write("ecsg_amd.rsf\n")
        for k, v in dims.items():
            if isinstance(v, str):
                v = '"%s"' % v
            elif isinstance(v, float):
                v = repr(v)
            f.write("\t%s=%s\n" % (k, v))
        f.write('\tesize=4\n\tdata_format="native_float"\n\tin="%s"\n' % dpath)


def _parse_header(text: str) -> Dict[str, str]:
    kv: Dict[str, str] = {}
    for tok in shlex.split(text, posix=True):
        if "=" in tok:
            k, v = tok.split("=", 1)
            kv[k] = v
    return kv


def read_rsf(path: str):
    """(header dict, float32 array shaped (..., n2, n1))."""
    raw = open(path, "rb").read()
    cut = raw.find(b"\x0c\x0c\x04")
    text = raw[: cut if cut >= 0 else len(raw)].decode("latin-1")
    kv = _parse_header(text)
    dims = []
    for k in range(1, 10):
        n = int(kv.get("n%d" % k, 1))
        dims.append(n)
    while len(dims) > 1 and dims[-1] == 1:
        dims.pop()
    count = int(np.prod(dims))
    fmt = kv.get("data_format", "native_float")
    if cut >= 0:
        buf = raw[cut + 3:]
    else:
        buf = open(kv["in"], "rb").read()
    if fmt == "native_float":
        a = np.frombuffer(buf, dtype="<f4", count=count)
    elif fmt == "xdr_float":
        a = np.frombuffer(buf, dtype=">f4", count=count).astype(np.float32)
    else:
        raise ValueError("unsupported data_format " + fmt)
    return kv, a.reshape(tuple(reversed(dims)))


def _movie_timing(m: Model):
    """(movie_dt, movie_nt) whose loop length ceil(movie_dt*movie_nt/dt)+1
    (main.cpp:165) equals m.nsteps."""
    if m.movie_dt is not None and m.movie_nt is not None:
        if int(math.ceil(m.movie_dt * m.movie_nt / m.dt)) + 1 == m.nsteps:
            return m.movie_dt, m.movie_nt
    for mdt, mnt in ((m.dt, m.nsteps - 1), (m.dt, m.nsteps - 2)):
        if mnt > 0 and int(math.ceil(mdt * mnt / m.dt)) + 1 == m.nsteps:
            return mdt, mnt
    raise ValueError("cannot express %d steps as a movie length" % m.nsteps)


def _position(m: Model, loc) -> tuple:
    """Metres of grid position (g, ix, iy, iz) that map back to it under the
    reference's strict-> stacking search: a point on a coarse grid's top plane
    is nudged a quarter step down (it would map to the finer grid above)."""
    g, ix, iy, iz = (int(v) for v in loc)
    G = m.grids[g]
    z0 = sum(m.grids[k].h[2] * m.grids[k].n[2] for k in range(g))
    z = z0 + (iz if (iz > 0 or g == 0) else 0.25) * G.h[2]
    return ix * G.h[0], iy * G.h[1], z


def write_flow_inputs(m: Model, folder: str, aniso: Optional[bool] = None) -> List[str]:
    """Writes the driver's input files for model m into folder and returns the
    ``key=value`` arguments (inputs, parameters and output names) of a run."""
    if m.src_type != 0:
        raise ValueError("the driver's source is the vx force (main.cpp:194-195)")
    os.makedirs(folder, exist_ok=True)
    aniso = (not m.iso) if aniso is None else aniso
    args = ["num_grids=%d" % m.num_grids, "radius=%d" % m.radius, "dt=%r" % m.dt, "top_bc=%d" % m.top_bc,
            "imode=%d" % m.imode]
    if aniso:
        args.append("aniso=y")
    for g, G in enumerate(m.grids):
        i = g + 1
        hdr = dict(d1=float(G.h[0]), d2=float(G.h[1]), d3=float(G.h[2]), o1=0.0, o2=0.0, o3=float(G.origin_z))
        p = os.path.join(folder, "buoy_%d.rsf" % i)
        write_rsf(p, G.buoy, **hdr)
        args.append("buoy_%d=%s" % (i, p))
        if aniso:
            for name, arr in zip(STIFF_FILES, m.stiff9(g)):
                p = os.path.join(folder, "%s_%d.rsf" % (name, i))
                write_rsf(p, arr, **hdr)
                args.append("%s_%d=%s" % (name, i, p))
        else:
            if not m.iso:
                raise ValueError("s11/s13 inputs need an isotropic model")
            for name, arr in zip(("s11", "s13"), G.stiff):
                p = os.path.join(folder, "%s_%d.rsf" % (name, i))
                write_rsf(p, arr, **hdr)
                args.append("%s_%d=%s" % (name, i, p))
    sx, sy, sz = m.src_xyz if m.src_xyz is not None else _position(m, (m.src_grid,) + tuple(m.src_loc))
    p = os.path.join(folder, "src.rsf")
    write_rsf(p, np.asarray(m.src_trace, np.float32).reshape(1, 1, 1, -1), d1=float(m.dt) / 1000.0, o1=0.0,
              o2=float(sx), o3=float(sy), o4=float(sz))
    args.append("src=%s" % p)
    if m.rec_xyz is not None:
        rec = np.asarray(m.rec_xyz, np.float32).reshape(-1, 3)
    else:
        rec = np.array([_position(m, loc) for loc in m.rec_loc], np.float32).reshape(-1, 3)
    p = os.path.join(folder, "rec_in.rsf")
    write_rsf(p, rec, n1=3, n2=int(rec.shape[0]))
    args.append("rec_in=%s" % p)
    mdt, mnt = _movie_timing(m)
    args += ["movie_dt=%r" % mdt, "movie_nt=%d" % mnt]
    args += ["rec_out=%s" % os.path.join(folder, "rec_out.rsf"),
             "energy_trace=%s" % os.path.join(folder, "energy_trace.rsf"),
             "res_trace=%s" % os.path.join(folder, "res_trace.rsf")]
    for v in ("vx", "vy", "vz"):
        for g in range(m.num_grids):
            args.append("movie_%s_%d=%s" % (v, g + 1, os.path.join(folder, "movie_%s_%d.rsf" % (v, g + 1))))
    return args


def driver_path() -> str:
    """The built ecsg_run executable (next to libecsg_hip.so)."""
    return os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "ecsg_run")
