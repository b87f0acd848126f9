"""Small synthetic models shared by the CPU and GPU tests."""
from __future__ import annotations

import math

import numpy as np

from ecsg_amd.model import Grid, Model, cfl_dt, ricker


def random_medium(n, rng, vp=(2000.0, 3000.0), vs_ratio=0.55, rho=(1800.0, 2400.0), pert=0.1):
    """Orthorhombic medium at primary nodes with random per-node perturbation
    of each stiffness (float32-representable values)."""
    nx, ny, nz = n
    shape = (nz + 1, ny + 1, nx + 1)
    vpf = rng.uniform(*vp, size=shape)
    vsf = vpf * vs_ratio
    rh = rng.uniform(*rho, size=shape)
    c11 = rh * vpf ** 2
    c44 = rh * vsf ** 2
    c12 = c11 - 2 * c44
    base = [c11, c11, c11, c12, c12, c12, 2 * c44, 2 * c44, 2 * c44]
    stiff = []
    for v in base:
        p = 1.0 + pert * rng.uniform(-1, 1, size=shape)
        stiff.append((v * p).astype(np.float32).astype(np.float64))
    buoy = (1.0 / rh).astype(np.float32).astype(np.float64)
    return buoy, stiff


def iso_medium(n, rng, vp=(2000.0, 3000.0), vs_ratio=0.55, rho=(1800.0, 2400.0)):
    nx, ny, nz = n
    shape = (nz + 1, ny + 1, nx + 1)
    vpf = rng.uniform(*vp, size=shape)
    rh = rng.uniform(*rho, size=shape)
    s11 = (rh * vpf ** 2).astype(np.float32).astype(np.float64)
    s13 = (s11 - 2 * rh * (vpf * vs_ratio) ** 2).astype(np.float32).astype(np.float64)
    return (1.0 / rh).astype(np.float32).astype(np.float64), [s11, s13]


def composite(ratios=(2,), coarse=(5, 4, 4), nz_fine=(4,), h0=10.0, nsteps=40, seed=0, iso=False,
              top_bc=1, src_len=None, fpeak=None, src_type=0, cfl=0.9):
    """Stack of grids, finest on top: grid k+1 is coarser by ratios[k].

    coarse: cells of the bottom (coarsest) grid; nz_fine: z cells of the finer grids.
    """
    rng = np.random.default_rng(seed)
    ng = len(ratios) + 1
    cells = [None] * ng
    steps = [None] * ng
    cells[-1] = tuple(coarse)
    hb = h0 * int(np.prod(ratios))
    steps[-1] = hb
    for k in range(ng - 2, -1, -1):
        r = ratios[k]
        cells[k] = (cells[k + 1][0] * r, cells[k + 1][1] * r, nz_fine[k])
        steps[k] = steps[k + 1] / r
    grids = []
    z0 = 0.0
    for k in range(ng):
        if iso:
            b, st = iso_medium(cells[k], rng)
        else:
            b, st = random_medium(cells[k], rng)
        grids.append(Grid(cells[k], (steps[k],) * 3, b, st, origin_z=z0))
        z0 += steps[k] * cells[k][2]
    dt = cfl_dt(grids, iso, frac=cfl)
    n0 = cells[0]
    src = (n0[0] // 2, n0[1] // 2, max(1, n0[2] // 2))
    if fpeak is None:
        fpeak = 0.25 / (dt * 6)
    L = src_len if src_len is not None else nsteps
    tr = ricker(L, dt, fpeak)
    rec = []
    for k in range(ng):
        n = cells[k]
        rec.append([k, n[0] // 3, n[1] // 2, 1])
        rec.append([k, (2 * n[0]) // 3, n[1] // 3, n[2] // 2])
    return Model(grids=grids, dt=dt, nsteps=nsteps, iso=iso, top_bc=top_bc, src_grid=0, src_loc=src,
                 src_type=src_type, src_trace=tr, rec_loc=np.array(rec, np.int32), name="test_composite")


def single(n=(8, 7, 6), h=10.0, nsteps=40, seed=0, iso=False, top_bc=1, src_len=None, src_type=1, cfl=0.9):
    rng = np.random.default_rng(seed)
    b, st = iso_medium(n, rng) if iso else random_medium(n, rng)
    G = Grid(tuple(n), (h, h, h), b, st)
    dt = cfl_dt([G], iso, frac=cfl)
    L = src_len if src_len is not None else nsteps
    tr = ricker(L, dt, 0.25 / (dt * 6))
    rec = np.array([[0, n[0] // 3, n[1] // 2, 1], [0, n[0] // 2, n[1] // 3, n[2] // 2]], np.int32)
    return Model(grids=[G], dt=dt, nsteps=nsteps, iso=iso, top_bc=top_bc, src_grid=0,
                 src_loc=(n[0] // 2, n[1] // 2, n[2] // 2), src_type=src_type, src_trace=tr, rec_loc=rec,
                 name="test_single")


def sxz_sparse(A):
    """The vx-sxz interface matrix as a sparse matrix from its 9 bands
    A[k, J, I] (band k couples unknown (I, J) to (I + k//3 - 1, J + k%3 - 1);
    unknowns numbered J * n0 + I)."""
    import scipy.sparse as sp
    _, n1, n0 = A.shape
    N = n0 * n1
    rows, cols, vals = [], [], []
    for k in range(9):
        dI, dJ = k // 3 - 1, k % 3 - 1
        for j in range(n1):
            for i in range(n0):
                ii, jj = i + dI, j + dJ
                if 0 <= ii < n0 and 0 <= jj < n1 and A[k, j, i] != 0:
                    rows.append(j * n0 + i)
                    cols.append(jj * n0 + ii)
                    vals.append(A[k, j, i])
    return sp.csr_matrix((vals, (rows, cols)), shape=(N, N))


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def _sbp_w(n_lo, n_hi, idx):
    return (1 - 0.5 * (idx == n_lo)) * (1 - 0.5 * (idx == n_hi))
def sbp_residuals(F, n, r=1):
    """The seven summation-by-parts sums of TestSBP (test_sbp.cpp:5-199) from
    the 9 wavefields F (Field layout with halo r, [z][y][x]) of a grid of n
    cells, in numpy.  Reads follow the Field's flat storage, so an index past
    a row continues in the next one, as the reference's do.  Returns the sums
    and the sums of the absolute terms."""
    nx, ny, nz = n
    def at(a, xs, ys, zs):
        # flat offsets in the Field layout (reads past a row continue in the next one)
        L2, L1, L0 = a.shape
        zz, yy, xx = np.meshgrid(np.asarray(zs) + r, np.asarray(ys) + r, np.asarray(xs) + r, indexing="ij")
        return a.ravel()[(zz * L1 + yy) * L0 + xx]
    vx, vy, vz, sxx, syy, szz, syz, sxz, sxy = [F[c] for c in range(9)]
    X = lambda lo, hi: np.arange(lo, hi + 1)
    def W(xs, ys, zs, wx, wy, wz):
        gx = _sbp_w(xs[0], xs[-1], xs) if wx else np.ones(len(xs))
        gy = _sbp_w(ys[0], ys[-1], ys) if wy else np.ones(len(ys))
        gz = _sbp_w(zs[0], zs[-1], zs) if wz else np.ones(len(zs))
        return gz[:, None, None] * gy[None, :, None] * gx[None, None, :]
    res, mag = [], []
    def two(tA, tB):
        res.append(tA.sum() + tB.sum()); mag.append(np.abs(tA).sum() + np.abs(tB).sum())
    # ranges
    vxr = (X(0, nx), X(0, ny), X(0, nz)); vyr = (X(0, nx - 1), X(0, ny - 1), X(0, nz)); vzr = (X(0, nx - 1), X(0, ny), X(0, nz - 1))
    snr = (X(0, nx - 1), X(0, ny), X(0, nz)); syzr = (X(0, nx - 1), X(0, ny - 1), X(0, nz - 1))
    sxzr = (X(0, nx), X(0, ny), X(0, nz - 1)); sxyr = (X(0, nx), X(0, ny - 1), X(0, nz))
    x, y, z = vxr
    m = W(x, y, z, 1, 1, 1)
    A = m * at(vx, x, y, z) * (at(sxz, x, y, z) - at(sxz, x, y, z - 1))
    x, y, z = sxzr
    B = W(x, y, z, 1, 1, 0) * (at(vx, x, y, z + 1) - at(vx, x, y, z)) * at(sxz, x, y, z)
    two(A, B)
    x, y, z = vxr
    A = m * at(vx, x, y, z) * (at(sxx, x, y, z) - at(sxx, x - 1, y, z))
    x, y, z = snr
    B = W(x, y, z, 0, 1, 1) * (at(vx, x + 1, y, z) - at(vx, x, y, z)) * at(sxx, x, y, z)
    two(A, B)
    x, y, z = vxr
    A = m * at(vx, x, y, z) * (at(sxy, x, y, z) - at(sxy, x, y - 1, z))
    x, y, z = sxyr
    B = W(x, y, z, 1, 0, 1) * (at(vx, x, y + 1, z) - at(vx, x, y, z)) * at(sxy, x, y, z)
    two(A, B)
    x, y, z = vyr
    A = W(x, y, z, 0, 0, 1) * at(vy, x, y, z) * (at(syz, x, y, z) - at(syz, x, y, z - 1))
    x, y, z = syzr
    B = (at(vy, x, y, z + 1) - at(vy, x, y, z)) * at(syz, x, y, z)
    two(A, B)
    x, y, z = vzr
    wy_ = W(x, y, z, 0, 1, 0)
    A = wy_ * at(vz, x, y, z) * (at(szz, x, y, z + 1) - at(szz, x, y, z))
    x2, y2, z2 = snr
    B = W(x2, y2, z2, 0, 1, 1) * (at(vz, x2, y2, z2) - at(vz, x2, y2, z2 - 1)) * at(szz, x2, y2, z2)
    two(A, B)
    A = wy_ * at(vz, x, y, z) * (at(sxz, x + 1, y, z) - at(sxz, x, y, z))
    x2, y2, z2 = sxzr
    B = W(x2, y2, z2, 1, 1, 0) * (at(vz, x2, y2, z2) - at(vz, x2 - 1, y2, z2)) * at(sxz, x2, y2, z2)
    two(A, B)
    A = wy_ * at(vz, x, y, z) * (at(syz, x, y + 1, z) - at(syz, x, y, z))
    x2, y2, z2 = syzr
    B = (at(vz, x2, y2, z2) - at(vz, x2, y2 - 1, z2)) * at(syz, x2, y2, z2)
    two(A, B)
    return np.array(res), np.array(mag)


HIP_ATTR_MULTIPROCESSOR_COUNT = 63   # hipDeviceAttributeMultiprocessorCount (hip_runtime_api.h, ROCm 7.2)


def cu_count(device: int = 0) -> int:
    """Compute units of a device, asked of the HIP runtime libecsg_hip.so
    runs on: the libamdhip64 already mapped into this process (ROCm's, or
    torch's bundled copy when torch was loaded first; a second runtime does
    not start in a process where one already runs)."""
    import ctypes
    import ecsg_amd
    ecsg_amd.load_library()
    path = None
    with open("/proc/self/maps") as fh:
        for line in fh:
            f = line.split()[-1]
            if "libamdhip64.so" in f:
                path = f
                break
    if path is None:
        raise RuntimeError("libamdhip64 is not loaded")
    hip = ctypes.CDLL(path)
    v = ctypes.c_int(0)
    rc = hip.hipDeviceGetAttribute(ctypes.byref(v), HIP_ATTR_MULTIPROCESSOR_COUNT, device)
    if rc != 0 or v.value <= 0:
        raise RuntimeError("hipDeviceGetAttribute(multiprocessor count) failed: %d" % rc)
    return v.value
