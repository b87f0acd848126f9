"""Model description and input generation for the ECSG time-stepping path.

Host-side setup that feeds the device path.  It restates what the
reference does before its time loop:

* grid geometry and z-stacking of the composite grids
  (init_del_par_field.cpp:21-72: sizes are RSF n-1, z-origins stacked);
* source and receiver location mapping (init_src.cpp:16-45,
  init_del_rec_trace.cpp:16-45: strict ``>`` z-stacking search, ``round``);
* the reference's own test-flow inputs
  (tests/test1-cmp-eci-ii/eci/SConstruct:28-195 with par{1,2,3}.py):
  layered ``buoy/s11/s13`` cubes, the Ricker wavelet, receiver lines;
* synthetic orthorhombic media for the larger workloads, and the
  energy-method CFL bound ``h / (vp * sqrt(3))`` that those flows use
  (eci/SConstruct:113).

Arrays at primary nodes are numpy ``(nz+1, ny+1, nx+1)`` (x fastest),
stiffnesses in the code convention of the reference: ``c44 = 2*C44``
(Voigt), because the shear update multiplies by ``0.5*c44``
(update_stress.cpp:107).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

# order of the 9 stiffness arrays, as stiffs_ptr (init_del_par_field.cpp:129-149)
STIFF_NAMES = ("c11", "c22", "c33", "c23", "c13", "c12", "c44", "c55", "c66")
# wavefield order, init_del_sim.cpp:161-173
FIELD_NAMES = ("vx", "vy", "vz", "sxx", "syy", "szz", "syz", "sxz", "sxy")
FIELD_POS = ((0, 0, 0), (1, 1, 0), (1, 0, 1), (1, 0, 0), (1, 0, 0), (1, 0, 0),
             (1, 1, 1), (0, 0, 1), (0, 1, 0))


@dataclass
class Grid:
    n: tuple            # cells (nx, ny, nz)
    h: tuple            # steps (hx, hy, hz)
    buoy: np.ndarray    # primary nodes, float64 holding float32-representable values
    stiff: List[np.ndarray]  # 9 arrays (code convention) or [s11, s13] when iso
    origin_z: float = 0.0

    @property
    def nodes(self) -> int:
        nx, ny, nz = self.n
        return (nx + 1) * (ny + 1) * (nz + 1)


@dataclass
class Model:
    grids: List[Grid]
    dt: float
    nsteps: int
    iso: bool = True
    top_bc: int = 1
    imode: int = 1
    radius: int = 1
    src_grid: int = 0
    src_loc: tuple = (0, 0, 0)
    src_type: int = 0           # 0: vx force x1000 (main.cpp:194-195); 1: explosive (main.cpp:197-202)
    src_trace: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float32))
    rec_loc: np.ndarray = field(default_factory=lambda: np.zeros((0, 4), np.int32))
    name: str = "model"
    # positions in metres and movie timing as a flow script hands them to the
    # command-line driver (None: derived from the grid indices and nsteps)
    src_xyz: Optional[tuple] = None
    rec_xyz: Optional[np.ndarray] = None
    movie_dt: Optional[float] = None
    movie_nt: Optional[int] = None

    @property
    def num_grids(self) -> int:
        return len(self.grids)

    @property
    def total_nodes(self) -> int:
        return sum(g.nodes for g in self.grids)

    def stiff9(self, g: int) -> List[np.ndarray]:
        """Nine primary-node stiffness arrays of grid g (isotropic inputs expanded)."""
        G = self.grids[g]
        if not self.iso:
            return list(G.stiff)
        s11, s13 = G.stiff
        s44 = s11 - s13
        return [s11, s11, s11, s13, s13, s13, s44, s44, s44]

    def ratios(self) -> List[int]:
        out = []
        for k in range(self.num_grids - 1):
            f, c = self.grids[k], self.grids[k + 1]
            r = f.n[0] // c.n[0]
            if f.n[0] != r * c.n[0] or f.n[1] != r * c.n[1]:
                raise ValueError("grid %d/%d: fine sizes must be the ratio times coarse sizes" % (k, k + 1))
            out.append(r)
        return out


# ---------------------------------------------------------------------------
# location mapping (init_src.cpp:16-45, init_del_rec_trace.cpp:16-45)

def map_location(grids: Sequence[Grid], x: float, y: float, z: float):
    """(x, y, z) in metres -> (grid, ix, iy, iz).  Strict ``>`` stacking search."""
    g = 0
    tmp = z
    while tmp > grids[g].h[2] * grids[g].n[2]:
        tmp -= grids[g].h[2] * grids[g].n[2]
        g += 1
        if g >= len(grids):
            raise ValueError("location below the last grid")
    ix = int(_cround(x / grids[g].h[0]))
    iy = int(_cround(y / grids[g].h[1]))
    iz = int(_cround(tmp / grids[g].h[2]))
    return g, ix, iy, iz


def _cround(v: float) -> float:
    """C ``round``: halves away from zero (Python's round is banker's)."""
    return math.floor(v + 0.5) if v >= 0 else -math.floor(-v + 0.5)


def ricker_sfmath(n: int, d1: float, fpeak: float, delrt: float) -> np.ndarray:
    """Ricker wavelet as Madagascar ``sfmath`` builds it in the reference flow
    (eci/SConstruct:118-147): constants printed with ``%f`` and evaluated in
    single precision.  ``(1-2a(t-t0)^2) exp(-a(t-t0)^2)``, a = pi^2 f^2."""
    pi = math.atan(1) * 4.0
    a = np.float32(float("%f" % ((pi ** 2) * (fpeak ** 2))))
    t0 = np.float32(float("%f" % delrt))
    d = np.float32(float("%f" % d1))
    t = np.arange(n, dtype=np.float32) * d
    x = (t - t0).astype(np.float32)
    two = np.float32(2.0)
    one = np.float32(1.0)
    p = ((two * a) * x) * x
    q = ((-a) * x) * x
    return ((one - p) * np.exp(q, dtype=np.float32)).astype(np.float32)


def ricker(n: int, dt: float, fpeak: float, t0: Optional[float] = None) -> np.ndarray:
    """Ricker wavelet sampled once per time step (float32)."""
    t0 = 1.0 / fpeak if t0 is None else t0
    t = np.arange(n, dtype=np.float64) * dt - t0
    a = (math.pi * fpeak) ** 2
    return ((1.0 - 2.0 * a * t * t) * np.exp(-a * t * t)).astype(np.float32)


# ---------------------------------------------------------------------------
# the reference's own composite test flow (tests/test1-cmp-eci-ii)

def _f6(v: float) -> float:
    """A value as the flow scripts hand it to Madagascar: printed with %f."""
    return float("%f" % v)


def _hdr(v: float) -> float:
    """A value the flow hands to a Madagascar program with %f (eci/SConstruct:73-103,
    138-148) and that program writes into its output header: parsed as float32,
    written with %g (six significant digits).  main.exe reads it back in double
    (init_del_par_field.cpp:34-40: `real step[3]`, `real` = double): the grid
    steps are 4.16667 and 8.33333 m, not the scripts' 4.166667 and 8.333333."""
    return float("%g" % float(np.float32(_f6(v))))


def reference_flow(test: int = 1, steps: Optional[int] = None, eci: bool = True) -> Model:
    """Inputs of tests/test1-cmp-eci-ii/{eci,ii}/SConstruct with par{test}.py.

    Two grids at ratio 2 (384x384x72 cells at h=4.16667 m over 192x192x59
    at 8.33333 m, the steps the input headers hold), top layer rho 2.5 cs 1 cp 2, lower layer rho 2.65 cs 2
    cp 4 (units m/ms, g/cm^3 as in the flow), free surface on top, Ricker
    15 Hz vx force, 4 receivers at z=58.333 m.  1172 steps by default.
    """
    n_extra = {1: 5, 2: 1, 3: 5}[test]
    src_option = {1: 0, 2: 0, 3: 1}[test]
    cfl, nspw, lx, ly, lz1, lz2, duration = 0.98, 16, 1600, 1600, 300, 500, 700
    fpeak = 15.0
    layers = [dict(rho=2.5, cs=1.0, cp=2.0), dict(rho=2.65, cs=2.0, cp=4.0)]
    for L in layers:
        L["buoy"] = 1 / L["rho"]
        L["s11"] = L["rho"] * L["cp"] ** 2
        L["s13"] = L["s11"] - 2 * L["rho"] * L["cs"] ** 2
    dx1 = layers[0]["cs"] * 1000 / fpeak / nspw
    dx2 = dx1 * 2
    nx2 = int(lx / dx2)
    ny2 = int(ly / dx2)
    nx1, ny1 = 2 * nx2, 2 * ny2
    nz1 = int(lz1 / dx1 / 2) * 2
    nz2 = int(lz2 / dx2)
    # grid steps as the buoyancy files' headers hold them (sfspike d#=%f -> %g)
    h1 = _hdr(dx1)
    h2 = _hdr(dx2)

    def cube(nx, ny, nz, key, split=None):
        out = np.empty((nz + 1, ny + 1, nx + 1), np.float64)
        v0 = np.float32(_f6(layers[0][key]))
        v1 = np.float32(_f6(layers[1][key]))
        if split is None:
            out[:] = v1
        else:
            out[:split] = v0
            out[split:] = v1
        return out

    top_split = nz1 + 1 - n_extra
    g1 = Grid((nx1, ny1, nz1), (h1, h1, h1), cube(nx1, ny1, nz1, "buoy", top_split),
              [cube(nx1, ny1, nz1, "s11", top_split), cube(nx1, ny1, nz1, "s13", top_split)])
    g2 = Grid((nx2, ny2, nz2), (h2, h2, h2), cube(nx2, ny2, nz2, "buoy"),
              [cube(nx2, ny2, nz2, "s11"), cube(nx2, ny2, nz2, "s13")])
    g2.origin_z = g1.origin_z + g1.n[2] * g1.h[2]
    grids = [g1, g2]

    max_vel = max(L["cp"] for L in layers)
    dt_exact = min(dx1, dx1, dx1) / max_vel / math.sqrt(3) * cfl
    dt = _f6(dt_exact)
    if src_option == 0:
        src_z = int(lz1 * 0.2 / dx2) * dx2
    else:
        src_z = (nz1 - 1) * dx1
    src_x = int(lx * 0.5 / dx2) * dx2
    src_y = int(ly * 0.5 / dx2) * dx2
    src_trace = ricker_sfmath(700, dt_exact / 1000.0, fpeak, 1 / fpeak)
    # source position: the o2..o4 keys sfmath writes into the wavelet's header (init_src.cpp:28-31)
    src_pos = (_hdr(src_x), _hdr(src_y), _hdr(src_z))
    sg, six, siy, siz = map_location(grids, *src_pos)

    rx = [int(lx * f / dx2) * dx2 for f in (0.55, 0.6, 0.65, 0.7)]
    rz = int(lz1 * 0.2 / dx2) * dx2
    rec = []
    rec_xyz = []
    for x in rx:
        # receiver file: ascii %f -> float32 (eci/SConstruct:169-175)
        xf, yf, zf = (float(np.float32(_f6(v))) for v in (x, src_y, rz))
        rec.append(map_location(grids, xf, yf, zf))
        rec_xyz.append((xf, yf, zf))

    movie_dt = _f6(30 * dt_exact)
    movie_nt = int(duration / (30 * dt_exact))
    nsteps = int(math.ceil(movie_dt * movie_nt / dt)) + 1 if steps is None else steps
    return Model(grids=grids, dt=dt, nsteps=nsteps, iso=True, top_bc=1, imode=1 if eci else 0,
                 src_grid=sg, src_loc=(six, siy, siz), src_type=0, src_trace=src_trace,
                 rec_loc=np.array(rec, np.int32), name="reference_test%d" % test,
                 src_xyz=src_pos, rec_xyz=np.array(rec_xyz, np.float32),
                 movie_dt=movie_dt, movie_nt=movie_nt)


# ---------------------------------------------------------------------------
# parameter staging (what InitParField leaves in stiffs_ptr)

# staggered position of each of the nine stiffnesses (init_del_par_field.cpp:82-88,
# the normal-stress moduli all at s11's position)
STIFF_POS = ((1, 0, 0),) * 6 + ((1, 1, 1), (0, 0, 1), (0, 1, 0))


def reshape_field(prim: np.ndarray, pos) -> np.ndarray:
    """ReshapeField (field.cpp:428-476): a primary-node array (nz+1, ny+1, nx+1)
    averaged onto staggered position pos = (px, py, pz); the eight terms are
    summed in the reference's order, then divided by 8."""
    px, py, pz = pos
    nz, ny, nx = (d - 1 for d in prim.shape)
    ez, ey, ex = nz + 1 - pz, ny + 1 - py, nx + 1 - px

    def f(dx, dy, dz):
        return prim[dz:dz + ez, dy:dy + ey, dx:dx + ex]

    s = f(0, 0, 0) + f(0, 0, pz)
    s = s + f(0, py, 0)
    s = s + f(0, py, pz)
    s = s + f(px, 0, 0)
    s = s + f(px, 0, pz)
    s = s + f(px, py, 0)
    s = s + f(px, py, pz)
    return s / 8


def staged_params(m: "Model", g: int) -> List[np.ndarray]:
    """The nine Fields stiffs_ptr[g][0..8] as InitParField builds them
    (init_del_par_field.cpp:94-149): for isotropic input three objects --
    s11 and s13 at (1,0,0), shared by entries 0..2 and 3..5 -- and s44 = s11 - s13
    at (1,1,1), (0,0,1), (0,1,0); otherwise each Cij staged on its own."""
    G = m.grids[g]
    if m.iso:
        s11, s13 = G.stiff
        s44 = s11 - s13
        a = reshape_field(s11, STIFF_POS[0])
        b = reshape_field(s13, STIFF_POS[4])
        return [a, a, a, b, b, b] + [reshape_field(s44, STIFF_POS[k]) for k in (6, 7, 8)]
    return [reshape_field(G.stiff[k], STIFF_POS[k]) for k in range(9)]


# ---------------------------------------------------------------------------
# synthetic media for the throughput workloads

def cfl_dt(model_grids: Sequence[Grid], iso: bool, frac: float = 0.9) -> float:
    """frac x min over grids of h_min / (vp_max sqrt(3)), the bound the
    reference flows use (eci/SConstruct:109-113); vp from max(C11,C22,C33)/rho."""
    best = math.inf
    for G in model_grids:
        if iso:
            c = G.stiff[0]
        else:
            c = np.maximum(np.maximum(G.stiff[0], G.stiff[1]), G.stiff[2])
        vp = float(np.sqrt(np.max(c * G.buoy)))
        best = min(best, min(G.h) / (vp * math.sqrt(3)))
    return frac * best


def _smooth_noise(shape, corr_cells: float, rng: np.random.Generator) -> np.ndarray:
    """Zero-mean, unit-max Gaussian-smoothed noise (separable FFT filter)."""
    a = rng.standard_normal(shape)
    if corr_cells > 0:
        for ax, n in enumerate(shape):
            k = np.fft.rfftfreq(n)
            filt = np.exp(-0.5 * (2 * math.pi * k * corr_cells) ** 2)
            a = np.fft.irfft(np.fft.rfft(a, axis=ax) * filt.reshape([-1 if i == ax else 1 for i in range(len(shape))]),
                             n=n, axis=ax)
    a -= a.mean()
    m = np.abs(a).max()
    return a / m if m > 0 else a


def orthorhombic_from_velocity(vp, vs, rho, pert=None):
    """Nine code-convention stiffnesses for an isotropic-valued orthorhombic
    medium (C11=C22=C33=rho vp^2, C44=C55=C66=rho vs^2, C12=C13=C23=C11-2C44)
    with optional per-component multiplicative perturbations."""
    c11 = rho * vp * vp
    c44 = rho * vs * vs
    c12 = c11 - 2 * c44
    vals = [c11, c11, c11, c12, c12, c12, 2 * c44, 2 * c44, 2 * c44]
    out = []
    for i, v in enumerate(vals):
        a = v * (1.0 + pert[i]) if pert is not None else v
        out.append(np.asarray(np.float32(a), np.float64) if np.ndim(a) == 0 else a.astype(np.float32).astype(np.float64))
    return out


def uniform_model(n=(80, 80, 80), h=10.0, vp=3000.0, vs=1732.0, rho=2000.0, nsteps=600,
                  src=(40, 40, 20), fpeak=15.0, explosive=True, iso=False) -> Model:
    """Single uniform grid (the reference's CPU-runnable case): homogeneous
    orthorhombic medium at isotropic values, free top, explosive source."""
    nx, ny, nz = n
    shape = (nz + 1, ny + 1, nx + 1)
    buoy = np.full(shape, np.float32(1.0 / rho), np.float64)
    if iso:
        s11 = np.full(shape, np.float32(rho * vp * vp), np.float64)
        s13 = np.full(shape, np.float32(rho * vp * vp - 2 * rho * vs * vs), np.float64)
        stiff = [s11, s13]
    else:
        stiff = [np.full(shape, float(v), np.float64) for v in
                 (np.asarray(x).item() for x in orthorhombic_from_velocity(vp, vs, rho))]
    G = Grid(n, (h, h, h), buoy, stiff)
    dt = cfl_dt([G], iso)
    rec = np.array([[0, ix, ny // 2, 2] for ix in range(nx + 1)], np.int32)
    return Model(grids=[G], dt=dt, nsteps=nsteps, iso=iso, top_bc=1, src_grid=0, src_loc=tuple(src),
                 src_type=1 if explosive else 0, src_trace=ricker(nsteps, dt, fpeak),
                 rec_loc=rec, name="uniform_%dx%dx%d" % n)


def layered_composite(sizes, steps_h, layers_fn, nsteps, seed=0, pert=0.1, corr_m=50.0,
                      fpeak=20.0, src_depth=30.0, nrec_side=None, top_bc=1) -> Model:
    """Composite model with per-grid orthorhombic media.

    sizes: list of cells (nx, ny, nz) per grid, top to bottom;
    steps_h: list of grid steps; layers_fn(z) -> (vp, vs, rho) at depth z (m);
    pert: relative amplitude of the smoothed random perturbation of each Cij and rho.
    """
    rng = np.random.default_rng(seed)
    grids = []
    z0 = 0.0
    for n, h in zip(sizes, steps_h):
        nx, ny, nz = n
        z = z0 + h * np.arange(nz + 1)
        vp = np.empty(nz + 1)
        vs = np.empty(nz + 1)
        rho = np.empty(nz + 1)
        for k, zz in enumerate(z):
            vp[k], vs[k], rho[k] = layers_fn(zz)
        shape = (nz + 1, ny + 1, nx + 1)
        corr = corr_m / h
        small = tuple(max(2, min(s, 64)) for s in shape)

        def field_noise():
            if pert == 0:
                return np.zeros(shape)
            base = _smooth_noise(small, corr * small[2] / shape[2], rng)
            # upsample the small smoothed field to the grid by linear interpolation
            out = base
            for ax in range(3):
                src_n = out.shape[ax]
                dst_n = shape[ax]
                x = np.linspace(0, src_n - 1, dst_n)
                i0 = np.floor(x).astype(int)
                i1 = np.minimum(i0 + 1, src_n - 1)
                w = (x - i0).reshape([-1 if i == ax else 1 for i in range(3)])
                out = np.take(out, i0, axis=ax) * (1 - w) + np.take(out, i1, axis=ax) * w
            return pert * out

        rho3 = rho[:, None, None] * (1.0 + field_noise())
        vp3 = np.broadcast_to(vp[:, None, None], shape)
        vs3 = np.broadcast_to(vs[:, None, None], shape)
        perts = [field_noise() for _ in range(9)]
        stiff = orthorhombic_from_velocity(vp3, vs3, rho[:, None, None], perts)
        buoy = (1.0 / rho3).astype(np.float32).astype(np.float64)
        grids.append(Grid(tuple(n), (h, h, h), buoy, stiff, origin_z=z0))
        z0 += h * nz
    dt = cfl_dt(grids, iso=False)
    g0 = grids[0]
    sx = g0.n[0] // 2 * g0.h[0]
    sy = g0.n[1] // 2 * g0.h[1]
    sg, six, siy, siz = map_location(grids, sx, sy, src_depth)
    if nrec_side is None:
        nrec_side = 2
    xs = np.linspace(0, g0.n[0], nrec_side + 2)[1:-1].round().astype(int)
    ys = np.linspace(0, g0.n[1], nrec_side + 2)[1:-1].round().astype(int)
    rec = np.array([[0, x, y, 1] for y in ys for x in xs], np.int32)
    return Model(grids=grids, dt=dt, nsteps=nsteps, iso=False, top_bc=top_bc, src_grid=sg,
                 src_loc=(six, siy, siz), src_type=0, src_trace=ricker(nsteps, dt, fpeak),
                 rec_loc=rec, name="composite_%dgrid" % len(grids))


def random_fields_like(model: Model, seed: int = 0, scale: float = 1.0):
    """Random wavefields (with halos) for kernel-level parity checks."""
    rng = np.random.default_rng(seed)
    out = []
    for G in model.grids:
        per = []
        for c in range(9):
            shape = field_shape(G.n, FIELD_POS[c], model.radius)
            per.append(rng.standard_normal(shape) * scale)
        out.append(per)
    return out


def field_shape(n, pos, r):
    """(len2, len1, len0) of a Field with halo r (field.cpp:85-98)."""
    lens = [(n[d] + 2 * r) if pos[d] else (n[d] + 1 + 2 * r) for d in range(3)]
    return (lens[2], lens[1], lens[0])


# ---------------------------------------------------------------------------
# the throughput workloads

def _layered_orthorhombic(sizes, steps_h, vp_fn, rng, pert, vs_over_vp=1 / 1.8):
    """Per-grid orthorhombic media: layered vp(z), vs = vp*vs_over_vp,
    rho = 1700 + 0.2 vp, each of the 9 stiffnesses and rho multiplied by an
    independent seeded uniform factor in [1-pert, 1+pert] per node."""
    grids = []
    z0 = 0.0
    for n, h in zip(sizes, steps_h):
        nx, ny, nz = n
        shape = (nz + 1, ny + 1, nx + 1)
        z = z0 + h * np.arange(nz + 1)
        vp = vp_fn(z)
        vs = vp * vs_over_vp
        rho = 1700.0 + 0.2 * vp
        c11 = rho * vp * vp
        c44 = rho * vs * vs
        c12 = c11 - 2 * c44
        base = [c11, c11, c11, c12, c12, c12, 2 * c44, 2 * c44, 2 * c44]

        def noisy(prof):
            f = rng.random(shape, dtype=np.float32)
            f *= np.float32(2 * pert)
            f += np.float32(1 - pert)
            f *= prof.astype(np.float32)[:, None, None]
            return f.astype(np.float64)

        stiff = [noisy(b) for b in base]
        rh = noisy(rho)
        buoy = (1.0 / rh).astype(np.float32).astype(np.float64)
        del rh
        grids.append(Grid(tuple(n), (h, h, h), buoy, stiff, origin_z=z0))
        z0 += h * nz
    return grids


def three_grid_workload(seed: int = 0, nsteps: int = 2000, nrec=(40, 30), fpeak: float = 15.0) -> Model:
    """Three-grid composite throughput workload: 1207x1207x81 nodes at h=5 m,
    403x403x101 at 15 m, 135x135x121 at 45 m (ratio 3 at each interface; the
    sizes that make the three grids align, 136.6 M nodes), layered
    orthorhombic medium with vp 1000 -> 5000 m/s linear in depth and +-10%
    seeded per-node perturbation of the 9 Cij and rho, explosive Ricker at
    50 m depth, nrec[0] x nrec[1] receivers on the surface grid."""
    sizes = [(1206, 1206, 80), (402, 402, 100), (134, 134, 120)]
    hs = [5.0, 15.0, 45.0]
    ztot = sum(n[2] * h for n, h in zip(sizes, hs))
    rng = np.random.default_rng(seed)
    grids = _layered_orthorhombic(sizes, hs, lambda z: 1000.0 + 4000.0 * z / ztot, rng, 0.10)
    dt = cfl_dt(grids, iso=False, frac=0.9)
    g0 = grids[0]
    src = map_location(grids, g0.n[0] // 2 * g0.h[0], g0.n[1] // 2 * g0.h[1], 50.0)
    xs = np.linspace(0, g0.n[0], nrec[0] + 2)[1:-1].round().astype(int)
    ys = np.linspace(0, g0.n[1], nrec[1] + 2)[1:-1].round().astype(int)
    rec = np.array([[0, x, y, 1] for y in ys for x in xs], np.int32)
    return Model(grids=grids, dt=dt, nsteps=nsteps, iso=False, top_bc=1, src_grid=src[0], src_loc=src[1:],
                 src_type=1, src_trace=ricker(min(nsteps, 400), dt, fpeak), rec_loc=rec,
                 name="three_grid_1207x1207x81_403x403x101_135x135x121")


def two_grid_workload(seed: int = 0, nsteps: int = 1500) -> Model:
    """Two-grid composite at ratio 3: 241x241x61 nodes at h=5 m over a
    low-velocity top layer (vp 1200, vs 600, rho 1800) and 81x81x121 at
    h=15 m below (vp 3500, vs 2000, rho 2300), orthorhombic stiffnesses with
    +-5% seeded Gaussian-smoothed perturbation (correlation 50 m), Ricker
    20 Hz at 30 m depth."""
    rng = np.random.default_rng(seed)
    sizes = [(240, 240, 60), (80, 80, 120)]
    hs = [5.0, 15.0]
    props = [(1200.0, 600.0, 1800.0), (3500.0, 2000.0, 2300.0)]
    grids = []
    z0 = 0.0
    for n, h, (vp, vs, rho) in zip(sizes, hs, props):
        nx, ny, nz = n
        shape = (nz + 1, ny + 1, nx + 1)
        c11 = rho * vp * vp
        c44 = rho * vs * vs
        c12 = c11 - 2 * c44
        base = [c11, c11, c11, c12, c12, c12, 2 * c44, 2 * c44, 2 * c44]
        corr = 50.0 / h
        stiff = [(v * (1 + 0.05 * _smooth_noise(shape, corr, rng))).astype(np.float32).astype(np.float64) for v in base]
        buoy = np.full(shape, np.float32(1.0 / rho), np.float64)
        grids.append(Grid(tuple(n), (h, h, h), buoy, stiff, origin_z=z0))
        z0 += h * nz
    dt = cfl_dt(grids, iso=False, frac=0.9)
    g0 = grids[0]
    src = map_location(grids, g0.n[0] // 2 * g0.h[0], g0.n[1] // 2 * g0.h[1], 30.0)
    rec = np.array([[0, x, g0.n[1] // 2, 1] for x in range(0, g0.n[0] + 1, 4)], np.int32)
    return Model(grids=grids, dt=dt, nsteps=nsteps, iso=False, top_bc=1, src_grid=src[0], src_loc=src[1:],
                 src_type=0, src_trace=ricker(min(nsteps, 300), dt, 20.0), rec_loc=rec,
                 name="two_grid_241x241x61_81x81x121")


def batch_workload(seed: int = 0, nmodels: int = 16, shots: int = 4, nsteps: int = 1000, pert: float = 0.05) -> List[Model]:
    """Batched shots over uneven two-grid models: nmodels ratio-3 composites
    whose fine grids have (nx, ny) cells drawn from multiples of 3 in
    [120, 240] and nz in [30, 60] at h = 5 m (vp 1200, vs 600, rho 1800), over
    coarse grids of (nx/3, ny/3, 2 nz) cells at 15 m (vp 3500, vs 2000,
    rho 2300); each of the 9 Cij and rho carries a seeded per-node
    +-pert perturbation.  Every model gets `shots` Ricker sources (10-25 Hz,
    seeded positions in the fine grid, vx force) and a receiver line along x.
    Returns nmodels * shots Models; the shots of one model share its media."""
    rng = np.random.default_rng(seed)
    props = [(1200.0, 600.0, 1800.0), (3500.0, 2000.0, 2300.0)]
    out = []
    for mi in range(nmodels):
        nx = 3 * int(rng.integers(40, 81))
        ny = 3 * int(rng.integers(40, 81))
        nz = int(rng.integers(30, 61))
        sizes = [(nx, ny, nz), (nx // 3, ny // 3, 2 * nz)]
        hs = [5.0, 15.0]
        grids = []
        z0 = 0.0
        for n, h, (vp, vs, rho) in zip(sizes, hs, props):
            shape = (n[2] + 1, n[1] + 1, n[0] + 1)
            c11 = rho * vp * vp
            c44 = rho * vs * vs
            c12 = c11 - 2 * c44
            base = [c11, c11, c11, c12, c12, c12, 2 * c44, 2 * c44, 2 * c44]

            def noisy(v):
                f = rng.random(shape, dtype=np.float32)
                f *= np.float32(2 * pert)
                f += np.float32(1 - pert)
                f *= np.float32(v)
                return f.astype(np.float64)

            stiff = [noisy(v) for v in base]
            buoy = (1.0 / noisy(rho)).astype(np.float32).astype(np.float64)
            grids.append(Grid(tuple(n), (h, h, h), buoy, stiff, origin_z=z0))
            z0 += h * n[2]
        dt = cfl_dt(grids, iso=False, frac=0.9)
        g0 = grids[0]
        rec = np.array([[0, x, g0.n[1] // 2, 1] for x in range(0, g0.n[0] + 1, 4)], np.int32)
        for si in range(shots):
            fpk = float(rng.uniform(10.0, 25.0))
            x = float(rng.uniform(0.2, 0.8)) * g0.n[0] * g0.h[0]
            y = float(rng.uniform(0.2, 0.8)) * g0.n[1] * g0.h[1]
            z = float(rng.uniform(10.0, 0.8 * g0.n[2] * g0.h[2]))
            src = map_location(grids, x, y, z)
            out.append(Model(grids=grids, dt=dt, nsteps=nsteps, iso=False, top_bc=1, src_grid=src[0], src_loc=src[1:],
                             src_type=0, src_trace=ricker(min(nsteps, 400), dt, fpk), rec_loc=rec,
                             name="batch_m%02d_s%d_%dx%dx%d" % (mi, si, nx + 1, ny + 1, nz + 1)))
    return out
