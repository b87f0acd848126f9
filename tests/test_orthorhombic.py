"""Nine distinct stiffnesses: the orthorhombic stress update
(update_stress.cpp:17-133, the free-surface rows :37-65) and the parameter
staging (init_del_par_field.cpp:77-150, ReshapeField field.cpp:428-476),
pinned by closed forms on the oracle (CPU) and on the HIP path (gpu).

* Uniform strain: a velocity field linear in position makes every staggered
  difference exact, so one stress update must give dt * C * eps at every
  stress node, with C the 6 x 6 orthorhombic stiffness (Voigt; the code's
  shear entries are 2*C44, 2*C55, 2*C66 and the update multiplies by 1/2),
  and on the free-surface plane the moduli condensed by szz = 0:
  ezz = -(C13 exx + C23 eyy) / C33.  Each of the 36 transpositions of two of
  the nine input arrays changes the result (checked on the oracle: a planted
  c12<->c13 or c55<->c66 swap in the arrays the update reads fails it).
* Standing modes: along each axis a, with velocity component b and stress
  s_ab, the 2nd-order staggered scheme evolves cos/sin modes as
  B' = B - b A G, A' = A + c B' G, G = 2 (dt/h) sin(k h / 2), with c = C11,
  C22, C33 for the three P modes and C44 (yz), C55 (xz), C66 (xy) for the six
  S modes: each modulus is read where the physics puts it.
* Staging: primary-node stiffnesses linear in position are averaged onto
  their staggered nodes, so the staged value is the linear function at the
  node's own position: c11..c12 at (x+h/2, y, z), c44 at (x+h/2, y+h/2,
  z+h/2), c55 at (x, y, z+h/2), c66 at (x, y+h/2, z).
"""
import itertools
import math

import numpy as np
import pytest

from ecsg_amd.model import FIELD_POS, Grid, Model, STIFF_NAMES
from oracle_lib import Oracle

RHO = 2000.0
# Voigt moduli (Pa); the normal block is diagonally dominant, so positive definite
VOIGT = dict(c11=9.1e9, c22=8.3e9, c33=7.7e9, c23=2.9e9, c13=3.3e9, c12=2.3e9, c44=2.1e9, c55=2.6e9, c66=3.1e9)


def code_values(v=VOIGT):
    """The nine input values in the code convention (c44 = 2*C44 etc.)."""
    return [v[k] * (2.0 if k in ("c44", "c55", "c66") else 1.0) for k in STIFF_NAMES]


def homogeneous(n, top_bc, values, h=10.0):
    nx, ny, nz = n
    shape = (nz + 1, ny + 1, nx + 1)
    g = Grid(n, (h, h, h), np.full(shape, 1.0 / RHO), [np.full(shape, c) for c in values])
    vp = math.sqrt(max(VOIGT["c11"], VOIGT["c22"], VOIGT["c33"]) / RHO)
    dt = 0.5 * h / (vp * math.sqrt(3))
    return Model(grids=[g], dt=dt, nsteps=8, iso=False, top_bc=top_bc, src_trace=np.zeros(0, np.float32),
                 rec_loc=np.zeros((0, 4), np.int32))


class Backend:
    """The oracle or the HIP path behind one small interface."""

    def __init__(self, kind, m):
        self.kind = kind
        if kind == "oracle":
            self.s = Oracle(m, eci_mode=1, check_energy=False)
        else:
            import ecsg_amd
            self.s = ecsg_amd.Simulation(m, precision="f64", check_energy=False)

    def shape(self, sel):
        d = self.s.dims(0, sel)
        return d[0] if self.kind == "oracle" else tuple(d)

    def get(self, sel):
        if self.kind != "oracle":
            self.s.sync()
        return self.s.get(0, sel)

    def set(self, sel, a):
        self.s.set(0, sel, a)

    def stress(self):
        self.s.update_stress(0)

    def step(self):
        if self.kind == "oracle":
            self.s.run(1)
        else:
            self.s.step(1, graph=False)

    def close(self):
        self.s.close()


def coords(shape, pos, h, halo):
    """Node coordinates (z, y, x) of a Field of this shape and position."""
    z = (np.arange(shape[0]) - halo + 0.5 * pos[2]) * h
    y = (np.arange(shape[1]) - halo + 0.5 * pos[1]) * h
    x = (np.arange(shape[2]) - halo + 0.5 * pos[0]) * h
    return np.meshgrid(z, y, x, indexing="ij")


# velocity gradient a[i][j] = d v_i / d x_j, generic values
GRAD = np.array([[0.31, -0.17, 0.23], [0.11, -0.29, 0.37], [-0.41, 0.19, 0.13]]) * 1e-3


def uniform_strain_errors(kind, inputs, top_bc, expect=code_values()):
    """Max relative error of one stress update of the linear velocity field
    GRAD.x against dt * C * eps computed from `expect` (the intended values),
    while the backend reads `inputs`."""
    n = (10, 9, 8)
    m = homogeneous(n, top_bc, inputs)
    b = Backend(kind, m)
    h, dt = m.grids[0].h[0], m.dt
    for c in range(3):
        sh = b.shape(c)
        Z, Y, X = coords(sh, FIELD_POS[c], h, 1)
        b.set(c, GRAD[c, 0] * X + GRAD[c, 1] * Y + GRAD[c, 2] * Z)
    for c in range(3, 9):
        b.set(c, np.zeros(b.shape(c)))
    b.stress()
    got = {c: b.get(c) for c in range(3, 9)}
    b.close()
    C = dict(zip(STIFF_NAMES, expect))
    exx, eyy, ezz = GRAD[0, 0], GRAD[1, 1], GRAD[2, 2]
    want = {
        3: dt * (C["c11"] * exx + C["c12"] * eyy + C["c13"] * ezz),
        4: dt * (C["c12"] * exx + C["c22"] * eyy + C["c23"] * ezz),
        5: dt * (C["c13"] * exx + C["c23"] * eyy + C["c33"] * ezz),
        6: dt * 0.5 * C["c44"] * (GRAD[1, 2] + GRAD[2, 1]),
        7: dt * 0.5 * C["c55"] * (GRAD[0, 2] + GRAD[2, 0]),
        8: dt * 0.5 * C["c66"] * (GRAD[0, 1] + GRAD[1, 0]),
    }
    # free surface: szz = 0 on the top plane, ezz condensed out
    ezz_fs = -(C["c13"] * exx + C["c23"] * eyy) / C["c33"]
    want_fs = {
        3: dt * (C["c11"] * exx + C["c12"] * eyy + C["c13"] * ezz_fs),
        4: dt * (C["c12"] * exx + C["c22"] * eyy + C["c23"] * ezz_fs),
        5: 0.0,
    }
    scale = max(abs(v) for v in want.values())
    errs = {}
    for c in range(3, 9):
        g = got[c][1:-1, 1:-1, 1:-1]  # the Field's own nodes (halo dropped)
        w = np.full(g.shape, want[c])
        if top_bc == 1 and c in want_fs:
            w[0] = want_fs[c]
        errs[c] = float(np.abs(g - w).max()) / scale
    return errs


@pytest.mark.parametrize("top_bc", [0, 1])
def test_uniform_strain_oracle(top_bc):
    errs = uniform_strain_errors("oracle", code_values(), top_bc)
    assert max(errs.values()) <= 1e-13, errs


def test_every_planted_stiffness_swap_fails_the_oracle():
    """A transposition of two of the nine arrays the update reads (e.g. c12
    and c13, or c55 and c66) is caught, on the interior and on the free
    surface."""
    vals = code_values()
    for i, j in itertools.combinations(range(9), 2):
        swapped = list(vals)
        swapped[i], swapped[j] = swapped[j], swapped[i]
        errs = uniform_strain_errors("oracle", swapped, 1, expect=vals)
        assert max(errs.values()) >= 1e-3, (STIFF_NAMES[i], STIFF_NAMES[j], errs)


MODES = [  # (axis, velocity component, stress selector, Voigt modulus)
    (0, 0, 3, "c11"), (1, 1, 4, "c22"), (2, 2, 5, "c33"),
    (0, 1, 8, "c66"), (1, 0, 8, "c66"),
    (0, 2, 7, "c55"), (2, 0, 7, "c55"),
    (1, 2, 6, "c44"), (2, 1, 6, "c44"),
]


def standing_mode_errors(kind, mode, inputs=None, nsteps=2):
    axis, vc, ssel, cname = mode
    L, W = 24, 14
    n = [W, W, W]
    n[axis] = L
    m = homogeneous(tuple(n), 0, code_values() if inputs is None else inputs)
    b = Backend(kind, m)
    h, dt = m.grids[0].h[0], m.dt
    k = 3 * math.pi / (L * h)
    for c in range(9):
        b.set(c, np.zeros(b.shape(c)))
    ss = b.shape(ssel)
    S = coords(ss, FIELD_POS[ssel], h, 1)[2 - axis]
    b.set(ssel, np.cos(k * S))
    G = 2 * (dt / h) * math.sin(k * h / 2)
    A, B = 1.0, 0.0
    for _ in range(nsteps):
        b.step()
        B = B - (1.0 / RHO) * A * G
        A = A + VOIGT[cname] * B * G
    v, s = b.get(vc), b.get(ssel)
    b.close()
    V = coords(v.shape, FIELD_POS[vc], h, 1)[2 - axis]
    want_v, want_s = B * np.sin(k * V), A * np.cos(k * S)
    # away from every wall by more than the 2 cells a step reaches
    lo = 2 * nsteps + 3

    def centre(a):
        return a[lo:a.shape[0] - lo, lo:a.shape[1] - lo, lo:a.shape[2] - lo]
    ev = float(np.abs(centre(v) - centre(want_v)).max()) / abs(B)
    es = float(np.abs(centre(s) - centre(want_s)).max()) / abs(A)
    return ev, es


@pytest.mark.parametrize("mode", MODES, ids=["%s-v%s" % (m[3], "xyz"[m[1]]) for m in MODES])
def test_standing_modes_oracle(mode):
    ev, es = standing_mode_errors("oracle", mode)
    assert ev <= 1e-12 and es <= 1e-12, (ev, es)


def test_standing_mode_detects_shear_swap_oracle():
    """c55 and c66 exchanged in the arrays: the x-propagating S modes run at the wrong rates."""
    vals = code_values()
    vals[7], vals[8] = vals[8], vals[7]
    for mode in (MODES[3], MODES[5]):
        ev, es = standing_mode_errors("oracle", mode, inputs=vals)
        assert max(ev, es) >= 1e-4, (mode, ev, es)


SLOPES = [(1.0 + 0.1 * m, 0.37 + 0.05 * m, -0.23 + 0.07 * m) for m in range(9)]


def linear_media(n, h):
    """Primary-node stiffnesses c_m = base_m (1 + (sx x + sy y + sz z) / 1000 m)."""
    nx, ny, nz = n
    Z, Y, X = coords((nz + 1, ny + 1, nx + 1), (0, 0, 0), h, 0)
    out = []
    for m, base in enumerate(code_values()):
        sx, sy, sz = SLOPES[m]
        out.append(base * (1 + (sx * X + sy * Y + sz * Z) / 1000.0))
    return out


def staging_errors(kind):
    n, h = (7, 6, 5), 10.0
    m = homogeneous(n, 1, code_values(), h=h)
    m.grids[0].stiff = linear_media(n, h)
    b = Backend(kind, m)
    pos = [(1, 0, 0)] * 6 + [(1, 1, 1), (0, 0, 1), (0, 1, 0)]
    errs = []
    for mi in range(9):
        got = b.get(10 + mi)
        Z, Y, X = coords(got.shape, pos[mi], h, 0)
        sx, sy, sz = SLOPES[mi]
        want = code_values()[mi] * (1 + (sx * X + sy * Y + sz * Z) / 1000.0)
        errs.append(float(np.abs(got - want).max() / np.abs(want).max()))
    b.close()
    return errs


def test_staging_positions_oracle():
    errs = staging_errors("oracle")
    assert max(errs) <= 1e-14, errs


# ---------------------------------------------------------------------------
# the HIP path, through the C ABI

@pytest.mark.gpu
@pytest.mark.parametrize("top_bc", [0, 1])
def test_uniform_strain_gpu(top_bc):
    errs = uniform_strain_errors("hip", code_values(), top_bc)
    assert max(errs.values()) <= 1e-13, errs


@pytest.mark.gpu
@pytest.mark.parametrize("mode", MODES, ids=["%s-v%s" % (m[3], "xyz"[m[1]]) for m in MODES])
def test_standing_modes_gpu(mode):
    ev, es = standing_mode_errors("hip", mode)
    assert ev <= 1e-12 and es <= 1e-12, (ev, es)


@pytest.mark.gpu
def test_staging_positions_gpu():
    errs = staging_errors("hip")
    assert max(errs) <= 1e-14, errs
