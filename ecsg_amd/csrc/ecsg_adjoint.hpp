// ecsg_adjoint.hpp -- adjoint (transposed) kernels of the composite time step.
// Included by ecsg_capi.hip only.
//
// The forward step is, for fixed parameters, a sequence of linear maps of the
// state (all wavefield entries including halos and interface ghosts):
//   UpdateVel (in-place stencil + ghost-velocity mirrors), receivers, source,
//   UpdateGhostVZ (tridiagonal solve + fine ghosts), UpdateStress (stencil +
//   free-surface mirrors), UpdateGhostSXZ (Jacobi solve + fine ghosts),
//   UpdateGhostSYZ (pointwise + fine ghosts).
// The adjoint step applies the exact transposes in reverse order.  Every
// transpose is written in gather form (each output entry summed by one
// thread in a fixed order), so results are deterministic; the interface maps
// are transposed through their restriction structure (see IfAdj below).  Gradient terms of
// each map with respect to the staged parameters are accumulated in the same
// kernels, using the forward state saved for that step.
//
// Notation: lam = adjoint wavefields (same layout as the forward fields);
// fv / fs = forward v^{n+1/2} / s^n read for the gradient terms.
#pragma once
#include "ecsg_kernels.hpp"

namespace ecsg {

// f[q] += v (adjoint fields; plain loads and stores: the stencils' cache
// hints made these sweeps slower)
template <typename T, typename IT>
__device__ __forceinline__ void acc_st(T *f, IT q, T v)
{
    *at(f, q) += v;
}
// The gradient-only sweeps' read-modify-writes of their accumulators
// (ECSG_GRAD_NT: with the cache hints) as a load (gld) and a store (gst)
// that the sweeps place apart: every load of a plane is issued before its
// first store, since the compiler may not move a load across a store that
// could alias it.
#ifndef ECSG_GRAD_NT
#define ECSG_GRAD_NT 1
#endif
template <typename T, typename IT>
__device__ __forceinline__ T gld(const T *f, IT q)
{
#if ECSG_GRAD_NT
    return ldc(f, q);
#else
    return *at(f, q);
#endif
}
template <typename T, typename IT>
__device__ __forceinline__ void gst(T *f, IT q, T v)
{
#if ECSG_GRAD_NT
    stc(f, q, v);
#else
    *at(f, q) = v;
#endif
}

template <typename T>
struct AdjGridDev {
    T *lam[9];          // adjoint vx vy vz sxx syy szz syz sxz sxy
    const T *fv[3];     // forward velocities of this step (stress transpose)
    const T *fs[6];     // forward stresses before this step (velocity transpose)
    T *gc[9];           // gradient w.r.t. the staged stiffnesses
    T *gb[3];           // gradient w.r.t. buoyancy as used by vx (node), vy (xy mean), vz (xz mean)
    const T *c[9];
    const T *b;
    const T *ac[12];    // masked coefficients (k_adj_coef): normal 11 22 33 23 13 12, 1/2 c44 c55 c66, b of vx vy vz
    int nx, ny, nz;
    long P, S2, o;
    T cx, cy, cz;
    int top_eff, bottom, free_top;
    int gx, gy, gz;     // workgroup grid
};

// Masked coefficients of the transposed stencils: the coefficient of each
// update at the points of that update's set, zero elsewhere (halos
// included), so that the adjoint weights need no set tests.  Normal block on
// the free-surface row: the condensed moduli of update_stress.cpp:37-65.
template <typename T>
__global__ __launch_bounds__(256) void k_adj_coef(GridDev<T> g, T *out, long astride)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const long nxy = (long)(nx + 1) * (ny + 1);
    if (t >= nxy * (nz + 1)) return;
    const int ix = (int)(t % (nx + 1)), iy = (int)((t / (nx + 1)) % (ny + 1)), iz = (int)(t / nxy);
    const long p = g.o + (long)iz * g.S2 + (long)iy * g.P + ix;
    T *o[12];
    for (int k = 0; k < 12; k++) o[k] = out + (long)k * astride;
    const T *b = g.b;
    if (ix <= nx - 1) {
        const T c11 = g.c[0][p], c22 = g.c[1][p], c33 = g.c[2][p], c23 = g.c[3][p], c13 = g.c[4][p], c12 = g.c[5][p];
        if (g.free_top && iz == 0) {
            o[0][p] = c11 - c13 * c13 / c33;
            o[1][p] = c22 - c23 * c23 / c33;
            o[5][p] = c12 - c13 * c23 / c33;
        } else {
            o[0][p] = c11; o[1][p] = c22; o[2][p] = c33; o[3][p] = c23; o[4][p] = c13; o[5][p] = c12;
        }
    }
    if (ix <= nx - 1 && iy <= ny - 1 && iz <= nz - 1) o[6][p] = T(0.5) * g.c[6][p];
    if (iz <= nz - 1) o[7][p] = T(0.5) * g.c[7][p];
    if (iy <= ny - 1) o[8][p] = T(0.5) * g.c[8][p];
    const bool zin = iz >= g.top_eff && iz <= nz - g.bottom;
    if (ix >= 1 && ix <= nx - 1 && iy >= 1 && iy <= ny - 1 && zin) o[9][p] = b[p];
    if (ix <= nx - 1 && iy <= ny - 1 && zin) o[10][p] = T(0.25) * (b[p] + b[p + 1] + b[p + g.P] + b[p + g.P + 1]);
    if (ix <= nx - 1 && iy >= 1 && iy <= ny - 1 && iz <= nz - 1)
        o[11][p] = T(0.25) * (b[p] + b[p + g.S2] + b[p + 1] + b[p + g.S2 + 1]);
}


// ---------------------------------------------------------------------------
// Transpose of UpdateStress (stress stencil) on the velocities (the gradient
// w.r.t. the nine staged stiffnesses is k_adj_grad's).  Forward:
//   normal:  sxx += c11 cx dvx + c12 cy dvy + c13 cz dvz (free row: condensed moduli)
//   syz += 1/2 c44 (cz dz vy + cy dy vz), sxz += 1/2 c55 (...), sxy += 1/2 c66 (...)
// The adjoint weights e = C lam (normal) and h = 1/2 c lam (shear) live on
// the stress points and vanish outside each stress field's update set.
// Adjoint stencil workgroups: 64 x AWY threads, each wavefront one x row of
// 64 lanes of which lanes 1..AXO produce outputs (the outer lanes only supply
// the x neighbours through wavefront shuffles); each thread marches AZ2 z
// planes, carrying its z neighbours and the y neighbours it needs as
// one-plane lookaheads in registers.  No barriers.
constexpr int AWY = 4;
constexpr int AXO = 62;
constexpr int AZ2 = 16;

template <typename T>
__device__ __forceinline__ T shfl_dn(T v) { return __shfl_down(v, 1, 64); }
template <typename T>
__device__ __forceinline__ T shfl_upw(T v) { return __shfl_up(v, 1, 64); }

// Workgroups take their tiles in the forward sweeps' banded XCD order
// (sweep_tile): k_adj_stress 2.51 -> 2.29 ms and k_adj_vel 2.04 -> 1.85 ms on
// the 1207x1207x81 grid against one contiguous tile range per XCD.
// The weights of the neighbouring rows (ey at iy+1; hxy, hyz at iy-1) come
// from the other wavefronts of the workgroup through LDS (xs: two buffers of
// three AWY x 64 planes, one barrier per z plane); only the workgroup's edge
// rows form them from memory.  Forming them from memory on every row
// re-read the y-neighbour rows: 1.41x the sweep's algorithmic bytes.
constexpr int AXS = 2 * 3 * AWY * 64;  // LDS entries of the exchange

template <typename T, typename IT>
__device__ __forceinline__ void adj_stress_body(const AdjGridDev<T> &a, const int tile, T *xs)
{
    const int bx = tile % a.gx;
    const int rest = tile / a.gx;
    const int by = rest % a.gy;
    const int bz = rest / a.gy;
    const int lane = threadIdx.x, ty = threadIdx.y;
    const int ix = bx * AXO + lane - 2;
    const int iy = by * AWY + ty - 1;
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    // rows past ny have no outputs but take part in the barriers
    const bool row = iy <= ny;
    const bool inx = row && ix >= -1 && ix <= nx + 1;
    const bool out = row && lane >= 1 && lane <= AXO && ix <= nx;
    const int z0 = bz * AZ2 - 1;
    const int z1 = min(z0 + AZ2, nz + 1);
    const IT P = (IT)a.P, S2 = (IT)a.S2;
    const T cx = a.cx, cy = a.cy, cz = a.cz;
    const T *a11 = a.ac[0], *a22 = a.ac[1], *a33 = a.ac[2], *a23 = a.ac[3], *a13 = a.ac[4], *a12 = a.ac[5];
    const T *h44 = a.ac[6], *h55 = a.ac[7], *h66 = a.ac[8];
    const T *lxx = a.lam[3], *lyy = a.lam[4], *lzz = a.lam[5], *lyz = a.lam[6], *lxz = a.lam[7], *lxy = a.lam[8];
    const bool ym_ok = iy - 1 >= -1;
    const IT qb = (IT)a.o + (IT)iy * P + (IT)ix;
    // adjoint weights of this thread's point at plane z (zero outside the array box)
    auto own = [&](int z, T &ex, T &ey, T &ez, T &hyz, T &hxz, T &hxy) {
        ex = ey = ez = hyz = hxz = hxy = T(0);
        if (!inx || z < -1 || z > nz + 1) return;
        const IT q = qb + (IT)z * S2;
        const T l1 = (*at(lxx, q)), l2 = (*at(lyy, q)), l3 = (*at(lzz, q));
        const T A11 = (*at(a11, q)), A22 = (*at(a22, q)), A33 = (*at(a33, q)), A23 = (*at(a23, q)), A13 = (*at(a13, q)), A12 = (*at(a12, q));
        ex = cx * (A11 * l1 + A12 * l2 + A13 * l3);
        ey = cy * (A12 * l1 + A22 * l2 + A23 * l3);
        ez = cz * (A13 * l1 + A23 * l2 + A33 * l3);
        hyz = (*at(h44, q)) * (*at(lyz, q));
        hxz = (*at(h55, q)) * (*at(lxz, q));
        hxy = (*at(h66, q)) * (*at(lxy, q));
    };
    // the neighbouring rows' weights at plane z
    int buf = 0;
    auto nbr = [&](int z, T ey, T hyz, T hxy, T &eyp, T &hxym, T &hyzm) {
        T *e = xs + buf * 3 * AWY * 64, *hz = e + AWY * 64, *hy = hz + AWY * 64;
        buf ^= 1;
        e[ty * 64 + lane] = ey;
        hz[ty * 64 + lane] = hyz;
        hy[ty * 64 + lane] = hxy;
        __syncthreads();
        eyp = hxym = hyzm = T(0);
        const bool zok = inx && z >= -1 && z <= nz + 1;
        if (ty + 1 < AWY) {
            eyp = e[(ty + 1) * 64 + lane];
        } else if (zok) {
            const IT qp = qb + (IT)z * S2 + P;
            eyp = cy * ((*at(a12, qp)) * (*at(lxx, qp)) + (*at(a22, qp)) * (*at(lyy, qp)) + (*at(a23, qp)) * (*at(lzz, qp)));
        }
        if (ty > 0) {
            hxym = hy[(ty - 1) * 64 + lane];
            hyzm = hz[(ty - 1) * 64 + lane];
        } else if (zok && ym_ok) {
            const IT qm = qb + (IT)z * S2 - P;
            hxym = (*at(h66, qm)) * (*at(lxy, qm));
            hyzm = (*at(h44, qm)) * (*at(lyz, qm));
        }
    };
    T ex, ey, ez, hyz, hxz, hxy, eyp, hxym, hyzm;
    T d0, d1, d2, d3, hyz_m, hxz_m;
    own(z0 - 1, d0, d1, d2, hyz_m, hxz_m, d3);
    own(z0, ex, ey, ez, hyz, hxz, hxy);
    nbr(z0, ey, hyz, hxy, eyp, hxym, hyzm);
    for (int iz = z0; iz < z1; iz++) {
        const IT q = qb + (IT)iz * S2;
        T ex_n, ey_n, ez_n, hyz_n, hxz_n, hxy_n, eyp_n, hxym_n, hyzm_n;
        own(iz + 1, ex_n, ey_n, ez_n, hyz_n, hxz_n, hxy_n);
        nbr(iz + 1, ey_n, hyz_n, hxy_n, eyp_n, hxym_n, hyzm_n);
        const T ex_xm = shfl_upw(ex);
        const T hxz_xp = shfl_dn(hxz), hxy_xp = shfl_dn(hxy);
        if (out) {
            if (ix >= 0 && iy >= 0 && iz >= 0) acc_st(a.lam[0], q, (ex_xm - ex) + cz * (hxz_m - hxz) + cy * (hxym - hxy));
            if (iz >= 0) acc_st(a.lam[1], q, (ey - eyp) + cz * (hyz_m - hyz) + cx * (hxy - hxy_xp));
            if (iy >= 0) acc_st(a.lam[2], q, (ez - ez_n) + cy * (hyzm - hyz) + cx * (hxz - hxz_xp));
        }
        hyz_m = hyz; hxz_m = hxz;
        ex = ex_n; ey = ey_n; ez = ez_n; hyz = hyz_n; hxz = hxz_n; hxy = hxy_n;
        eyp = eyp_n; hxym = hxym_n; hyzm = hyzm_n;
    }
}
template <typename T, typename IT>
__global__ __launch_bounds__(64 *AWY) void k_adj_stress(AdjGridDev<T> a)
{
    __shared__ T xs[AXS];
    adj_stress_body<T, IT>(a, sweep_tile(blockIdx.x, a.gx * a.gy * a.gz, a.gx), xs);
}


// Gradient w.r.t. the nine staged stiffnesses at the stress points:
// lam_s (after the transposed stress mirrors) times the derivative of the
// stress update w.r.t. each coefficient, from this step's forward velocities.
// Own sweep (64 x GBY threads, GZC planes per thread, the forward velocities'
// z neighbours carried in registers): apart from the transposed stencil its
// register needs are small, and lam_s is not written by k_adj_stress.
constexpr int GBY = 4;
constexpr int GZC = 8;

// workgroups of a gradient sweep over one grid
template <typename T>
__host__ __device__ __forceinline__ int adj_grad_tiles(const AdjGridDev<T> &a)
{
    return ((a.nx + 64) / 64) * ((a.ny + GBY) / GBY) * ((a.nz + GZC) / GZC);
}

template <typename T, typename IT>
__device__ __forceinline__ void adj_grad_body(const AdjGridDev<T> &a, const int tile)
{
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const int gx = (nx + 64) / 64, gy = (ny + GBY) / GBY;
    const int bx = tile % gx, by = (tile / gx) % gy, bz = tile / (gx * gy);
    const int ix = bx * 64 + threadIdx.x, iy = by * GBY + threadIdx.y;
    if (ix > nx || iy > ny) return;
    const int z0 = bz * GZC, z1 = min(z0 + GZC, nz + 1);
    const IT P = (IT)a.P, S2 = (IT)a.S2;
    const T cx = a.cx, cy = a.cy, cz = a.cz;
    const T *lxx = a.lam[3], *lyy = a.lam[4], *lzz = a.lam[5], *lyz = a.lam[6], *lxz = a.lam[7], *lxy = a.lam[8];
    const T *fvx = a.fv[0], *fvy = a.fv[1], *fvz = a.fv[2];
    IT q = (IT)a.o + (IT)z0 * S2 + (IT)iy * P + (IT)ix;
    T fvz_m = (*at(fvz, q - S2)), fvx_c = (*at(fvx, q)), fvy_c = (*at(fvy, q));
    T *const *gc = a.gc;
    for (int iz = z0; iz < z1; iz++, q += S2) {
        // Every operand of the plane, the nine accumulators included, is
        // loaded before any accumulator is stored (all lie inside the
        // halo-padded arrays for ix <= nx, iy <= ny, iz <= nz): one group of
        // loads in flight per plane instead of one round trip per
        // read-modify-write (1207 x 1207 x 81 grid, FP32: 2.93 -> 2.53 ms).
        const T fvx_n = (*at(fvx, q + S2)), fvy_n = (*at(fvy, q + S2)), fvz_c = (*at(fvz, q));
        const T l1 = (*at(lxx, q)), l2 = (*at(lyy, q)), l3 = (*at(lzz, q)), l4 = (*at(lyz, q)), l5 = (*at(lxz, q)), l6 = (*at(lxy, q));
        const T fvx_e = (*at(fvx, q + 1)), fvy_s = (*at(fvy, q - P)), fvz_n = (*at(fvz, q + P)), fvz_w = (*at(fvz, q - 1));
        const T fvx_nn = (*at(fvx, q + P)), fvy_w = (*at(fvy, q - 1));
        T g[9];
#pragma unroll
        for (int k = 0; k < 9; k++) g[k] = gld(gc[k], q);
        if (ix <= nx - 1) {
            const T X = cx * (fvx_e - fvx_c);
            const T Y = cy * (fvy_c - fvy_s);
            if (a.free_top && iz == 0) {
                const T c33 = (*at(a.c[2], q)), c23 = (*at(a.c[3], q)), c13 = (*at(a.c[4], q));
                const T u = c13 / c33, v = c23 / c33;
                gst(gc[0], q, g[0] + l1 * X);
                gst(gc[1], q, g[1] + l2 * Y);
                gst(gc[5], q, g[5] + (l1 * Y + l2 * X));
                gst(gc[4], q, g[4] + (l1 * (T(-2) * u * X - v * Y) + l2 * (-v * X)));
                gst(gc[3], q, g[3] + (l1 * (-u * Y) + l2 * (-u * X - T(2) * v * Y)));
                gst(gc[2], q, g[2] + (l1 * (u * u * X + u * v * Y) + l2 * (v * u * X + v * v * Y)));
            } else {
                const T Z = cz * (fvz_c - fvz_m);
                gst(gc[0], q, g[0] + l1 * X);
                gst(gc[1], q, g[1] + l2 * Y);
                gst(gc[2], q, g[2] + l3 * Z);
                gst(gc[3], q, g[3] + (l2 * Z + l3 * Y));
                gst(gc[4], q, g[4] + (l1 * Z + l3 * X));
                gst(gc[5], q, g[5] + (l1 * Y + l2 * X));
            }
        }
        if (ix <= nx - 1 && iy <= ny - 1 && iz <= nz - 1)
            gst(gc[6], q, g[6] + l4 * (T(0.5) * (cz * (fvy_n - fvy_c) + cy * (fvz_n - fvz_c))));
        if (iz <= nz - 1)
            gst(gc[7], q, g[7] + l5 * (T(0.5) * (cz * (fvx_n - fvx_c) + cx * (fvz_c - fvz_w))));
        if (iy <= ny - 1)
            gst(gc[8], q, g[8] + l6 * (T(0.5) * (cy * (fvx_nn - fvx_c) + cx * (fvy_c - fvy_w))));
        fvz_m = fvz_c;
        fvx_c = fvx_n;
        fvy_c = fvy_n;
    }
}
template <typename T, typename IT>
__global__ __launch_bounds__(64 * GBY) void k_adj_grad(AdjGridDev<T> a)
{
    adj_grad_body<T, IT>(a, sweep_tile(blockIdx.x, adj_grad_tiles(a), (a.nx + 64) / 64));
}


// ---------------------------------------------------------------------------
// Transpose of UpdateVel (velocity stencil) on the stresses (the gradient
// w.r.t. buoyancy is k_adj_gradb's).  Forward:
//   vx += b (cx dx sxx + cy dy sxy + cz dz sxz)                on the vx update set
//   vy += mean_xy(b) (cx dx sxy + cz dz syz + cy dy syy)       on the vy update set
//   vz += mean_xz(b) (cx dx sxz + cz dz szz + cy dy syz)       on the vz update set
// with the adjoint weights w = (masked buoyancy) * lam_v.
template <typename T, typename IT>
__device__ __forceinline__ void adj_vel_body(const AdjGridDev<T> &a, const int tile, T *xs)
{
    const int bx = tile % a.gx;
    const int rest = tile / a.gx;
    const int by = rest % a.gy;
    const int bz = rest / a.gy;
    const int lane = threadIdx.x, ty = threadIdx.y;
    const int ix = bx * AXO + lane - 1;
    const int iy = by * AWY + ty;
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    // rows past ny have no outputs but take part in the barriers
    const bool row = iy <= ny;
    const bool inx = row && ix >= -1 && ix <= nx + 1;
    const bool out = row && lane >= 1 && lane <= AXO && ix <= nx;
    const int z0 = bz * AZ2 - 1;
    const int z1 = min(z0 + AZ2, nz + 1);
    const IT P = (IT)a.P, S2 = (IT)a.S2;
    const T cx = a.cx, cy = a.cy, cz = a.cz;
    const T *mbx = a.ac[9], *mby = a.ac[10], *mbz = a.ac[11];
    const T *lvx = a.lam[0], *lvy = a.lam[1], *lvz = a.lam[2];
    const bool ym_ok = iy - 1 >= -1;
    const IT qb = (IT)a.o + (IT)iy * P + (IT)ix;
    auto own = [&](int z, T &wx, T &wy, T &wz) {
        wx = wy = wz = T(0);
        if (!inx || z < -1 || z > nz + 1) return;
        const IT q = qb + (IT)z * S2;
        wx = (*at(mbx, q)) * (*at(lvx, q));
        wy = (*at(mby, q)) * (*at(lvy, q));
        wz = (*at(mbz, q)) * (*at(lvz, q));
    };
    // the neighbouring rows' weights (wx, wz at iy+1; wy at iy-1) through
    // LDS, as in adj_stress_body
    int buf = 0;
    auto nbr = [&](int z, T wx, T wy, T wz, T &wxp, T &wzp, T &wym) {
        T *ex = xs + buf * 3 * AWY * 64, *ey = ex + AWY * 64, *ez = ey + AWY * 64;
        buf ^= 1;
        ex[ty * 64 + lane] = wx;
        ey[ty * 64 + lane] = wy;
        ez[ty * 64 + lane] = wz;
        __syncthreads();
        wxp = wzp = wym = T(0);
        const bool zok = inx && z >= -1 && z <= nz + 1;
        if (ty + 1 < AWY) {
            wxp = ex[(ty + 1) * 64 + lane];
            wzp = ez[(ty + 1) * 64 + lane];
        } else if (zok) {
            const IT qp = qb + (IT)z * S2 + P;
            wxp = (*at(mbx, qp)) * (*at(lvx, qp));
            wzp = (*at(mbz, qp)) * (*at(lvz, qp));
        }
        if (ty > 0) wym = ey[(ty - 1) * 64 + lane];
        else if (zok && ym_ok) {
            const IT qm = qb + (IT)z * S2 - P;
            wym = (*at(mby, qm)) * (*at(lvy, qm));
        }
    };
    T wx, wy, wz, wxp, wzp, wym, d0, d1, wz_m;
    own(z0 - 1, d0, d1, wz_m);
    own(z0, wx, wy, wz);
    nbr(z0, wx, wy, wz, wxp, wzp, wym);
    for (int iz = z0; iz < z1; iz++) {
        const IT q = qb + (IT)iz * S2;
        T wx_n, wy_n, wz_n, wxp_n, wzp_n, wym_n;
        own(iz + 1, wx_n, wy_n, wz_n);
        nbr(iz + 1, wx_n, wy_n, wz_n, wxp_n, wzp_n, wym_n);
        const T wx_xp = shfl_dn(wx);
        const T wy_xm = shfl_upw(wy), wz_xm = shfl_upw(wz);
        if (out) {
            if (iz >= 0) {
                if (ix <= nx - 1) {
                    acc_st(a.lam[3], q, cx * (wx - wx_xp));
                    acc_st(a.lam[4], q, cy * (wym - wy));
                    acc_st(a.lam[5], q, cz * (wz_m - wz));
                }
                if (iy <= ny - 1) acc_st(a.lam[8], q, cy * (wx - wxp) + cx * (wy_xm - wy));
            }
            acc_st(a.lam[7], q, cz * (wx - wx_n) + cx * (wz_xm - wz));
            if (ix <= nx - 1 && iy <= ny - 1) acc_st(a.lam[6], q, cz * (wy - wy_n) + cy * (wz - wzp));
        }
        wz_m = wz;
        wx = wx_n; wy = wy_n; wz = wz_n; wxp = wxp_n; wzp = wzp_n; wym = wym_n;
    }
}
template <typename T, typename IT>
__global__ __launch_bounds__(64 *AWY) void k_adj_vel(AdjGridDev<T> a)
{
    __shared__ T xs[AXS];
    adj_vel_body<T, IT>(a, sweep_tile(blockIdx.x, a.gx * a.gy * a.gz, a.gx), xs);
}


// Gradient w.r.t. buoyancy at the velocity points (vx node, vy xy mean, vz
// xz mean): lam_v (after the transposed velocity mirrors) times the stress
// differences of this step's forward stresses s^n.  Own sweep, like
// k_adj_grad; lam_v is not written by k_adj_vel.
template <typename T, typename IT>
__device__ __forceinline__ void adj_gradb_body(const AdjGridDev<T> &a, const int tile)
{
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const int gx = (nx + 64) / 64, gy = (ny + GBY) / GBY;
    const int bx = tile % gx, by = (tile / gx) % gy, bz = tile / (gx * gy);
    const int ix = bx * 64 + threadIdx.x, iy = by * GBY + threadIdx.y;
    if (ix > nx || iy > ny) return;
    const int z0 = bz * GZC, z1 = min(z0 + GZC, nz + 1);
    const IT P = (IT)a.P, S2 = (IT)a.S2;
    const T cx = a.cx, cy = a.cy, cz = a.cz;
    const T *lvx = a.lam[0], *lvy = a.lam[1], *lvz = a.lam[2];
    const T *sxx = a.fs[0], *syy = a.fs[1], *szz = a.fs[2], *syz = a.fs[3], *sxz = a.fs[4], *sxy = a.fs[5];
    IT q = (IT)a.o + (IT)z0 * S2 + (IT)iy * P + (IT)ix;
    T sxz_m = (*at(sxz, q - S2)), syz_m = (*at(syz, q - S2)), szz_c = (*at(szz, q));
    for (int iz = z0; iz < z1; iz++, q += S2) {
        // all loads of the plane before the accumulators are stored
        const T sxz_c = (*at(sxz, q)), syz_c = (*at(syz, q)), szz_n = (*at(szz, q + S2));
        const T l1 = (*at(lvx, q)), l2 = (*at(lvy, q)), l3 = (*at(lvz, q));
        const T sxx_c = (*at(sxx, q)), sxx_w = (*at(sxx, q - 1)), sxy_c = (*at(sxy, q)), sxy_s = (*at(sxy, q - P)), sxy_e = (*at(sxy, q + 1));
        const T syy_c = (*at(syy, q)), syy_n = (*at(syy, q + P)), sxz_e = (*at(sxz, q + 1)), syz_s = (*at(syz, q - P));
        const T g0 = gld(a.gb[0], q), g1 = gld(a.gb[1], q), g2 = gld(a.gb[2], q);
        const bool zin = iz >= a.top_eff && iz <= nz - a.bottom;
        if (ix >= 1 && ix <= nx - 1 && iy >= 1 && iy <= ny - 1 && zin)
            gst(a.gb[0], q, g0 + l1 * (cx * (sxx_c - sxx_w) + cy * (sxy_c - sxy_s) + cz * (sxz_c - sxz_m)));
        if (ix <= nx - 1 && iy <= ny - 1 && zin)
            gst(a.gb[1], q, g1 + l2 * (cx * (sxy_e - sxy_c) + cz * (syz_c - syz_m) + cy * (syy_n - syy_c)));
        if (ix <= nx - 1 && iy >= 1 && iy <= ny - 1 && iz <= nz - 1)
            gst(a.gb[2], q, g2 + l3 * (cx * (sxz_e - sxz_c) + cz * (szz_n - szz_c) + cy * (syz_c - syz_s)));
        sxz_m = sxz_c;
        syz_m = syz_c;
        szz_c = szz_n;
    }
}
template <typename T, typename IT>
__global__ __launch_bounds__(64 * GBY) void k_adj_gradb(AdjGridDev<T> a)
{
    adj_gradb_body<T, IT>(a, sweep_tile(blockIdx.x, adj_grad_tiles(a), (a.nx + 64) / 64));
}


// ---------------------------------------------------------------------------
// Transposed ghost mirrors h = -i: lam[i] -= sum over the mirrors of i of
// lam[h] (gather per source entry), then lam[h] = 0 (k_zero_list).
template <typename T>
__device__ __forceinline__ void mirror_T_body(T *lam, const long *__restrict__ src, const int *__restrict__ start,
                                                  const long *__restrict__ halo, int n, const int t)
{
    if (t >= n) return;
    T acc = lam[src[t]];
    for (int q = start[t]; q < start[t + 1]; q++) acc -= lam[halo[q]];
    lam[src[t]] = acc;
}
template <typename T>
__global__ __launch_bounds__(256) void k_mirror_T(T *lam, const long *__restrict__ src, const int *__restrict__ start,
                                                  const long *__restrict__ halo, int n)
{
    mirror_T_body(lam, src, start, halo, n, (int)(blockIdx.x * blockDim.x + threadIdx.x));
}


template <typename T>
__device__ __forceinline__ void zero_list_body(T *lam, const long *__restrict__ idx, int n, const int t)
{
    if (t < n) lam[idx[t]] = T(0);
}
template <typename T>
__global__ __launch_bounds__(256) void k_zero_list(T *lam, const long *__restrict__ idx, int n)
{
    zero_list_body(lam, idx, n, (int)(blockIdx.x * blockDim.x + threadIdx.x));
}


// Tridiagonal solve with the prefactored elimination of the vz-szz system
// (symmetric, so it also solves the transposed system); solution in B.
__device__ __forceinline__ void tri_solve_body(double *__restrict__ B, const double *__restrict__ M,
                                                  const double *__restrict__ Dp, const double *__restrict__ U, int nx2, int n, const int I)
{
    if (I >= nx2) return;
    constexpr int CH = 16;
    double prev = 0;
    for (int q0 = 0; q0 < n; q0 += CH) {
        double bq[CH], mq[CH];
#pragma unroll
        for (int t = 0; t < CH; t++)
            if (q0 + t < n) {
                bq[t] = B[(long)(q0 + t) * nx2 + I];
                mq[t] = M[(long)(q0 + t) * nx2 + I];
            }
#pragma unroll
        for (int t = 0; t < CH; t++)
            if (q0 + t < n) {
                const double v = (q0 + t == 0) ? bq[t] : bq[t] - mq[t] * prev;
                B[(long)(q0 + t) * nx2 + I] = v;
                prev = v;
            }
    }
    double x = prev / Dp[(long)(n - 1) * nx2 + I];
    B[(long)(n - 1) * nx2 + I] = x;
    for (int q1 = n - 2; q1 >= 0; q1 -= CH) {
        double bq[CH], uq[CH], dq[CH];
#pragma unroll
        for (int t = 0; t < CH; t++)
            if (q1 - t >= 0) {
                bq[t] = B[(long)(q1 - t) * nx2 + I];
                uq[t] = U[(long)(q1 - t) * nx2 + I];
                dq[t] = Dp[(long)(q1 - t) * nx2 + I];
            }
#pragma unroll
        for (int t = 0; t < CH; t++)
            if (q1 - t >= 0) {
                x = (bq[t] - uq[t] * x) / dq[t];
                B[(long)(q1 - t) * nx2 + I] = x;
            }
    }
}
__global__ __launch_bounds__(64) void k_tri_solve(double *__restrict__ B, const double *__restrict__ M,
                                                  const double *__restrict__ Dp, const double *__restrict__ U, int nx2, int n)
{
    tri_solve_body(B, M, Dp, U, nx2, n, (int)(blockIdx.x * blockDim.x + threadIdx.x));
}


// The same with one wavefront per line (lines of at most VZ_WAVE_LDS / 32
// entries): the line staged in LDS by the lanes, the recurrence by lane 0 in
// tri_solve_body's operation order, the solution written back by the lanes.
__device__ __forceinline__ void tri_solve_wave(double *__restrict__ B, const double *__restrict__ M,
                                               const double *__restrict__ Dp, const double *__restrict__ U, int nx2, int n,
                                               const int I, double *sh)
{
    double *sb = sh, *sm = sh + n, *su = sh + 2 * n, *sd = sh + 3 * n;
    const int lane = threadIdx.x;
    for (int q = lane; q < n; q += 64) {
        const long r = (long)q * nx2 + I;
        sb[q] = B[r];
        sm[q] = M[r];
        su[q] = U[r];
        sd[q] = Dp[r];
    }
    __syncthreads();
    if (lane == 0) {
        double prev = sb[0];
        for (int q = 1; q < n; q++) {
            prev = sb[q] - sm[q] * prev;
            sb[q] = prev;
        }
        double x = prev / sd[n - 1];
        sb[n - 1] = x;
        for (int q = n - 2; q >= 0; q--) {
            x = (sb[q] - su[q] * x) / sd[q];
            sb[q] = x;
        }
    }
    __syncthreads();
    for (int q = lane; q < n; q += 64) B[(long)q * nx2 + I] = sb[q];
}

__global__ __launch_bounds__(64) void k_tri_solve_wave(double *__restrict__ B, const double *__restrict__ M,
                                                       const double *__restrict__ Dp, const double *__restrict__ U, int nx2, int n)
{
    extern __shared__ double tri_sh[];
    tri_solve_wave(B, M, Dp, U, nx2, n, (int)blockIdx.x, tri_sh);
}

// Denominator of the pointwise vy-syz ghost of each coarse (I, J), as the
// forward kernel k_gsyz forms it.
template <typename T>
__global__ __launch_bounds__(256) void k_gsyz_den(IfaceDev<T> a, double *__restrict__ den)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int nx2 = a.nx2, ny2 = a.ny2, r = a.r, nz1 = a.nz1;
    if (t >= nx2 * ny2) return;
    const int I = t % nx2, J = t / nx2;
    const T *b1 = a.F.b, *b2 = a.C.b;
    const long qc = a.C.o + (long)J * a.C.P + I;
    const double c0 = 0.25 * ((((double)b2[qc] + (double)b2[qc + 1]) + (double)b2[qc + a.C.P]) + (double)b2[qc + a.C.P + 1]);
    const double r2 = (double)(r * r);
    double dd = r2 * c0 * a.k2z;
    for (int bo = 0; bo < r; bo++)
        for (int ao = 0; ao < r; ao++) {
            const long q = a.F.o + (long)nz1 * a.F.S2 + (long)(r * J + bo) * a.F.P + r * I + ao;
            const double c = 0.25 * ((((double)b1[q] + (double)b1[q + 1]) + (double)b1[q + a.F.P]) + (double)b1[q + a.F.P + 1]);
            dd += c * a.k1z;
        }
    den[t] = dd;
}

// Receiver misfit injection (transpose of the receiver gather): for each
// distinct receiver location, lam_vx += sum of the residuals recorded there.
template <typename T>
struct RecAdj {
    T *lamvx[MAXG];
    const long *loc;      // per distinct location: grid << 40 | flat index into vx
    const int *start;     // receivers of each location
    const int *rec;
    const T *resid;       // nrec x nsteps
    int nloc, nsteps;
};

template <typename T>
__device__ __forceinline__ void adj_rec_body(const RecAdj<T> &a, int step, const int t)
{
    if (t >= a.nloc) return;
    T acc = T(0);
    for (int q = a.start[t]; q < a.start[t + 1]; q++) acc += a.resid[(long)a.rec[q] * a.nsteps + step];
    const long k = a.loc[t];
    T *v = a.lamvx[k >> 40];
    v[k & ((1L << 40) - 1)] += acc;
}
template <typename T>
__global__ __launch_bounds__(256) void k_adj_rec(RecAdj<T> a, int step)
{
    adj_rec_body(a, step, (int)(blockIdx.x * blockDim.x + threadIdx.x));
}


__global__ void k_set_int(int *p, int v)
{
    if (threadIdx.x == 0) *p = v;
}

// ---------------------------------------------------------------------------
// Reduced-precision checkpoints with one max-abs scale per array.
template <typename T>
__global__ __launch_bounds__(256) void k_maxabs(const T *__restrict__ x, long n, float *__restrict__ part)
{
    __shared__ float red[4];
    float m = 0.f;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        m = fmaxf(m, fabsf((float)x[i]));
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
}

// one array's scale: the maximum of the partial maxima
__global__ __launch_bounds__(64) void k_ck_scale(const float *__restrict__ part, int np, float *__restrict__ scale)
{
    float m = 0.f;
    for (int i = threadIdx.x; i < np; i += 64) m = fmaxf(m, part[i]);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
    if (threadIdx.x == 0) *scale = m;
}

// float -> bfloat16 bits, round to nearest even (finite inputs)
__device__ __forceinline__ unsigned short bf16_bits(float f)
{
    const unsigned int u = __float_as_uint(f);
    const unsigned int lsb = (u >> 16) & 1u;
    return (unsigned short)((u + 0x7fffu + lsb) >> 16);
}

// fmt 1: IEEE half of x / s; fmt 2: bfloat16 of x / s; s = max |x| of the array
template <typename T>
__global__ __launch_bounds__(256) void k_ck_pack(const T *__restrict__ x, long n, unsigned short *__restrict__ out,
                                                 const float *__restrict__ scale, int fmt)
{
    const float s = *scale;
    const float inv = s > 0.f ? 1.0f / s : 0.f;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const float v = (float)x[i] * inv;
        if (fmt == 1) {
            const _Float16 h = (_Float16)v;
            out[i] = *reinterpret_cast<const unsigned short *>(&h);
        } else {
            out[i] = bf16_bits(v);
        }
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_ck_unpack(T *__restrict__ x, long n, const unsigned short *__restrict__ in,
                                                   const float *__restrict__ scale, int fmt)
{
    const float s = *scale;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        float v;
        if (fmt == 1) {
            const unsigned short u = in[i];
            v = (float)*reinterpret_cast<const _Float16 *>(&u);
        } else {
            v = __uint_as_float(((unsigned int)in[i]) << 16);
        }
        x[i] = (T)(v * s);
    }
}

}  // namespace ecsg

namespace ecsg {

// ---------------------------------------------------------------------------
// Structured transposes of the three interface maps.  Each interface system
// restricts per-fine-node quantities kf(i, j) with window weights w_q(i, j)
// onto the coarse rows q:  rhs_q = r^m kc_q - sum_(i,j) w_q(i, j) kf(i, j),
// and its matrix is A = diag(coarse) + k1z R diag(param) P.  With the
// multipliers mu = A^-1 lam_x (A symmetric) prolonged to the fine nodes,
//   nu(i, j) = sum_q w_q(i, j) mu_q,
// the transpose and the parameter derivatives become local:
//   lam[kf inputs] -= nu * d kf / d input,   grad[param] -= nu * d kf / d param,
// and the matrix terms add  -k1z nu(i, j) (P x)(i, j)  to the gradient.
template <typename T>
struct IfAdj {
    T *lf[9], *lc[9];            // adjoint wavefields, fine / coarse grid
    const T *ff[9], *fc[9];      // forward state read by the gradient terms
    T *gf[12], *gc[12];          // gradient accumulators
    const T *cf[9], *cc[9];      // staged stiffness
    const T *bf, *bc;            // buoyancy
    long oF, PF, SF, oC, PC, SC;
    int r, nx2, ny2, nz1;
    double k1x, k1y, k1z, k2x, k2y, k2z;
};

#define AF_(i, j, k) (a.oF + (long)(k) * a.SF + (long)(j) * a.PF + (i))
#define AC_(i, j, k) (a.oC + (long)(k) * a.SC + (long)(j) * a.PC + (i))

template <typename T>
__device__ __forceinline__ void acc_to(T *p, double v) { *p = (T)((double)*p + v); }

template <typename T>
__device__ __forceinline__ double xy_mean(const T *b, long q, long P)
{
    return 0.25 * ((((double)b[q] + (double)b[q + 1]) + (double)b[q + P]) + (double)b[q + P + 1]);
}

// ----- vx-sxz --------------------------------------------------------------
// multiplier of interior coarse node (I, J) from the iterate layout (zero border)
__device__ __forceinline__ double sxz_mu(const double *X, int I, int J, int nx2, int ny2)
{
    if (I < 0 || I > nx2 || J < 0 || J > ny2) return 0.0;
    return X[(long)J * (nx2 + 1) + I];
}
__device__ __forceinline__ double sxz_nu(const double *X, int i, int j, int r, int nx2, int ny2)
{
    double s = 0;
    const int I0 = i / r, J0 = j / r;
    for (int dJ = 0; dJ <= 1; dJ++)
        for (int dI = 0; dI <= 1; dI++) {
            const int Ic = I0 + dI, Jc = J0 + dJ;
            const int ai = i - r * Ic, bj = j - r * Jc;
            if (ai <= -r || ai >= r || bj <= -r || bj >= r) continue;
            s += (double)((r - (ai < 0 ? -ai : ai)) * (r - (bj < 0 ? -bj : bj))) * sxz_mu(X, Ic, Jc, nx2, ny2);
        }
    return s;
}

// fine-ghost transpose gathered on interior coarse nodes; rhs of the solve
template <typename T>
__device__ __forceinline__ void asxz_ghost_body(const IfAdj<T> &a, double *__restrict__ rhs, const int t)
{
    const int n0 = a.nx2 - 1, n1 = a.ny2 - 1, r = a.r;
    if (t >= n0 * n1) return;
    const int I = t % n0 + 1, J = t / n0 + 1;
    double R = 0;
    for (int bo = -(r - 1); bo <= r - 1; bo++)
        for (int ao = -(r - 1); ao <= r - 1; ao++)
            R += hatw(ao, r) * hatw(bo, r) * (double)a.lf[7][AF_(r * I + ao, r * J + bo, a.nz1)];
    T *s2 = a.lc[7];
    const long qc = AC_(I, J, 0);
    acc_to(&s2[qc], R);
    rhs[t] = (double)s2[qc - a.SC] + R;
    s2[qc - a.SC] = T(0);
}
template <typename T>
__global__ __launch_bounds__(256) void k_asxz_ghost(IfAdj<T> a, double *__restrict__ rhs)
{
    asxz_ghost_body(a, rhs, (int)(blockIdx.x * blockDim.x + threadIdx.x));
}


template <typename T>
__device__ __forceinline__ void asxz_fine0_body(const IfAdj<T> &a, const int t)
{
    const int nx1 = a.r * a.nx2, ny1 = a.r * a.ny2;
    if (t >= (nx1 - 1) * (ny1 - 1)) return;
    const int i = t % (nx1 - 1) + 1, j = t / (nx1 - 1) + 1;
    const long q = AF_(i, j, a.nz1);
    T *s1 = a.lf[7];
    s1[q - a.SF] -= s1[q];
    s1[q] = T(0);
}
template <typename T>
__global__ __launch_bounds__(256) void k_asxz_fine0(IfAdj<T> a)
{
    asxz_fine0_body(a, (int)(blockIdx.x * blockDim.x + threadIdx.x));
}


template <typename T>
__device__ __forceinline__ void asxz_fine_body(const IfAdj<T> &a, const double *__restrict__ X, const int t)
{
    const int r = a.r, nx2 = a.nx2, ny2 = a.ny2, nx1 = r * nx2, ny1 = r * ny2;
    if (t >= (nx1 + 1) * (ny1 + 1)) return;
    const int i = t % (nx1 + 1), j = t / (nx1 + 1);
    const double nu0 = sxz_nu(X, i, j, r, nx2, ny2);
    const double nux = (i + 1 <= nx1) ? sxz_nu(X, i + 1, j, r, nx2, ny2) : 0.0;
    const double nuy = (j + 1 <= ny1) ? sxz_nu(X, i, j + 1, r, nx2, ny2) : 0.0;
    const long q = AF_(i, j, a.nz1);
    const double b = (double)a.bf[q], bx = (double)a.bf[q + 1], by = (double)a.bf[q + a.PF];
    if (nu0 != 0.0) {
        acc_to(&a.lf[0][q], -nu0);
        acc_to(&a.lf[7][q - a.SF], 2 * a.k1z * b * nu0);
        // gradient w.r.t. b1 (node): B terms and the matrix terms (P x)
        const int I0 = i / r, J0 = j / r;
        double S = 0;
        for (int dJ = 0; dJ <= 1; dJ++)
            for (int dI = 0; dI <= 1; dI++) {
                const int Ic = I0 + dI, Jc = J0 + dJ;
                if (Ic < 1 || Ic > nx2 - 1 || Jc < 1 || Jc > ny2 - 1) continue;
                const double pw = hatw(i - r * Ic, r) * hatw(j - r * Jc, r);
                if (pw == 0) continue;
                const long q2 = AC_(Ic, Jc, 0);
                S += pw * ((double)a.fc[7][q2] + (double)a.fc[7][q2 - a.SC]);
            }
        const T *sxx = a.ff[3], *sxy = a.ff[8], *sxz = a.ff[7];
        const double k = a.k1x * ((double)sxx[q] - (double)sxx[q - 1]) + a.k1y * ((double)sxy[q] - (double)sxy[q - a.PF])
                       + a.k1z * (S - 2 * (double)sxz[q - a.SF]);
        acc_to(&a.gf[9][q], -nu0 * k);
    }
    if (i <= nx1 - 1 && (nu0 != 0.0 || nux != 0.0)) acc_to(&a.lf[3][q], -a.k1x * (b * nu0 - bx * nux));
    if (j <= ny1 - 1 && (nu0 != 0.0 || nuy != 0.0)) acc_to(&a.lf[8][q], -a.k1y * (b * nu0 - by * nuy));
}
template <typename T>
__global__ __launch_bounds__(256) void k_asxz_fine(IfAdj<T> a, const double *__restrict__ X)
{
    asxz_fine_body(a, X, (int)(blockIdx.x * blockDim.x + threadIdx.x));
}


template <typename T>
__device__ __forceinline__ void asxz_coarse_body(const IfAdj<T> &a, const double *__restrict__ X, const int t)
{
    const int r = a.r, nx2 = a.nx2, ny2 = a.ny2;
    if (t >= (nx2 + 1) * (ny2 + 1)) return;
    const int I = t % (nx2 + 1), J = t / (nx2 + 1);
    const double r2 = (double)(r * r), r4 = r2 * r2;
    const double mu0 = sxz_mu(X, I, J, nx2, ny2), mux = sxz_mu(X, I + 1, J, nx2, ny2), muy = sxz_mu(X, I, J + 1, nx2, ny2);
    const long qc = AC_(I, J, 0);
    const double b2 = (double)a.bc[qc], b2x = (double)a.bc[qc + 1], b2y = (double)a.bc[qc + a.PC];
    if (mu0 != 0.0) acc_to(&a.lc[0][qc], r4 * mu0);
    if (I <= nx2 - 1 && (mu0 != 0.0 || mux != 0.0)) acc_to(&a.lc[3][qc], r4 * a.k2x * (b2 * mu0 - b2x * mux));
    if (J <= ny2 - 1 && (mu0 != 0.0 || muy != 0.0)) acc_to(&a.lc[8][qc], r4 * a.k2y * (b2 * mu0 - b2y * muy));
    if (I >= 1 && I <= nx2 - 1 && J >= 1 && J <= ny2 - 1) {
        // sxz2(I, J, 0): coarse term and the prolongation inside every fine kf
        double sfine = 0;
        for (int bo = -(r - 1); bo <= r - 1; bo++)
            for (int ao = -(r - 1); ao <= r - 1; ao++) {
                const int i = r * I + ao, j = r * J + bo;
                const double nu = sxz_nu(X, i, j, r, nx2, ny2);
                if (nu != 0.0) sfine += hatw(ao, r) * hatw(bo, r) * (double)a.bf[AF_(i, j, a.nz1)] * nu;
            }
        acc_to(&a.lc[7][qc], r4 * a.k2z * b2 * mu0 - a.k1z * sfine);
        if (mu0 != 0.0) {
            const T *sxx = a.fc[3], *sxy = a.fc[8], *sxz = a.fc[7];
            const double k = a.k2x * ((double)sxx[qc] - (double)sxx[qc - 1]) + a.k2y * ((double)sxy[qc] - (double)sxy[qc - a.PC])
                           + a.k2z * ((double)sxz[qc] - (double)sxz[qc - a.SC]);
            acc_to(&a.gc[9][qc], mu0 * r4 * k);
        }
    }
}
template <typename T>
__global__ __launch_bounds__(256) void k_asxz_coarse(IfAdj<T> a, const double *__restrict__ X)
{
    asxz_coarse_body(a, X, (int)(blockIdx.x * blockDim.x + threadIdx.x));
}


// ----- vz-szz --------------------------------------------------------------
__device__ __forceinline__ double vz_mu(const double *M, int I, int J, int nx2, int ny2)
{
    if (I < 0 || I > nx2 - 1 || J < 1 || J > ny2 - 1) return 0.0;
    return M[(long)(J - 1) * nx2 + I];
}
__device__ __forceinline__ double vz_nu(const double *M, int i, int j, int r, int nx2, int ny2)
{
    if (i < 0 || i > r * nx2 - 1) return 0.0;
    const int I = i / r, J0 = j / r;
    double s = 0;
    for (int dJ = 0; dJ <= 1; dJ++) {
        const int J = J0 + dJ, bo = j - r * J;
        if (bo <= -r || bo >= r) continue;
        s += (double)(r - (bo < 0 ? -bo : bo)) * vz_mu(M, I, J, nx2, ny2);
    }
    return s;
}

template <typename T>
__device__ __forceinline__ void avz_ghost_body(const IfAdj<T> &a, double *__restrict__ rhs, const int t)
{
    const int r = a.r, nx2 = a.nx2, ny2 = a.ny2, ny1 = r * ny2;
    if (t >= nx2 * (ny2 + 1)) return;
    const int I = t % nx2, J = t / nx2;
    double R = 0;
    for (int bo = -(r - 1); bo <= r - 1; bo++) {
        const int j = r * J + bo;
        if (j < 1 || j > ny1 - 1) continue;
        const double h = hatw(bo, r);
        for (int ao = 0; ao < r; ao++) R += h * (double)a.lf[2][AF_(r * I + ao, j, a.nz1)];
    }
    T *v2 = a.lc[2];
    const long qc = AC_(I, J, 0);
    acc_to(&v2[qc], R);
    if (J >= 1 && J <= ny2 - 1) {
        rhs[(long)(J - 1) * nx2 + I] = (double)v2[qc - a.SC] + R;
        v2[qc - a.SC] = T(0);
    } else {
        acc_to(&v2[qc - a.SC], R);
    }
}
template <typename T>
__global__ __launch_bounds__(256) void k_avz_ghost(IfAdj<T> a, double *__restrict__ rhs)
{
    avz_ghost_body(a, rhs, (int)(blockIdx.x * blockDim.x + threadIdx.x));
}


template <typename T>
__device__ __forceinline__ void avz_fine0_body(const IfAdj<T> &a, const int t)
{
    const int nx1 = a.r * a.nx2, ny1 = a.r * a.ny2;
    if (t >= nx1 * (ny1 - 1)) return;
    const int i = t % nx1, j = t / nx1 + 1;
    const long q = AF_(i, j, a.nz1);
    T *v1 = a.lf[2];
    v1[q - a.SF] -= v1[q];
    v1[q] = T(0);
}
template <typename T>
__global__ __launch_bounds__(256) void k_avz_fine0(IfAdj<T> a)
{
    avz_fine0_body(a, (int)(blockIdx.x * blockDim.x + threadIdx.x));
}


template <typename T>
__device__ __forceinline__ void avz_fine_body(const IfAdj<T> &a, const double *__restrict__ M, const int t)
{
    const int r = a.r, nx2 = a.nx2, ny2 = a.ny2, nx1 = r * nx2, ny1 = r * ny2;
    if (t >= (nx1 + 1) * (ny1 + 1)) return;
    const int i = t % (nx1 + 1), j = t / (nx1 + 1);
    const double nu0 = vz_nu(M, i, j, r, nx2, ny2);
    const double nuxm = vz_nu(M, i - 1, j, r, nx2, ny2);
    const double nuyp = (j + 1 <= ny1) ? vz_nu(M, i, j + 1, r, nx2, ny2) : 0.0;
    const long q = AF_(i, j, a.nz1);
    const T *c13 = a.cf[4], *c23 = a.cf[3], *c33 = a.cf[2];
    if (nu0 != 0.0 || nuxm != 0.0) {
        const double cur = nu0 != 0.0 ? (double)c13[q] * nu0 : 0.0;
        const double prv = nuxm != 0.0 ? (double)c13[q - 1] * nuxm : 0.0;
        acc_to(&a.lf[0][q], a.k1x * (cur - prv));
    }
    if (i <= nx1 - 1 && j <= ny1 - 1 && (nu0 != 0.0 || nuyp != 0.0)) {
        const double cur = nu0 != 0.0 ? (double)c23[q] * nu0 : 0.0;
        const double nxt = nuyp != 0.0 ? (double)c23[q + a.PF] * nuyp : 0.0;
        acc_to(&a.lf[1][q], -a.k1y * (cur - nxt));
    }
    if (nu0 != 0.0) {
        acc_to(&a.lf[5][q], -nu0);
        acc_to(&a.lf[2][q - a.SF], 2 * a.k1z * (double)c33[q] * nu0);
        const T *vx = a.ff[0], *vy = a.ff[1], *vz1 = a.ff[2], *vz2 = a.fc[2];
        const int I = i / r, J0 = j / r;
        double pv = 0, px = 0;
        for (int dJ = 0; dJ <= 1; dJ++) {
            const int J = J0 + dJ, bo = j - r * J;
            if (bo <= -r || bo >= r || J > ny2) continue;
            const double h = hatw(bo, r);
            const long q2 = AC_(I, J, 0);
            pv += h * (double)vz2[q2];
            if (J >= 1 && J <= ny2 - 1) px += h * (double)vz2[q2 - a.SC];
        }
        acc_to(&a.gf[4][q], -nu0 * a.k1x * ((double)vx[q + 1] - (double)vx[q]));
        acc_to(&a.gf[3][q], -nu0 * a.k1y * ((double)vy[q] - (double)vy[q - a.PF]));
        acc_to(&a.gf[2][q], -nu0 * a.k1z * (pv + px - 2 * (double)vz1[q - a.SF]));
    }
}
template <typename T>
__global__ __launch_bounds__(256) void k_avz_fine(IfAdj<T> a, const double *__restrict__ M)
{
    avz_fine_body(a, M, (int)(blockIdx.x * blockDim.x + threadIdx.x));
}


template <typename T>
__device__ __forceinline__ void avz_coarse_body(const IfAdj<T> &a, const double *__restrict__ M, const int t)
{
    const int r = a.r, nx2 = a.nx2, ny2 = a.ny2, ny1 = r * ny2;
    if (t >= (nx2 + 1) * (ny2 + 1)) return;
    const int I = t % (nx2 + 1), J = t / (nx2 + 1);
    const double r3 = (double)(r * r * r);
    const double mu0 = vz_mu(M, I, J, nx2, ny2), muxm = vz_mu(M, I - 1, J, nx2, ny2), muyp = vz_mu(M, I, J + 1, nx2, ny2);
    const long qc = AC_(I, J, 0);
    const T *c13 = a.cc[4], *c23 = a.cc[3], *c33 = a.cc[2];
    if (mu0 != 0.0) acc_to(&a.lc[5][qc], r3 * mu0);
    if (mu0 != 0.0 || muxm != 0.0) {
        const double cur = mu0 != 0.0 ? (double)c13[qc] * mu0 : 0.0;
        const double prv = muxm != 0.0 ? (double)c13[qc - 1] * muxm : 0.0;
        acc_to(&a.lc[0][qc], r3 * a.k2x * (prv - cur));
    }
    if (I <= nx2 - 1 && J <= ny2 - 1 && (mu0 != 0.0 || muyp != 0.0)) {
        const double cur = mu0 != 0.0 ? (double)c23[qc] * mu0 : 0.0;
        const double nxt = muyp != 0.0 ? (double)c23[qc + a.PC] * muyp : 0.0;
        acc_to(&a.lc[1][qc], r3 * a.k2y * (cur - nxt));
    }
    if (I <= nx2 - 1) {
        // vz2(I, J, 0): coarse term and the y interpolation inside every fine kf
        double sfine = 0;
        for (int bo = -(r - 1); bo <= r - 1; bo++) {
            const int j = r * J + bo;
            if (j < 1 || j > ny1 - 1) continue;
            const double h = hatw(bo, r);
            for (int ao = 0; ao < r; ao++) {
                const int i = r * I + ao;
                const double nu = vz_nu(M, i, j, r, nx2, ny2);
                if (nu != 0.0) sfine += h * (double)a.cf[2][AF_(i, j, a.nz1)] * nu;
            }
        }
        const double own = mu0 != 0.0 ? r3 * a.k2z * (double)c33[qc] * mu0 : 0.0;
        if (own != 0.0 || sfine != 0.0) acc_to(&a.lc[2][qc], own - a.k1z * sfine);
        if (mu0 != 0.0) {
            const T *vx = a.fc[0], *vy = a.fc[1], *vz = a.fc[2];
            acc_to(&a.gc[4][qc], mu0 * r3 * a.k2x * ((double)vx[qc + 1] - (double)vx[qc]));
            acc_to(&a.gc[3][qc], mu0 * r3 * a.k2y * ((double)vy[qc] - (double)vy[qc - a.PC]));
            acc_to(&a.gc[2][qc], mu0 * r3 * a.k2z * ((double)vz[qc] - (double)vz[qc - a.SC]));
        }
    }
}
template <typename T>
__global__ __launch_bounds__(256) void k_avz_coarse(IfAdj<T> a, const double *__restrict__ M)
{
    avz_coarse_body(a, M, (int)(blockIdx.x * blockDim.x + threadIdx.x));
}


// ----- vy-syz --------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void asyz_ghost_body(const IfAdj<T> &a, const double *__restrict__ den, double *__restrict__ mu, const int t)
{
    const int r = a.r, nx2 = a.nx2, ny2 = a.ny2;
    if (t >= nx2 * ny2) return;
    const int I = t % nx2, J = t / nx2;
    double R = 0;
    for (int bo = 0; bo < r; bo++)
        for (int ao = 0; ao < r; ao++) R += (double)a.lf[6][AF_(r * I + ao, r * J + bo, a.nz1)];
    T *s2 = a.lc[6];
    const long qc = AC_(I, J, 0);
    acc_to(&s2[qc], R);
    mu[t] = ((double)s2[qc - a.SC] + R) / den[t];
    s2[qc - a.SC] = T(0);
}
template <typename T>
__global__ __launch_bounds__(256) void k_asyz_ghost(IfAdj<T> a, const double *__restrict__ den, double *__restrict__ mu)
{
    asyz_ghost_body(a, den, mu, (int)(blockIdx.x * blockDim.x + threadIdx.x));
}


template <typename T>
__device__ __forceinline__ void asyz_fine0_body(const IfAdj<T> &a, const int t)
{
    const int nx1 = a.r * a.nx2, ny1 = a.r * a.ny2;
    if (t >= nx1 * ny1) return;
    const int i = t % nx1, j = t / nx1;
    const long q = AF_(i, j, a.nz1);
    T *s1 = a.lf[6];
    s1[q - a.SF] -= s1[q];
    s1[q] = T(0);
}
template <typename T>
__global__ __launch_bounds__(256) void k_asyz_fine0(IfAdj<T> a)
{
    asyz_fine0_body(a, (int)(blockIdx.x * blockDim.x + threadIdx.x));
}


__device__ __forceinline__ double syz_mu(const double *mu, int I, int J, int nx2, int ny2)
{
    if (I < 0 || I > nx2 - 1 || J < 0 || J > ny2 - 1) return 0.0;
    return mu[(long)J * nx2 + I];
}

template <typename T>
__device__ __forceinline__ void asyz_fine_body(const IfAdj<T> &a, const double *__restrict__ mu, const int t)
{
    const int r = a.r, nx2 = a.nx2, ny2 = a.ny2, nx1 = r * nx2, ny1 = r * ny2;
    if (t >= (nx1 + 1) * (ny1 + 1)) return;
    const int i = t % (nx1 + 1), j = t / (nx1 + 1);
    auto nu = [&](int ii, int jj) -> double {
        if (ii < 0 || ii > nx1 - 1 || jj < 0 || jj > ny1 - 1) return 0.0;
        return syz_mu(mu, ii / r, jj / r, nx2, ny2);
    };
    const double nu0 = nu(i, j), nuxm = nu(i - 1, j), nuym = nu(i, j - 1);
    const long q = AF_(i, j, a.nz1);
    const T *b = a.bf;
    const long P = a.PF;
    if (j <= ny1 - 1 && (nu0 != 0.0 || nuxm != 0.0)) {
        const double cur = nu0 != 0.0 ? xy_mean(b, q, P) * nu0 : 0.0;
        const double prv = nuxm != 0.0 ? xy_mean(b, q - 1, P) * nuxm : 0.0;
        acc_to(&a.lf[8][q], -a.k1x * (prv - cur));
    }
    if (i <= nx1 - 1 && (nu0 != 0.0 || nuym != 0.0)) {
        const double cur = nu0 != 0.0 ? xy_mean(b, q, P) * nu0 : 0.0;
        const double prv = nuym != 0.0 ? xy_mean(b, q - P, P) * nuym : 0.0;
        acc_to(&a.lf[4][q], -a.k1y * (prv - cur));
    }
    if (nu0 != 0.0) {
        const double c = xy_mean(b, q, P);
        acc_to(&a.lf[1][q], -nu0);
        acc_to(&a.lf[6][q - a.SF], 2 * a.k1z * c * nu0);
        const T *sxy = a.ff[8], *syy = a.ff[4], *syz1 = a.ff[6], *syz2 = a.fc[6];
        const long qc = AC_(i / r, j / r, 0);
        const double k = a.k1x * ((double)sxy[q + 1] - (double)sxy[q]) + a.k1y * ((double)syy[q + P] - (double)syy[q])
                       + a.k1z * ((double)syz2[qc] + (double)syz2[qc - a.SC] - 2 * (double)syz1[q - a.SF]);
        acc_to(&a.gf[10][q], -nu0 * k);
    }
}
template <typename T>
__global__ __launch_bounds__(256) void k_asyz_fine(IfAdj<T> a, const double *__restrict__ mu)
{
    asyz_fine_body(a, mu, (int)(blockIdx.x * blockDim.x + threadIdx.x));
}


template <typename T>
__device__ __forceinline__ void asyz_coarse_body(const IfAdj<T> &a, const double *__restrict__ mu, const int t)
{
    const int r = a.r, nx2 = a.nx2, ny2 = a.ny2;
    if (t >= (nx2 + 1) * (ny2 + 1)) return;
    const int I = t % (nx2 + 1), J = t / (nx2 + 1);
    const double r2 = (double)(r * r);
    const double mu0 = syz_mu(mu, I, J, nx2, ny2), muxm = syz_mu(mu, I - 1, J, nx2, ny2), muym = syz_mu(mu, I, J - 1, nx2, ny2);
    const long qc = AC_(I, J, 0);
    const T *b = a.bc;
    const long P = a.PC;
    if (J <= ny2 - 1 && (mu0 != 0.0 || muxm != 0.0)) {
        const double cur = mu0 != 0.0 ? xy_mean(b, qc, P) * mu0 : 0.0;
        const double prv = muxm != 0.0 ? xy_mean(b, qc - 1, P) * muxm : 0.0;
        acc_to(&a.lc[8][qc], r2 * a.k2x * (prv - cur));
    }
    if (I <= nx2 - 1 && (mu0 != 0.0 || muym != 0.0)) {
        const double cur = mu0 != 0.0 ? xy_mean(b, qc, P) * mu0 : 0.0;
        const double prv = muym != 0.0 ? xy_mean(b, qc - P, P) * muym : 0.0;
        acc_to(&a.lc[4][qc], r2 * a.k2y * (prv - cur));
    }
    if (mu0 != 0.0) {
        const double c0 = xy_mean(b, qc, P);
        double cs = 0;
        for (int bo = 0; bo < r; bo++)
            for (int ao = 0; ao < r; ao++) cs += xy_mean(a.bf, AF_(r * I + ao, r * J + bo, a.nz1), a.PF);
        acc_to(&a.lc[1][qc], r2 * mu0);
        acc_to(&a.lc[6][qc], (r2 * c0 * a.k2z - cs * a.k1z) * mu0);
        const T *sxy = a.fc[8], *syy = a.fc[4], *syz = a.fc[6];
        const double k = a.k2x * ((double)sxy[qc + 1] - (double)sxy[qc]) + a.k2y * ((double)syy[qc + P] - (double)syy[qc])
                       + a.k2z * ((double)syz[qc] - (double)syz[qc - a.SC]);
        acc_to(&a.gc[10][qc], mu0 * r2 * k);
    }
}
template <typename T>
__global__ __launch_bounds__(256) void k_asyz_coarse(IfAdj<T> a, const double *__restrict__ mu)
{
    asyz_coarse_body(a, mu, (int)(blockIdx.x * blockDim.x + threadIdx.x));
}


#undef AF_
#undef AC_

}  // namespace ecsg
