// ecsg_fused.hpp -- UpdateVel and UpdateStress of one grid in one pass.
// Included by ecsg_capi.hip only.
//
// The state is double-buffered: the pass reads the fields before the step
// (A) and writes the fields after it (B), so a workgroup may evaluate
// v^{n+1/2} on a halo ring around its tile without racing the neighbouring
// workgroups.  Each workgroup is 64 x (FBY + 2) threads: one x row of 64
// lanes per wavefront; lanes 1..62 of rows 1..FBY own outputs, the outer
// lanes and rows only evaluate the velocities their neighbours need.  A
// thread marches z; per plane it evaluates v at the next plane (lookahead)
// into one of two LDS planes, then updates the stresses of the current
// plane from the LDS neighbours and its own registers.  Algorithmic traffic
// per node: read v, s, b and the 9 stiffnesses, write v and s (28 accesses
// instead of 13 + 24 for the two separate passes).
//
// Every arithmetic expression is the one of k_vel / k_stress (same operand
// order), so the result equals the two-pass step bit for bit; the ghost
// velocity mirrors are evaluated from their source entries.  The normal
// stresses of a plane that reads interface ghost velocities (iz = nz above
// an interface, iz = 0 below one) are left to k_stress_defer, which runs
// after the vz-szz interface solve.  The source is folded in: a vx force
// enters the velocities the stresses see (the stored vx excludes it; the
// receiver/source kernel records and adds it afterwards, in the two-pass
// order), an explosive source enters the stresses of its point.
#pragma once
#include "ecsg_kernels.hpp"

namespace ecsg {

constexpr int FXO = 62;  // output lanes per wavefront row

template <typename T>
struct FusedDev {
    GridDev<T> g;          // f[]: the state before the step; parameters, geometry
    T *o[9];               // the state after the step
    int defer0, defern;    // normal stresses of plane 0 / plane nz wait for the interface ghosts
    int gx, gy, gz, zc;    // fused tiling
    long src_p;            // source point (-1: none in this grid)
    int src_type, src_len;
    const float *src_tr;
    const int *step;
};

#define FI(ix, iy, iz) (g.o + (long)(iz) * g.S2 + (long)(iy) * g.P + (ix))

// v^{n+1/2} on the field's own index range (update_vel.cpp:30-69)
template <typename T>
__device__ __forceinline__ T fv_vx(const GridDev<T> &g, int ix, int iy, int iz)
{
    const long p = FI(ix, iy, iz);
    const T vo = g.f[0][p];
    if (ix >= 1 && ix <= g.nx - 1 && iy >= 1 && iy <= g.ny - 1 && iz >= g.top_eff && iz <= g.nz - g.bottom) {
        const T *sxx = g.f[3], *sxz = g.f[7], *sxy = g.f[8];
        return vo + g.b[p] * (g.cx * (sxx[p] - sxx[p - 1]) + g.cy * (sxy[p] - sxy[p - g.P]) + g.cz * (sxz[p] - sxz[p - g.S2]));
    }
    return vo;
}
template <typename T>
__device__ __forceinline__ T fv_vy(const GridDev<T> &g, int ix, int iy, int iz)
{
    const long p = FI(ix, iy, iz);
    const T vo = g.f[1][p];
    if (iz >= g.top_eff && iz <= g.nz - g.bottom) {
        const T *syy = g.f[4], *syz = g.f[6], *sxy = g.f[8], *b = g.b;
        const T bav = T(0.25) * (b[p] + b[p + 1] + b[p + g.P] + b[p + g.P + 1]);
        return vo + bav * (g.cx * (sxy[p + 1] - sxy[p]) + g.cz * (syz[p] - syz[p - g.S2]) + g.cy * (syy[p + g.P] - syy[p]));
    }
    return vo;
}
template <typename T>
__device__ __forceinline__ T fv_vz(const GridDev<T> &g, int ix, int iy, int iz)
{
    const long p = FI(ix, iy, iz);
    const T vo = g.f[2][p];
    if (iy >= 1 && iy <= g.ny - 1) {
        const T *szz = g.f[5], *syz = g.f[6], *sxz = g.f[7], *b = g.b;
        const T bav = T(0.25) * (b[p] + b[p + g.S2] + b[p + 1] + b[p + g.S2 + 1]);
        return vo + bav * (g.cx * (sxz[p + 1] - sxz[p]) + g.cz * (szz[p + g.S2] - szz[p]) + g.cy * (syz[p] - syz[p - g.P]));
    }
    return vo;
}

// the values the stresses read at any entry a workgroup touches: the update
// on the field's range, the ghost mirrors (update_vel.cpp:75-246, radius 1)
// from their sources; entries no stress reads return 0
template <typename T>
__device__ __forceinline__ T fm_vx(const GridDev<T> &g, int ix, int iy, int iz)
{
    if (ix < 0 || ix > g.nx || iy < 0 || iy > g.ny || iz < 0 || iz > g.nz) return T(0);
    return fv_vx(g, ix, iy, iz);
}
template <typename T>
__device__ __forceinline__ T fm_vy(const GridDev<T> &g, int ix, int iy, int iz)
{
    const int nx = g.nx, ny = g.ny;
    if (iz < 0 || iz > g.nz) return T(0);
    const bool xin = ix >= 0 && ix <= nx - 1, yin = iy >= 0 && iy <= ny - 1;
    if (xin && yin) return fv_vy(g, ix, iy, iz);
    if (yin) {
        if (ix == -1) return -fv_vy(g, 0, iy, iz);
        if (ix == nx) return -fv_vy(g, nx - 1, iy, iz);
        return T(0);
    }
    if (xin) {
        if (iy == -1) return iz <= g.zy ? -fv_vy(g, ix, 0, iz) : g.f[1][FI(ix, iy, iz)];
        if (iy == ny) return iz <= g.zy ? -fv_vy(g, ix, ny - 1, iz) : g.f[1][FI(ix, iy, iz)];
    }
    return T(0);
}
template <typename T>
__device__ __forceinline__ T fm_vz(const GridDev<T> &g, int ix, int iy, int iz)
{
    const int nx = g.nx, nz = g.nz;
    if (iy < 0 || iy > g.ny) return T(0);
    const bool xin = ix >= 0 && ix <= nx - 1, zin = iz >= 0 && iz <= nz - 1;
    if (xin && zin) return fv_vz(g, ix, iy, iz);
    if (zin) {
        if (ix == -1) return -fv_vz(g, 0, iy, iz);
        if (ix == nx) return -fv_vz(g, nx - 1, iy, iz);
        return T(0);
    }
    if (xin) {
        if (iz == -1) return g.top_eff ? -fv_vz(g, ix, iy, 0) : g.f[2][FI(ix, iy, iz)];
        if (iz == nz) return g.bottom ? -fv_vz(g, ix, iy, nz - 1) : g.f[2][FI(ix, iy, iz)];
    }
    return T(0);
}

// Velocities of one column carried up the z planes: the stresses and
// buoyancies of the neighbouring planes stay in registers (as in k_vel).
// Valid for columns with ix in [0, nx-1] and iy in [0, ny-1] (all three
// velocity fields on their own index range); other columns use fm_*.
template <typename T>
struct ColVel {
    T sxz_m, syz_m, szz_c, b0, b1, vz_prev;
    __device__ __forceinline__ void init(const GridDev<T> &g, long p)
    {
        sxz_m = g.f[7][p - g.S2];
        syz_m = g.f[6][p - g.S2];
        szz_c = g.f[5][p];
        b0 = g.b[p];
        b1 = g.b[p + 1];
        vz_prev = T(0);
    }
    // v^{n+1/2} at plane iz of column (ix, iy), p its index; iz in [0, nz]
    __device__ __forceinline__ void eval(const GridDev<T> &g, int ix, int iy, int iz, long p, T &vx, T &vy, T &vz)
    {
        const int nx = g.nx, ny = g.ny, nz = g.nz;
        const long P = g.P, S2 = g.S2;
        const T cx = g.cx, cy = g.cy, cz = g.cz;
        const T *sxx = g.f[3], *syy = g.f[4], *szz = g.f[5], *syz = g.f[6], *sxz = g.f[7], *sxy = g.f[8], *b = g.b;
        const T sxz_c = sxz[p];
        const T syz_c = syz[p];
        const T szz_n = szz[p + S2];
        const T b0n = b[p + S2];
        const T b1n = b[p + S2 + 1];
        const bool zin = iz >= g.top_eff && iz <= nz - g.bottom;
        vx = g.f[0][p];
        if (ix >= 1 && ix <= nx - 1 && iy >= 1 && iy <= ny - 1 && zin)
            vx = vx + b0 * (cx * (sxx[p] - sxx[p - 1]) + cy * (sxy[p] - sxy[p - P]) + cz * (sxz_c - sxz_m));
        vy = g.f[1][p];
        if (zin) {
            const T bav = T(0.25) * (b0 + b1 + b[p + P] + b[p + P + 1]);
            vy = vy + bav * (cx * (sxy[p + 1] - sxy[p]) + cz * (syz_c - syz_m) + cy * (syy[p + P] - syy[p]));
        }
        if (iz <= nz - 1) {
            vz = g.f[2][p];
            if (iy >= 1 && iy <= ny - 1) {
                const T bav = T(0.25) * (b0 + b0n + b1 + b1n);
                vz = vz + bav * (cx * (sxz[p + 1] - sxz_c) + cz * (szz_n - szz_c) + cy * (syz_c - syz[p - P]));
            }
        } else {
            // plane nz: bottom mirror of plane nz-1, or the interface ghost
            vz = g.bottom ? -vz_prev : g.f[2][p];
        }
        vz_prev = vz;
        sxz_m = sxz_c;
        syz_m = syz_c;
        szz_c = szz_n;
        b0 = b0n;
        b1 = b1n;
    }
};

template <typename T, int FBY>
__global__ __launch_bounds__(64 * (FBY + 2)) void k_fused(FusedDev<T> a)
{
    __shared__ T sv[2][3][FBY + 2][64];
    const GridDev<T> &g = a.g;
    const int tile = xcd_tile(blockIdx.x, a.gx * a.gy * a.gz);
    const int bx = tile % a.gx;
    const int rest = tile / a.gx;
    const int by = rest % a.gy;
    const int bz = rest / a.gy;
    const int lane = threadIdx.x, w = threadIdx.y;
    const int ix = bx * FXO - 1 + lane;
    const int iy = by * FBY - 1 + w;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const bool out = lane >= 1 && lane <= FXO && w >= 1 && w <= FBY && ix <= nx && iy <= ny;
    const bool fast = ix >= 0 && ix <= nx - 1 && iy >= 0 && iy <= ny - 1;
    const int z0 = bz * a.zc;
    const int z1 = min(z0 + a.zc, nz + 1);
    const long P = g.P, S2 = g.S2;
    const T cx = g.cx, cy = g.cy, cz = g.cz;
    const long pcol = g.o + (long)iy * P + ix;   // index of (ix, iy, 0)
    // the source of this step
    T sadd = T(0);
    long sp = -1;
    if (a.src_p >= 0) {
        const int step = *a.step;
        if (step < a.src_len) {
            sp = a.src_p;
            if (a.src_type == 0) sadd = (T)(1000 * a.src_tr[step]);
            else sadd = (T)a.src_tr[step];
        }
    }
    const bool vsrc = sp >= 0 && a.src_type == 0;
    ColVel<T> cv;
    T vx_c, vy_c, vz_c, vz_m = T(0);
    if (fast) {
        cv.init(g, pcol + (long)z0 * S2);
        if (z0 > 0) cv.vz_prev = fm_vz(g, ix, iy, z0 - 1);
        cv.eval(g, ix, iy, z0, pcol + (long)z0 * S2, vx_c, vy_c, vz_c);
        if (out) vz_m = z0 > 0 ? fm_vz(g, ix, iy, z0 - 1) : (g.top_eff ? -vz_c : g.f[2][pcol - S2]);
    } else {
        vx_c = fm_vx(g, ix, iy, z0);
        vy_c = fm_vy(g, ix, iy, z0);
        vz_c = fm_vz(g, ix, iy, z0);
        if (out) vz_m = fm_vz(g, ix, iy, z0 - 1);
    }
    T vx_cs = vx_c;  // vx without the source, as stored
    if (vsrc && pcol + (long)z0 * S2 == sp && ix >= 0 && ix <= nx && iy >= 0 && iy <= ny) vx_c = vx_c + sadd;
    sv[z0 & 1][0][w][lane] = vx_c;
    sv[z0 & 1][1][w][lane] = vy_c;
    sv[z0 & 1][2][w][lane] = vz_c;
    T *__restrict__ ovx = a.o[0];
    T *__restrict__ ovy = a.o[1];
    T *__restrict__ ovz = a.o[2];
    for (int iz = z0; iz < z1; iz++) {
        __syncthreads();
        // lookahead: v at plane iz + 1
        T vx_n = T(0), vy_n = T(0), vz_n = T(0);
        const long pn = pcol + (long)(iz + 1) * S2;
        if (iz + 1 <= nz) {
            if (fast) {
                cv.eval(g, ix, iy, iz + 1, pn, vx_n, vy_n, vz_n);
            } else {
                vx_n = fm_vx(g, ix, iy, iz + 1);
                vy_n = fm_vy(g, ix, iy, iz + 1);
                vz_n = fm_vz(g, ix, iy, iz + 1);
            }
        }
        const T vx_ns = vx_n;
        if (vsrc && pn == sp && ix >= 0 && ix <= nx && iy >= 0 && iy <= ny) vx_n = vx_n + sadd;
        const int nb = (iz + 1) & 1, cb = iz & 1;
        sv[nb][0][w][lane] = vx_n;
        sv[nb][1][w][lane] = vy_n;
        sv[nb][2][w][lane] = vz_n;
        if (out) {
            const long p = pcol + (long)iz * S2;
            // ---- v^{n+1/2} of this entry and its ghost images (k_vel)
            ovx[p] = vx_cs;
            if (ix <= nx - 1 && iy <= ny - 1) ovy[p] = vy_c;
            if (ix <= nx - 1 && iz <= nz - 1) ovz[p] = vz_c;
            if (g.top_eff && iz == 0 && ix <= nx - 1) ovz[p - S2] = -vz_c;
            if (g.bottom && iz == nz - 1 && ix <= nx - 1) ovz[p + S2] = -vz_c;
            if (ix == 0) {
                if (iz <= nz - 1) ovz[p - 1] = -vz_c;
                if (iy <= ny - 1) ovy[p - 1] = -vy_c;
            }
            if (ix == nx - 1) {
                if (iz <= nz - 1) ovz[p + 1] = -vz_c;
                if (iy <= ny - 1) ovy[p + 1] = -vy_c;
            }
            if (ix <= nx - 1) {
                if (iy == 0 && iz <= g.zy) ovy[p - P] = -vy_c;
                if (iy == ny - 1) {
                    if (ix == 0) ovy[p + P - 1] = -g.f[1][p + P];
                    if (ix == nx - 1) ovy[p + P + 1] = -g.f[1][p + P];
                    if (iz <= g.zy) ovy[p + P] = -vy_c;
                }
            }
            // ---- stresses (k_stress), neighbours from LDS
            const T vx_xp = sv[cb][0][w][lane + 1];
            const T vx_yp = sv[cb][0][w + 1][lane];
            const T vy_xm = sv[cb][1][w][lane - 1];
            const T vy_ym = sv[cb][1][w - 1][lane];
            const T vz_xm = sv[cb][2][w][lane - 1];
            const T vz_yp = sv[cb][2][w + 1][lane];
            const T *sxx = g.f[3], *syy = g.f[4], *szz = g.f[5], *syz = g.f[6], *sxz = g.f[7], *sxy = g.f[8];
            const bool here = (p == sp) && !vsrc && sp >= 0;
            if (ix <= nx - 1) {
                const bool defer = (a.defer0 && iz == 0) || (a.defern && iz == nz);
                if (!defer) {
                    T sxx_o = sxx[p], syy_o = syy[p], szz_o = szz[p];
                    if (here) {
                        sxx_o = sxx_o + sadd;
                        syy_o = syy_o + sadd;
                        szz_o = szz_o + sadd;
                    }
                    const T c11 = g.c[0][p], c22 = g.c[1][p], c33 = g.c[2][p], c23 = g.c[3][p], c13 = g.c[4][p], c12 = g.c[5][p];
                    const T dvx = vx_xp - vx_c;
                    const T dvy = vy_c - vy_ym;
                    if (g.free_top && iz == 0) {
                        a.o[3][p] = sxx_o + ((c11 - c13 * c13 / c33) * cx * dvx + (c12 - c13 * c23 / c33) * cy * dvy);
                        a.o[4][p] = syy_o + ((c12 - c23 * c13 / c33) * cx * dvx + (c22 - c23 * c23 / c33) * cy * dvy);
                        a.o[5][p] = szz_o;
                    } else {
                        const T dvz = vz_c - vz_m;
                        a.o[3][p] = sxx_o + (c11 * cx * dvx + c12 * cy * dvy + c13 * cz * dvz);
                        a.o[4][p] = syy_o + (c12 * cx * dvx + c22 * cy * dvy + c23 * cz * dvz);
                        a.o[5][p] = szz_o + (c13 * cx * dvx + c23 * cy * dvy + c33 * cz * dvz);
                    }
                }
            }
            if (ix <= nx - 1 && iy <= ny - 1 && iz <= nz - 1) {
                const T syz_n = syz[p] + T(0.5) * g.c[6][p] * (cz * (vy_n - vy_c) + cy * (vz_yp - vz_c));
                a.o[6][p] = syz_n;
                if (g.free_top && iz == 0) a.o[6][p - S2] = -syz_n;
            }
            if (iz <= nz - 1) {
                const T sxz_n = sxz[p] + T(0.5) * g.c[7][p] * (cz * (vx_n - vx_c) + cx * (vz_c - vz_xm));
                a.o[7][p] = sxz_n;
                if (g.free_top && iz == 0 && ix >= 1 && ix <= nx - 1 && iy >= 1 && iy <= ny - 1) a.o[7][p - S2] = -sxz_n;
            }
            if (iy <= ny - 1) a.o[8][p] = sxy[p] + T(0.5) * g.c[8][p] * (cy * (vx_yp - vx_c) + cx * (vy_c - vy_xm));
        }
        vz_m = vz_c;
        vx_c = vx_n;
        vx_cs = vx_ns;
        vy_c = vy_n;
        vz_c = vz_n;
    }
}

// Normal stresses of one plane whose update reads interface ghost
// velocities (update_stress.cpp:69-97): velocities after the step and the
// ghost solve from o[], stresses before the step from g.f[].
template <typename T>
__global__ __launch_bounds__(256) void k_stress_defer(GridDev<T> g, T *const o0, T *const o1, T *const o2, T *const o3,
                                                      T *const o4, T *const o5, int iz)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int nx = g.nx, ny = g.ny;
    if (t >= nx * (ny + 1)) return;
    const int ix = t % nx, iy = t / nx;
    const long p = FI(ix, iy, iz);
    const T cx = g.cx, cy = g.cy, cz = g.cz;
    const T c11 = g.c[0][p], c22 = g.c[1][p], c33 = g.c[2][p], c23 = g.c[3][p], c13 = g.c[4][p], c12 = g.c[5][p];
    const T vx_c = o0[p];
    const T dvx = o0[p + 1] - vx_c;
    const T dvy = o1[p] - o1[p - g.P];
    const T dvz = o2[p] - o2[p - g.S2];
    const T sxx_o = g.f[3][p], syy_o = g.f[4][p], szz_o = g.f[5][p];
    o3[p] = sxx_o + (c11 * cx * dvx + c12 * cy * dvy + c13 * cz * dvz);
    o4[p] = syy_o + (c12 * cx * dvx + c22 * cy * dvy + c23 * cz * dvz);
    o5[p] = szz_o + (c13 * cx * dvx + c23 * cy * dvy + c33 * cz * dvz);
}

#undef FI

}  // namespace ecsg
