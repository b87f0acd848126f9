// ecsg_fused.hpp -- UpdateVel and UpdateStress of one grid in one pass.
// Included by ecsg_capi.hip only.
//
// The state is double-buffered: the pass reads the fields before the step
// (A) and writes the fields after it (B), so a wavefront may evaluate
// v^{n+1/2} outside its own columns without racing the neighbouring
// wavefronts.  A wavefront owns 64 line-aligned x columns of a few rows and
// marches z, evaluating v at the next plane (lookahead) before updating the
// stresses of the current plane; x neighbours come from adjacent lanes, y
// neighbours from its own rows, and it evaluates the few velocities beyond
// its columns itself.  k_fused_lean covers the tiles whose halos lie inside
// the walls and the top and bottom planes (no branch but the source),
// k_fused_w the rest (mirrors, ghost images, boundary planes).  Algorithmic
// traffic per node: read v, s, b and the 9 stiffnesses, write v and s (28
// accesses instead of 13 + 24 for the two separate passes).
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

template <typename T>
struct FusedDev {
    GridDev<T> g;          // f[]: the state before the step; parameters, geometry
    T *o[9];               // the state after the step
    int defer0, defern;    // normal stresses of plane 0 / plane nz wait for the interface ghosts
    int gx, gy, gz, zc;    // fused tiling
    // k_fused_w: z segments [0, zlo), chunks of zc over [zlo, zhi), [zhi, nz+1)
    // (gz = nzc + 2); wavefronts whose columns [cx0, cx1), rows [ry0, ry1)
    // and chunk lie in the box k_fused_lean covers return at once
    int zlo, zhi, nzc;
    int cx0, cx1, ry0, ry1;
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

// the values the stresses read at any entry a wavefront touches: the update
// on the field's range, the ghost mirrors (update_vel.cpp:75-246, radius 1)
// from their sources; entries no stress reads return 0.  The entry whose
// value (or mirror image) is wanted is chosen first, then evaluated once.
template <typename T>
__device__ __forceinline__ T fr_vx(const GridDev<T> &g, int ix, int iy, int iz)
{
    if (ix < 0 || ix > g.nx || iy < 0 || iy > g.ny || iz < 0 || iz > g.nz) return T(0);
    return fv_vx(g, ix, iy, iz);
}
template <typename T>
__device__ __forceinline__ T fr_vy(const GridDev<T> &g, int ix, int iy, int iz)
{
    const int nx = g.nx, ny = g.ny;
    if (iz < 0 || iz > g.nz) return T(0);
    const bool xin = ix >= 0 && ix <= nx - 1, yin = iy >= 0 && iy <= ny - 1;
    int jx = ix, jy = iy;
    bool neg = false;
    if (!(xin && yin)) {
        if (yin && (ix == -1 || ix == nx)) {
            jx = ix == -1 ? 0 : nx - 1;
            neg = true;
        } else if (xin && (iy == -1 || iy == ny)) {
            if (iz > g.zy) return g.f[1][FI(ix, iy, iz)];
            jy = iy == -1 ? 0 : ny - 1;
            neg = true;
        } else {
            return T(0);
        }
    }
    const T v = fv_vy(g, jx, jy, iz);
    return neg ? -v : v;
}
template <typename T>
__device__ __forceinline__ T fr_vz(const GridDev<T> &g, int ix, int iy, int iz)
{
    const int nx = g.nx, nz = g.nz;
    if (iy < 0 || iy > g.ny) return T(0);
    const bool xin = ix >= 0 && ix <= nx - 1, zin = iz >= 0 && iz <= nz - 1;
    int jx = ix, jz = iz;
    bool neg = false;
    if (!(xin && zin)) {
        if (zin && (ix == -1 || ix == nx)) {
            jx = ix == -1 ? 0 : nx - 1;
            neg = true;
        } else if (xin && (iz == -1 || iz == nz)) {
            if (iz == -1 ? !g.top_eff : !g.bottom) return g.f[2][FI(ix, iy, iz)];
            jz = iz == -1 ? 0 : nz - 1;
            neg = true;
        } else {
            return T(0);
        }
    }
    const T v = fv_vz(g, jx, iy, jz);
    return neg ? -v : v;
}

// Velocities of one column carried up the z planes: the stresses and
// buoyancies of the neighbouring planes stay in registers (as in k_vel).
// Valid for columns with ix in [0, nx-1] and iy in [0, ny-1] (all three
// velocity fields on their own index range); other columns use fr_*.
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

// Wavefront-owned form of the same pass: no LDS and no workgroup barrier.  A
// wavefront owns the 64 line-aligned columns ix = 64*bx + lane of FR
// consecutive rows and marches them up its z chunk, evaluating v^{n+1/2} one
// plane ahead of the stresses.  x neighbours come from the adjacent lanes
// (shuffles), y neighbours from its own rows (registers); the velocities it
// needs outside its columns it evaluates itself at the plane being updated:
// vy of the row above, vx and vz of the row below, and, spread over lanes
// 0 .. 3*FR-1, vx right of lane 63 and vy, vz left of lane 0 for each row.
// The rows of a plane are taken one after the other (lookahead, then the
// stresses), which bounds the registers a wavefront holds.  Every value is
// the expression k_fused evaluates (same operand order), so the results are
// the same bit for bit.  Workgroups of 4 wavefronts stack their rows: a tile
// is 64 x 4*FR columns by a.zc planes.
#ifndef FUSED_W_WAVES
#define FUSED_W_WAVES 4
#endif
template <typename T, int FR>
__global__ __launch_bounds__(256, FUSED_W_WAVES) void k_fused_w(FusedDev<T> a)
{
    const GridDev<T> &g = a.g;
    const int tile = xcd_tile(blockIdx.x, a.gx * a.gy * a.gz);
    const int bx = tile % a.gx;
    const int rest = tile / a.gx;
    const int by = rest % a.gy;
    const int bz = rest / a.gy;
    const int lane = threadIdx.x & 63;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const int x0 = bx * 64, ix = x0 + lane;
    const int y0 = (by * 4 + (int)(threadIdx.x >> 6)) * FR;  // first row of this wavefront
    if (y0 > ny) return;                                    // the whole wavefront
    int z0, z1;
    if (bz == 0) {
        z0 = 0;
        z1 = a.zlo;
    } else if (bz <= a.nzc) {
        if (x0 >= a.cx0 && x0 + 64 <= a.cx1 && y0 >= a.ry0 && y0 + FR <= a.ry1) return;  // k_fused_lean's
        z0 = a.zlo + (bz - 1) * a.zc;
        z1 = min(z0 + a.zc, a.zhi);
    } else {
        z0 = a.zhi;
        z1 = nz + 1;
    }
    if (z0 >= z1) return;
    const long P = g.P, S2 = g.S2;
    const T cx = g.cx, cy = g.cy, cz = g.cz;
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
    // a vx force enters the velocities the stresses read at its point; the
    // stored vx excludes it (src_raw keeps that value for the store)
    T src_raw = T(0);
    ColVel<T> cv[FR];
    T vx[FR], vy[FR], vz[FR], vzm[FR];
#pragma unroll
    for (int r = 0; r < FR; r++) {
        const int iy = y0 + r;
        const long pc = g.o + (long)iy * P + ix;
        const bool fast = ix <= nx - 1 && iy <= ny - 1;
        const bool out = ix <= nx && iy <= ny;
        vzm[r] = T(0);
        if (fast) {
            cv[r].init(g, pc + (long)z0 * S2);
            if (z0 > 0) cv[r].vz_prev = fr_vz(g, ix, iy, z0 - 1);
            cv[r].eval(g, ix, iy, z0, pc + (long)z0 * S2, vx[r], vy[r], vz[r]);
            if (out) vzm[r] = z0 > 0 ? fr_vz(g, ix, iy, z0 - 1) : (g.top_eff ? -vz[r] : g.f[2][pc - S2]);
        } else {
            vx[r] = fr_vx(g, ix, iy, z0);
            vy[r] = fr_vy(g, ix, iy, z0);
            vz[r] = fr_vz(g, ix, iy, z0);
            if (out) vzm[r] = fr_vz(g, ix, iy, z0 - 1);
        }
        if (vsrc && pc + (long)z0 * S2 == sp && ix >= 0 && ix <= nx && iy >= 0 && iy <= ny) {
            src_raw = vx[r];
            vx[r] = vx[r] + sadd;
        }
    }
    T *__restrict__ ovx = a.o[0];
    T *__restrict__ ovy = a.o[1];
    T *__restrict__ ovz = a.o[2];
    const T *sxx = g.f[3], *syy = g.f[4], *szz = g.f[5], *syz = g.f[6], *sxz = g.f[7], *sxy = g.f[8];
    for (int iz = z0; iz < z1; iz++) {
        // the neighbours outside the wavefront's columns, at plane iz
        T vy_above = fr_vy(g, ix, y0 - 1, iz);
        T hvx = fr_vx(g, ix, y0 + FR, iz);
        if (vsrc && FI(ix, y0 + FR, iz) == sp && ix <= nx && y0 + FR <= ny) hvx = hvx + sadd;
        const T hvz = fr_vz(g, ix, y0 + FR, iz);
        T xh = T(0);
        if (lane < FR) {
            xh = fr_vx(g, x0 + 64, y0 + lane, iz);
            if (vsrc && FI(x0 + 64, y0 + lane, iz) == sp && x0 + 64 <= nx && y0 + lane <= ny) xh = xh + sadd;
        } else if (lane < 2 * FR) {
            xh = fr_vy(g, x0 - 1, y0 + lane - FR, iz);
        } else if (lane < 3 * FR) {
            xh = fr_vz(g, x0 - 1, y0 + lane - 2 * FR, iz);
        }
#pragma unroll
        for (int r = 0; r < FR; r++) {
            __builtin_amdgcn_sched_barrier(0);
            const int iy = y0 + r;
            const long p = g.o + (long)iz * S2 + (long)iy * P + ix;
            // lookahead: v at plane iz + 1 of this row
            T nvx = T(0), nvy = T(0), nvz = T(0);
            if (iz + 1 <= nz) {
                if (ix <= nx - 1 && iy <= ny - 1) {
                    cv[r].eval(g, ix, iy, iz + 1, p + S2, nvx, nvy, nvz);
                } else {
                    nvx = fr_vx(g, ix, iy, iz + 1);
                    nvy = fr_vy(g, ix, iy, iz + 1);
                    nvz = fr_vz(g, ix, iy, iz + 1);
                }
            }
            T nraw = src_raw;
            if (vsrc && p + S2 == sp && ix >= 0 && ix <= nx && iy >= 0 && iy <= ny) {
                nraw = nvx;
                nvx = nvx + sadd;
            }
            const T vx_c = vx[r], vy_c = vy[r], vz_c = vz[r], vz_m = vzm[r];
            const T sh_xp = __shfl_down(vx_c, 1, 64);
            const T sh_ym = __shfl_up(vy_c, 1, 64);
            const T sh_zm = __shfl_up(vz_c, 1, 64);
            const T h_xp = __shfl(xh, r, 64);
            const T h_ym = __shfl(xh, FR + r, 64);
            const T h_zm = __shfl(xh, 2 * FR + r, 64);
            const T vy_ym = vy_above;
            vy_above = vy_c;
            if (ix <= nx && iy <= ny) {
                const T vx_xp = lane == 63 ? h_xp : sh_xp;
                const T vy_xm = lane == 0 ? h_ym : sh_ym;
                const T vz_xm = lane == 0 ? h_zm : sh_zm;
                const T vx_yp = r + 1 < FR ? vx[r + 1 < FR ? r + 1 : r] : hvx;
                const T vz_yp = r + 1 < FR ? vz[r + 1 < FR ? r + 1 : r] : hvz;
                // ---- v^{n+1/2} of this entry and its ghost images (k_vel)
                ovx[p] = (vsrc && p == sp) ? src_raw : vx_c;
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
                // ---- stresses (k_stress)
                const bool here = (p == sp) && !vsrc && sp >= 0;
                if (ix <= nx - 1) {
                    const bool defer = (a.defer0 && iz == 0) || (a.defern && iz == nz);
                    if (!defer) {
                        T sxx_o = ldc(sxx, p), syy_o = ldc(syy, p), szz_o = ldc(szz, p);
                        if (here) {
                            sxx_o = sxx_o + sadd;
                            syy_o = syy_o + sadd;
                            szz_o = szz_o + sadd;
                        }
                        const T c11 = ldc(g.c[0], p), c22 = ldc(g.c[1], p), c33 = ldc(g.c[2], p), c23 = ldc(g.c[3], p),
                                c13 = ldc(g.c[4], p), c12 = ldc(g.c[5], p);
                        const T dvx = vx_xp - vx_c;
                        const T dvy = vy_c - vy_ym;
                        if (g.free_top && iz == 0) {
                            stc(a.o[3], p, sxx_o + ((c11 - c13 * c13 / c33) * cx * dvx + (c12 - c13 * c23 / c33) * cy * dvy));
                            stc(a.o[4], p, syy_o + ((c12 - c23 * c13 / c33) * cx * dvx + (c22 - c23 * c23 / c33) * cy * dvy));
                            stc(a.o[5], p, szz_o);
                        } else {
                            const T dvz = vz_c - vz_m;
                            stc(a.o[3], p, sxx_o + (c11 * cx * dvx + c12 * cy * dvy + c13 * cz * dvz));
                            stc(a.o[4], p, syy_o + (c12 * cx * dvx + c22 * cy * dvy + c23 * cz * dvz));
                            stc(a.o[5], p, szz_o + (c13 * cx * dvx + c23 * cy * dvy + c33 * cz * dvz));
                        }
                    }
                }
                if (ix <= nx - 1 && iy <= ny - 1 && iz <= nz - 1) {
                    const T syz_n = ldc(syz, p) + T(0.5) * ldc(g.c[6], p) * (cz * (nvy - vy_c) + cy * (vz_yp - vz_c));
                    stc(a.o[6], p, syz_n);
                    if (g.free_top && iz == 0) a.o[6][p - S2] = -syz_n;
                }
                if (iz <= nz - 1) {
                    const T sxz_n = ldc(sxz, p) + T(0.5) * ldc(g.c[7], p) * (cz * (nvx - vx_c) + cx * (vz_c - vz_xm));
                    stc(a.o[7], p, sxz_n);
                    if (g.free_top && iz == 0 && ix >= 1 && ix <= nx - 1 && iy >= 1 && iy <= ny - 1) a.o[7][p - S2] = -sxz_n;
                }
                if (iy <= ny - 1)
                    stc(a.o[8], p, ldc(sxy, p) + T(0.5) * ldc(g.c[8], p) * (cy * (vx_yp - vx_c) + cx * (vy_c - vy_xm)));
            }
            src_raw = nraw;
            vzm[r] = vz_c;
            vx[r] = nvx;
            vy[r] = nvy;
            vz[r] = nvz;
        }
    }
}

// loads / stores at a 32-bit element offset from a uniform base: the byte
// offset is formed in 32 bits, so the address is base (scalar) + offset
template <typename T>
__device__ __forceinline__ T l32(const T *a, unsigned i)
{
    return *(const T *)((const char *)a + i * (unsigned)sizeof(T));
}
template <typename T>
__device__ __forceinline__ T l32n(const T *a, unsigned i)
{
    return __builtin_nontemporal_load((const T *)((const char *)a + i * (unsigned)sizeof(T)));
}
template <typename T>
__device__ __forceinline__ void s32n(T *a, unsigned i, T v)
{
    __builtin_nontemporal_store(v, (T *)((char *)a + i * (unsigned)sizeof(T)));
}

// The interior tiles of the fused pass (same results as k_fused / k_fused_w): every column of the tile, its halo
// rows and columns are away from the grid's walls, and its planes (with the
// plane below and the lookahead above) away from the top and bottom, so no
// mirror, ghost image, free-surface row or deferred stress occurs and every
// update is taken; the kernel is branch-free but for the source.  Offsets
// are 32-bit (the host checks that every index fits).
template <typename T>
struct FusedLean {
    const T *A;      // state before the step: array k at A + k*as
    T *B;            // state after the step
    const T *Pm;     // parameters: c11..c66 at Pm + k*as (k = 0..8), buoyancy at Pm + 9*as
    unsigned as, P, S2, o;
    T cx, cy, cz;
    int bx0, by0, gxl, gyl;  // interior tile box (x, y)
    int zlo, zhi, zc, nzc;   // planes [zlo, zhi) in chunks of zc
    unsigned src_p;          // source point (0xffffffff: none in this grid)
    int src_type, src_len;
    const float *src_tr;
    const int *step;
};

#ifndef FUSED_L_WAVES
#define FUSED_L_WAVES 3
#endif
template <typename T, int FR>
__global__ __launch_bounds__(256, FUSED_L_WAVES) void k_fused_lean(FusedLean<T> a)
{
    const int tile = xcd_tile(blockIdx.x, a.gxl * a.gyl * a.nzc);
    const int tx = tile % a.gxl;
    const int rest = tile / a.gxl;
    const int ty = rest % a.gyl;
    const int tz = rest / a.gyl;
    const int lane = threadIdx.x & 63;
    const int x0 = (a.bx0 + tx) * 64, ix = x0 + lane;
    const int y0 = ((a.by0 + ty) * 4 + (int)(threadIdx.x >> 6)) * FR;
    const int z0 = a.zlo + tz * a.zc;
    const int z1 = min(z0 + a.zc, a.zhi);
    const unsigned P = a.P, S2 = a.S2, as = a.as;
    const T cx = a.cx, cy = a.cy, cz = a.cz;
    const T *__restrict__ vxA = a.A;
    const T *__restrict__ vyA = a.A + as;
    const T *__restrict__ vzA = a.A + 2 * (size_t)as;
    const T *__restrict__ sxx = a.A + 3 * (size_t)as;
    const T *__restrict__ syy = a.A + 4 * (size_t)as;
    const T *__restrict__ szz = a.A + 5 * (size_t)as;
    const T *__restrict__ syz = a.A + 6 * (size_t)as;
    const T *__restrict__ sxz = a.A + 7 * (size_t)as;
    const T *__restrict__ sxy = a.A + 8 * (size_t)as;
    const T *__restrict__ b = a.Pm + 9 * (size_t)as;
    // the source of this step (a vx force enters the velocities the stresses
    // read, the stored vx excludes it; an explosive source the stresses)
    unsigned sp = 0xffffffffu;
    T sadd = T(0);
    if (a.src_p != 0xffffffffu) {
        const int step = *a.step;
        if (step < a.src_len) {
            sp = a.src_p;
            sadd = a.src_type == 0 ? (T)(1000 * a.src_tr[step]) : (T)a.src_tr[step];
        }
    }
    const bool vsrc = a.src_type == 0, ssrc = a.src_type == 1;
    // v^{n+1/2} at one entry (update_vel.cpp:30-69, every update taken)
    auto ev_vx = [&](unsigned p) -> T {
        return l32(vxA, p) + l32(b, p) * (cx * (l32(sxx, p) - l32(sxx, p - 1)) + cy * (l32(sxy, p) - l32(sxy, p - P)) + cz * (l32(sxz, p) - l32(sxz, p - S2)));
    };
    auto ev_vy = [&](unsigned p) -> T {
        const T bav = T(0.25) * (l32(b, p) + l32(b, p + 1) + l32(b, p + P) + l32(b, p + P + 1));
        return l32(vyA, p) + bav * (cx * (l32(sxy, p + 1) - l32(sxy, p)) + cz * (l32(syz, p) - l32(syz, p - S2)) + cy * (l32(syy, p + P) - l32(syy, p)));
    };
    auto ev_vz = [&](unsigned p) -> T {
        const T bav = T(0.25) * (l32(b, p) + l32(b, p + S2) + l32(b, p + 1) + l32(b, p + S2 + 1));
        return l32(vzA, p) + bav * (cx * (l32(sxz, p + 1) - l32(sxz, p)) + cz * (l32(szz, p + S2) - l32(szz, p)) + cy * (l32(syz, p) - l32(syz, p - P)));
    };
    // The velocities outside the wavefront's columns at plane z (row above:
    // vy; row below: vx, vz; lanes 0..FR-1: vx right of lane 63, vy and vz
    // left of lane 0 for row `lane`), evaluated one plane ahead, together
    // with the lookahead whose loads bring the same lines.
    const int hr = lane < FR ? lane : 0;
    const unsigned pc0 = a.o + (unsigned)y0 * P + (unsigned)ix;   // (ix, y0, 0)
    const unsigned phc = pc0 - (unsigned)lane + (unsigned)hr * P;  // (x0, y0 + hr, 0)
    struct Halo { T vy_up, vx_dn, vz_dn, vx_r, vy_l, vz_l; };
    auto halo = [&](int z) -> Halo {
        Halo h;
        const unsigned pz = pc0 + (unsigned)z * S2, ph = phc + (unsigned)z * S2;
        h.vy_up = ev_vy(pz - P);
        h.vx_dn = ev_vx(pz + FR * P);
        if (vsrc && pz + FR * P == sp) h.vx_dn = h.vx_dn + sadd;
        h.vz_dn = ev_vz(pz + FR * P);
        h.vx_r = ev_vx(ph + 64);
        if (vsrc && ph + 64 == sp) h.vx_r = h.vx_r + sadd;
        h.vy_l = ev_vy(ph - 1);
        h.vz_l = ev_vz(ph - 1);
        return h;
    };
    // carried per row: the stresses and buoyancies the next lookahead shares
    // with the last one (as ColVel), v at plane iz (as the stresses read it),
    // vz at plane iz-1, and the stresses of plane iz the lookahead loaded
    T sxz_m[FR], syz_m[FR], szz_c[FR], b0[FR], b1[FR];
    T vx[FR], vy[FR], vz[FR], vzm[FR];
    T oxx[FR], oyy[FR], ozz[FR], oxy[FR], oyz[FR], oxz[FR];
    T src_raw = T(0);
#pragma unroll
    for (int r = 0; r < FR; r++) {
        const unsigned p = pc0 + (unsigned)r * P + (unsigned)z0 * S2;
        // the carries of the first lookahead (plane z0 + 1): stresses of
        // plane z0 below it, its own szz and buoyancies
        sxz_m[r] = l32(sxz, p);
        syz_m[r] = l32(syz, p);
        szz_c[r] = l32(szz, p + S2);
        b0[r] = l32(b, p + S2);
        b1[r] = l32(b, p + S2 + 1);
        oxx[r] = l32(sxx, p); oyy[r] = l32(syy, p); ozz[r] = l32(szz, p);
        oxy[r] = l32(sxy, p); oyz[r] = syz_m[r]; oxz[r] = sxz_m[r];
        vx[r] = ev_vx(p);
        vy[r] = ev_vy(p);
        vz[r] = ev_vz(p);
        vzm[r] = ev_vz(p - S2);
        if (vsrc && p == sp) { src_raw = vx[r]; vx[r] = vx[r] + sadd; }
    }
    Halo h = halo(z0);
    T *__restrict__ ovx = a.B;
    T *__restrict__ ovy = a.B + as;
    T *__restrict__ ovz = a.B + 2 * (size_t)as;
    const T *Pm = a.Pm;
    for (int iz = z0; iz < z1; iz++) {
        const unsigned pz = pc0 + (unsigned)iz * S2;
        __builtin_amdgcn_sched_barrier(0);
        const Halo hn = halo(iz + 1);
        const T h_xp_all = h.vx_r, h_ym_all = h.vy_l, h_zm_all = h.vz_l;
        T vy_above = h.vy_up;
#pragma unroll
        for (int r = 0; r < FR; r++) {
            __builtin_amdgcn_sched_barrier(0);
            const unsigned p = pz + (unsigned)r * P;
            // lookahead: v at plane iz + 1 of this row (carried as in k_vel)
            const unsigned pn = p + S2;
            const T sxx_n = l32(sxx, pn), syy_n = l32(syy, pn), sxy_n = l32(sxy, pn);
            const T sxz_c = l32(sxz, pn), syz_c = l32(syz, pn), szz_n = l32(szz, pn + S2), b0n = l32(b, pn + S2), b1n = l32(b, pn + S2 + 1);
            T nvx = l32(vxA, pn) + b0[r] * (cx * (sxx_n - l32(sxx, pn - 1)) + cy * (sxy_n - l32(sxy, pn - P)) + cz * (sxz_c - sxz_m[r]));
            const T bavy = T(0.25) * (b0[r] + b1[r] + l32(b, pn + P) + l32(b, pn + P + 1));
            const T nvy = l32(vyA, pn) + bavy * (cx * (l32(sxy, pn + 1) - sxy_n) + cz * (syz_c - syz_m[r]) + cy * (l32(syy, pn + P) - syy_n));
            const T bavz = T(0.25) * (b0[r] + b0n + b1[r] + b1n);
            const T nvz = l32(vzA, pn) + bavz * (cx * (l32(sxz, pn + 1) - sxz_c) + cz * (szz_n - szz_c[r]) + cy * (syz_c - l32(syz, pn - P)));
            const T szz_nc = szz_c[r];
            sxz_m[r] = sxz_c; syz_m[r] = syz_c; szz_c[r] = szz_n; b0[r] = b0n; b1[r] = b1n;
            T nraw = src_raw;
            if (vsrc && pn == sp) { nraw = nvx; nvx = nvx + sadd; }
            const T vx_c = vx[r], vy_c = vy[r], vz_c = vz[r], vz_m = vzm[r];
            const T sh_xp = __shfl_down(vx_c, 1, 64);
            const T sh_ym = __shfl_up(vy_c, 1, 64);
            const T sh_zm = __shfl_up(vz_c, 1, 64);
            const T h_xp = __shfl(h_xp_all, r, 64);
            const T h_ym = __shfl(h_ym_all, r, 64);
            const T h_zm = __shfl(h_zm_all, r, 64);
            const T vx_xp = lane == 63 ? h_xp : sh_xp;
            const T vy_xm = lane == 0 ? h_ym : sh_ym;
            const T vz_xm = lane == 0 ? h_zm : sh_zm;
            const T vx_yp = r + 1 < FR ? vx[r + 1 < FR ? r + 1 : r] : h.vx_dn;
            const T vz_yp = r + 1 < FR ? vz[r + 1 < FR ? r + 1 : r] : h.vz_dn;
            const T vy_ym = vy_above;
            vy_above = vy_c;
            // velocities
            s32n(ovx, p, (vsrc && p == sp) ? src_raw : vx_c);
            s32n(ovy, p, vy_c);
            s32n(ovz, p, vz_c);
            // stresses (update_stress.cpp:69-133)
            T sxx_o = oxx[r], syy_o = oyy[r], szz_o = ozz[r];
            if (ssrc && p == sp) {
                sxx_o = sxx_o + sadd;
                syy_o = syy_o + sadd;
                szz_o = szz_o + sadd;
            }
            const T c11 = l32n(Pm, p), c22 = l32n(Pm + as, p), c33 = l32n(Pm + 2 * (size_t)as, p);
            const T c23 = l32n(Pm + 3 * (size_t)as, p), c13 = l32n(Pm + 4 * (size_t)as, p), c12 = l32n(Pm + 5 * (size_t)as, p);
            const T dvx = vx_xp - vx_c;
            const T dvy = vy_c - vy_ym;
            const T dvz = vz_c - vz_m;
            s32n(a.B + 3 * (size_t)as, p, sxx_o + (c11 * cx * dvx + c12 * cy * dvy + c13 * cz * dvz));
            s32n(a.B + 4 * (size_t)as, p, syy_o + (c12 * cx * dvx + c22 * cy * dvy + c23 * cz * dvz));
            s32n(a.B + 5 * (size_t)as, p, szz_o + (c13 * cx * dvx + c23 * cy * dvy + c33 * cz * dvz));
            s32n(a.B + 6 * (size_t)as, p, oyz[r] + T(0.5) * l32n(Pm + 6 * (size_t)as, p) * (cz * (nvy - vy_c) + cy * (vz_yp - vz_c)));
            s32n(a.B + 7 * (size_t)as, p, oxz[r] + T(0.5) * l32n(Pm + 7 * (size_t)as, p) * (cz * (nvx - vx_c) + cx * (vz_c - vz_xm)));
            s32n(a.B + 8 * (size_t)as, p, oxy[r] + T(0.5) * l32n(Pm + 8 * (size_t)as, p) * (cy * (vx_yp - vx_c) + cx * (vy_c - vy_xm)));
            oxx[r] = sxx_n; oyy[r] = syy_n; ozz[r] = szz_nc; oxy[r] = sxy_n; oyz[r] = syz_c; oxz[r] = sxz_c;
            src_raw = nraw;
            vzm[r] = vz_c;
            vx[r] = nvx;
            vy[r] = nvy;
            vz[r] = nvz;
        }
        h = hn;
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
