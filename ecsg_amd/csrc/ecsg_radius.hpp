// ecsg_radius.hpp -- single-grid stencils of radius R = 2..6
// (update_vel.cpp / update_stress.cpp with their `ir` loops and the wider
// ghost mirrors; FD coefficients init_del_fdcoeff.cpp:38-85).  Included by
// ecsg_capi.hip only.  Composite grids stay radius 1 (the interface
// formulas assume it).
//
// A step is: k_vel_r (every updated velocity point, one thread per (x, y)
// column chunk), the velocity ghost mirrors in two launches, k_stress_r,
// the stress ghost mirrors in one launch.  Each update accumulates its
// `radius` terms one after another, as the reference's loops over ir do, so
// FP64 storage reproduces it bit for bit.  The mirror loops are transcribed
// one to one into MirrorLoop descriptors (same index ranges, same sources);
// within one launch no loop reads what another writes.  The velocity x
// mirrors of vy read the y-halo row iy = ny, which the y mirrors rewrite, so
// those run in a second launch, after them, as in the reference.
#pragma once
#include "ecsg_kernels.hpp"

namespace ecsg {

constexpr int RMAX = 6;
constexpr int RBY = 4;    // rows per workgroup
constexpr int RZC = 8;    // z planes per thread

template <typename T>
struct RCoef {
    T c[3][RMAX];  // coeff[d][ir] = fd[ir] * dt / h[d] (init_del_fdcoeff.cpp:19)
};

// One ghost-mirror loop: for ir in [1, nir], a in [a0, a1], b in [b0, b1]
// (the two other axes, a the lower of them), f[dst] = sign * f[src] with
// dst/src = base + step * ir on the mirrored axis.
struct MirrorLoop {
    int f;            // field 0..8
    int axis;         // mirrored axis 0 x, 1 y, 2 z
    int dbase, dstep; // destination coordinate on the axis: dbase + dstep * ir
    int sbase, sstep; // source coordinate on the axis
    int a0, a1, b0, b1;
    int nir;
    int sign;
};

constexpr int MAXML = 20;

struct MirrorSet {
    MirrorLoop m[MAXML];
    long start[MAXML + 1];  // prefix sums of the loop sizes
    int n;
};

template <typename T>
__global__ __launch_bounds__(256) void k_mirror(GridDev<T> g, MirrorSet ms)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ms.start[ms.n]) return;
    int k = 0;
    while (t >= ms.start[k + 1]) k++;
    const MirrorLoop &L = ms.m[k];
    long u = t - ms.start[k];
    const int na = L.a1 - L.a0 + 1, nb = L.b1 - L.b0 + 1;
    const int ia = L.a0 + (int)(u % na);
    u /= na;
    const int ib = L.b0 + (int)(u % nb);
    const int ir = 1 + (int)(u / nb);
    int d[3], s[3];
    const int ax = L.axis, oa = ax == 0 ? 1 : 0, ob = ax == 2 ? 1 : 2;
    d[oa] = s[oa] = ia;
    d[ob] = s[ob] = ib;
    d[ax] = L.dbase + L.dstep * ir;
    s[ax] = L.sbase + L.sstep * ir;
    T *f = g.f[L.f];
    const long pd = g.o + (long)d[2] * g.S2 + (long)d[1] * g.P + d[0];
    const long ps = g.o + (long)s[2] * g.S2 + (long)s[1] * g.P + s[0];
    f[pd] = L.sign > 0 ? f[ps] : -f[ps];
}

// UpdateVel interior updates (update_vel.cpp:30-69), radius R.
template <typename T, int R, bool EN>
__global__ __launch_bounds__(BX * RBY) void k_vel_r(GridDev<T> g, RCoef<T> co, double *__restrict__ part,
                                                    const int *__restrict__ src, T *__restrict__ src_old)
{
    __shared__ double lds[RBY];
    const int gx = (g.nx + BX) / BX, gy = (g.ny + RBY) / RBY;
    const int tile = blockIdx.x;
    const int bx = tile % gx, by = (tile / gx) % gy, bz = tile / (gx * gy);
    const int ix = bx * BX + threadIdx.x, iy = by * RBY + threadIdx.y;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const int z0 = bz * RZC, z1 = min(z0 + RZC, nz + 1);
    double acc = 0.0;
    if (ix <= nx && iy <= ny) {
        const long S2 = g.S2, P = g.P;
        T *__restrict__ vx = g.f[0];
        T *__restrict__ vy = g.f[1];
        T *__restrict__ vz = g.f[2];
        const T *__restrict__ sxx = g.f[3];
        const T *__restrict__ syy = g.f[4];
        const T *__restrict__ szz = g.f[5];
        const T *__restrict__ syz = g.f[6];
        const T *__restrict__ sxz = g.f[7];
        const T *__restrict__ sxy = g.f[8];
        const T *__restrict__ b = g.b;
        const bool ok_x = ix >= 1 && ix <= nx - 1 && iy >= 1 && iy <= ny - 1;
        const bool ok_y = ix <= nx - 1 && iy <= ny - 1;
        const bool ok_z = ix <= nx - 1 && iy >= 1 && iy <= ny - 1;
        const int zlo = g.top_eff, zhi = nz - g.bottom;
        for (int iz = z0; iz < z1; iz++) {
            const long p = g.o + (long)iz * S2 + (long)iy * P + ix;
            const bool zin = iz >= zlo && iz <= zhi;
            const T vx_o = vx[p];
            T vx_c = vx_o;
            const T b0 = b[p];
            if (ok_x && zin) {
#pragma unroll
                for (int ir = 0; ir < R; ir++)
                    vx_c += b0 * (co.c[0][ir] * (sxx[p + ir] - sxx[p - ir - 1])
                                  + co.c[1][ir] * (sxy[p + ir * P] - sxy[p - (ir + 1) * P])
                                  + co.c[2][ir] * (sxz[p + ir * S2] - sxz[p - (ir + 1) * S2]));
                vx[p] = vx_c;
            }
            const T vy_o = vy[p];
            T vy_c = vy_o;
            if (ok_y && zin) {
                const T bav = T(0.25) * (((b0 + b[p + 1]) + b[p + P]) + b[p + P + 1]);
#pragma unroll
                for (int ir = 0; ir < R; ir++)
                    vy_c += bav * (co.c[0][ir] * (sxy[p + ir + 1] - sxy[p - ir])
                                   + co.c[2][ir] * (syz[p + ir * S2] - syz[p - (ir + 1) * S2])
                                   + co.c[1][ir] * (syy[p + (ir + 1) * P] - syy[p - ir * P]));
                vy[p] = vy_c;
            }
            const T vz_o = vz[p];
            T vz_c = vz_o;
            if (ok_z && iz <= nz - 1) {
                const T bav = T(0.25) * (((b0 + b[p + S2]) + b[p + 1]) + b[p + S2 + 1]);
#pragma unroll
                for (int ir = 0; ir < R; ir++)
                    vz_c += bav * (co.c[0][ir] * (sxz[p + ir + 1] - sxz[p - ir])
                                   + co.c[2][ir] * (szz[p + (ir + 1) * S2] - szz[p - ir * S2])
                                   + co.c[1][ir] * (syz[p + ir * P] - syz[p - (ir + 1) * P]));
                vz[p] = vz_c;
            }
            if (EN) {
                // kinetic part of ComputeEnergy, as in vel_tile
                const double m2 = (iz == 0 || iz == nz) ? 0.5 : 1.0;
                const double m1 = (iy == 0 || iy == ny) ? 0.5 : 1.0;
                const double m0 = (ix == 0 || ix == nx) ? 0.5 : 1.0;
                acc += m2 * m1 * m0 * (1.0 / (double)b0) * (double)vx_c * (double)vx_o;
                if (src && ix == src[0] && iy == src[1] && iz == src[2]) *src_old = vx_o;
                const double d0 = b0, d1 = b[p + 1];
                if (ok_y) {
                    const double dP = b[p + P], dP1 = b[p + P + 1];
                    const double rb = ((((((((d0 + d0) + dP) + dP) + d1) + d1) + dP1) + dP1)) / 8;
                    acc += m2 * (1.0 / rb) * (double)vy_c * (double)vy_o;
                }
                if (ix <= nx - 1 && iz <= nz - 1) {
                    const double dn = b[p + S2], d1n = b[p + S2 + 1];
                    const double rb = ((((((((d0 + dn) + d0) + dn) + d1) + d1n) + d1) + d1n)) / 8;
                    acc += m1 * (1.0 / rb) * (double)vz_c * (double)vz_o;
                }
            }
        }
    }
    if (EN) {
        const double s = block_sum<RBY>(acc, lds);
        if (threadIdx.x == 0 && threadIdx.y == 0) part[tile] = s;
    }
}

// UpdateStress interior updates (update_stress.cpp:34-133), radius R.
template <typename T, int R, bool EN>
__global__ __launch_bounds__(BX * RBY) void k_stress_r(GridDev<T> g, RCoef<T> co, double *__restrict__ part)
{
    __shared__ double lds[RBY];
    const int gx = (g.nx + BX) / BX, gy = (g.ny + RBY) / RBY;
    const int tile = blockIdx.x;
    const int bx = tile % gx, by = (tile / gx) % gy, bz = tile / (gx * gy);
    const int ix = bx * BX + threadIdx.x, iy = by * RBY + threadIdx.y;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const int z0 = bz * RZC, z1 = min(z0 + RZC, nz + 1);
    double acc = 0.0;
    if (ix <= nx && iy <= ny) {
        const long S2 = g.S2, P = g.P;
        const T *__restrict__ vx = g.f[0];
        const T *__restrict__ vy = g.f[1];
        const T *__restrict__ vz = g.f[2];
        T *__restrict__ sxx = g.f[3];
        T *__restrict__ syy = g.f[4];
        T *__restrict__ szz = g.f[5];
        T *__restrict__ syz = g.f[6];
        T *__restrict__ sxz = g.f[7];
        T *__restrict__ sxy = g.f[8];
        for (int iz = z0; iz < z1; iz++) {
            const long p = g.o + (long)iz * S2 + (long)iy * P + ix;
            if (ix <= nx - 1) {
                const T c11 = g.c[0][p], c22 = g.c[1][p], c33 = g.c[2][p], c23 = g.c[3][p], c13 = g.c[4][p], c12 = g.c[5][p];
                const T sxx_o = sxx[p], syy_o = syy[p], szz_o = szz[p];
                T a = sxx_o, bq = syy_o, c = szz_o;
                if (g.free_top && iz == 0) {
                    // update_stress.cpp:37-65
                    const T f11 = c11 - c13 * c13 / c33, f12 = c12 - c13 * c23 / c33;
                    const T f21 = c12 - c23 * c13 / c33, f22 = c22 - c23 * c23 / c33;
#pragma unroll
                    for (int ir = 0; ir < R; ir++) {
                        const T dvx = vx[p + ir + 1] - vx[p - ir];
                        const T dvy = vy[p + ir * P] - vy[p - (ir + 1) * P];
                        a += f11 * co.c[0][ir] * dvx + f12 * co.c[1][ir] * dvy;
                        bq += f21 * co.c[0][ir] * dvx + f22 * co.c[1][ir] * dvy;
                    }
                    sxx[p] = a;
                    syy[p] = bq;
                } else {
#pragma unroll
                    for (int ir = 0; ir < R; ir++) {
                        const T dvx = vx[p + ir + 1] - vx[p - ir];
                        const T dvy = vy[p + ir * P] - vy[p - (ir + 1) * P];
                        const T dvz = vz[p + ir * S2] - vz[p - (ir + 1) * S2];
                        a += c11 * co.c[0][ir] * dvx + c12 * co.c[1][ir] * dvy + c13 * co.c[2][ir] * dvz;
                        bq += c12 * co.c[0][ir] * dvx + c22 * co.c[1][ir] * dvy + c23 * co.c[2][ir] * dvz;
                        c += c13 * co.c[0][ir] * dvx + c23 * co.c[1][ir] * dvy + c33 * co.c[2][ir] * dvz;
                    }
                    sxx[p] = a;
                    syy[p] = bq;
                    szz[p] = c;
                }
                if (EN) {
                    double s11, s22, s33, s23, s13, s12;
                    compliance(g, c11, c22, c33, c23, c13, c12, s11, s22, s33, s23, s13, s12);
                    const double m = ((iz == 0 || iz == nz) ? 0.5 : 1.0) * ((iy == 0 || iy == ny) ? 0.5 : 1.0);
                    const double x = sxx_o, y = syy_o, z = szz_o;
                    acc += m * s11 * x * x + m * s22 * y * y + m * s33 * z * z
                         + 2 * (m * s23 * y * z) + 2 * (m * s13 * x * z) + 2 * (m * s12 * x * y);
                }
            }
            if (ix <= nx - 1 && iy <= ny - 1 && iz <= nz - 1) {
                const T c44 = g.c[6][p];
                const T o = syz[p];
                T v = o;
#pragma unroll
                for (int ir = 0; ir < R; ir++)
                    v += T(0.5) * c44 * (co.c[2][ir] * (vy[p + (ir + 1) * S2] - vy[p - ir * S2])
                                         + co.c[1][ir] * (vz[p + (ir + 1) * P] - vz[p - ir * P]));
                syz[p] = v;
                if (EN) acc += 2 * ((1.0 / (double)c44) * (double)o * (double)o);
            }
            if (iz <= nz - 1) {
                const T c55 = g.c[7][p];
                const T o = sxz[p];
                T v = o;
#pragma unroll
                for (int ir = 0; ir < R; ir++)
                    v += T(0.5) * c55 * (co.c[2][ir] * (vx[p + (ir + 1) * S2] - vx[p - ir * S2])
                                         + co.c[0][ir] * (vz[p + ir] - vz[p - ir - 1]));
                sxz[p] = v;
                if (EN) {
                    const double m = ((iy == 0 || iy == ny) ? 0.5 : 1.0) * ((ix == 0 || ix == nx) ? 0.5 : 1.0);
                    acc += 2 * (m * (1.0 / (double)c55) * (double)o * (double)o);
                }
            }
            if (iy <= ny - 1) {
                const T c66 = g.c[8][p];
                const T o = sxy[p];
                T v = o;
#pragma unroll
                for (int ir = 0; ir < R; ir++)
                    v += T(0.5) * c66 * (co.c[1][ir] * (vx[p + (ir + 1) * P] - vx[p - ir * P])
                                         + co.c[0][ir] * (vy[p + ir] - vy[p - ir - 1]));
                sxy[p] = v;
                if (EN) {
                    const double m = ((iz == 0 || iz == nz) ? 0.5 : 1.0) * ((ix == 0 || ix == nx) ? 0.5 : 1.0);
                    acc += 2 * (m * (1.0 / (double)c66) * (double)o * (double)o);
                }
            }
        }
    }
    if (EN) {
        const double s = block_sum<RBY>(acc, lds);
        if (threadIdx.x == 0 && threadIdx.y == 0) part[tile] = s;
    }
}

// FD coefficients of the staggered first derivative (init_del_fdcoeff.cpp:38-85)
inline int fd_table(int radius, double *c)
{
    switch (radius) {
    case 1: c[0] = 1.0; return 0;
    case 2: c[0] = 9.0 / 8; c[1] = -1.0 / 24; return 0;
    case 3: c[0] = 75.0 / 64; c[1] = -25.0 / 384; c[2] = 3.0 / 640; return 0;
    case 4: c[0] = 1225.0 / 1024; c[1] = -245.0 / 3072; c[2] = 49.0 / 5120; c[3] = -5.0 / 7168; return 0;
    case 5: c[0] = 711.0 / 587; c[1] = -158.0 / 1761; c[2] = 296.0 / 21383; c[3] = -25.0 / 14159; c[4] = 17.0 / 143243; return 0;
    case 6:
        c[0] = 1225.0 / 1003; c[1] = -338.0 / 3487; c[2] = 35.0 / 2006; c[3] = -127.0 / 42800; c[4] = 19.0 / 52924;
        c[5] = -6.0 / 274627;
        return 0;
    default: return -1;
    }
}

}  // namespace ecsg
