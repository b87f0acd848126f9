// ecsg_diag.hpp -- the summation-by-parts diagnostic of test_sbp.cpp:5-199.
// Included by ecsg_capi.hip only.
//
// Seven sums of the form (u, D s)_u + (D u, s)_s over the velocity and
// stress nodes of one grid, with weight 1/2 per boundary index in the
// dimensions the reference weights.  They vanish (to rounding) on states of
// a rigid or free-surface single grid; a composite grid's interface flux
// shows in the z ones.  One thread per (x, y) column chunk accumulates all
// seven in FP64; per-workgroup partials, then one fixed-order sum each.
#pragma once
#include "ecsg_kernels.hpp"

namespace ecsg {

constexpr int SBP_BY = 4, SBP_ZC = 8;

template <typename T>
__global__ __launch_bounds__(BX * SBP_BY) void k_sbp_part(GridDev<T> g, int halo, double *__restrict__ part, int nblk)
{
    __shared__ double lds[7][SBP_BY];
    const int gx = (g.nx + BX) / BX, gy = (g.ny + SBP_BY) / SBP_BY;
    const int tile = blockIdx.x;
    const int bx = tile % gx, by = (tile / gx) % gy, bz = tile / (gx * gy);
    const int ix = bx * BX + threadIdx.x, iy = by * SBP_BY + threadIdx.y;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const int z0 = bz * SBP_ZC, z1 = min(z0 + SBP_ZC, nz + 1);
    double r[7] = {0, 0, 0, 0, 0, 0, 0};
    if (ix <= nx && iy <= ny) {
        const long S2 = g.S2, P = g.P;
        const T *vx = g.f[0], *vy = g.f[1], *vz = g.f[2], *sxx = g.f[3], *szz = g.f[5], *syz = g.f[6], *sxz = g.f[7],
                *sxy = g.f[8];
        auto wt = [](int i, int hi) { return (1 - 0.5 * (i == 0)) * (1 - 0.5 * (i == hi)); };
        for (int iz = z0; iz < z1; iz++) {
            const long p = g.o + (long)iz * S2 + (long)iy * P + ix;
            // vx nodes: x 0..nx, y 0..ny, z 0..nz
            {
                const double m = wt(iz, nz) * wt(iy, ny) * wt(ix, nx);
                const double u = vx[p];
                r[0] += m * u * ((double)sxz[p] - (double)sxz[p - S2]);
                r[1] += m * u * ((double)sxx[p] - (double)sxx[p - 1]);
                r[2] += m * u * ((double)sxy[p] - (double)sxy[p - P]);
            }
            // sxz nodes: x 0..nx, y 0..ny, z 0..nz-1
            if (iz <= nz - 1) {
                const double m = wt(iy, ny) * wt(ix, nx);
                r[0] += m * ((double)vx[p + S2] - (double)vx[p]) * (double)sxz[p];
                r[5] += m * ((double)vz[p] - (double)vz[p - 1]) * (double)sxz[p];
            }
            // sxx / szz nodes: x 0..nx-1, y 0..ny, z 0..nz
            if (ix <= nx - 1) {
                const double m = wt(iz, nz) * wt(iy, ny);
                r[1] += m * ((double)vx[p + 1] - (double)vx[p]) * (double)sxx[p];
                r[4] += m * ((double)vz[p] - (double)vz[p - S2]) * (double)szz[p];
            }
            // sxy nodes: x 0..nx, y 0..ny-1, z 0..nz
            if (iy <= ny - 1) {
                const double m = wt(iz, nz) * wt(ix, nx);
                r[2] += m * ((double)vx[p + P] - (double)vx[p]) * (double)sxy[p];
            }
            // vy nodes: x 0..nx-1, y 0..ny-1, z 0..nz
            if (ix <= nx - 1 && iy <= ny - 1) {
                const double m = wt(iz, nz);
                r[3] += m * (double)vy[p] * ((double)syz[p] - (double)syz[p - S2]);
            }
            // syz nodes: x 0..nx-1, y 0..ny-1, z 0..nz-1
            if (ix <= nx - 1 && iy <= ny - 1 && iz <= nz - 1) {
                r[3] += ((double)vy[p + S2] - (double)vy[p]) * (double)syz[p];
                r[6] += ((double)vz[p] - (double)vz[p - P]) * (double)syz[p];
            }
            // vz nodes: x 0..nx-1, y 0..ny, z 0..nz-1
            if (ix <= nx - 1 && iz <= nz - 1) {
                const double m = wt(iy, ny);
                const double u = vz[p];
                r[4] += m * u * ((double)szz[p + S2] - (double)szz[p]);
                r[5] += m * u * ((double)sxz[p + 1] - (double)sxz[p]);
                // syz at y = ny+1 lies past the field's rows for halo 1; the
                // reference's flat Field storage (rows of ny+2, field.cpp:
                // 123-136) reads it as (x, -1) of the next z plane
                const long qn = (iy + 1 >= ny + halo) ? p + S2 + (long)(1 - ny - 2 * halo) * P : p + P;
                r[6] += m * u * ((double)syz[qn] - (double)syz[p]);
            }
        }
    }
    for (int k = 0; k < 7; k++) {
        double v = r[k];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        const int lane = threadIdx.x + threadIdx.y * blockDim.x;
        if ((lane & 63) == 0) lds[k][lane >> 6] = v;
    }
    __syncthreads();
    if (threadIdx.x < 7 && threadIdx.y == 0) {
        double s = 0;
        for (int q = 0; q < SBP_BY; q++) s += lds[threadIdx.x][q];
        part[(long)threadIdx.x * nblk + tile] = s;
    }
}

__global__ __launch_bounds__(64) void k_sbp_final(const double *__restrict__ part, int nblk, double *__restrict__ out)
{
    const double s = wave_sum_fixed(part + (long)blockIdx.x * nblk, nblk);
    if (threadIdx.x == 0) out[blockIdx.x] = s;
}

// ---------------------------------------------------------------------------
// CFL bound of the flows (eci/SConstruct:107-114): the largest
// max(c11, c22, c33) * b over the primary nodes of a grid, in FP64.  A max of
// correctly rounded products: the same value in any order.  c1 == nullptr:
// the isotropic input, c0 = s11 = rho vp^2.
__global__ __launch_bounds__(256) void k_cfl_part(const double *__restrict__ b, const double *__restrict__ c0,
                                                  const double *__restrict__ c1, const double *__restrict__ c2,
                                                  long n, double *__restrict__ part)
{
    __shared__ double lds[4];
    double m = 0.0;
    for (long q = (long)blockIdx.x * 256 + threadIdx.x; q < n; q += (long)gridDim.x * 256) {
        double c = c0[q];
        if (c1) {
            c = fmax(c, c1[q]);
            c = fmax(c, c2[q]);
        }
        m = fmax(m, c * b[q]);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = fmax(fmax(lds[0], lds[1]), fmax(lds[2], lds[3]));
}

__global__ __launch_bounds__(64) void k_cfl_final(const double *__restrict__ part, int n, double *__restrict__ out)
{
    double m = 0.0;
    for (int q = threadIdx.x; q < n; q += 64) m = fmax(m, part[q]);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
    if (threadIdx.x == 0) out[0] = m;
}

// dt = frac * min over grids of h_min / (sqrt(vp2_g) sqrt 3), from the
// per-grid maxima vp2[g * stride], in the host formula's operation order.
struct CflDtArgs {
    double hmin[8];
    int ng;
    long stride;
    double frac;
};
__global__ void k_cfl_dt(const double *__restrict__ vp2, CflDtArgs a, double *__restrict__ dt)
{
    if (threadIdx.x != 0) return;
    double best = INFINITY;
    for (int g = 0; g < a.ng; g++) {
        const double vp = sqrt(vp2[(long)g * a.stride]);
        best = fmin(best, a.hmin[g] / (vp * sqrt(3.0)));
    }
    dt[0] = a.frac * best;
}

}  // namespace ecsg
