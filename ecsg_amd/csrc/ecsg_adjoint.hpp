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
// thread in a fixed order), so results are deterministic.  Gradient terms of
// each map with respect to the staged parameters are accumulated in the same
// kernels, using the forward state saved for that step.
//
// Notation: lam = adjoint wavefields (same layout as the forward fields);
// fv / fs = forward v^{n+1/2} / s^n read for the gradient terms.
#pragma once
#include "ecsg_kernels.hpp"

namespace ecsg {

template <typename T>
struct AdjGridDev {
    T *lam[9];          // adjoint vx vy vz sxx syy szz syz sxz sxy
    const T *fv[3];     // forward velocities of this step (stress transpose)
    const T *fs[6];     // forward stresses before this step (velocity transpose)
    T *gc[9];           // gradient w.r.t. the staged stiffnesses
    T *gb[3];           // gradient w.r.t. buoyancy as used by vx (node), vy (xy mean), vz (xz mean)
    const T *c[9];
    const T *b;
    int nx, ny, nz;
    long P, S2, o;
    T cx, cy, cz;
    int top_eff, bottom, free_top;
    int gx, gy, gz;     // tiling of the index box [-1, n] x [-1, n] x [-1, n]
};

constexpr int ABY = 4;   // adjoint stencil workgroup: 64 x ABY threads
constexpr int AZC = 8;   // z planes per thread

// ---------------------------------------------------------------------------
// Transpose of UpdateStress (stress stencil) on the velocities, with the
// gradient w.r.t. the nine staged stiffnesses.  Forward:
//   normal:  sxx += c11 cx dvx + c12 cy dvy + c13 cz dvz (free row: condensed moduli)
//   syz += 1/2 c44 (cz dz vy + cy dy vz), sxz += 1/2 c55 (...), sxy += 1/2 c66 (...)
// The adjoint weights e (normal) and h (shear) live on the stress points and
// vanish outside each stress field's update set.
template <typename T>
__device__ __forceinline__ void adj_normal_w(const AdjGridDev<T> &a, int ix, int iy, int iz, long p, T &ex, T &ey, T &ez)
{
    ex = T(0); ey = T(0); ez = T(0);
    if (ix < 0 || ix > a.nx - 1 || iy < 0 || iy > a.ny || iz < 0 || iz > a.nz) return;
    const T lx = a.lam[3][p], ly = a.lam[4][p];
    const T c11 = a.c[0][p], c22 = a.c[1][p], c33 = a.c[2][p], c23 = a.c[3][p], c13 = a.c[4][p], c12 = a.c[5][p];
    if (a.free_top && iz == 0) {
        ex = a.cx * ((c11 - c13 * c13 / c33) * lx + (c12 - c23 * c13 / c33) * ly);
        ey = a.cy * ((c12 - c13 * c23 / c33) * lx + (c22 - c23 * c23 / c33) * ly);
    } else {
        const T lz = a.lam[5][p];
        ex = a.cx * (c11 * lx + c12 * ly + c13 * lz);
        ey = a.cy * (c12 * lx + c22 * ly + c23 * lz);
        ez = a.cz * (c13 * lx + c23 * ly + c33 * lz);
    }
}
template <typename T>
__device__ __forceinline__ T adj_hyz(const AdjGridDev<T> &a, int ix, int iy, int iz, long p)
{
    if (ix < 0 || ix > a.nx - 1 || iy < 0 || iy > a.ny - 1 || iz < 0 || iz > a.nz - 1) return T(0);
    return T(0.5) * a.c[6][p] * a.lam[6][p];
}
template <typename T>
__device__ __forceinline__ T adj_hxz(const AdjGridDev<T> &a, int ix, int iy, int iz, long p)
{
    if (ix < 0 || ix > a.nx || iy < 0 || iy > a.ny || iz < 0 || iz > a.nz - 1) return T(0);
    return T(0.5) * a.c[7][p] * a.lam[7][p];
}
template <typename T>
__device__ __forceinline__ T adj_hxy(const AdjGridDev<T> &a, int ix, int iy, int iz, long p)
{
    if (ix < 0 || ix > a.nx || iy < 0 || iy > a.ny - 1 || iz < 0 || iz > a.nz) return T(0);
    return T(0.5) * a.c[8][p] * a.lam[8][p];
}

template <typename T>
__global__ __launch_bounds__(BX *ABY) void k_adj_stress(AdjGridDev<T> a)
{
    const int bx = blockIdx.x % a.gx;
    const int rest = blockIdx.x / a.gx;
    const int by = rest % a.gy;
    const int bz = rest / a.gy;
    const int ix = bx * BX + threadIdx.x - 1;
    const int iy = by * ABY + threadIdx.y - 1;
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    if (ix > nx || iy > ny) return;
    const int z0 = bz * AZC - 1;
    const int z1 = min(z0 + AZC, nz + 1);
    const long P = a.P, S2 = a.S2;
    const T cx = a.cx, cy = a.cy, cz = a.cz;
    for (int iz = z0; iz < z1; iz++) {
        const long q = a.o + (long)iz * S2 + (long)iy * P + ix;
        T ex, ey, ez, exm, eyp, ezp, dum;
        adj_normal_w(a, ix, iy, iz, q, ex, ey, ez);
        adj_normal_w(a, ix - 1, iy, iz, q - 1, exm, dum, dum);
        adj_normal_w(a, ix, iy + 1, iz, q + P, dum, eyp, dum);
        adj_normal_w(a, ix, iy, iz + 1, q + S2, dum, dum, ezp);
        const T hyz = adj_hyz(a, ix, iy, iz, q);
        const T hxz = adj_hxz(a, ix, iy, iz, q);
        const T hxy = adj_hxy(a, ix, iy, iz, q);
        // vx on [0, nx] x [0, ny] x [0, nz]
        if (ix >= 0 && iy >= 0 && iz >= 0 && iz <= nz) {
            const T d = (exm - ex) + cz * (adj_hxz(a, ix, iy, iz - 1, q - S2) - hxz) + cy * (adj_hxy(a, ix, iy - 1, iz, q - P) - hxy);
            a.lam[0][q] += d;
        }
        // vy on [-1, nx] x [-1, ny] x [0, nz]
        if (iz >= 0 && iz <= nz) {
            const T d = (ey - eyp) + cz * (adj_hyz(a, ix, iy, iz - 1, q - S2) - hyz) + cx * (hxy - adj_hxy(a, ix + 1, iy, iz, q + 1));
            a.lam[1][q] += d;
        }
        // vz on [-1, nx] x [0, ny] x [-1, nz]
        if (iy >= 0) {
            const T d = (ez - ezp) + cy * (adj_hyz(a, ix, iy - 1, iz, q - P) - hyz) + cx * (hxz - adj_hxz(a, ix + 1, iy, iz, q + 1));
            a.lam[2][q] += d;
        }
        // gradient w.r.t. the staged stiffnesses at this stress point
        if (ix >= 0 && iy >= 0 && iz >= 0) {
            const T *fvx = a.fv[0], *fvy = a.fv[1], *fvz = a.fv[2];
            if (ix <= nx - 1) {
                const T lx = a.lam[3][q], ly = a.lam[4][q];
                const T X = cx * (fvx[q + 1] - fvx[q]);
                const T Y = cy * (fvy[q] - fvy[q - P]);
                if (a.free_top && iz == 0) {
                    const T c33 = a.c[2][q], c23 = a.c[3][q], c13 = a.c[4][q];
                    const T u = c13 / c33, v = c23 / c33;
                    a.gc[0][q] += lx * X;
                    a.gc[1][q] += ly * Y;
                    a.gc[5][q] += lx * Y + ly * X;
                    a.gc[4][q] += lx * (T(-2) * u * X - v * Y) + ly * (-v * X);
                    a.gc[3][q] += lx * (-u * Y) + ly * (-u * X - T(2) * v * Y);
                    a.gc[2][q] += lx * (u * u * X + u * v * Y) + ly * (v * u * X + v * v * Y);
                } else {
                    const T lz = a.lam[5][q];
                    const T Z = cz * (fvz[q] - fvz[q - S2]);
                    a.gc[0][q] += lx * X;
                    a.gc[1][q] += ly * Y;
                    a.gc[2][q] += lz * Z;
                    a.gc[3][q] += ly * Z + lz * Y;
                    a.gc[4][q] += lx * Z + lz * X;
                    a.gc[5][q] += lx * Y + ly * X;
                }
            }
            if (ix <= nx - 1 && iy <= ny - 1 && iz <= nz - 1)
                a.gc[6][q] += a.lam[6][q] * (T(0.5) * (cz * (fvy[q + S2] - fvy[q]) + cy * (fvz[q + P] - fvz[q])));
            if (iz <= nz - 1)
                a.gc[7][q] += a.lam[7][q] * (T(0.5) * (cz * (fvx[q + S2] - fvx[q]) + cx * (fvz[q] - fvz[q - 1])));
            if (iy <= ny - 1)
                a.gc[8][q] += a.lam[8][q] * (T(0.5) * (cy * (fvx[q + P] - fvx[q]) + cx * (fvy[q] - fvy[q - 1])));
        }
    }
}

// ---------------------------------------------------------------------------
// Transpose of UpdateVel (velocity stencil) on the stresses, with the gradient
// w.r.t. buoyancy.  Forward:
//   vx += b (cx dx sxx + cy dy sxy + cz dz sxz)                on the vx update set
//   vy += mean_xy(b) (cx dx sxy + cz dz syz + cy dy syy)       on the vy update set
//   vz += mean_xz(b) (cx dx sxz + cz dz szz + cy dy syz)       on the vz update set
template <typename T>
__device__ __forceinline__ T adj_wx(const AdjGridDev<T> &a, int ix, int iy, int iz, long p)
{
    if (ix < 1 || ix > a.nx - 1 || iy < 1 || iy > a.ny - 1 || iz < a.top_eff || iz > a.nz - a.bottom) return T(0);
    return a.b[p] * a.lam[0][p];
}
template <typename T>
__device__ __forceinline__ T adj_wy(const AdjGridDev<T> &a, int ix, int iy, int iz, long p)
{
    if (ix < 0 || ix > a.nx - 1 || iy < 0 || iy > a.ny - 1 || iz < a.top_eff || iz > a.nz - a.bottom) return T(0);
    const T *b = a.b;
    return T(0.25) * (b[p] + b[p + 1] + b[p + a.P] + b[p + a.P + 1]) * a.lam[1][p];
}
template <typename T>
__device__ __forceinline__ T adj_wz(const AdjGridDev<T> &a, int ix, int iy, int iz, long p)
{
    if (ix < 0 || ix > a.nx - 1 || iy < 1 || iy > a.ny - 1 || iz < 0 || iz > a.nz - 1) return T(0);
    const T *b = a.b;
    return T(0.25) * (b[p] + b[p + a.S2] + b[p + 1] + b[p + a.S2 + 1]) * a.lam[2][p];
}

template <typename T>
__global__ __launch_bounds__(BX *ABY) void k_adj_vel(AdjGridDev<T> a)
{
    const int bx = blockIdx.x % a.gx;
    const int rest = blockIdx.x / a.gx;
    const int by = rest % a.gy;
    const int bz = rest / a.gy;
    const int ix = bx * BX + threadIdx.x;
    const int iy = by * ABY + threadIdx.y;
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    if (ix > nx || iy > ny) return;
    const int z0 = bz * AZC - 1;
    const int z1 = min(z0 + AZC, nz + 1);
    const long P = a.P, S2 = a.S2;
    const T cx = a.cx, cy = a.cy, cz = a.cz;
    for (int iz = z0; iz < z1; iz++) {
        const long q = a.o + (long)iz * S2 + (long)iy * P + ix;
        const T wx = adj_wx(a, ix, iy, iz, q);
        const T wy = adj_wy(a, ix, iy, iz, q);
        const T wz = adj_wz(a, ix, iy, iz, q);
        if (iz >= 0) {
            if (ix <= nx - 1) {
                a.lam[3][q] += cx * (wx - adj_wx(a, ix + 1, iy, iz, q + 1));
                a.lam[4][q] += cy * (adj_wy(a, ix, iy - 1, iz, q - P) - wy);
                a.lam[5][q] += cz * (adj_wz(a, ix, iy, iz - 1, q - S2) - wz);
            }
            if (iy <= ny - 1)
                a.lam[8][q] += cy * (wx - adj_wx(a, ix, iy + 1, iz, q + P)) + cx * (adj_wy(a, ix - 1, iy, iz, q - 1) - wy);
        }
        a.lam[7][q] += cz * (wx - adj_wx(a, ix, iy, iz + 1, q + S2)) + cx * (adj_wz(a, ix - 1, iy, iz, q - 1) - wz);
        if (ix <= nx - 1 && iy <= ny - 1)
            a.lam[6][q] += cz * (wy - adj_wy(a, ix, iy, iz + 1, q + S2)) + cy * (wz - adj_wz(a, ix, iy + 1, iz, q + P));
        // gradient w.r.t. buoyancy at this velocity point, forward s^n
        if (iz >= 0) {
            const T *sxx = a.fs[0], *syy = a.fs[1], *szz = a.fs[2], *syz = a.fs[3], *sxz = a.fs[4], *sxy = a.fs[5];
            const bool zin = iz >= a.top_eff && iz <= nz - a.bottom;
            if (ix >= 1 && ix <= nx - 1 && iy >= 1 && iy <= ny - 1 && zin)
                a.gb[0][q] += a.lam[0][q] * (cx * (sxx[q] - sxx[q - 1]) + cy * (sxy[q] - sxy[q - P]) + cz * (sxz[q] - sxz[q - S2]));
            if (ix <= nx - 1 && iy <= ny - 1 && zin)
                a.gb[1][q] += a.lam[1][q] * (cx * (sxy[q + 1] - sxy[q]) + cz * (syz[q] - syz[q - S2]) + cy * (syy[q + P] - syy[q]));
            if (ix <= nx - 1 && iy >= 1 && iy <= ny - 1 && iz <= nz - 1)
                a.gb[2][q] += a.lam[2][q] * (cx * (sxz[q + 1] - sxz[q]) + cz * (szz[q + S2] - szz[q]) + cy * (syz[q] - syz[q - P]));
        }
    }
}

// ---------------------------------------------------------------------------
// Transposed ghost mirrors h = -i: lam[i] -= sum over the mirrors of i of
// lam[h] (gather per source entry), then lam[h] = 0 (k_zero_list).
template <typename T>
__global__ __launch_bounds__(256) void k_mirror_T(T *lam, const long *__restrict__ src, const int *__restrict__ start,
                                                  const long *__restrict__ halo, int n)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    T acc = lam[src[t]];
    for (int q = start[t]; q < start[t + 1]; q++) acc -= lam[halo[q]];
    lam[src[t]] = acc;
}

template <typename T>
__global__ __launch_bounds__(256) void k_zero_list(T *lam, const long *__restrict__ idx, int n)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) lam[idx[t]] = T(0);
}

// ---------------------------------------------------------------------------
// Interface maps as term lists.  A key names one entry of one array of the
// fine (side 0) or coarse (side 1) grid: key = ((side*16 + a) << KSH) | idx,
// a = wavefield 0..8, or for parameters 0..8 staged stiffness, 9 buoyancy at
// the node, 10 its xy mean, 11 its xz mean.
constexpr int KSH = 36;

template <typename T>
struct TLDev {
    T *lam[2];            // adjoint wavefield blocks (9 arrays each)
    long astr[2];         // array stride
    const T *fwd[2][9];   // forward state read by the gradient terms
    const T *par[2];      // staged stiffness arrays 0..8, buoyancy as array 9
    T *grad[2];           // 12 gradient arrays
    long P[2], S2[2];
};

__device__ __forceinline__ int tl_side(long k) { return (int)(k >> (KSH + 4)); }
__device__ __forceinline__ int tl_arr(long k) { return (int)((k >> KSH) & 15); }
__device__ __forceinline__ long tl_idx(long k) { return k & ((1L << KSH) - 1); }

template <typename T>
__device__ __forceinline__ T &tl_lam(const TLDev<T> &d, long k)
{
    const int s = tl_side(k);
    return d.lam[s][(long)tl_arr(k) * d.astr[s] + tl_idx(k)];
}

template <typename T>
__device__ __forceinline__ double tl_param(const TLDev<T> &d, long k)
{
    const int s = tl_side(k), a = tl_arr(k);
    const long i = tl_idx(k);
    const T *base = d.par[s];
    const long as = d.astr[s];
    if (a < 9) return (double)base[(long)a * as + i];
    const T *b = base + 9 * as;
    if (a == 9) return (double)b[i];
    if (a == 10) return 0.25 * ((((double)b[i] + (double)b[i + 1]) + (double)b[i + d.P[s]]) + (double)b[i + d.P[s] + 1]);
    return 0.25 * ((((double)b[i] + (double)b[i + d.S2[s]]) + (double)b[i + 1]) + (double)b[i + d.S2[s] + 1]);
}

// Transposed fine-ghost assignments out = sum cst * in: lam[in] += sum cst * lam[out]
template <typename T>
__global__ __launch_bounds__(256) void k_tl_fine(TLDev<T> d, const long *__restrict__ ukey, const int *__restrict__ start,
                                                 const long *__restrict__ okey, const double *__restrict__ cst, int n)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    double acc = 0;
    for (int q = start[t]; q < start[t + 1]; q++) acc += cst[q] * (double)tl_lam(d, okey[q]);
    T &x = tl_lam(d, ukey[t]);
    x = (T)((double)x + acc);
}

template <typename T>
__global__ __launch_bounds__(256) void k_tl_zero(TLDev<T> d, const long *__restrict__ key, int n)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) tl_lam(d, key[t]) = T(0);
}

// Adjoint of the solve's output assignment: rhs[row] = lam[x entry] (divided
// by the pointwise denominator when den != null), lam[x entry] = 0.
template <typename T>
__global__ __launch_bounds__(256) void k_tl_readx(TLDev<T> d, const long *__restrict__ xkey, const int *__restrict__ row,
                                                  const double *__restrict__ den, double *__restrict__ rhs, int n)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    T &x = tl_lam(d, xkey[t]);
    const double v = (double)x;
    rhs[row[t]] = den ? v / den[row[t]] : v;
    x = T(0);
}

// lam[entry] += sum over terms of cst * param * mu[row]   (B^T mu)
template <typename T>
__global__ __launch_bounds__(256) void k_tl_prop(TLDev<T> d, const long *__restrict__ ukey, const int *__restrict__ start,
                                                 const int *__restrict__ row, const long *__restrict__ pkey,
                                                 const double *__restrict__ cst, const double *__restrict__ mu, int n)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    double acc = 0;
    for (int q = start[t]; q < start[t + 1]; q++) {
        const double p = pkey[q] < 0 ? 1.0 : tl_param(d, pkey[q]);
        acc += cst[q] * p * mu[row[q]];
    }
    T &x = tl_lam(d, ukey[t]);
    x = (T)((double)x + acc);
}

// grad[param] += sum over terms of cst * mu[row] * forward[entry]
// (mu^T (dB/dp u - dA/dp x); the A terms carry the negated constant)
template <typename T>
__global__ __launch_bounds__(256) void k_tl_grad(TLDev<T> d, const long *__restrict__ ukey, const int *__restrict__ start,
                                                 const int *__restrict__ row, const long *__restrict__ fkey,
                                                 const double *__restrict__ cst, const double *__restrict__ mu, int n)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    double acc = 0;
    for (int q = start[t]; q < start[t + 1]; q++) {
        const long k = fkey[q];
        const double u = (double)d.fwd[tl_side(k)][tl_arr(k)][tl_idx(k)];
        acc += cst[q] * mu[row[q]] * u;
    }
    const long k = ukey[t];
    const int s = tl_side(k);
    T &g = d.grad[s][(long)tl_arr(k) * d.astr[s] + tl_idx(k)];
    g = (T)((double)g + acc);
}

// Tridiagonal solve with the prefactored elimination of the vz-szz system
// (symmetric, so it also solves the transposed system); solution in B.
__global__ __launch_bounds__(64) void k_tri_solve(double *__restrict__ B, const double *__restrict__ M,
                                                  const double *__restrict__ Dp, const double *__restrict__ U, int nx2, int n)
{
    const int I = blockIdx.x * blockDim.x + threadIdx.x;
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
__global__ __launch_bounds__(256) void k_adj_rec(RecAdj<T> a, int step)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.nloc) return;
    T acc = T(0);
    for (int q = a.start[t]; q < a.start[t + 1]; q++) acc += a.resid[(long)a.rec[q] * a.nsteps + step];
    const long k = a.loc[t];
    T *v = a.lamvx[k >> 40];
    v[k & ((1L << 40) - 1)] += acc;
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
