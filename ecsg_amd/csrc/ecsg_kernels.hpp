// ecsg_kernels.hpp -- CDNA4 (gfx950) kernels of the ECSG composite
// staggered-grid time step.  Included by ecsg_capi.hip only.
//
// Device layout of every array of a grid (wavefields, staged stiffness,
// buoyancy): x fastest, index range [-1, n+1] in each dimension,
//   idx(ix, iy, iz) = o + iz*S2 + iy*P + ix,
// where the row pitch P is padded so that ix = 0 starts a 128-byte line.
//
// Stencil kernels: one thread owns an (ix, iy) column and marches a chunk of
// z planes, keeping the z-neighbours of its own column in registers; x/y
// neighbours come from the L1/L2-resident rows of the same and adjacent
// workgroups.  Workgroups are 64 x 4 threads (one wavefront per x row of
// 64 nodes); the tile order is remapped so that consecutive tiles (which
// share y halo rows) land on the same XCD and its L2.
//
// All floating-point expressions keep the reference's operand order and the
// build disables contraction, so the double-precision path reproduces the
// reference arithmetic operation for operation.
#pragma once
#include <hip/hip_runtime.h>

namespace ecsg {

constexpr int BX = 64;
constexpr int BY = 4;
constexpr int NTHR = BX * BY;
constexpr int MAXG = 8;

template <typename T>
struct GridDev {
    T *f[9];        // vx vy vz sxx syy szz syz sxz sxy
    T *st[9];       // when set: every value written to f (and every owned entry)
                    // is also written here -- the saved state of this step
    const T *c[9];  // c11 c22 c33 c23 c13 c12 c44 c55 c66 at the stress positions
    const T *b;     // buoyancy, primary nodes
    int nx, ny, nz;
    long P, S2, o;
    T cx, cy, cz;   // dt/h per dimension (init_del_fdcoeff.cpp:19)
    int top_eff;    // 1: vx/vy not updated on iz = 0 (rigid top, update_vel.cpp:17-18,30)
    int bottom;     // 1: last grid, rigid bottom
    int free_top;   // 1: free-surface top (update_stress.cpp:28-30)
    int iso;        // isotropic compliance formula for the energy
    int zy;         // last z plane on which vy is mirrored across the y walls
    int zc, gx, gy, nblk;  // tiling
    int zb, ze;            // z planes [zb, ze) covered by a stencil launch
};

// Workgroup id -> tile id such that each XCD (ids dispatched round robin over
// 8 XCDs) receives one contiguous range of tiles.
__device__ __forceinline__ int xcd_tile(int b, int n)
{
    const int q = n >> 3, rem = n & 7, x = b & 7, k = b >> 3;
    return x * q + (x < rem ? x : rem) + k;
}

// Tile order of the forward stencil sweeps: bands of ECSG_TILE_BAND tile
// rows (gx x band tiles, one z chunk) dealt to the 8 XCDs in turn, so that
// the XCDs sweep neighbouring bands of the same planes at the same time;
// the tiles past the last whole group of 8 bands keep their own index.
// ECSG_TILE_BAND = 0: one contiguous range of tiles per XCD (xcd_tile).
#ifndef ECSG_TILE_BAND
#define ECSG_TILE_BAND 4
#endif
__device__ __forceinline__ int sweep_tile(int b, int n, int gx)
{
#if ECSG_TILE_BAND > 0
    const int gs = gx * ECSG_TILE_BAND, full = n / (8 * gs) * (8 * gs);
    if (b >= full) return b;
    const int x = b & 7, k = b >> 3;
    return ((k / gs) * 8 + x) * gs + k % gs;
#else
    (void)gx;
    return xcd_tile(b, n);
#endif
}

// Cache hints of the stencil sweeps (ECSG_NT bit 0: streaming stores of the
// updated fields; bit 1: non-temporal loads of the single-use streams, the
// stiffnesses and the values being updated).  Both together took the
// 118 M-node FP32 stress sweep from 2.42 to 2.28 ms and the velocity sweep
// from 1.38 to 1.33 ms on MI355X; either alone gained less or nothing.
#ifndef ECSG_NT
#define ECSG_NT 3
#endif
template <typename T>
__device__ __forceinline__ T ldc(const T *a, long i)
{
#if ECSG_NT & 2
    return __builtin_nontemporal_load(a + i);
#else
    return a[i];
#endif
}
template <typename T>
__device__ __forceinline__ void stc(T *a, long i, T v)
{
#if ECSG_NT & 1
    __builtin_nontemporal_store(v, a + i);
#else
    a[i] = v;
#endif
}

// Element access of the stencil sweeps.  With IT = unsigned (every array of
// the grid within 4 GiB) the byte offset is formed in 32 bits, so a load is
// a scalar base plus a 32-bit vector offset: the velocity sweep then holds
// 31 instead of 75 VGPRs, the stress sweep 34 instead of 95 (FP32), and the
// 1207x1207x81 velocity sweep takes 1.40 instead of 1.45 ms.  IT = long
// indexes as usual (larger grids).
template <typename T>
__device__ __forceinline__ T *at(T *a, unsigned i)
{
    return (T *)((char *)a + i * (unsigned)sizeof(T));
}
template <typename T>
__device__ __forceinline__ const T *at(const T *a, unsigned i)
{
    return (const T *)((const char *)a + i * (unsigned)sizeof(T));
}
template <typename T>
__device__ __forceinline__ T *at(T *a, long i)
{
    return a + i;
}
template <typename T>
__device__ __forceinline__ const T *at(const T *a, long i)
{
    return a + i;
}
template <typename T>
__device__ __forceinline__ T ldc(const T *a, unsigned i)
{
#if ECSG_NT & 2
    return __builtin_nontemporal_load(at(a, i));
#else
    return *at(a, i);
#endif
}
template <typename T>
__device__ __forceinline__ void stc(T *a, unsigned i, T v)
{
#if ECSG_NT & 1
    __builtin_nontemporal_store(v, at(a, i));
#else
    *at(a, i) = v;
#endif
}

// Fixed-order workgroup reduction of one double per thread (NW wavefronts).
template <int NW>
__device__ __forceinline__ double block_sum(double v, double *lds)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const int lane = threadIdx.x + threadIdx.y * blockDim.x;
    if ((lane & 63) == 0) lds[lane >> 6] = v;
    __syncthreads();
    double r = 0.0;
    if (lane == 0)
        for (int q = 0; q < NW; q++) r += lds[q];
    return r;  // valid in lane 0
}

// ---------------------------------------------------------------------------
// UpdateVel (update_vel.cpp:4-249) with its radius-1 ghost mirrors fused.
// EN: also the kinetic part of ComputeEnergy (compute_energy.cpp:38-39,
// TriMulFields field.cpp:482-518) from the pre- and post-update velocities.
template <typename T, bool EN, int BYv, typename IT>
__device__ __forceinline__ void vel_tile(const GridDev<T> &g, const int tile, double *__restrict__ part,
                                         const int *__restrict__ src, T *__restrict__ src_old, double *lds)
{
    const int bx = tile % g.gx;
    const int rest = tile / g.gx;
    const int by = rest % g.gy;
    const int bz = rest / g.gy;
    const int ix = bx * BX + threadIdx.x;
    const int iy = by * BYv + threadIdx.y;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const int z0 = g.zb + bz * g.zc;
    const int z1 = min(z0 + g.zc, g.ze);
    double acc = 0.0;
    if (ix <= nx && iy <= ny) {
        const IT S2 = (IT)g.S2, P = (IT)g.P;
        IT p = (IT)g.o + (IT)z0 * S2 + (IT)iy * P + (IT)ix;
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
        const T cx = g.cx, cy = g.cy, cz = g.cz;
        const bool sv = g.st[0] != nullptr;
        T sxz_m = (*at(sxz, p - S2));
        T syz_m = (*at(syz, p - S2));
        T szz_c = (*at(szz, p));
        T b0 = (*at(b, p));
        T b1 = (*at(b, p + 1));
        // the x-shifted operands come from the neighbouring lanes; the
        // wavefront's end lanes (and the last node of a row, whose right
        // neighbour has no lane) load them (1207x1207x81 velocity interior
        // beside the Jacobi solve 1.27 -> 1.19 ms, three-grid step 4.178 ->
        // 4.128 ms)
        const bool xe = threadIdx.x == BX - 1 || ix == nx, xw = threadIdx.x == 0;
        for (int iz = z0; iz < z1; ++iz, p += S2) {
            // Every operand of the plane is loaded before any store (all of
            // them lie inside the halo-padded arrays for ix <= nx, iy <= ny,
            // iz <= nz), so the loads of a plane are in flight together
            // instead of in three dependent groups behind the stores, which
            // the compiler may not move loads across.
            const T sxz_c = (*at(sxz, p));
            const T syz_c = (*at(syz, p));
            const T szz_n = (*at(szz, p + S2));
            const T b0n = (*at(b, p + S2));
            const T vx_o = ldc(vx, p);
            const T vy_o = ldc(vy, p);
            const T vz_o = ldc(vz, p);
            const T sxx_c = (*at(sxx, p));
            const T sxy_c = (*at(sxy, p)), sxy_s = (*at(sxy, p - P));
            const T syy_c = (*at(syy, p)), syy_n = (*at(syy, p + P));
            const T syz_s = (*at(syz, p - P));
            const T bP = (*at(b, p + P));
            T b1n = T(0), sxx_w = T(0), sxy_e = T(0), sxz_e = T(0), bP1 = T(0);
            if (xe) {
                b1n = (*at(b, p + S2 + 1));
                sxy_e = (*at(sxy, p + 1));
                sxz_e = (*at(sxz, p + 1));
                bP1 = (*at(b, p + P + 1));
            }
            if (xw) sxx_w = (*at(sxx, p - 1));
            {
                const T t0 = __shfl_down(b0n, 1, 64), t1 = __shfl_down(sxy_c, 1, 64);
                const T t2 = __shfl_down(sxz_c, 1, 64), t3 = __shfl_down(bP, 1, 64);
                const T t4 = __shfl_up(sxx_c, 1, 64);
                if (!xe) { b1n = t0; sxy_e = t1; sxz_e = t2; bP1 = t3; }
                if (!xw) sxx_w = t4;
            }
            const bool zin = iz >= zlo && iz <= zhi;
            // vx, update_vel.cpp:30-37
            T vx_c = vx_o;
            if (ok_x && zin) {
                vx_c = vx_o + b0 * (cx * (sxx_c - sxx_w) + cy * (sxy_c - sxy_s) + cz * (sxz_c - sxz_m));
                stc(vx, p, vx_c);
            }
            // vy, update_vel.cpp:43-53
            T vy_c = vy_o;
            if (ok_y && zin) {
                const T bav = T(0.25) * (b0 + b1 + bP + bP1);
                vy_c = vy_o + bav * (cx * (sxy_e - sxy_c) + cz * (syz_c - syz_m) + cy * (syy_n - syy_c));
                stc(vy, p, vy_c);
            }
            // vz, update_vel.cpp:59-69
            T vz_c = vz_o;
            if (ok_z && iz <= nz - 1) {
                const T bav = T(0.25) * (b0 + b0n + b1 + b1n);
                vz_c = vz_o + bav * (cx * (sxz_e - sxz_c) + cz * (szz_n - szz_c) + cy * (syz_c - syz_s));
                stc(vz, p, vz_c);
            }
            // ghost velocities for radius 1 (update_vel.cpp:75-246), same
            // index ranges as the reference loops
#define PUT(f, arr, i, v)                      \
    {                                          \
        const T v_ = (v);                      \
        (*at(arr, i)) = v_;                           \
        if (sv) (*at(g.st[f], i)) = v_;               \
    }
            if (g.top_eff && iz == 0 && ix <= nx - 1) PUT(2, vz, p - S2, -vz_c)
            if (g.bottom && iz == nz - 1 && ix <= nx - 1) PUT(2, vz, p + S2, -vz_c)
            if (ix == 0) {
                if (iz <= nz - 1) PUT(2, vz, p - 1, -vz_c)
                if (iy <= ny - 1) PUT(1, vy, p - 1, -vy_c)
            }
            if (ix == nx - 1) {
                if (iz <= nz - 1) PUT(2, vz, p + 1, -vz_c)
                if (iy <= ny - 1) PUT(1, vy, p + 1, -vy_c)
            }
            if (ix <= nx - 1) {
                if (iy == 0 && iz <= g.zy) PUT(1, vy, p - P, -vy_c)
                if (iy == ny - 1) {
                    // the x mirrors of row iy = ny copy that halo row before
                    // the y mirror below overwrites it (reference loop order)
                    if (ix == 0) PUT(1, vy, p + P - 1, -(*at(vy, p + P)))
                    if (ix == nx - 1) PUT(1, vy, p + P + 1, -(*at(vy, p + P)))
                    if (iz <= g.zy) PUT(1, vy, p + P, -vy_c)
                }
            }
#undef PUT
            if (sv) {
                // the entries this thread owns, updated or not
                stc(g.st[0], p, vx_c);
                if (ok_y) stc(g.st[1], p, vy_c);
                if (ix <= nx - 1 && iz <= nz - 1) stc(g.st[2], p, vz_c);
            }
            if (EN) {
                const double m2 = (iz == 0 || iz == nz) ? 0.5 : 1.0;
                const double m1 = (iy == 0 || iy == ny) ? 0.5 : 1.0;
                const double m0 = (ix == 0 || ix == nx) ? 0.5 : 1.0;
                // density for vx: 1/buoy (init_del_par_field.cpp:231)
                acc += m2 * m1 * m0 * (1.0 / (double)b0) * (double)vx_c * (double)vx_o;
                if (src && ix == src[0] && iy == src[1] && iz == src[2]) *src_old = vx_o;
                if (ok_y) {
                    // density for vy: 1/ReshapeField(buoy,(1,1,0)) (field.cpp:465-472)
                    const double rb = ((((((((double)b0 + (double)b0) + (double)bP) + (double)bP) + (double)b1) + (double)b1)
                                        + (double)bP1) + (double)bP1) / 8;
                    acc += m2 * 1.0 * 1.0 * (1.0 / rb) * (double)vy_c * (double)vy_o;
                }
                if (ix <= nx - 1 && iz <= nz - 1) {
                    const double rb = ((((((((double)b0 + (double)b0n) + (double)b0) + (double)b0n) + (double)b1)
                                         + (double)b1n) + (double)b1) + (double)b1n) / 8;
                    acc += 1.0 * m1 * 1.0 * (1.0 / rb) * (double)vz_c * (double)vz_o;
                }
            }
            sxz_m = sxz_c;
            syz_m = syz_c;
            szz_c = szz_n;
            b0 = b0n;
            b1 = b1n;
        }
    }
    if (EN) {
        const double s = block_sum<BYv>(acc, lds);
        if (threadIdx.x == 0 && threadIdx.y == 0) part[tile] = s;
    }
}

template <typename T, bool EN, int BYv, typename IT>
__global__ __launch_bounds__(BX * BYv) void k_vel(GridDev<T> g, double *__restrict__ part,
                                                  const int *__restrict__ src, T *__restrict__ src_old)
{
    __shared__ double lds[BYv];
    vel_tile<T, EN, BYv, IT>(g, sweep_tile(blockIdx.x, g.nblk, g.gx), part, src, src_old, lds);
}

// compliance of the normal block at one point (energy weights)
template <typename T>
__device__ __forceinline__ void compliance(const GridDev<T> &g, T c11, T c22, T c33, T c23, T c13, T c12,
                                           double &s11, double &s22, double &s33, double &s23, double &s13,
                                           double &s12)
{
    if (g.iso) {
        // init_del_par_field.cpp:259-278
        const double a = c11, b = c13;
        double base = a * b * b * 3 - b * b * b * 2 - a * a * a;
        base = 1 / base;
        s11 = s22 = s33 = (b * b - a * a) * base;
        s23 = s13 = s12 = (b * a - b * b) * base;
    } else {
        const double m11 = c11, m22 = c22, m33 = c33, m23 = c23, m13 = c13, m12 = c12;
        const double a11 = m22 * m33 - m23 * m23, a22 = m11 * m33 - m13 * m13, a33 = m11 * m22 - m12 * m12;
        const double a23 = m12 * m13 - m11 * m23, a13 = m12 * m23 - m13 * m22, a12 = m13 * m23 - m12 * m33;
        const double det = m11 * a11 + m12 * a12 + m13 * a13;
        s11 = a11 / det; s22 = a22 / det; s33 = a33 / det;
        s23 = a23 / det; s13 = a13 / det; s12 = a12 / det;
    }
}

// ---------------------------------------------------------------------------
// UpdateStress (update_stress.cpp:4-311), free surface and its radius-1 ghost
// stresses fused.  EN: strain part of ComputeEnergy (compute_energy.cpp:40-49)
// from the pre-update stresses.
template <typename T, bool EN, int BYv, typename IT>
__device__ __forceinline__ void stress_tile(const GridDev<T> &g, const int tile, double *__restrict__ part, double *lds)
{
    const int bx = tile % g.gx;
    const int rest = tile / g.gx;
    const int by = rest % g.gy;
    const int bz = rest / g.gy;
    const int ix = bx * BX + threadIdx.x;
    const int iy = by * BYv + threadIdx.y;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const int z0 = g.zb + bz * g.zc;
    const int z1 = min(z0 + g.zc, g.ze);
    double acc = 0.0;
    if (ix <= nx && iy <= ny) {
        const IT S2 = (IT)g.S2, P = (IT)g.P;
        IT p = (IT)g.o + (IT)z0 * S2 + (IT)iy * P + (IT)ix;
        const T *__restrict__ vx = g.f[0];
        const T *__restrict__ vy = g.f[1];
        const T *__restrict__ vz = g.f[2];
        T *__restrict__ sxx = g.f[3];
        T *__restrict__ syy = g.f[4];
        T *__restrict__ szz = g.f[5];
        T *__restrict__ syz = g.f[6];
        T *__restrict__ sxz = g.f[7];
        T *__restrict__ sxy = g.f[8];
        const T *__restrict__ C11 = g.c[0];
        const T *__restrict__ C22 = g.c[1];
        const T *__restrict__ C33 = g.c[2];
        const T *__restrict__ C23 = g.c[3];
        const T *__restrict__ C13 = g.c[4];
        const T *__restrict__ C12 = g.c[5];
        const T *__restrict__ C44 = g.c[6];
        const T *__restrict__ C55 = g.c[7];
        const T *__restrict__ C66 = g.c[8];
        const bool ok_n = ix <= nx - 1;
        const bool ok_yz = ix <= nx - 1 && iy <= ny - 1;
        const bool ok_xy = iy <= ny - 1;
        const T cx = g.cx, cy = g.cy, cz = g.cz;
        const bool sv = g.st[0] != nullptr;
        T vz_m = (*at(vz, p - S2));
        T vx_c = (*at(vx, p));
        T vy_c = (*at(vy, p));
        for (int iz = z0; iz < z1; ++iz, p += S2) {
            // Every operand of the plane is loaded before any store (all of
            // them lie inside the halo-padded arrays, parameters included,
            // for ix <= nx, iy <= ny, iz <= nz), so the plane's loads are in
            // flight together instead of in four groups behind the stores.
            const T vx_n = (*at(vx, p + S2));
            const T vy_n = (*at(vy, p + S2));
            const T vz_c = (*at(vz, p));
            const T sxx_o = ldc(sxx, p), syy_o = ldc(syy, p), szz_o = ldc(szz, p);
            const T c11 = ldc(C11, p), c22 = ldc(C22, p), c33 = ldc(C33, p), c23 = ldc(C23, p), c13 = ldc(C13, p),
                    c12 = ldc(C12, p);
            const T vx_e = (*at(vx, p + 1)), vy_s = (*at(vy, p - P));
            const T syz_o = ldc(syz, p), c44 = ldc(C44, p), vz_n = (*at(vz, p + P));
            const T sxz_o = ldc(sxz, p), c55 = ldc(C55, p), vz_w = (*at(vz, p - 1));
            const T sxy_o = ldc(sxy, p), c66 = ldc(C66, p), vx_nn = (*at(vx, p + P)), vy_w = (*at(vy, p - 1));
            if (ok_n) {
                const T dvx = vx_e - vx_c;
                const T dvy = vy_c - vy_s;
                if (g.free_top && iz == 0) {
                    // update_stress.cpp:37-65
                    const T a = sxx_o + ((c11 - c13 * c13 / c33) * cx * dvx + (c12 - c13 * c23 / c33) * cy * dvy);
                    const T bq = syy_o + ((c12 - c23 * c13 / c33) * cx * dvx + (c22 - c23 * c23 / c33) * cy * dvy);
                    stc(sxx, p, a);
                    stc(syy, p, bq);
                    if (sv) { stc(g.st[3], p, a); stc(g.st[4], p, bq); stc(g.st[5], p, szz_o); }
                } else {
                    // update_stress.cpp:69-97
                    const T dvz = vz_c - vz_m;
                    const T a = sxx_o + (c11 * cx * dvx + c12 * cy * dvy + c13 * cz * dvz);
                    const T bq = syy_o + (c12 * cx * dvx + c22 * cy * dvy + c23 * cz * dvz);
                    const T cc = szz_o + (c13 * cx * dvx + c23 * cy * dvy + c33 * cz * dvz);
                    stc(sxx, p, a);
                    stc(syy, p, bq);
                    stc(szz, p, cc);
                    if (sv) { stc(g.st[3], p, a); stc(g.st[4], p, bq); stc(g.st[5], p, cc); }
                }
                if (EN) {
                    double s11, s22, s33, s23, s13, s12;
                    compliance(g, c11, c22, c33, c23, c13, c12, s11, s22, s33, s23, s13, s12);
                    const double m = ((iz == 0 || iz == nz) ? 0.5 : 1.0) * ((iy == 0 || iy == ny) ? 0.5 : 1.0);
                    const double a = sxx_o, bb = syy_o, c = szz_o;
                    acc += m * s11 * a * a + m * s22 * bb * bb + m * s33 * c * c
                         + 2 * (m * s23 * bb * c) + 2 * (m * s13 * a * c) + 2 * (m * s12 * a * bb);
                }
            }
            // syz, update_stress.cpp:103-109
            if (ok_yz && iz <= nz - 1) {
                const T syz_n = syz_o + T(0.5) * c44 * (cz * (vy_n - vy_c) + cy * (vz_n - vz_c));
                stc(syz, p, syz_n);
                if (g.free_top && iz == 0) (*at(syz, p - S2)) = -syz_n;  // update_stress.cpp:264-267
                if (sv) {
                    stc(g.st[6], p, syz_n);
                    if (g.free_top && iz == 0) (*at(g.st[6], p - S2)) = -syz_n;
                }
                if (EN) acc += 2 * ((1.0 / (double)c44) * (double)syz_o * (double)syz_o);
            }
            // sxz, update_stress.cpp:115-121
            if (iz <= nz - 1) {
                const T sxz_n = sxz_o + T(0.5) * c55 * (cz * (vx_n - vx_c) + cx * (vz_c - vz_w));
                stc(sxz, p, sxz_n);
                if (g.free_top && iz == 0 && ix >= 1 && ix <= nx - 1 && iy >= 1 && iy <= ny - 1) {
                    (*at(sxz, p - S2)) = -sxz_n;  // update_stress.cpp:274-277
                    if (sv) (*at(g.st[7], p - S2)) = -sxz_n;
                }
                if (sv) stc(g.st[7], p, sxz_n);
                if (EN) {
                    const double m = ((iy == 0 || iy == ny) ? 0.5 : 1.0) * ((ix == 0 || ix == nx) ? 0.5 : 1.0);
                    acc += 2 * (m * (1.0 / (double)c55) * (double)sxz_o * (double)sxz_o);
                }
            }
            // sxy, update_stress.cpp:127-133
            if (ok_xy) {
                const T sxy_n = sxy_o + T(0.5) * c66 * (cy * (vx_nn - vx_c) + cx * (vy_c - vy_w));
                stc(sxy, p, sxy_n);
                if (sv) stc(g.st[8], p, sxy_n);
                if (EN) {
                    const double m = ((iz == 0 || iz == nz) ? 0.5 : 1.0) * ((ix == 0 || ix == nx) ? 0.5 : 1.0);
                    acc += 2 * (m * (1.0 / (double)c66) * (double)sxy_o * (double)sxy_o);
                }
            }
            vz_m = vz_c;
            vx_c = vx_n;
            vy_c = vy_n;
        }
    }
    if (EN) {
        const double s = block_sum<BYv>(acc, lds);
        if (threadIdx.x == 0 && threadIdx.y == 0) part[tile] = s;
    }
}

template <typename T, bool EN, int BYv, typename IT>
__global__ __launch_bounds__(BX * BYv) void k_stress(GridDev<T> g, double *__restrict__ part)
{
    __shared__ double lds[BYv];
    stress_tile<T, EN, BYv, IT>(g, sweep_tile(blockIdx.x, g.nblk, g.gx), part, lds);
}

// Batched stencil launches: the workgroups of many grids (of many models) in
// one launch; start[m] is the first workgroup of grid m (prefix sums over n
// grids, ascending).  Tiles are taken in XCD order over the whole launch, so
// the tiles of one grid stay together on an XCD.
__device__ __forceinline__ int batch_find(const int *__restrict__ start, int n, int b)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (start[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// A launch of fewer workgroups than tiles is persistent: workgroup w takes
// the tiles w, w + gridDim.x, ... of the XCD-ordered list (gridDim.x a
// multiple of 8 keeps each workgroup within its XCD's range of tiles).
template <typename T, int BYv, typename IT>
__global__ __launch_bounds__(BX * BYv) void k_vel_batch(const GridDev<T> *__restrict__ gs, const int *__restrict__ start, int n,
                                                        int total)
{
    __shared__ double lds[BYv];
    for (int v = blockIdx.x; v < total; v += gridDim.x) {
        const int b = xcd_tile(v, total);
        const int m = batch_find(start, n, b);
        vel_tile<T, false, BYv, IT>(gs[m], b - start[m], nullptr, nullptr, nullptr, lds);
    }
}

template <typename T, int BYv, typename IT>
__global__ __launch_bounds__(BX * BYv) void k_stress_batch(const GridDev<T> *__restrict__ gs, const int *__restrict__ start, int n,
                                                           int total)
{
    __shared__ double lds[BYv];
    for (int v = blockIdx.x; v < total; v += gridDim.x) {
        const int b = xcd_tile(v, total);
        const int m = batch_find(start, n, b);
        stress_tile<T, false, BYv, IT>(gs[m], b - start[m], nullptr, lds);
    }
}

// ---------------------------------------------------------------------------
// Interface kernels.  Fine grid F (upper, z index nz1 on the interface plane)
// and coarse grid C (lower, z index 0 on the plane); refinement ratio r.
template <typename T>
struct IfaceDev {
    GridDev<T> F, C;
    int r;
    int nx2, ny2, nz1;
    double k1x, k1y, k1z, k2x, k2y, k2z;  // coefficients as the stencils use them
};

__device__ __forceinline__ double hatw(int a, int r)
{
    return (double)(r - (a < 0 ? -a : a)) / (double)r;
}

template <typename T>
__device__ __forceinline__ double ld(const T *a, long i) { return (double)a[i]; }

#define FIDX(G, i, j, k) ((G).o + (long)(k) * (G).S2 + (long)(j) * (G).P + (i))

// vz-szz pair: right-hand side of the tridiagonal system of each coarse
// x-line (derived ratio-r form of update_ghost_vz.cpp:152-203).
// B layout [J-1][I].
template <typename T>
__device__ __forceinline__ void gvz_rhs_body(const IfaceDev<T> &a, const int t, double *__restrict__ B)
{
    const int nx2 = a.nx2, ny2 = a.ny2, r = a.r, nz1 = a.nz1;
    if (t >= nx2 * (ny2 - 1)) return;
    const int I = t % nx2;
    const int J = t / nx2 + 1;
    const GridDev<T> &F = a.F;
    const GridDev<T> &C = a.C;
    const T *vx1 = F.f[0], *vy1 = F.f[1], *vz1 = F.f[2], *szz1 = F.f[5];
    const T *vx2 = C.f[0], *vy2 = C.f[1], *vz2 = C.f[2], *szz2 = C.f[5];
    const T *c33_1 = F.c[2], *c23_1 = F.c[3], *c13_1 = F.c[4];
    const T *c33_2 = C.c[2], *c23_2 = C.c[3], *c13_2 = C.c[4];
    double fs = 0;
    for (int bo = -(r - 1); bo <= r - 1; bo++) {
        const int j = r * J + bo;
        const double wj = (double)(r - (bo < 0 ? -bo : bo));
        const int Jl = (bo >= 0) ? J : J - 1;
        const int bb = j - r * Jl;
        const double pv = hatw(bb, r) * ld(vz2, FIDX(C, I, Jl, 0)) + hatw(bb - r, r) * ld(vz2, FIDX(C, I, Jl + 1, 0));
        for (int ao = 0; ao < r; ao++) {
            const int i = r * I + ao;
            const long q = FIDX(F, i, j, nz1);
            const double cc = ld(c33_1, q);
            const double kf = ld(szz1, q) + ld(c13_1, q) * a.k1x * (ld(vx1, q + 1) - ld(vx1, q))
                            + ld(c23_1, q) * a.k1y * (ld(vy1, q) - ld(vy1, q - F.P))
                            + cc * a.k1z * (pv - 2 * ld(vz1, q - F.S2));
            fs += wj * kf;
        }
    }
    const long qc = FIDX(C, I, J, 0);
    const double kc = ld(szz2, qc) + ld(c13_2, qc) * a.k2x * (ld(vx2, qc + 1) - ld(vx2, qc))
                    + ld(c23_2, qc) * a.k2y * (ld(vy2, qc) - ld(vy2, qc - C.P))
                    + ld(c33_2, qc) * a.k2z * ld(vz2, qc);
    B[(long)(J - 1) * nx2 + I] = (double)(r * r * r) * kc - fs;
}

template <typename T>
__global__ __launch_bounds__(256) void k_gvz_rhs(IfaceDev<T> a, double *__restrict__ B)
{
    gvz_rhs_body(a, blockIdx.x * blockDim.x + threadIdx.x, B);
}

// Prefactored Thomas solve, one thread per coarse x-line (update_ghost_vz.cpp:
// 211-220 with the time-invariant elimination factors computed once).
// M: multipliers, Dp: eliminated diagonal (Rp = 1/Dp), U: super-diagonal, all [q][I].
template <typename T>
__device__ __forceinline__ void gvz_solve_body(const IfaceDev<T> &a, const int I, double *__restrict__ B,
                                               const double *__restrict__ M, const double *__restrict__ Rp,
                                               const double *__restrict__ U)
{
    const int nx2 = a.nx2, n = a.ny2 - 1;
    if (I >= nx2) return;
    // Each line is one long dependent chain; its loads come in chunks of CH
    // rows, the next chunk issued before the current one is consumed, so the
    // memory latency overlaps the arithmetic (software pipelining).
    constexpr int CH = 16;
    double prev = 0;
    double bq[CH], mq[CH], bn[CH], mn[CH];
#pragma unroll
    for (int t = 0; t < CH; t++)
        if (t < n) {
            bq[t] = B[(long)t * nx2 + I];
            mq[t] = M[(long)t * nx2 + I];
        }
    for (int q0 = 0; q0 < n; q0 += CH) {
        const int q1 = q0 + CH;
#pragma unroll
        for (int t = 0; t < CH; t++)
            if (q1 + t < n) {
                bn[t] = B[(long)(q1 + t) * nx2 + I];
                mn[t] = M[(long)(q1 + t) * nx2 + I];
            }
#pragma unroll
        for (int t = 0; t < CH; t++)
            if (q0 + t < n) {
                const double v = (q0 + t == 0) ? bq[t] : bq[t] - mq[t] * prev;
                B[(long)(q0 + t) * nx2 + I] = v;
                prev = v;
            }
#pragma unroll
        for (int t = 0; t < CH; t++) {
            bq[t] = bn[t];
            mq[t] = mn[t];
        }
    }
    T *vz2 = a.C.f[2];
    double x = prev * Rp[(long)(n - 1) * nx2 + I];
    T *vz2s = a.C.st[2];
    vz2[FIDX(a.C, I, n, -1)] = (T)x;
    if (vz2s) vz2s[FIDX(a.C, I, n, -1)] = (T)x;
    // back substitution, rows n-2 .. 0, same pipelining (B rows were written
    // by this thread above, so they are read back after its own stores)
    double uq[CH], dq[CH], un[CH], dn[CH];
#pragma unroll
    for (int t = 0; t < CH; t++)
        if (n - 2 - t >= 0) {
            const long r = (long)(n - 2 - t) * nx2 + I;
            bq[t] = B[r];
            uq[t] = U[r];
            dq[t] = Rp[r];
        }
    for (int q1 = n - 2; q1 >= 0; q1 -= CH) {
        const int q2 = q1 - CH;
#pragma unroll
        for (int t = 0; t < CH; t++)
            if (q2 - t >= 0) {
                const long r = (long)(q2 - t) * nx2 + I;
                bn[t] = B[r];
                un[t] = U[r];
                dn[t] = Rp[r];
            }
#pragma unroll
        for (int t = 0; t < CH; t++)
            if (q1 - t >= 0) {
                x = (bq[t] - uq[t] * x) * dq[t];
                vz2[FIDX(a.C, I, q1 - t + 1, -1)] = (T)x;
                if (vz2s) vz2s[FIDX(a.C, I, q1 - t + 1, -1)] = (T)x;
            }
#pragma unroll
        for (int t = 0; t < CH; t++) {
            bq[t] = bn[t];
            uq[t] = un[t];
            dq[t] = dn[t];
        }
    }
}

template <typename T>
__global__ __launch_bounds__(64) void k_gvz_solve(IfaceDev<T> a, double *__restrict__ B, const double *__restrict__ M,
                                                  const double *__restrict__ Rp, const double *__restrict__ U)
{
    gvz_solve_body(a, blockIdx.x * blockDim.x + threadIdx.x, B, M, Rp, U);
}

// The same solve with one wavefront per coarse x-line I (the form used
// whenever the line fits in LDS, 32 (ny2-1) bytes <= VZ_WAVE_LDS): the 64
// lanes stage the line's right-hand side and elimination factors in LDS, lane
// 0 runs the two sweeps of the recurrence there -- operation for operation as
// gvz_solve_body -- and the lanes write the ghosts vz2(I, J, -1) back.  The
// line's chain then waits on LDS instead of HBM latencies.
constexpr int VZ_WAVE_LDS = 65536;
template <typename T>
__device__ __forceinline__ void gvz_solve_wave(const IfaceDev<T> &a, const int I, const double *__restrict__ B,
                                               const double *__restrict__ M, const double *__restrict__ Rp,
                                               const double *__restrict__ U, double *sh)
{
    const int nx2 = a.nx2, n = a.ny2 - 1;
    double *sb = sh, *sm = sh + n, *su = sh + 2 * n, *sr = sh + 3 * n;
    const int lane = threadIdx.x;
    for (int q = lane; q < n; q += 64) {
        const long r = (long)q * nx2 + I;
        sb[q] = B[r];
        sm[q] = M[r];
        su[q] = U[r];
        sr[q] = Rp[r];
    }
    __syncthreads();
    if (lane == 0) {
        double prev = sb[0];
        for (int q = 1; q < n; q++) {
            prev = sb[q] - sm[q] * prev;
            sb[q] = prev;
        }
        double x = prev * sr[n - 1];
        sb[n - 1] = x;
        for (int q = n - 2; q >= 0; q--) {
            x = (sb[q] - su[q] * x) * sr[q];
            sb[q] = x;
        }
    }
    __syncthreads();
    T *vz2 = a.C.f[2];
    T *vz2s = a.C.st[2];
    for (int q = lane; q < n; q += 64) {
        const T v = (T)sb[q];
        vz2[FIDX(a.C, I, q + 1, -1)] = v;
        if (vz2s) vz2s[FIDX(a.C, I, q + 1, -1)] = v;
    }
}

template <typename T>
__global__ __launch_bounds__(64) void k_gvz_solve_wave(IfaceDev<T> a, const double *__restrict__ B,
                                                       const double *__restrict__ M, const double *__restrict__ Rp,
                                                       const double *__restrict__ U)
{
    extern __shared__ double vz_sh[];
    gvz_solve_wave(a, (int)blockIdx.x, B, M, Rp, U, vz_sh);
}

// Fine ghosts vz1(., ., nz1): injection in x, hat in y (update_ghost_vz.cpp:
// 231-248 generalised).  One thread per fine (i, j) with j in [1, ny1-1].
template <typename T>
__device__ __forceinline__ void gvz_fine_body(const IfaceDev<T> &a, const int t)
{
    const int r = a.r, nx1 = r * a.nx2, ny1 = r * a.ny2;
    if (t >= nx1 * (ny1 - 1)) return;
    const int i = t % nx1;
    const int j = t / nx1 + 1;
    const int I = i / r, J = j / r, bo = j - r * J;
    const T *vz2 = a.C.f[2];
    T *vz1 = a.F.f[2];
    const double t0 = ld(vz2, FIDX(a.C, I, J, 0)) + ld(vz2, FIDX(a.C, I, J, -1));
    double tmp;
    if (bo == 0) {
        tmp = t0;
    } else {
        const double t1 = ld(vz2, FIDX(a.C, I, J + 1, 0)) + ld(vz2, FIDX(a.C, I, J + 1, -1));
        tmp = hatw(bo, r) * t0 + hatw(bo - r, r) * t1;
    }
    const long q = FIDX(a.F, i, j, a.nz1);
    const T v = (T)(-ld(vz1, q - a.F.S2) + tmp);
    vz1[q] = v;
    if (a.F.st[2]) a.F.st[2][q] = v;
}

template <typename T>
__global__ __launch_bounds__(256) void k_gvz_fine(IfaceDev<T> a)
{
    gvz_fine_body(a, blockIdx.x * blockDim.x + threadIdx.x);
}

// vx-sxz pair: right-hand side on interior coarse nodes (derived ratio-r form
// of update_ghost_sxz.cpp:139-226) and per-workgroup sums of |B|
// (AvgAbs, array.cpp:188-197).  B layout [J-1][I-1].
template <typename T>
__device__ __forceinline__ void gsxz_rhs_body(const IfaceDev<T> &a, const int lb, double *__restrict__ B,
                                              double *__restrict__ abs_part, double *lds)
{
    const int t = lb * 256 + threadIdx.x;
    const int nx2 = a.nx2, ny2 = a.ny2, r = a.r, nz1 = a.nz1;
    const int n0 = nx2 - 1, n1 = ny2 - 1;
    double v = 0;
    if (t < n0 * n1) {
        const int I = t % n0 + 1;
        const int J = t / n0 + 1;
        const GridDev<T> &F = a.F;
        const GridDev<T> &C = a.C;
        const T *vx1 = F.f[0], *sxx1 = F.f[3], *sxz1 = F.f[7], *sxy1 = F.f[8], *b1 = F.b;
        const T *vx2 = C.f[0], *sxx2 = C.f[3], *sxz2 = C.f[7], *sxy2 = C.f[8], *b2 = C.b;
        double fs = 0;
        for (int bo = -(r - 1); bo <= r - 1; bo++)
            for (int ao = -(r - 1); ao <= r - 1; ao++) {
                const int i = r * I + ao, j = r * J + bo;
                const double w = (double)((r - (ao < 0 ? -ao : ao)) * (r - (bo < 0 ? -bo : bo)));
                const long q = FIDX(F, i, j, nz1);
                const double bb = ld(b1, q);
                double S = 0;
                for (int dJ = -1; dJ <= 1; dJ++)
                    for (int dI = -1; dI <= 1; dI++) {
                        const int Ic = I + dI, Jc = J + dJ;
                        const int ai = i - r * Ic, bj = j - r * Jc;
                        if (ai >= r || ai <= -r || bj >= r || bj <= -r) continue;
                        if (Ic < 1 || Ic > nx2 - 1 || Jc < 1 || Jc > ny2 - 1) continue;
                        const double pw = hatw(ai, r) * hatw(bj, r);
                        S += pw * ld(sxz2, FIDX(C, Ic, Jc, 0));
                    }
                const double kf = ld(vx1, q) + bb * (a.k1x * (ld(sxx1, q) - ld(sxx1, q - 1))
                                                   + a.k1y * (ld(sxy1, q) - ld(sxy1, q - F.P))
                                                   + a.k1z * (S - 2 * ld(sxz1, q - F.S2)));
                fs += w * kf;
            }
        const long qc = FIDX(C, I, J, 0);
        const double kc = ld(vx2, qc) + ld(b2, qc) * (a.k2x * (ld(sxx2, qc) - ld(sxx2, qc - 1))
                                                     + a.k2y * (ld(sxy2, qc) - ld(sxy2, qc - C.P))
                                                     + a.k2z * ld(sxz2, qc));
        const double r2 = (double)(r * r);
        v = (r2 * r2) * kc - fs;
        B[t] = v;
    }
    // 1-D block of 256: reuse the 2-D helper with threadIdx.y == 0
    double s = v < 0 ? -v : v;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) abs_part[lb] = ((lds[0] + lds[1]) + lds[2]) + lds[3];
}

template <typename T>
__global__ __launch_bounds__(256) void k_gsxz_rhs(IfaceDev<T> a, double *__restrict__ B, double *__restrict__ abs_part)
{
    __shared__ double lds[4];
    gsxz_rhs_body(a, blockIdx.x, B, abs_part, lds);
}

struct JacobiArgs {
    const double *A;      // 9 bands, band k at A + k*nn, [J-1][I-1]
    const double *B;
    double *X0, *X1;      // iterates (nx2+1)(ny2+1), zero border; X0 holds the warm start
    const double *abs_part;
    int nabs;
    double *part;         // 2 parities x 2 sums x gridDim.x
    long long *stats;     // total, last, capped
    double tol_cap, tol_fac;
    double floor;         // rounding-floor factor of the stagnation exit
    int cap;
};

// Sum of v[0..n-1] by one wavefront in a fixed order; all lanes get it.
__device__ __forceinline__ double wave_sum_fixed(const double *v, int n)
{
    const int lane = threadIdx.x & 63;
    double s = 0;
    for (int i = lane; i < n; i += 64) s += v[i];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    return s;
}

// Loads and stores of the data the workgroups of a grid-barrier kernel pass
// to each other (iterates, norm partials): device-scope relaxed atomics,
// which gfx950 performs coherently across the XCDs' L2 caches without a
// fence.  The barrier (grid_barrier) then needs no L2 writeback/invalidate,
// which would also flush the lines of the stencil kernels running beside
// the solve: measured, a streaming kernel beside a barrier kernel with
// device-scope fences ran at 0.73x of its speed alone, with workgroup-scope
// fences at 0.92x.
__device__ __forceinline__ double ld_co(const double *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_co(double *p, double v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// wave_sum_fixed over values other workgroups of the same launch wrote
__device__ __forceinline__ double wave_sum_co(const double *v, int n)
{
    const int lane = threadIdx.x & 63;
    double s = 0;
    for (int i = lane; i < n; i += 64) s += ld_co(v + i);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    return s;
}

constexpr int JT = 1024;   // threads per Jacobi workgroup
constexpr int JPPT = 2;    // unknowns owned per thread
constexpr int MAXI = 4;    // interfaces solved by one launch

// One interface of a batched Jacobi solve.
template <typename T>
struct JacItf {
    const double *A;        // 9 bands, band k at A + k*nn, [J-1][I-1]
    const double *B;
    double *X0, *X1;        // iterates (nx2+1)(ny2+1), zero border; X0 holds the warm start
    const double *abs_part; // per-workgroup sums of |B|
    int nabs;
    double *part;           // 2 parities x 2 sums x nblk
    long long *stats;       // total, last, capped
    T *sxz1, *sxz2;
    T *sxz1s, *sxz2s;       // optional second destinations of the ghost writes
    long o1, P1, S21, o2, P2, S22;
    int nx2, ny2, nz1, r;
    int blk0, nblk;         // workgroups [blk0, blk0+nblk) work on this interface
    double rho2;            // square of the Jacobi spectral bound (Chebyshev)
};

template <typename T>
struct JacBatch {
    JacItf<T> it[MAXI];
    int n;
    double tol_cap, tol_fac;  // tol = min(tol_cap, tol_fac * AvgAbs(B))
    double floor;             // rounding-floor factor of the stagnation exit
    double rel_tol;           // also stop when the update norm <= rel_tol * ||X|| (0: off)
    int cap;
    int cheby;                // 1: Chebyshev acceleration of the Jacobi sweeps
    int solve_only;           // 1: leave the solution in X0 only (no ghost scatter)
    unsigned *bar;            // grid barrier state: arrivals, exits, timeout flag
    unsigned long long bar_limit;  // longest wait at one barrier, wall-clock ticks
};

// Grid barrier of the Jacobi kernels, with a bounded wait.  bar[0] counts
// arrivals over the launch (barrier number g is passed once it reaches
// g * nblk), bar[1] counts workgroups that have finished, bar[2] is a timeout
// flag.  A workgroup that waits more than `limit` ticks of the constant-rate
// wall clock raises the flag; every waiting or later-arriving workgroup sees
// it and gets false back, and the kernel ends early with the flag left for
// the host to report, instead of waiting for ever on a workgroup that cannot
// be placed.  limit = 0 raises the flag at the first barrier (the host's test
// of the error path).  What crosses the barrier goes through ld_co / st_co
// (coherent at device scope), so the arrival only has to follow the
// completion of every wave's stores (s_waitcnt before the workgroup
// barrier) and the loads after it are ordered by workgroup-scope fences.
__device__ __forceinline__ bool grid_barrier(unsigned *bar, unsigned nblk, unsigned &gen, unsigned long long limit)
{
    __shared__ int s_ok;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_s_waitcnt(0);  // this wave's coherent stores are complete
    __syncthreads();
    gen++;
    if (threadIdx.x == 0) {
        int ok = 1;
        const unsigned target = gen * nblk;
        __hip_atomic_fetch_add(&bar[0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (limit == 0) __hip_atomic_store(&bar[2], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);  // forced
        const unsigned long long t0 = wall_clock64();
        // poll every ~1024 cycles (s_sleep 16); the timeout flag and the
        // clock every 8th poll (three-grid step 4.178 -> 4.168 ms against
        // polling every ~64 cycles with both read each time)
        unsigned it = 0;
        while (__hip_atomic_load(&bar[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            if ((++it & 7u) == 0) {
                if (__hip_atomic_load(&bar[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) { ok = 0; break; }
                if (wall_clock64() - t0 > limit) {
                    __hip_atomic_store(&bar[2], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    ok = 0;
                    break;
                }
            }
            __builtin_amdgcn_s_sleep(16);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        s_ok = ok;
    }
    __syncthreads();
    return s_ok != 0;
}

// End of a kernel that used grid_barrier: the last workgroup to finish
// (every other one has passed its last barrier by then) zeroes the counters
// for the next launch.  The timeout flag stays until the host clears it.
__device__ __forceinline__ void grid_barrier_done(unsigned *bar, unsigned nblk)
{
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned old = __hip_atomic_fetch_add(&bar[1], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (old == nblk - 1) {
            __hip_atomic_store(&bar[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&bar[1], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// Jacobi2D + Norm2D (update_ghost_sxz.cpp:314-388) for all interfaces of a
// composite model in one cooperative persistent kernel: the iterations of the
// interfaces run side by side.  Each thread owns up to JPPT unknowns for the
// whole solve and keeps their nine matrix bands and right-hand side in
// registers, so a sweep only reads the (L2-resident) previous iterate.  Every
// iteration ends with a grid barrier, after which every workgroup sums the
// per-workgroup norms of every interface in the same order, so all take the
// same decisions.  With cheby = 1 the sweeps are combined by the Chebyshev
// semi-iterative recurrence (same fixed point and exit test).  Then the
// scatter of the coarse ghosts and the fine ghosts (update_ghost_sxz.cpp:
// 247-295, generalised to ratio r).
template <typename T>
__global__ __launch_bounds__(JT) void k_jacobi(JacBatch<T> jb)
{
    unsigned bgen = 0;  // grid barriers passed
    __shared__ double lds[JT / 64];
    __shared__ double lds2[JT / 64];
    __shared__ double s_tol[MAXI], s_n2[MAXI], s_x2[MAXI];
    int me = 0;
    for (int k = 1; k < jb.n; k++)
        if ((int)blockIdx.x >= jb.it[k].blk0) me = k;
    const JacItf<T> &J = jb.it[me];
    const int nx2 = J.nx2, ny2 = J.ny2;
    const int n0 = nx2 - 1, n1 = ny2 - 1;
    const long nn = (long)n0 * n1;
    const long w = nx2 + 1;
    const int lb = blockIdx.x - J.blk0;
    const long gtid = (long)lb * blockDim.x + threadIdx.x;
    const long gstride = (long)J.nblk * blockDim.x;
    const int nw = blockDim.x >> 6;
    if (threadIdx.x < 64)
        for (int k = 0; k < jb.n; k++) {
            const JacItf<T> &K = jb.it[k];
            const double sb = wave_sum_fixed(K.abs_part, K.nabs);
            if (threadIdx.x == 0) {
                const double avg = sb / (double)((long)(K.nx2 - 1) * (K.ny2 - 1));
                const double t2 = avg * jb.tol_fac;
                s_tol[k] = (jb.tol_cap < t2) ? jb.tol_cap : t2;
            }
        }
    __syncthreads();
    double Ar[JPPT][9], Br[JPPT], xprev[JPPT];
    long xi[JPPT];
    bool own[JPPT];
#pragma unroll
    for (int k = 0; k < JPPT; k++) {
        const long q = gtid + k * gstride;
        own[k] = q < nn;
        const long qq = own[k] ? q : 0;
        const long i0 = qq % n0 + 1, i1 = qq / n0 + 1;
        xi[k] = i1 * w + i0;
#pragma unroll
        for (int b = 0; b < 9; b++) Ar[k][b] = J.A[b * nn + qq];
        Br[k] = J.B[qq];
        xprev[k] = J.X0[xi[k]];
    }
    double *Xo = J.X0, *Xn = J.X1;
    // per-interface iteration state, replicated identically in every workgroup
    int it_k[MAXI];
    bool done_k[MAXI];
    double prev_k[MAXI];
#pragma unroll
    for (int k = 0; k < MAXI; k++) { it_k[k] = 0; done_k[k] = (k >= jb.n); prev_k[k] = 0; }
    double omega = 1.0;
    int my_it = 0;
    while (true) {
        bool active = false;
#pragma unroll
        for (int k = 0; k < MAXI; k++)
            if (k == me) { active = !done_k[k]; my_it = it_k[k]; }
        if (active) {
            // Chebyshev weights: omega_1 = 1, omega_2 = 2/(2-rho^2),
            // omega_{k+1} = 1/(1 - rho^2 omega_k / 4)
            if (jb.cheby) omega = (my_it == 0) ? 1.0 : (my_it == 1 ? 2.0 / (2.0 - J.rho2) : 1.0 / (1.0 - 0.25 * J.rho2 * omega));
            double acc = 0, ax = 0;
#pragma unroll
            for (int k = 0; k < JPPT; k++) {
                if (!own[k]) continue;
                const long c = xi[k];
                const double xo = ld_co(Xo + c);
                double xn = (Br[k]
                             - Ar[k][0] * ld_co(Xo + (c - w - 1))
                             - Ar[k][1] * ld_co(Xo + (c - 1))
                             - Ar[k][2] * ld_co(Xo + (c + w - 1))
                             - Ar[k][3] * ld_co(Xo + (c - w))
                             - Ar[k][5] * ld_co(Xo + (c + w))
                             - Ar[k][6] * ld_co(Xo + (c - w + 1))
                             - Ar[k][7] * ld_co(Xo + (c + 1))
                             - Ar[k][8] * ld_co(Xo + (c + w + 1))) / Ar[k][4];
                if (jb.cheby) {
                    xn = omega * (xn - xprev[k]) + xprev[k];
                    xprev[k] = xo;
                }
                const double d = xn - xo;
                acc += d * d;
                ax += xn * xn;
                st_co(Xn + c, xn);
            }
            double s = acc, sx = ax;
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                s += __shfl_xor(s, off, 64);
                sx += __shfl_xor(sx, off, 64);
            }
            if ((threadIdx.x & 63) == 0) {
                lds[threadIdx.x >> 6] = s;
                lds2[threadIdx.x >> 6] = sx;
            }
            __syncthreads();
            if (threadIdx.x == 0) {
                double t = 0, tx = 0;
                for (int q = 0; q < nw; q++) { t += lds[q]; tx += lds2[q]; }
                double *part = J.part + (my_it & 1) * 2 * J.nblk;
                st_co(part + lb, t);
                st_co(part + J.nblk + lb, tx);
            }
            double *tsw = Xo; Xo = Xn; Xn = tsw;
        }
        if (!grid_barrier(jb.bar, gridDim.x, bgen, jb.bar_limit)) {
            grid_barrier_done(jb.bar, gridDim.x);  // timed out: the host reports it
            return;
        }
        // every workgroup evaluates every active interface's exit test from
        // the same partial sums in the same order
        if (threadIdx.x < 64)
            for (int k = 0; k < jb.n; k++) {
                bool dk = false;
                int ik = 0;
#pragma unroll
                for (int kk = 0; kk < MAXI; kk++)
                    if (kk == k) { dk = done_k[kk]; ik = it_k[kk]; }
                if (dk) continue;
                const JacItf<T> &K = jb.it[k];
                const double *part = K.part + (ik & 1) * 2 * K.nblk;
                const double n2 = wave_sum_co(part, K.nblk);
                const double x2 = wave_sum_co(part + K.nblk, K.nblk);
                if (threadIdx.x == 0) { s_n2[k] = n2; s_x2[k] = x2; }
            }
        __syncthreads();
        bool all_done = true;
#pragma unroll
        for (int k = 0; k < MAXI; k++) {
            if (done_k[k]) continue;
            const double norm = sqrt(s_n2[k]);
            const double xnorm = sqrt(s_x2[k]);
            it_k[k]++;
            // the reference's exit, a rounding-floor stagnation exit, and a cap
            if (!(norm > s_tol[k])) done_k[k] = true;
            else if (it_k[k] > 1 && norm >= prev_k[k] && norm <= jb.floor * xnorm) done_k[k] = true;
            else if (it_k[k] >= jb.cap) done_k[k] = true;
            prev_k[k] = norm;
            if (!done_k[k]) all_done = false;
        }
        if (all_done) break;
    }
    int iters = 0;
#pragma unroll
    for (int k = 0; k < MAXI; k++)
        if (k == me) iters = it_k[k];
    // Xo holds the final iterate; keep it in X0 for the next warm start and
    // write the coarse ghosts sxz2(I, J, -1).
#pragma unroll
    for (int k = 0; k < JPPT; k++) {
        if (!own[k]) continue;
        const long c = xi[k];
        const double x = ld_co(Xo + c);
        if (Xo != J.X0) J.X0[c] = x;
        const long i1 = c / w, i0 = c - i1 * w;
        J.sxz2[J.o2 + (long)(-1) * J.S22 + i1 * J.P2 + i0] = (T)x;
        if (J.sxz2s) J.sxz2s[J.o2 + (long)(-1) * J.S22 + i1 * J.P2 + i0] = (T)x;
    }
    // the fine ghosts follow in k_jacobi_fine
    if (lb == 0 && threadIdx.x == 0) {
        J.stats[0] += iters;
        J.stats[1] = iters;
        if (iters >= jb.cap) J.stats[2] = 1;
    }
    grid_barrier_done(jb.bar, gridDim.x);
}

// ---------------------------------------------------------------------------
// Tiled s-step form of the same Jacobi solve.  Workgroup = one TJ x TJ tile of
// unknowns of one interface.  Per grid barrier it loads the current iterate on
// the tile plus a halo of width SJ into LDS and performs SJ Jacobi sweeps
// there, the updated region shrinking by one ring per sweep; the values it
// produces on the tile are those of SJ global Jacobi sweeps, operation for
// operation.  Norm partials are kept per sweep and the tile after each sweep
// is stashed in LDS, so the exit happens at exactly the sweep where the
// reference's test (and the stagnation test) first holds.
constexpr int TJ = 32;                   // tile edge (unknowns)
#ifndef ECSG_SJ
#define ECSG_SJ 6
#endif
constexpr int SJ = ECSG_SJ;               // sweeps per grid barrier
constexpr int EJ = TJ + 2 * SJ;          // LDS edge of tile + halo
constexpr int RJ = TJ + 2 * (SJ - 1);    // edge of the first sweep's region
// Threads per workgroup: JTT, or JTT_FEW for launches of at most
// JTT_FEW_TILES tiles.  With few tiles the solve is the longest path of the
// step and more waves per tile hide its LDS and FP64 latencies (ratio-3
// two-grid model, 9 tiles: 0.354 against 0.382 ms per step); with many tiles
// the extra waves slow the stencils beside the solve (194 tiles: 4.59 against
// 4.53 ms per step).
constexpr int JTT = 256;
constexpr int JTT_FEW = 512;
constexpr int JTT_FEW_TILES = 64;

template <typename T, int NT>
__global__ __launch_bounds__(NT) void k_jacobi_tiled(JacBatch<T> jb)
{
    constexpr int KJ = (RJ * RJ + NT - 1) / NT;  // region points per thread
    unsigned bgen = 0;  // grid barriers passed
    __shared__ double xb[2][EJ * EJ];
    __shared__ double stash[SJ][TJ * TJ];
    __shared__ double red[NT / 64][2 * SJ];
    __shared__ double s_tol[MAXI];
    __shared__ double s_n2[MAXI][SJ], s_x2[MAXI][SJ];
    int me = 0;
    for (int k = 1; k < jb.n; k++)
        if ((int)blockIdx.x >= jb.it[k].blk0) me = k;
    const JacItf<T> &J = jb.it[me];
    const int nx2 = J.nx2, ny2 = J.ny2;
    const int n0 = nx2 - 1, n1 = ny2 - 1;
    const long w = nx2 + 1;
    const int lb = blockIdx.x - J.blk0;
    const int ntx = (n0 + TJ - 1) / TJ;
    const int I0 = 1 + (lb % ntx) * TJ, J0 = 1 + (lb / ntx) * TJ;  // tile origin (unknown indices)
    const int tid = threadIdx.x;
    if (tid < 64)
        for (int k = 0; k < jb.n; k++) {
            const JacItf<T> &K = jb.it[k];
            const double sb = wave_sum_fixed(K.abs_part, K.nabs);
            if (tid == 0) {
                const double avg = sb / (double)((long)(K.nx2 - 1) * (K.ny2 - 1));
                const double t2 = avg * jb.tol_fac;
                s_tol[k] = (jb.tol_cap < t2) ? jb.tol_cap : t2;
            }
        }
    // this thread's points of the first sweep's region, bands and rhs in registers
    const long nn = (long)n0 * n1;
    double Ar[KJ][9], Br[KJ];
    int li[KJ];          // LDS index, -1 when not an unknown
    int sh[KJ];          // ring distance outside the tile (0 inside)
    int ti[KJ];          // tile index, -1 outside the tile
#pragma unroll
    for (int k = 0; k < KJ; k++) {
        const int p = tid + k * NT;
        li[k] = -1; sh[k] = 0; ti[k] = -1;
        if (p < RJ * RJ) {
            const int ri = p % RJ, rj = p / RJ;
            const int I = I0 - (SJ - 1) + ri, Jg = J0 - (SJ - 1) + rj;
            const int di = max(max(I0 - I, I - (I0 + TJ - 1)), 0), dj = max(max(J0 - Jg, Jg - (J0 + TJ - 1)), 0);
            sh[k] = max(di, dj);
            if (I >= 1 && I <= n0 && Jg >= 1 && Jg <= n1) {
                li[k] = (rj + 1) * EJ + (ri + 1);
                if (sh[k] == 0) ti[k] = (Jg - J0) * TJ + (I - I0);
                const long q = (long)(Jg - 1) * n0 + (I - 1);
#pragma unroll
                for (int b = 0; b < 9; b++) Ar[k][b] = J.A[b * nn + q];
                Br[k] = J.B[q];
            }
        }
    }
    double *Xo = J.X0, *Xn = J.X1;
    int it_k[MAXI];
    bool done_k[MAXI];
    double prev_k[MAXI];
    int stop_k[MAXI];    // sweep (1..SJ) of the last batch at which interface k stopped
#pragma unroll
    for (int k = 0; k < MAXI; k++) { it_k[k] = 0; done_k[k] = (k >= jb.n); prev_k[k] = 0; stop_k[k] = SJ; }
    int batch = 0;
    while (true) {
        bool active = false;
#pragma unroll
        for (int k = 0; k < MAXI; k++)
            if (k == me) active = !done_k[k];
        if (active) {
            // load tile + halo of the current iterate (zero outside the unknowns)
            for (int p = tid; p < EJ * EJ; p += NT) {
                const int ei = p % EJ, ej = p / EJ;
                const int I = I0 - SJ + ei, Jg = J0 - SJ + ej;
                double v = 0;
                if (I >= 1 && I <= n0 && Jg >= 1 && Jg <= n1) v = ld_co(Xo + ((long)Jg * w + I));
                xb[0][p] = v;
                xb[1][p] = v;
            }
            __syncthreads();
            double acc[SJ], ax[SJ];
#pragma unroll
            for (int t = 0; t < SJ; t++) { acc[t] = 0; ax[t] = 0; }
#pragma unroll
            for (int t = 1; t <= SJ; t++) {
                const double *xo = xb[(t - 1) & 1];
                double *xn = xb[t & 1];
#pragma unroll
                for (int k = 0; k < KJ; k++) {
                    const int c = li[k];
                    if (c < 0 || sh[k] > SJ - t) continue;
                    const double v = (Br[k]
                                      - Ar[k][0] * xo[c - EJ - 1]
                                      - Ar[k][1] * xo[c - 1]
                                      - Ar[k][2] * xo[c + EJ - 1]
                                      - Ar[k][3] * xo[c - EJ]
                                      - Ar[k][5] * xo[c + EJ]
                                      - Ar[k][6] * xo[c - EJ + 1]
                                      - Ar[k][7] * xo[c + 1]
                                      - Ar[k][8] * xo[c + EJ + 1]) / Ar[k][4];
                    xn[c] = v;
                    if (ti[k] >= 0) {
                        const double d = v - xo[c];
                        acc[t - 1] += d * d;
                        ax[t - 1] += v * v;
                        stash[t - 1][ti[k]] = v;
                    }
                }
                __syncthreads();
            }
            // per-sweep workgroup sums, fixed order
#pragma unroll
            for (int t = 0; t < SJ; t++) {
                double a = acc[t], b = ax[t];
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) {
                    a += __shfl_xor(a, off, 64);
                    b += __shfl_xor(b, off, 64);
                }
                if ((tid & 63) == 0) { red[tid >> 6][2 * t] = a; red[tid >> 6][2 * t + 1] = b; }
            }
            __syncthreads();
            if (tid < 2 * SJ) {
                double v = 0;
                for (int q = 0; q < NT / 64; q++) v += red[q][tid];
                st_co(J.part + (((long)(batch & 1) * J.nblk + lb) * 2 * SJ + tid), v);
            }
            // the tile after SJ sweeps is the next batch's starting iterate
            for (int p = tid; p < TJ * TJ; p += NT) {
                const int I = I0 + p % TJ, Jg = J0 + p / TJ;
                if (I <= n0 && Jg <= n1) st_co(Xn + ((long)Jg * w + I), stash[SJ - 1][p]);
            }
        }
        if (!grid_barrier(jb.bar, gridDim.x, bgen, jb.bar_limit)) {
            grid_barrier_done(jb.bar, gridDim.x);  // timed out: the host reports it
            return;
        }
        // every workgroup replays every active interface's per-sweep exit test;
        // one pass over the workgroups' partials gathers all 2*SJ sums at once
        // (lane-strided sums in increasing workgroup order, then the butterfly)
        if (tid < 64)
            for (int k = 0; k < jb.n; k++) {
                bool dk = false;
#pragma unroll
                for (int kk = 0; kk < MAXI; kk++)
                    if (kk == k) dk = done_k[kk];
                if (dk) continue;
                const JacItf<T> &K = jb.it[k];
                const double *part = K.part + (long)(batch & 1) * K.nblk * 2 * SJ;
                double a[2 * SJ];
#pragma unroll
                for (int t = 0; t < 2 * SJ; t++) a[t] = 0;
#pragma unroll 4
                for (int q = tid; q < K.nblk; q += 64) {
                    const double *pq = part + (long)q * 2 * SJ;
#pragma unroll
                    for (int t = 0; t < 2 * SJ; t++) a[t] += ld_co(pq + t);
                }
#pragma unroll
                for (int t = 0; t < 2 * SJ; t++)
#pragma unroll
                    for (int off = 32; off > 0; off >>= 1) a[t] += __shfl_xor(a[t], off, 64);
                if (tid == 0)
#pragma unroll
                    for (int t = 0; t < SJ; t++) { s_n2[k][t] = a[2 * t]; s_x2[k][t] = a[2 * t + 1]; }
            }
        __syncthreads();
        bool all_done = true;
#pragma unroll
        for (int k = 0; k < MAXI; k++) {
            if (done_k[k]) continue;
            for (int t = 0; t < SJ && !done_k[k]; t++) {
                const double norm = sqrt(s_n2[k][t]);
                const double xnorm = sqrt(s_x2[k][t]);
                it_k[k]++;
                if (!(norm > s_tol[k])) done_k[k] = true;
                else if (it_k[k] > 1 && norm >= prev_k[k] && norm <= jb.floor * xnorm) done_k[k] = true;
                else if (norm <= jb.rel_tol * xnorm) done_k[k] = true;
                else if (it_k[k] >= jb.cap) done_k[k] = true;
                prev_k[k] = norm;
                if (done_k[k]) stop_k[k] = t + 1;
            }
            if (!done_k[k]) all_done = false;
        }
        if (active) { double *tsw = Xo; Xo = Xn; Xn = tsw; }
        batch++;
        if (all_done) break;
    }
    int iters = 0, stop = SJ;
#pragma unroll
    for (int k = 0; k < MAXI; k++)
        if (k == me) { iters = it_k[k]; stop = stop_k[k]; }
    // the tile at the exit sweep: next warm start and coarse ghosts sxz2(I,J,-1)
    for (int p = tid; p < TJ * TJ; p += NT) {
        const int I = I0 + p % TJ, Jg = J0 + p / TJ;
        if (I <= n0 && Jg <= n1) {
            const double x = stash[stop - 1][p];
            J.X0[(long)Jg * w + I] = x;
            if (!jb.solve_only) {
                J.sxz2[J.o2 - J.S22 + (long)Jg * J.P2 + I] = (T)x;
                if (J.sxz2s) J.sxz2s[J.o2 - J.S22 + (long)Jg * J.P2 + I] = (T)x;
            }
        }
    }
    // the fine ghosts follow in k_jacobi_fine
    if (lb == 0 && tid == 0) {
        J.stats[0] += iters;
        J.stats[1] = iters;
        if (iters >= jb.cap) J.stats[2] = 1;
    }
    grid_barrier_done(jb.bar, gridDim.x);
}

// Fine ghosts sxz1(i, j, nz1) of the vx-sxz interfaces from the coarse ghosts
// the Jacobi kernels wrote (update_ghost_sxz.cpp:247-295, ratio-r form): one
// thread per fine interior node (i, j) in [1, nx1-1] x [1, ny1-1] of
// interface blockIdx.y, i fastest.  A launch of its own after the cooperative
// solve, so that it runs at full occupancy.
template <typename T>
__global__ __launch_bounds__(256) void k_jacobi_fine(JacBatch<T> jb)
{
    const JacItf<T> &J = jb.it[blockIdx.y];
    const int r = J.r, nx2 = J.nx2, ny2 = J.ny2, nx1 = r * nx2, ny1 = r * ny2;
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (q >= (nx1 - 1) * (ny1 - 1)) return;
    const int i = q % (nx1 - 1) + 1, jj = q / (nx1 - 1) + 1;
    const int Ia = i / r, Ja = jj / r, ao = i - r * Ia, bo = jj - r * Ja;
    double S = 0;
    for (int dJ = 0; dJ <= 1; dJ++)
        for (int dI = 0; dI <= 1; dI++) {
            const int Ic = Ia + dI, Jc = Ja + dJ;
            if (Ic < 1 || Ic > nx2 - 1 || Jc < 1 || Jc > ny2 - 1) continue;
            const double pw = hatw(ao - r * dI, r) * hatw(bo - r * dJ, r);
            if (pw == 0) continue;
            const long q2 = J.o2 + (long)Jc * J.P2 + Ic;
            S += pw * ((double)J.sxz2[q2 - J.S22] + (double)J.sxz2[q2]);
        }
    const long qf = J.o1 + (long)J.nz1 * J.S21 + (long)jj * J.P1 + i;
    const T v = (T)(-(double)J.sxz1[qf - J.S21] + S);
    J.sxz1[qf] = v;
    if (J.sxz1s) J.sxz1s[qf] = v;
}

// ---------------------------------------------------------------------------
// Direct solve of the vx-sxz system (the reference solves it only by the
// Jacobi sweeps of update_ghost_sxz.cpp:229-236, 314-350; the matrix is the
// one those sweeps use, update_ghost_sxz.cpp:96-133 in ratio-r form).  The
// unknowns of coarse row J (n0 = nx2-1 of them) form block J; the 9 bands
// make the matrix block tridiagonal: diagonal blocks D_J (bands 1, 4, 7),
// blocks L_J coupling row J to row J-1 (bands 0, 3, 6) and U_J coupling it
// to row J+1 (bands 2, 5, 8), each tridiagonal in I.  The host forms the
// Schur complements S_0 = D_0, S_J = D_J - L_J S_{J-1}^-1 U_{J-1} once per
// parameter set and stores W_J = S_J^-1 (dense, column-major: entry (I, k)
// at W + J*n0*n0 + k*n0 + I).  Per solve, one workgroup per interface:
//   y_0 = b_0,             y_J = b_J - L_J (W_{J-1} y_{J-1}),
//   x_{n1-1} = W y_{n1-1}, x_J = W_J (y_J - U_J x_{J+1}),
// block elimination without pivoting between blocks, which the system's
// symmetry and diagonal dominance make stable.
template <typename T>
struct SxzDirect {
    const double *A;     // 9 bands, band k at A + k*nn, [J-1][I-1]
    const double *W;     // inverse Schur complements
    const double *B;     // right-hand side [J-1][I-1]
    double *X0;          // solution out, (nx2+1)(ny2+1), zero border
    double *Y;           // nn entries: the forward sweep's y
    T *sxz2, *sxz2s;     // coarse ghosts sxz2(I, J, -1), optional second destination
    long o2, P2, S22;
    int nx2, ny2;
};

template <typename T>
struct SxzDirectSet {
    SxzDirect<T> it[MAXG - 1];
};

constexpr int JD = 1024;            // threads of the direct-solve workgroup
constexpr int JD_MAXN0 = JD;        // longest coarse row (unknowns per block)

// part[p][I] = sum over k of part p of W(I, k) v[k], W column-major n0 x n0;
// thread (p, I) = (tid / n0, tid % n0), parts of KS consecutive k
__device__ __forceinline__ void dense_mv_parts(const double *__restrict__ W, const double *v, double *part, int n0, int P,
                                               int KS)
{
    const int tid = threadIdx.x;
    const int p = tid / n0, I = tid - p * n0;
    if (p < P) {
        const int k0 = p * KS, k1 = min(n0, k0 + KS);
        double acc = 0;
        for (int k = k0; k < k1; k++) acc += W[(long)k * n0 + I] * v[k];
        part[p * n0 + I] = acc;
    }
}

// the parts of row I summed in order p = 0, 1, ...
__device__ __forceinline__ double dense_mv_sum(const double *part, int n0, int P, int I)
{
    double s = part[I];
    for (int p = 1; p < P; p++) s += part[p * n0 + I];
    return s;
}

template <typename T>
__global__ __launch_bounds__(JD) void k_sxz_direct(SxzDirectSet<T> set)
{
    __shared__ double va[JD_MAXN0], vb[JD_MAXN0], part[JD];
    const SxzDirect<T> &D = set.it[blockIdx.x];
    const int n0 = D.nx2 - 1, n1 = D.ny2 - 1;
    const long nn = (long)n0 * n1, w = D.nx2 + 1, bs = (long)n0 * n0;
    const int P = min(JD / n0, n0), KS = (n0 + P - 1) / P;
    const int tid = threadIdx.x;
    const double *__restrict__ A = D.A;
    const double *__restrict__ B = D.B;
    double *Y = D.Y;
    // forward sweep; vb holds y_{J-1}, va the product W_{J-1} y_{J-1}
    if (tid < n0) {
        const double y = B[tid];
        Y[tid] = y;
        vb[tid] = y;
    }
    for (int J = 1; J < n1; J++) {
        __syncthreads();
        dense_mv_parts(D.W + (long)(J - 1) * bs, vb, part, n0, P, KS);
        __syncthreads();
        if (tid < n0) va[tid] = dense_mv_sum(part, n0, P, tid);
        __syncthreads();
        if (tid < n0) {
            const long q = (long)J * n0 + tid;
            double y = B[q] - A[3 * nn + q] * va[tid];
            if (tid > 0) y -= A[q] * va[tid - 1];
            if (tid + 1 < n0) y -= A[6 * nn + q] * va[tid + 1];
            Y[q] = y;
            vb[tid] = y;
        }
    }
    // backward sweep; vb holds y_J - U_J x_{J+1}, va x_J
    for (int J = n1 - 1; J >= 0; J--) {
        __syncthreads();
        dense_mv_parts(D.W + (long)J * bs, vb, part, n0, P, KS);
        __syncthreads();
        if (tid < n0) {
            const double x = dense_mv_sum(part, n0, P, tid);
            va[tid] = x;
            const long c = (long)(J + 1) * w + tid + 1;
            D.X0[c] = x;
            const long idx = D.o2 - D.S22 + (long)(J + 1) * D.P2 + tid + 1;
            D.sxz2[idx] = (T)x;
            if (D.sxz2s) D.sxz2s[idx] = (T)x;
        }
        __syncthreads();
        if (J > 0 && tid < n0) {
            const long q = (long)(J - 1) * n0 + tid;
            double t = Y[q] - A[5 * nn + q] * va[tid];
            if (tid > 0) t -= A[2 * nn + q] * va[tid - 1];
            if (tid + 1 < n0) t -= A[8 * nn + q] * va[tid + 1];
            vb[tid] = t;
        }
    }
}

// vy-syz pair: pointwise coarse ghost and its fine children (derived ratio-r
// form of update_ghost_syz.cpp:162-238).
template <typename T>
__device__ __forceinline__ void gsyz_body(const IfaceDev<T> &a, const int t)
{
    const int nx2 = a.nx2, ny2 = a.ny2, r = a.r, nz1 = a.nz1;
    if (t >= nx2 * ny2) return;
    const int I = t % nx2, J = t / nx2;
    const GridDev<T> &F = a.F;
    const GridDev<T> &C = a.C;
    const T *syy1 = F.f[4], *sxy1 = F.f[8], *vy1 = F.f[1], *b1 = F.b;
    T *syz1 = F.f[6];
    const T *syy2 = C.f[4], *sxy2 = C.f[8], *vy2 = C.f[1], *b2 = C.b;
    T *syz2 = C.f[6];
    const long qc = FIDX(C, I, J, 0);
    const double c0 = 0.25 * (((ld(b2, qc) + ld(b2, qc + 1)) + ld(b2, qc + C.P)) + ld(b2, qc + C.P + 1));
    const double s2 = ld(syz2, qc);
    const double r2 = (double)(r * r);
    double num = r2 * (ld(vy2, qc) + c0 * (a.k2x * (ld(sxy2, qc + 1) - ld(sxy2, qc))
                                         + a.k2y * (ld(syy2, qc + C.P) - ld(syy2, qc))
                                         + a.k2z * s2));
    double den = r2 * c0 * a.k2z;
    for (int bo = 0; bo < r; bo++)
        for (int ao = 0; ao < r; ao++) {
            const long q = FIDX(F, r * I + ao, r * J + bo, nz1);
            const double c = 0.25 * (((ld(b1, q) + ld(b1, q + 1)) + ld(b1, q + F.P)) + ld(b1, q + F.P + 1));
            num -= ld(vy1, q) + c * (a.k1x * (ld(sxy1, q + 1) - ld(sxy1, q))
                                    + a.k1y * (ld(syy1, q + F.P) - ld(syy1, q))
                                    + a.k1z * (s2 - 2 * ld(syz1, q - F.S2)));
            den += c * a.k1z;
        }
    const T g2 = (T)(num / den);
    syz2[qc - C.S2] = g2;
    if (C.st[6]) C.st[6][qc - C.S2] = g2;
    const double tmp = (double)g2 + s2;
    for (int bo = 0; bo < r; bo++)
        for (int ao = 0; ao < r; ao++) {
            const long q = FIDX(F, r * I + ao, r * J + bo, nz1);
            const T v = (T)(-ld(syz1, q - F.S2) + tmp);
            syz1[q] = v;
            if (F.st[6]) F.st[6][q] = v;
        }
}

template <typename T>
__global__ __launch_bounds__(256) void k_gsyz(IfaceDev<T> a)
{
    gsyz_body(a, blockIdx.x * blockDim.x + threadIdx.x);
}

// ---------------------------------------------------------------------------
// Receivers, source, energy finalisation and the step counter.
template <typename T>
struct PointArgs {
    T *vx[MAXG];
    T *sxx[MAXG], *syy[MAXG], *szz[MAXG];
    T *st[MAXG][4];      // optional second destinations of vx, sxx, syy, szz
    const T *b[MAXG];
    long o[MAXG], P[MAXG], S2[MAXG];
    int n[MAXG][3];
    int nrec;
    const int *rec;      // nrec x 4
    T *traces;           // nrec x max_steps
    int max_steps;
    int *step;           // device step counter
    int src_g, src_i, src_j, src_k, src_type, src_len;
    const float *src_tr;
    const T *src_old;    // pre-update vx at the source point (energy)
    double *fix;         // kinetic-energy correction of the injected source
    int energy;
    int skip_s;          // 1: an explosive source is not added here (the fused pass folded it in)
};

// main.cpp:183-195: receivers sample vx after UpdateVel, then the source is added.
template <typename T>
__device__ __forceinline__ void recsrc_body(const PointArgs<T> &a)
{
    const int step = *a.step;
    if (step < a.max_steps)
        for (int q = threadIdx.x; q < a.nrec; q += blockDim.x) {
            const int g = a.rec[4 * q], i = a.rec[4 * q + 1], jj = a.rec[4 * q + 2], k = a.rec[4 * q + 3];
            a.traces[(long)q * a.max_steps + step] = a.vx[g][a.o[g] + (long)k * a.S2[g] + (long)jj * a.P[g] + i];
        }
    __syncthreads();
    if (threadIdx.x == 0) {
        double fix = 0;
        if (step < a.src_len) {
            const int g = a.src_g;
            const long q = a.o[g] + (long)a.src_k * a.S2[g] + (long)a.src_j * a.P[g] + a.src_i;
            if (a.src_type == 0) {
                const float add = 1000 * a.src_tr[step];
                a.vx[g][q] = a.vx[g][q] + (T)add;
                if (a.st[g][0]) a.st[g][0][q] = a.vx[g][q];
                if (a.energy) {
                    const int *n = a.n[g];
                    const double m = ((a.src_k == 0 || a.src_k == n[2]) ? 0.5 : 1.0)
                                   * ((a.src_j == 0 || a.src_j == n[1]) ? 0.5 : 1.0)
                                   * ((a.src_i == 0 || a.src_i == n[0]) ? 0.5 : 1.0);
                    fix = m * (1.0 / (double)a.b[g][q]) * (double)(T)add * (double)(*a.src_old);
                }
            } else if (!a.skip_s) {
                const T add = (T)a.src_tr[step];
                a.sxx[g][q] = a.sxx[g][q] + add;
                a.syy[g][q] = a.syy[g][q] + add;
                a.szz[g][q] = a.szz[g][q] + add;
                if (a.st[g][1]) {
                    a.st[g][1][q] = a.sxx[g][q];
                    a.st[g][2][q] = a.syy[g][q];
                    a.st[g][3][q] = a.szz[g][q];
                }
            }
        }
        *a.fix = fix;
        if (!a.energy) *a.step = step + 1;
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_recsrc(PointArgs<T> a)
{
    recsrc_body(a);
}

struct EnergyArgs {
    const double *ke[MAXG];
    const double *se[MAXG];
    int nblk[MAXG];
    double vol[MAXG];   // 0.5*hx*hy*hz
    int ng;
    int src_g;
    const double *fix;
    double *energy;
    int *step;
    int max_steps;
};

// compute_energy.cpp:34-54 summed over grids (main.cpp:219-224); ends the step.
__global__ __launch_bounds__(64) void k_energy(EnergyArgs a)
{
    double e = 0;
    for (int g = 0; g < a.ng; g++) {
        double sk = wave_sum_fixed(a.ke[g], a.nblk[g]);
        double ss = wave_sum_fixed(a.se[g], a.nblk[g]);
        double kin = sk;
        if (g == a.src_g) kin += *a.fix;
        e += (kin + ss) * a.vol[g];
    }
    if (threadIdx.x == 0) {
        const int step = *a.step;
        if (step < a.max_steps) a.energy[step] = e;
        *a.step = step + 1;
    }
}

__global__ void k_end_step(int *step)
{
    if (threadIdx.x == 0) *step = *step + 1;
}

// compute_res.cpp:4-95 with r^2: fine-plane sum minus r^2 times the coarse
// sum, one workgroup, fixed order.
template <typename T>
__global__ __launch_bounds__(256) void k_res(IfaceDev<T> a, int which, double *out)
{
    __shared__ double lds[4];
    const int r = a.r, nx1 = r * a.nx2, ny1 = r * a.ny2, nx2 = a.nx2, ny2 = a.ny2, nz1 = a.nz1;
    const T *vf, *sf, *vc, *sc;
    int i0s, j0s, i0e, j0e;  // fine ranges [s, e)
    if (which == 0) { vf = a.F.f[0]; sf = a.F.f[7]; vc = a.C.f[0]; sc = a.C.f[7]; i0s = 1; j0s = 1; i0e = nx1; j0e = ny1; }
    else if (which == 1) { vf = a.F.f[1]; sf = a.F.f[6]; vc = a.C.f[1]; sc = a.C.f[6]; i0s = 0; j0s = 0; i0e = nx1; j0e = ny1; }
    else { vf = a.F.f[5]; sf = a.F.f[2]; vc = a.C.f[5]; sc = a.C.f[2]; i0s = 0; j0s = 1; i0e = nx1; j0e = ny1; }
    double s1 = 0, s2 = 0;
    const long nfx = i0e - i0s, nf = nfx * (j0e - j0s);
    for (long q = threadIdx.x; q < nf; q += blockDim.x) {
        const int i = (int)(q % nfx) + i0s, jj = (int)(q / nfx) + j0s;
        const long p = FIDX(a.F, i, jj, nz1);
        s1 += ld(vf, p) * (ld(sf, p - a.F.S2) + ld(sf, p));
    }
    const int ics = i0s, jcs = j0s, ice = (which == 0) ? nx2 : nx2, jce = ny2;
    const long ncx = ice - ics, nc = ncx * (jce - jcs);
    const double rr = (double)(r * r);
    for (long q = threadIdx.x; q < nc; q += blockDim.x) {
        const int i = (int)(q % ncx) + ics, jj = (int)(q / ncx) + jcs;
        const long p = FIDX(a.C, i, jj, 0);
        s2 += rr * ld(vc, p) * (ld(sc, p - a.C.S2) + ld(sc, p));
    }
    double t = s1 - s2;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off, 64);
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = t;
    __syncthreads();
    if (threadIdx.x == 0) *out = ((lds[0] + lds[1]) + lds[2]) + lds[3];
}

#undef FIDX

// Field::AvgAbs (field.cpp:215-228) over the field's own range (no halo):
// per-workgroup sums of |f| over x-rows, then one fixed-order sum.
template <typename T>
__global__ __launch_bounds__(256) void k_avgabs_part(const T *__restrict__ f, long o, long P, long S2, int lx, int ly, int lz,
                                                     double *__restrict__ part)
{
    __shared__ double lds[4];
    const long rows = (long)ly * lz;
    double acc = 0;
    for (long row = blockIdx.x; row < rows; row += gridDim.x) {
        const long iy = row % ly, iz = row / ly;
        const T *p = f + o + iz * S2 + iy * P;
        for (int ix = threadIdx.x; ix < lx; ix += 256) acc += fabs((double)p[ix]);
    }
    const double s = block_sum<4>(acc, lds);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ __launch_bounds__(64) void k_avgabs_final(const double *__restrict__ part, int n, double cz, double cy, double cx,
                                                    double *__restrict__ out)
{
    const double s = wave_sum_fixed(part, n);
    if (threadIdx.x == 0) *out = s / cz / cy / cx;
}

}  // namespace ecsg
