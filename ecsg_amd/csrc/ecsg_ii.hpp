// ecsg_ii.hpp -- the reference's "intuitive" interpolation (imode 0, ratio 2)
// of the interface ghosts: UpdateGhostVZ / SXZ / SYZ mode 0
// (update_ghost_vz.cpp:36-123, update_ghost_sxz.cpp:36-84,
// update_ghost_syz.cpp:23-150).  Included by ecsg_capi.hip only.
//
// Each formula reads the fine planes nz1-2 / nz1-1 and the coarse plane 0
// only and writes the coarse ghost plane -1 or the fine plane nz1, so one
// launch per function computes every written point independently.  The
// arithmetic is in double with the reference's operation order, so FP64
// storage reproduces it bit for bit.
#pragma once
#include "ecsg_kernels.hpp"

namespace ecsg {

#define IIX(G, i, j, k) ((G).o + (long)(k) * (G).S2 + (long)(j) * (G).P + (i))

template <typename T>
__device__ __forceinline__ void put2(T *f, T *fs, long q, double v)
{
    f[q] = (T)v;
    if (fs) fs[q] = (T)v;
}

// UpdateGhostVZ mode 0.  Threads: coarse ghosts nx2*(ny2-1), then fine
// quadruples (nx2-1)*ny2, then the x edges ny2.  xend selects the form of the
// x-end edge (update_ghost_vz.cpp:98-107): 0 leaves it unwritten, as the
// committed version of that line does (its stores fall outside the field);
// 1 is the working copy's form, whose coarse reads use the fine index nx1-1
// on vz2 and so, in the reference's flat Field storage (row length nx2+2,
// field.cpp:123-136), read the element nx2+2 further on: x index
// (nx1-1) - (nx2+2) of the next y row.
template <typename T>
__global__ __launch_bounds__(256) void k_ii_vz(IfaceDev<T> a, int xend)
{
    const int nx2 = a.nx2, ny2 = a.ny2, nz1 = a.nz1;
    const GridDev<T> &F = a.F;
    const GridDev<T> &C = a.C;
    T *vz1 = F.f[2], *vz2 = C.f[2];
    T *vz1s = F.st[2], *vz2s = C.st[2];
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long na = (long)nx2 * (ny2 - 1), nb = (long)(nx2 - 1) * ny2;
    if (t < na) {
        const int i0 = (int)(t % nx2), i1 = (int)(t / nx2) + 1;
        const long p2 = IIX(F, 2 * i0, 2 * i1, nz1 - 2), p1 = p2 + F.S2;
        const double v = 0.25 * (((ld(vz1, p2) + ld(vz1, p2 + 1)) + ld(vz1, p1)) + ld(vz1, p1 + 1));
        put2(vz2, vz2s, IIX(C, i0, i1, -1), v);
        return;
    }
    t -= na;
    if (t < nb) {
        const int i0 = (int)(t % (nx2 - 1)), i1 = (int)(t / (nx2 - 1));
        const long qc = IIX(C, i0, i1, 0);
        const double A = ld(vz2, qc), B = ld(vz2, qc + 1), Cc = ld(vz2, qc + C.P), D = ld(vz2, qc + C.P + 1);
        const long q0 = IIX(F, 2 * i0 + 1, 2 * i1, nz1), q1 = q0 + F.P;  // rows 2*i1 and 2*i1+1
        put2(vz1, vz1s, q0, ld(vz1, q0 - F.S2) / 3 + A / 2 + B / 6);
        put2(vz1, vz1s, q0 + 1, ld(vz1, q0 + 1 - F.S2) / 3 + A / 6 + B / 2);
        put2(vz1, vz1s, q1, ld(vz1, q1 - F.S2) / 3 + A / 4 + B / 12 + Cc / 4 + D / 12);
        put2(vz1, vz1s, q1 + 1, ld(vz1, q1 + 1 - F.S2) / 3 + A / 12 + B / 4 + Cc / 12 + D / 4);
        return;
    }
    t -= nb;
    if (t >= ny2) return;
    const int i1 = (int)t;
    {
        const long q0 = IIX(F, 0, 2 * i1, nz1), q1 = q0 + F.P;
        const long qc = IIX(C, 0, i1, 0);
        const double e0 = ld(vz1, q0 - F.S2 + 1);  // vz1(1, 2*i1, nz1-1), used by both rows
        put2(vz1, vz1s, q0, ld(vz1, q0 - F.S2) * 2 / 3 - e0 / 3 + ld(vz2, qc) * 2 / 3);
        put2(vz1, vz1s, q1, ld(vz1, q1 - F.S2) * 2 / 3 - e0 / 3 + ld(vz2, qc) / 3 + ld(vz2, qc + C.P) / 3);
    }
    if (xend) {
        const int xe = 2 * nx2 - 1;
        const long q0 = IIX(F, xe, 2 * i1, nz1), q1 = q0 + F.P;
        const int len0 = nx2 + 2, w = xe + 1;
        const int xa = w % len0 - 1, ya = i1 + w / len0;  // vz2(xe, i1) in the flat storage
        const long qc = IIX(C, xa, ya, 0);
        const double e0 = ld(vz1, q0 - F.S2 - 1);  // vz1(xe-1, 2*i1, nz1-1)
        put2(vz1, vz1s, q0, ld(vz1, q0 - F.S2) * 2 / 3 - e0 / 3 + ld(vz2, qc) * 2 / 3);
        put2(vz1, vz1s, q1, ld(vz1, q1 - F.S2) * 2 / 3 - e0 / 3 + ld(vz2, qc) / 3 + ld(vz2, qc + C.P) / 3);
    }
}

// UpdateGhostSXZ mode 0: one thread per coarse node (i0, i1) in
// [0,nx2] x [0,ny2]: the coarse ghost and the fine children it owns.
template <typename T>
__global__ __launch_bounds__(256) void k_ii_sxz(IfaceDev<T> a)
{
    const int nx2 = a.nx2, ny2 = a.ny2, nz1 = a.nz1;
    const GridDev<T> &F = a.F;
    const GridDev<T> &C = a.C;
    T *sxz1 = F.f[7], *sxz2 = C.f[7];
    T *sxz1s = F.st[7], *sxz2s = C.st[7];
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)(nx2 + 1) * (ny2 + 1)) return;
    const int i0 = (int)(t % (nx2 + 1)), i1 = (int)(t / (nx2 + 1));
    const bool xin = i0 >= 1 && i0 < nx2, yin = i1 >= 1 && i1 < ny2;
    const long qf = IIX(F, 2 * i0, 2 * i1, nz1);
    const long qc = IIX(C, i0, i1, 0);
    if (xin && yin) {
        put2(sxz2, sxz2s, qc - C.S2, 0.5 * (ld(sxz1, qf - 2 * F.S2) + ld(sxz1, qf - F.S2)));
        put2(sxz1, sxz1s, qf, ld(sxz1, qf - F.S2) / 3 + ld(sxz2, qc) * 2 / 3.0);
    }
    if (i0 < nx2 && yin)
        put2(sxz1, sxz1s, qf + 1, ld(sxz1, qf + 1 - F.S2) / 3 + ld(sxz2, qc) / 3 + ld(sxz2, qc + 1) / 3);
    if (xin && i1 < ny2)
        put2(sxz1, sxz1s, qf + F.P, ld(sxz1, qf + F.P - F.S2) / 3 + ld(sxz2, qc) / 3 + ld(sxz2, qc + C.P) / 3);
    if (i0 < nx2 && i1 < ny2)
        put2(sxz1, sxz1s, qf + F.P + 1,
             ld(sxz1, qf + F.P + 1 - F.S2) / 3 + ld(sxz2, qc) / 6 + ld(sxz2, qc + C.P) / 6 + ld(sxz2, qc + 1) / 6
                 + ld(sxz2, qc + C.P + 1) / 6);
}

// UpdateGhostSYZ mode 0.  Threads: coarse cells nx2*ny2 (ghost + interior
// children), then the y = 0 / y = ny1-1 edges (nx1), then the x = 0 /
// x = nx1-1 edges (ny1-2).
template <typename T>
__global__ __launch_bounds__(256) void k_ii_syz(IfaceDev<T> a)
{
    const int nx2 = a.nx2, ny2 = a.ny2, nz1 = a.nz1, nx1 = 2 * nx2, ny1 = 2 * ny2;
    const GridDev<T> &F = a.F;
    const GridDev<T> &C = a.C;
    T *syz1 = F.f[6], *syz2 = C.f[6];
    T *syz1s = F.st[6], *syz2s = C.st[6];
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long na = (long)nx2 * ny2;
    if (t < na) {
        const int i0 = (int)(t % nx2), i1 = (int)(t / nx2);
        const long p2 = IIX(F, 2 * i0, 2 * i1, nz1 - 2), p1 = p2 + F.S2;
        const double g = 0.125 * (((((((ld(syz1, p2) + ld(syz1, p2 + 1)) + ld(syz1, p2 + F.P)) + ld(syz1, p2 + F.P + 1))
                                     + ld(syz1, p1)) + ld(syz1, p1 + 1)) + ld(syz1, p1 + F.P)) + ld(syz1, p1 + F.P + 1));
        const long qc = IIX(C, i0, i1, 0);
        put2(syz2, syz2s, qc - C.S2, g);
        if (i0 < nx2 - 1 && i1 < ny2 - 1) {
            const double A = ld(syz2, qc), B = ld(syz2, qc + 1), Cc = ld(syz2, qc + C.P), D = ld(syz2, qc + C.P + 1);
            const long q = IIX(F, 2 * i0 + 1, 2 * i1 + 1, nz1);
            put2(syz1, syz1s, q, ld(syz1, q - F.S2) / 3 + 0.375 * A + 0.125 * B + 0.125 * Cc + D / 24);
            put2(syz1, syz1s, q + 1, ld(syz1, q + 1 - F.S2) / 3 + 0.125 * A + 0.375 * B + Cc / 24 + 0.125 * D);
            put2(syz1, syz1s, q + F.P, ld(syz1, q + F.P - F.S2) / 3 + 0.125 * A + B / 24 + 0.375 * Cc + 0.125 * D);
            put2(syz1, syz1s, q + F.P + 1, ld(syz1, q + F.P + 1 - F.S2) / 3 + A / 24 + 0.125 * B + 0.125 * Cc + 0.375 * D);
        }
        return;
    }
    t -= na;
    const int xe = nx1 - 1, ye = ny1 - 1;
    if (t < nx1) {
        const int i0 = (int)t;
        for (int side = 0; side < 2; side++) {
            const int j = side ? ye : 0, jn = side ? ye - 1 : 1;
            const long q = IIX(F, i0, j, nz1 - 1);  // plane nz1-1
            const double c2 = ld(syz2, IIX(C, i0 / 2, j / 2, 0)) * 2 / 3;
            double v;
            if (i0 == 0)
                v = ld(syz1, q) - ld(syz1, q + 1) / 3 - ld(syz1, IIX(F, i0, jn, nz1 - 1)) / 3 + c2;
            else if (i0 == xe)
                v = ld(syz1, q) - ld(syz1, q - 1) / 3 - ld(syz1, IIX(F, i0, jn, nz1 - 1)) / 3 + c2;
            else
                v = ld(syz1, q) / 3 + ld(syz1, q - 1 + 2 * (i0 % 2)) / 3 - ld(syz1, IIX(F, i0, jn, nz1 - 1)) / 3 + c2;
            put2(syz1, syz1s, q + F.S2, v);
        }
        return;
    }
    t -= nx1;
    if (t >= ny1 - 2) return;
    const int i1 = (int)t + 1, jm = i1 - 1 + 2 * (i1 % 2);
    {
        const long q = IIX(F, 0, i1, nz1 - 1);
        const double v = ld(syz1, q) / 3 + ld(syz1, IIX(F, 0, jm, nz1 - 1)) / 3 - ld(syz1, q + 1) / 3
                         + ld(syz2, IIX(C, 0, i1 / 2, 0)) * 2 / 3;
        put2(syz1, syz1s, q + F.S2, v);
    }
    {
        const long q = IIX(F, xe, i1, nz1 - 1);
        const double v = ld(syz1, q) / 3 + ld(syz1, IIX(F, xe, jm, nz1 - 1)) / 3 - ld(syz1, q - 1) / 3
                         + ld(syz2, IIX(C, xe / 2, i1 / 2, 0)) * 2 / 3;
        put2(syz1, syz1s, q + F.S2, v);
    }
}

#undef IIX

}  // namespace ecsg
