// ecsg_batch.hpp -- batched launches over many composite models, and the
// single-workgroup Jacobi solve of small vx-sxz interface systems.
// Included by ecsg_capi.hip only.
//
// A batch advances many models (shots, parameter sets, sizes) with one
// launch per phase of the time step: the stencil kernels take the tiles of
// every grid of every model from one flat work list (k_vel_batch,
// k_stress_batch), the interface kernels take the blocks of every interface
// (below), the receivers/source one workgroup per model, and each small
// interface system is solved by one workgroup holding its iterate in LDS.
// Nothing here is cooperative, so a batched step can be captured in a
// hipGraph and replayed.
#pragma once
#include "ecsg_kernels.hpp"

namespace ecsg {

// One interface of a batched launch.
template <typename T>
struct IfWork {
    IfaceDev<T> a;
    double *B, *M, *Rp, *U;  // vz-szz right-hand side, prefactored elimination, pivot reciprocals
    double *Bs, *absp;       // vx-sxz right-hand side and per-block |B| sums
};

template <typename T>
__global__ __launch_bounds__(256) void k_gvz_rhs_batch(const IfWork<T> *__restrict__ w, const int *__restrict__ start, int n)
{
    const int m = batch_find(start, n, blockIdx.x);
    gvz_rhs_body(w[m].a, (blockIdx.x - start[m]) * 256 + (int)threadIdx.x, w[m].B);
}

template <typename T>
__global__ __launch_bounds__(64) void k_gvz_solve_batch(const IfWork<T> *__restrict__ w, const int *__restrict__ start, int n)
{
    const int m = batch_find(start, n, blockIdx.x);
    const IfWork<T> &W = w[m];
    gvz_solve_body(W.a, (blockIdx.x - start[m]) * 64 + (int)threadIdx.x, W.B, W.M, W.Rp, W.U);
}

// one wavefront per coarse x-line of every interface (gvz_solve_wave)
template <typename T>
__global__ __launch_bounds__(64) void k_gvz_solve_wave_batch(const IfWork<T> *__restrict__ w, const int *__restrict__ start, int n)
{
    extern __shared__ double vz_sh[];
    const int m = batch_find(start, n, blockIdx.x);
    const IfWork<T> &W = w[m];
    gvz_solve_wave(W.a, (int)blockIdx.x - start[m], W.B, W.M, W.Rp, W.U, vz_sh);
}

template <typename T>
__global__ __launch_bounds__(256) void k_gvz_fine_batch(const IfWork<T> *__restrict__ w, const int *__restrict__ start, int n)
{
    const int m = batch_find(start, n, blockIdx.x);
    gvz_fine_body(w[m].a, (blockIdx.x - start[m]) * 256 + (int)threadIdx.x);
}

template <typename T>
__global__ __launch_bounds__(256) void k_gsxz_rhs_batch(const IfWork<T> *__restrict__ w, const int *__restrict__ start, int n)
{
    __shared__ double lds[4];
    const int m = batch_find(start, n, blockIdx.x);
    gsxz_rhs_body(w[m].a, blockIdx.x - start[m], w[m].Bs, w[m].absp, lds);
}

template <typename T>
__global__ __launch_bounds__(256) void k_gsyz_batch(const IfWork<T> *__restrict__ w, const int *__restrict__ start, int n)
{
    const int m = batch_find(start, n, blockIdx.x);
    gsyz_body(w[m].a, (blockIdx.x - start[m]) * 256 + (int)threadIdx.x);
}

// receivers and source of every model: one workgroup per model
template <typename T>
__global__ __launch_bounds__(256) void k_recsrc_batch(const PointArgs<T> *__restrict__ pa)
{
    recsrc_body(pa[blockIdx.x]);
}

// ---------------------------------------------------------------------------
// Jacobi2D + Norm2D (update_ghost_sxz.cpp:314-388) of one interface per
// workgroup, both iterates in LDS.  Same sweeps (same operand order) and the
// same exit tests as the tiled solver; the norms are summed in one fixed
// order per workgroup.  Then the coarse ghosts and the fine ghosts
// (update_ghost_sxz.cpp:247-295 generalised to ratio r), unless solve_only.
template <typename T>
struct JacLds {
    const double *A;          // 9 bands, band k at A + k*nn, [J-1][I-1]
    const double *B;
    const double *abs_part;   // per-block sums of |B|
    double *X0;               // warm start in, solution out; (nx2+1)(ny2+1), zero border
    long long *stats;         // total, last, capped
    T *sxz1, *sxz2, *sxz1s, *sxz2s;
    long o1, P1, S21, o2, P2, S22;
    int nx2, ny2, nz1, r, nabs;
};

struct JacParams {
    double tol_cap, tol_fac;  // tol = min(tol_cap, tol_fac * AvgAbs(B))
    double floor;             // rounding-floor factor of the stagnation exit
    double rel_tol;           // also stop when the update norm <= rel_tol * ||X|| (0: off)
    int cap;
    int solve_only;
};

constexpr int JL = 1024;          // threads of the LDS Jacobi workgroup
constexpr int JL_MAXW = 9728;     // largest (nx2+1)(ny2+1): two iterates in 152 KiB of LDS

template <typename P>
struct JacLdsSet {                // the interfaces of one model, passed by value
    JacLds<P> it[MAXG - 1];
};

template <typename T>
__device__ __forceinline__ void jacobi_lds_body(const JacLds<T> &J, const JacParams &jp, double *xs)
{
    __shared__ double red[JL / 64][2];
    __shared__ double s_tol;
    __shared__ int s_done, s_it;
    const int nx2 = J.nx2, ny2 = J.ny2, n0 = nx2 - 1, n1 = ny2 - 1;
    const long w = nx2 + 1, W = w * (ny2 + 1), nn = (long)n0 * n1;
    const int tid = threadIdx.x;
    double *xo = xs, *xn = xs + W;
    for (long i = tid; i < W; i += JL) {
        const double v = J.X0[i];
        xo[i] = v;
        xn[i] = v;
    }
    if (tid < 64) {
        const double sb = wave_sum_fixed(J.abs_part, J.nabs);
        if (tid == 0) {
            const double t2 = sb / (double)nn * jp.tol_fac;
            s_tol = (jp.tol_cap < t2) ? jp.tol_cap : t2;
            s_it = 0;
        }
    }
    __syncthreads();
    const double *A = J.A, *B = J.B;
    double prev = 0;
    while (true) {
        double acc = 0, ax = 0;
        for (long q = tid; q < nn; q += JL) {
            const long c = (q / n0 + 1) * w + (q % n0 + 1);
            const double v = (B[q]
                              - A[q] * xo[c - w - 1]
                              - A[nn + q] * xo[c - 1]
                              - A[2 * nn + q] * xo[c + w - 1]
                              - A[3 * nn + q] * xo[c - w]
                              - A[5 * nn + q] * xo[c + w]
                              - A[6 * nn + q] * xo[c - w + 1]
                              - A[7 * nn + q] * xo[c + 1]
                              - A[8 * nn + q] * xo[c + w + 1]) / A[4 * nn + q];
            const double d = v - xo[c];
            acc += d * d;
            ax += v * v;
            xn[c] = v;
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            acc += __shfl_xor(acc, off, 64);
            ax += __shfl_xor(ax, off, 64);
        }
        if ((tid & 63) == 0) {
            red[tid >> 6][0] = acc;
            red[tid >> 6][1] = ax;
        }
        __syncthreads();
        if (tid == 0) {
            double a2 = 0, x2 = 0;
            for (int k = 0; k < JL / 64; k++) {
                a2 += red[k][0];
                x2 += red[k][1];
            }
            const double norm = sqrt(a2), xnorm = sqrt(x2);
            const int it = ++s_it;
            int done = 0;
            if (!(norm > s_tol)) done = 1;
            else if (it > 1 && norm >= prev && norm <= jp.floor * xnorm) done = 1;
            else if (norm <= jp.rel_tol * xnorm) done = 1;
            else if (it >= jp.cap) done = 1;
            prev = norm;
            s_done = done;
        }
        __syncthreads();
        double *tsw = xo; xo = xn; xn = tsw;  // xo: the iterate of this sweep
        if (s_done) break;
    }
    for (long i = tid; i < W; i += JL) J.X0[i] = xo[i];
    if (!jp.solve_only) {
        for (long q = tid; q < nn; q += JL) {
            const long I = q % n0 + 1, Jr = q / n0 + 1;
            const T x = (T)xo[Jr * w + I];
            const long idx = J.o2 - J.S22 + Jr * J.P2 + I;
            J.sxz2[idx] = x;
            if (J.sxz2s) J.sxz2s[idx] = x;
        }
        const int r = J.r, nx1 = r * nx2, ny1 = r * ny2;
        const long nf = (long)(nx1 - 1) * (ny1 - 1);
        for (long q = tid; q < nf; q += JL) {
            const int i = (int)(q % (nx1 - 1)) + 1, jj = (int)(q / (nx1 - 1)) + 1;
            const int Ia = i / r, Ja = jj / r, ao = i - r * Ia, bo = jj - r * Ja;
            double S = 0;
            for (int dJ = 0; dJ <= 1; dJ++)
                for (int dI = 0; dI <= 1; dI++) {
                    const int Ic = Ia + dI, Jc = Ja + dJ;
                    if (Ic < 1 || Ic > nx2 - 1 || Jc < 1 || Jc > ny2 - 1) continue;
                    const double pw = hatw(ao - r * dI, r) * hatw(bo - r * dJ, r);
                    if (pw == 0) continue;
                    const long q2 = J.o2 + (long)Jc * J.P2 + Ic;
                    S += pw * ((double)(T)xo[(long)Jc * w + Ic] + (double)J.sxz2[q2]);
                }
            const long qf = J.o1 + (long)J.nz1 * J.S21 + (long)jj * J.P1 + i;
            const T v = (T)(-(double)J.sxz1[qf - J.S21] + S);
            J.sxz1[qf] = v;
            if (J.sxz1s) J.sxz1s[qf] = v;
        }
    }
    if (tid == 0) {
        const int it = s_it;
        J.stats[0] += it;
        J.stats[1] = it;
        if (it >= jp.cap) J.stats[2] = 1;
    }
}

// the interfaces of one model (by value; workgroup k solves interface k)
template <typename T>
__global__ __launch_bounds__(JL) void k_jacobi_lds(JacLdsSet<T> set, JacParams jp)
{
    extern __shared__ double xs[];
    jacobi_lds_body(set.it[blockIdx.x], jp, xs);
}

// the interfaces of a batch (device array)
template <typename T>
__global__ __launch_bounds__(JL) void k_jacobi_lds_batch(const JacLds<T> *__restrict__ its, JacParams jp)
{
    extern __shared__ double xs[];
    jacobi_lds_body(its[blockIdx.x], jp, xs);
}

}  // namespace ecsg
