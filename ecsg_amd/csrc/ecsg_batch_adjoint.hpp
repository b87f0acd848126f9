// ecsg_batch_adjoint.hpp -- the transposed time step of many models with one
// launch per phase.  Included by ecsg_capi.hip only.
//
// Each launch takes the work of every member from a flat list: the adjoint
// stencil and gradient sweeps the tiles of every grid (prefix sums of the
// per-grid tile counts, XCD order over the whole launch, as the forward
// batch), the interface transposes the thread blocks of every interface of
// one interface level, the transposed ghost mirrors the entries of every
// grid's lists, the misfit injection the receiver locations of every member.
// The per-entry work is the single-model kernels' bodies (ecsg_adjoint.hpp),
// so a member's gradient is the one its own ecsg_gradient computes, bit for
// bit; members of one launch share no data.
#pragma once
#include "ecsg_adjoint.hpp"
#include "ecsg_batch.hpp"

namespace ecsg {

// One interface of a batched transposed step.  a: the forward state views
// of the vy-syz and vx-sxz transposes (state after the step), ap: those of
// the vz-szz transpose (velocities after, stresses before the step).
template <typename T>
struct AdjIfWork {
    IfAdj<T> a, ap;
    double *Ba;                   // vx-sxz transposed right-hand side
    const double *X0a;            // vx-sxz multipliers (solution of the LDS Jacobi)
    double *mu_vz, *mu_syz;
    const double *den;            // vy-syz pointwise denominators
    const double *M, *Dp, *U;     // prefactored vz-szz elimination
};

enum {
    AB_SYZ_GHOST, AB_SYZ_FINE0, AB_SYZ_FINE, AB_SYZ_COARSE,
    AB_SXZ_GHOST, AB_SXZ_FINE0, AB_SXZ_FINE, AB_SXZ_COARSE,
    AB_VZ_GHOST, AB_VZ_FINE0, AB_VZ_FINE, AB_VZ_COARSE, AB_VZ_TRI,
    AB_NPHASE
};

// threads of phase ph on one interface (AB_VZ_TRI: coarse x lines)
__host__ __device__ inline long adj_if_threads(int ph, int r, int nx2, int ny2)
{
    const long nx1 = (long)r * nx2, ny1 = (long)r * ny2;
    switch (ph) {
    case AB_SYZ_GHOST: return (long)nx2 * ny2;
    case AB_SYZ_FINE0: return nx1 * ny1;
    case AB_SXZ_GHOST: return (long)(nx2 - 1) * (ny2 - 1);
    case AB_SXZ_FINE0: return (nx1 - 1) * (ny1 - 1);
    case AB_VZ_GHOST: return (long)nx2 * (ny2 + 1);
    case AB_VZ_FINE0: return nx1 * (ny1 - 1);
    case AB_SYZ_FINE: case AB_SXZ_FINE: case AB_VZ_FINE: return (nx1 + 1) * (ny1 + 1);
    case AB_VZ_TRI: return nx2;  // one workgroup per line
    default: return (long)(nx2 + 1) * (ny2 + 1);  // the three coarse phases
    }
}

template <typename T, int PH>
__global__ __launch_bounds__(256) void k_adj_if_batch(const AdjIfWork<T> *__restrict__ w, const int *__restrict__ start, int n)
{
    const int m = batch_find(start, n, blockIdx.x);
    const AdjIfWork<T> &W = w[m];
    const int t = (blockIdx.x - start[m]) * 256 + (int)threadIdx.x;
    if (PH == AB_SYZ_GHOST) asyz_ghost_body(W.a, W.den, W.mu_syz, t);
    else if (PH == AB_SYZ_FINE0) asyz_fine0_body(W.a, t);
    else if (PH == AB_SYZ_FINE) asyz_fine_body(W.a, (const double *)W.mu_syz, t);
    else if (PH == AB_SYZ_COARSE) asyz_coarse_body(W.a, (const double *)W.mu_syz, t);
    else if (PH == AB_SXZ_GHOST) asxz_ghost_body(W.a, W.Ba, t);
    else if (PH == AB_SXZ_FINE0) asxz_fine0_body(W.a, t);
    else if (PH == AB_SXZ_FINE) asxz_fine_body(W.a, W.X0a, t);
    else if (PH == AB_SXZ_COARSE) asxz_coarse_body(W.a, W.X0a, t);
    else if (PH == AB_VZ_GHOST) avz_ghost_body(W.ap, W.mu_vz, t);
    else if (PH == AB_VZ_FINE0) avz_fine0_body(W.ap, t);
    else if (PH == AB_VZ_FINE) avz_fine_body(W.ap, (const double *)W.mu_vz, t);
    else if (PH == AB_VZ_COARSE) avz_coarse_body(W.ap, (const double *)W.mu_vz, t);
}

// the tridiagonal solves of the transposed vz-szz systems: one wavefront per
// coarse x-line (tri_solve_wave)
template <typename T>
__global__ __launch_bounds__(64) void k_adj_tri_batch(const AdjIfWork<T> *__restrict__ w, const int *__restrict__ start, int n)
{
    extern __shared__ double tri_sh[];
    const int m = batch_find(start, n, blockIdx.x);
    const AdjIfWork<T> &W = w[m];
    tri_solve_wave(W.mu_vz, W.M, W.Dp, W.U, W.a.nx2, W.a.ny2 - 1, (int)blockIdx.x - start[m], tri_sh);
}

// adjoint stencils and gradient sweeps over the tiles of every grid
enum { AG_STRESS, AG_GRAD, AG_VEL, AG_GRADB };

template <typename T, int PH, typename IT>
__global__ __launch_bounds__(64 * (AWY > GBY ? AWY : GBY)) void k_adj_grid_batch(const AdjGridDev<T> *__restrict__ g, const int *__restrict__ start, int n,
                                                        int total)
{
    const int b = xcd_tile(blockIdx.x, total);
    const int m = batch_find(start, n, b);
    if constexpr (PH == AG_STRESS || PH == AG_VEL) {
        __shared__ T xs[AXS];
        if constexpr (PH == AG_STRESS) adj_stress_body<T, IT>(g[m], b - start[m], xs);
        else adj_vel_body<T, IT>(g[m], b - start[m], xs);
    } else if constexpr (PH == AG_GRAD) {
        adj_grad_body<T, IT>(g[m], b - start[m]);
    } else {
        adj_gradb_body<T, IT>(g[m], b - start[m]);
    }
}

// transposed ghost mirrors of every grid: gathers, then the halo zeroing
template <typename T>
struct MirrorWork {
    T *lam;
    const long *src;
    const int *start;
    const long *halo;
    int n, nh;
};

template <typename T, bool ZERO>
__global__ __launch_bounds__(256) void k_mirror_batch(const MirrorWork<T> *__restrict__ w, const int *__restrict__ start, int n)
{
    const int m = batch_find(start, n, blockIdx.x);
    const MirrorWork<T> &W = w[m];
    const int t = (blockIdx.x - start[m]) * 256 + (int)threadIdx.x;
    if (ZERO) zero_list_body(W.lam, W.halo, W.nh, t);
    else mirror_T_body(W.lam, W.src, W.start, W.halo, W.n, t);
}

// misfit injection of every member at step `step`
template <typename T>
__global__ __launch_bounds__(256) void k_adj_rec_batch(const RecAdj<T> *__restrict__ r, const int *__restrict__ start, int n,
                                                       int step)
{
    const int m = batch_find(start, n, blockIdx.x);
    adj_rec_body(r[m], step, (blockIdx.x - start[m]) * 256 + (int)threadIdx.x);
}

// dst += src over n entries (the gradients of a model's shots, summed in
// member order)
template <typename T>
__global__ __launch_bounds__(256) void k_add_into(T *__restrict__ dst, const T *__restrict__ src, long n)
{
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) dst[i] = dst[i] + src[i];
}

}  // namespace ecsg
