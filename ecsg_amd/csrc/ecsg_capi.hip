// ecsg_capi.hip -- host side of libecsg_hip.so: device state, parameter
// staging, time-invariant interface matrices, kernel launches, hipGraph
// capture, and the C ABI declared in include/ecsg_hip.h.
#include "ecsg_hip.h"
#include "ecsg_kernels.hpp"
#include "ecsg_adjoint.hpp"
#include "ecsg_batch.hpp"
#include "ecsg_batch_adjoint.hpp"
#include "ecsg_fused.hpp"
#include "ecsg_ii.hpp"
#include "ecsg_radius.hpp"
#include "ecsg_diag.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

using namespace ecsg;

// Grid-barrier launches (the tiled and register Jacobi kernels).  Their
// barrier is the library's own (grid_barrier: device-scope atomics with a
// bounded wait), so they are ordinary launches: the runtime's cooperative
// launch would route them through the global-wave-sync queue, which one
// process holds at a time (a second process then gets "illegal state").  The
// callers check that every workgroup of the launch can be resident at once
// (occupancy x CUs), which is what the barrier needs.  Launches of every
// handle on one device are issued in one order, each behind the previous
// one: two such kernels from different streams could otherwise each hold
// part of the GPU while waiting at a barrier for workgroups that the other
// keeps from being placed.  The per-device events live until the process
// ends or the last handle is destroyed (coop_release), so that no HIP
// object of the library outlives the handles into the runtime's teardown.
static std::mutex coop_mu;
static hipEvent_t coop_last[64];
static int live_handles = 0;

static void coop_acquire()
{
    std::lock_guard<std::mutex> lock(coop_mu);
    live_handles++;
}
static void coop_release()
{
    std::lock_guard<std::mutex> lock(coop_mu);
    if (--live_handles > 0) return;
    for (auto &ev : coop_last)
        if (ev) {
            hipEventDestroy(ev);
            ev = nullptr;
        }
}

static hipError_t launch_cooperative(const void *fn, dim3 grid, dim3 block, void **args, hipStream_t s)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= 64) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(coop_mu);
    hipEvent_t &ev = coop_last[dev];
    if (!ev) {
        e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
        if (e != hipSuccess) { ev = nullptr; return e; }
    } else {
        e = hipStreamWaitEvent(s, ev, 0);
        if (e != hipSuccess) return e;
    }
    e = hipLaunchKernel(fn, grid, block, args, 0, s);
    if (e != hipSuccess) return e;
    return hipEventRecord(ev, s);
}

// Device layout of one grid's arrays: x fastest, ix = 0 at the start of a
// 128-byte line, a halo of H entries on every side, rows padded to whole
// lines; `astride` entries per array (a whole number of lines).
struct Layout {
    long P, S2, o, astride;
};
static Layout grid_layout(const int n[3], long H, long elem_bytes)
{
    const long line = 128 / elem_bytes;
    const long xoff = line;  // line >= H for every supported radius
    Layout L;
    L.P = ((xoff + n[0] + 1 + H) + line - 1) / line * line;
    L.S2 = L.P * (n[1] + 1 + 2 * H);
    L.o = H * L.S2 + H * L.P + xoff;
    L.astride = ((L.S2 * (n[2] + 1 + 2 * H)) + line - 1) / line * line;
    return L;
}

namespace {

// every index a radius-1 stencil sweep of this grid forms (halo planes and
// the one-entry neighbours included) has a byte offset below 2^32: the
// sweeps then use 32-bit offsets (k_vel / k_stress with IT = unsigned)
template <typename T>
bool fits32(const ecsg::GridDev<T> &v)
{
    return (double)(v.o + (long)(v.nz + 3) * v.S2) * sizeof(T) < 4294967296.0;
}
// the same for the adjoint sweeps of a grid (every array they touch has the
// grid's layout)
template <typename T>
bool fits32(const ecsg::AdjGridDev<T> &a)
{
    return (double)(a.o + (long)(a.nz + 3) * a.S2) * sizeof(T) < 4294967296.0;
}

// wavefield / stiffness positions (init_del_sim.cpp:163-173,
// init_del_par_field.cpp:82-88 extended to the nine stiffnesses)
const int SIM_POS[9][3] = {{0, 0, 0}, {1, 1, 0}, {1, 0, 1}, {1, 0, 0}, {1, 0, 0},
                           {1, 0, 0}, {1, 1, 1}, {0, 0, 1}, {0, 1, 0}};
const int ST_POS[9][3] = {{1, 0, 0}, {1, 0, 0}, {1, 0, 0}, {1, 0, 0}, {1, 0, 0},
                          {1, 0, 0}, {1, 1, 1}, {0, 0, 1}, {0, 1, 0}};

constexpr int ZCHUNK = 8;        // planes per lane of the batched sweeps
constexpr int ZCHUNK_MODEL = 1;  // planes per lane of a single model's sweeps (banded tile order)
constexpr int JACOBI_CAP = 100000;
constexpr size_t DIRECT_MAX_BYTES = (size_t)512 << 20;  // direct vx-sxz factors per interface

// Inverse of the dense n x n row-major matrix a (destroyed), Gauss-Jordan
// elimination with partial pivoting; false when a pivot column is zero.
bool invert_dense(std::vector<double> &a, std::vector<double> &inv, int n)
{
    inv.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; i++) inv[(size_t)i * n + i] = 1.0;
    for (int c = 0; c < n; c++) {
        int p = c;
        double best = std::fabs(a[(size_t)c * n + c]);
        for (int r = c + 1; r < n; r++)
            if (std::fabs(a[(size_t)r * n + c]) > best) { best = std::fabs(a[(size_t)r * n + c]); p = r; }
        if (!(best > 0) || !std::isfinite(best)) return false;
        if (p != c)
            for (int j = 0; j < n; j++) {
                std::swap(a[(size_t)p * n + j], a[(size_t)c * n + j]);
                std::swap(inv[(size_t)p * n + j], inv[(size_t)c * n + j]);
            }
        const double piv = a[(size_t)c * n + c];
        for (int j = c; j < n; j++) a[(size_t)c * n + j] /= piv;
        for (int j = 0; j < n; j++) inv[(size_t)c * n + j] /= piv;
        for (int r = 0; r < n; r++) {
            if (r == c) continue;
            const double f = a[(size_t)r * n + c];
            if (f == 0) continue;
            for (int j = c; j < n; j++) a[(size_t)r * n + j] -= f * a[(size_t)c * n + j];
            for (int j = 0; j < n; j++) inv[(size_t)r * n + j] -= f * inv[(size_t)c * n + j];
        }
    }
    return true;
}

// Factors of the direct vx-sxz solve (k_sxz_direct): the inverse Schur
// complements W_J = S_J^-1 of the block-tridiagonal 9-band matrix A (band m
// at A + m*nn couples unknown (I, J) to (I + m/3 - 1, J + m%3 - 1), rows of
// n0 unknowns, n1 rows): S_0 = D_0, S_J = D_J - L_J W_{J-1} U_{J-1}, each W_J
// stored column-major, entry (i, k) at W + J*n0*n0 + k*n0 + i.
bool sxz_schur_inverses(const std::vector<double> &A, int n0, int n1, std::vector<double> &W)
{
    const size_t nn = (size_t)n0 * n1, bs = (size_t)n0 * n0;
    W.assign((size_t)n1 * bs, 0.0);
    auto band = [&](int m, int J, int i) { return A[m * nn + (size_t)J * n0 + i]; };
    std::vector<double> S(bs), M(bs), inv;
    for (int J = 0; J < n1; J++) {
        std::fill(S.begin(), S.end(), 0.0);
        for (int i = 0; i < n0; i++) {
            if (i > 0) S[(size_t)i * n0 + i - 1] = band(1, J, i);
            S[(size_t)i * n0 + i] = band(4, J, i);
            if (i + 1 < n0) S[(size_t)i * n0 + i + 1] = band(7, J, i);
        }
        if (J > 0) {
            const double *Wp = &W[(size_t)(J - 1) * bs];
            // M = W_{J-1} U_{J-1}, U_{J-1}(k, k + d) = band(3(d+1)+2, J-1, k)
            for (int i = 0; i < n0; i++)
                for (int c = 0; c < n0; c++) {
                    double acc = 0;
                    for (int d = -1; d <= 1; d++) {
                        const int k = c - d;
                        if (k < 0 || k >= n0) continue;
                        acc += Wp[(size_t)k * n0 + i] * band(3 * (d + 1) + 2, J - 1, k);
                    }
                    M[(size_t)i * n0 + c] = acc;
                }
            // S -= L_J M, L_J(i, i + d) = band(3(d+1), J, i)
            for (int i = 0; i < n0; i++)
                for (int d = -1; d <= 1; d++) {
                    const int k = i + d;
                    if (k < 0 || k >= n0) continue;
                    const double l = band(3 * (d + 1), J, i);
                    if (l == 0) continue;
                    for (int c = 0; c < n0; c++) S[(size_t)i * n0 + c] -= l * M[(size_t)k * n0 + c];
                }
        }
        if (!invert_dense(S, inv, n0)) return false;
        double *Wj = &W[(size_t)J * bs];
        for (int i = 0; i < n0; i++)
            for (int k = 0; k < n0; k++) Wj[(size_t)k * n0 + i] = inv[(size_t)i * n0 + k];
    }
    return true;
}

struct Status {
    int code = ECSG_OK;
    std::string msg;
};

class EngineBase {
public:
    EngineBase() { coop_acquire(); }
    virtual ~EngineBase() { coop_release(); }
    Status st;
    int fail(int code, const std::string &m)
    {
        st.code = code;
        st.msg = m;
        return code;
    }
    virtual int init(const ecsg_config &c) = 0;
    virtual int set_params(int g, const double *buoy, const double *const stiff[9], int iso) = 0;
    virtual int set_params_staged(int g, const double *buoy, const double *const stiff[9]) = 0;
    virtual int set_source(int g, int ix, int iy, int iz, int type, const float *tr, int len) = 0;
    virtual int set_receivers(int n, const int *loc) = 0;
    virtual int update_vel(int g) = 0;
    virtual int record_receivers() = 0;
    virtual int add_source() = 0;
    virtual int ghost_vz(int k) = 0;
    virtual int energy(double *e) = 0;
    virtual int update_stress(int g) = 0;
    virtual int ghost_sxz(int k) = 0;
    virtual int ghost_syz(int k) = 0;
    virtual int end_step() = 0;
    virtual int residual(int which, double *out) = 0;
    virtual int avg_abs(int g, int sel, double *out) = 0;
    virtual int sbp(int g, double *out) = 0;
    virtual int step_n(int n, int use_graph) = 0;
    virtual int steps_done() const = 0;
    virtual int sync() = 0;
    virtual int dims(int g, int sel, int len[3]) const = 0;
    virtual int get_field(int g, int sel, double *out) = 0;
    virtual int set_field(int g, int sel, const double *in) = 0;
    virtual int get_traces(double *out) = 0;
    virtual int get_energy(double *out) = 0;
    virtual int jacobi_stats(int k, long long *tot, int *last, int *capped) = 0;
    virtual int profile(int nsteps, double ms[6], int cnt[6]) = 0;
    virtual int step_timed(int nsteps, double ms[2]) = 0;
    virtual int reset() = 0;
    virtual int set_checkpointing(int every, int format) = 0;
    virtual int set_ring_steps(int k) = 0;
    virtual int checkpoint_interval() const = 0;
    virtual void checkpoint_levels(int *k1, int *k2, long long *recomputed) const = 0;
    virtual int gradient(int nsteps, const double *obs, double *misfit) = 0;
    virtual int get_gradient(int g, int sel, double *out) = 0;
    virtual int adjoint_step() = 0;
    virtual int adj_field(int g, int sel, double *out, const double *in) = 0;
    virtual int adjoint_stats(long long *sweeps, double *ms) = 0;
    virtual int autotune(int reps, int *by_out, int *zc_out) = 0;
    virtual int set_fused(int on) = 0;
    virtual int set_overlap(int on) = 0;
    virtual int overlap_active() const = 0;
    virtual int set_barrier_timeout(int ms) = 0;
    virtual int set_jacobi_tol(double cap, double fac) = 0;
};

#define HIPCHK(x)                                                                                  \
    do {                                                                                           \
        hipError_t e_ = (x);                                                                       \
        if (e_ != hipSuccess) return fail(ECSG_EHIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <typename T>
struct DevGrid {
    int n[3] = {0, 0, 0};
    double h[3] = {0, 0, 0};
    long P = 0, S2 = 0, o = 0, astride = 0;
    T *base = nullptr;
    GridDev<T> dev{};
    bool params = false;
    bool staged = false;  // parameters given at their staggered positions (ecsg_set_params_staged)
    int iso = 0;
    double *ke = nullptr, *se = nullptr;  // energy partials, nblk each
    int by = 4;                           // workgroup height of the stencil tiles
    int byv = 0;                          // of the overlapped step's velocity sweeps (0: by)
    // host copies (double of the T-rounded values) of the planes the
    // interface matrices need: buoyancy and c33 on iz = 0 and iz = nz
    std::vector<double> b_top, b_bot, c33_top, c33_bot;
    // arrays of base: 0..8 field set 0, 9..17 stiffness, 18 buoyancy; field
    // set 1 (alt, allocated when the fused pass is enabled) holds the other
    // buffer of the fused step; the fields of the current state are set `cur`
    int cur = 0;
    T *alt = nullptr;
    // radius > 1 (single grid): coefficients, mirror loops, stencil workgroups
    RCoef<T> rco{};
    MirrorSet mva{}, mvb{}, mst{};
    int rblk = 0;
    T *fset(int k) const { return k ? alt : base; }
    T *fields() const { return fset(cur); }
    T *arr(int a) const { return a < 9 ? fields() + (long)a * astride : base + (long)a * astride; }
};

struct DevIface {
    int r = 0, nx2 = 0, ny2 = 0, nz1 = 0;
    double *B = nullptr, *M = nullptr, *Dp = nullptr, *U = nullptr;  // vz system
    double *Rp = nullptr;                                            // 1 / Dp
    double *Bs = nullptr, *A = nullptr, *absp = nullptr;             // sxz system
    double *X0 = nullptr, *X1 = nullptr, *jpart = nullptr;
    double *W = nullptr;  // direct sxz mode: inverse Schur complements (k_sxz_direct)
    long long *stats = nullptr;
    int nabs = 0, coop = 0, tiles = 0;
    double rho = 0;  // Gershgorin bound of the Jacobi iteration matrix
    bool ready = false;
};

template <typename T>
class Engine : public EngineBase {
public:
    ecsg_config cfg{};
    hipStream_t s = nullptr;
    std::vector<DevGrid<T>> G;
    std::vector<DevIface> I;
    int nrec = 0;
    int *d_rec = nullptr;
    T *d_traces = nullptr;
    int *d_step = nullptr;
    double *d_energy = nullptr, *d_fix = nullptr;
    double *d_diag = nullptr;  // residual / AvgAbs reductions (DIAG_N doubles)
    T *e_save = nullptr;       // stress snapshot of the per-phase energy call
    size_t e_save_n = 0;
    T *d_src_old = nullptr;
    int *d_src_loc = nullptr;
    float *d_src_tr = nullptr;
    int src_g = -1, src_loc[3] = {0, 0, 0}, src_type = 0, src_len = 0;
    int host_step = 0;
    double tol_cap = 0, tol_fac = 0;
    int max_coop = 0;  // co-resident workgroups of the register Jacobi kernel
    int max_tiled = 0; // co-resident workgroups of the tiled Jacobi kernel
    int max_tiled_few = 0;  // the same for its JTT_FEW-thread form
    int ncu = 0;          // compute units of the device
    // grid barrier state of the cooperative Jacobi kernels (arrivals, exits,
    // timeout flag) and the longest wait at one barrier, in ticks of the
    // device's constant-rate wall clock
    unsigned *d_bar = nullptr;
    unsigned long long bar_limit = 0;
    int bar_ms = 10000;
    bool lds_ok = false;  // every interface fits the single-workgroup k_jacobi_lds
    // With sxz_solver 0 the single-model step uses k_jacobi_lds
    // (graph-capturable) for iterates of up to LDS_PREF_MAXW entries and the
    // tiled cooperative solve above: one workgroup runs ~45 sweeps of a
    // 41 x 41 iterate in about the time 4 tiles take (0.32 vs 0.35 ms per
    // FP32 step, ratio-3 two-grid models), of a 51 x 51 one slower (0.39 vs
    // 0.38 ms), of an 81 x 81 one in 0.76 vs 0.48 ms.  sxz_solver 2 raises the
    // limit to all that fits in LDS (JL_MAXW): models stepped concurrently on
    // their own streams then share the GPU, where each cooperative solve
    // would hold ~200 CUs and run eagerly (64 two-grid models of 1-4 M nodes,
    // 50 steps each: 8.6 against 3.1 Gpt/s).
    static constexpr long LDS_PREF_MAXW = 2048;
    long lds_pref_maxw = LDS_PREF_MAXW;
    // the single-model step solves the vx-sxz systems with k_jacobi_lds
    bool use_lds() const { return lds_ok && lds_maxw <= lds_pref_maxw; }
    // the step's vx-sxz solves hold no grid barrier (one workgroup per
    // interface: LDS Jacobi or the direct solve)
    bool no_grid_barrier() const { return cfg.sxz_solver == 3 || (use_lds() && cfg.sxz_solver == 0); }
    long lds_maxw = 0;
    // fused velocity+stress pass on double-buffered fields
    int use_fused = 0;
    int fused_form = 4;          // rows per wavefront of k_fused_w (2, 4 or 8)
    int fused_zc = 0;            // z planes per fused tile (0: the stencil tiling's)
    int fused_lean = 1;          // interior tiles by k_fused_lean
    static constexpr int FRL = sizeof(T) == 4 ? 4 : 2;  // rows per wavefront of k_fused_lean
    bool sets_dirty = false;     // field sets may differ in entries no kernel writes
    hipGraphExec_t gexec_f = nullptr;
    hipGraph_t graph_f = nullptr;
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    bool graph_failed = false;
    // profiling
    bool prof = false;
    bool live = false;  // step_timed: events around the grid-0 stencil launches only
    std::vector<hipEvent_t> evs;
    std::vector<int> ev_kind;
    // interface updates on a second stream beside the stencil planes that do
    // not touch the interfaces (enqueue_step_overlap)
    hipStream_t s2 = nullptr;
    hipEvent_t ov_ev[4] = {nullptr, nullptr, nullptr, nullptr};
    int overlap_on = -1;  // 1 on, 0 off, -1 with the tiled Jacobi solve
    bool ov_open = false;  // s has not yet waited for the last sxz/syz ghost update
    std::vector<int> h_rec;  // receiver locations (grid, ix, iy, iz)

    ~Engine() override
    {
        // let enqueued steps finish before their graph and arrays go
        if (s) hipStreamSynchronize(s);
        if (s2) hipStreamSynchronize(s2);
        if (gexec) hipGraphExecDestroy(gexec);
        if (graph) hipGraphDestroy(graph);
        for (auto &g : G) {
            hipFree(g.base);
            hipFree(g.alt);
            hipFree(g.ke);
            hipFree(g.se);
        }
        for (auto &f : I) {
            hipFree(f.B); hipFree(f.M); hipFree(f.Dp); hipFree(f.U); hipFree(f.Rp);
            hipFree(f.Bs); hipFree(f.A); hipFree(f.absp); hipFree(f.X0); hipFree(f.X1);
            hipFree(f.jpart); hipFree(f.stats); hipFree(f.W);
        }
        hipFree(d_bar);
        hipFree(d_rec); hipFree(d_traces); hipFree(d_step); hipFree(d_energy); hipFree(d_fix); hipFree(d_diag); hipFree(e_save); hipFree(d_sbp);
        hipFree(d_src_old); hipFree(d_src_loc); hipFree(d_src_tr);
        adj_free();
        for (auto e : evs) hipEventDestroy(e);
        for (auto e : ov_ev)
            if (e) hipEventDestroy(e);
        if (s2) hipStreamDestroy(s2);
        if (s) hipStreamDestroy(s);
    }

    template <typename P>
    int dalloc(P **p, size_t n)
    {
        hipError_t e = hipMalloc((void **)p, n * sizeof(P) + 64);
        if (e != hipSuccess) return fail(ECSG_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
        e = hipMemset(*p, 0, n * sizeof(P) + 64);
        if (e != hipSuccess) return fail(ECSG_EHIP, std::string("hipMemset: ") + hipGetErrorString(e));
        return ECSG_OK;
    }

    int init(const ecsg_config &c) override
    {
        cfg = c;
        if (c.sxz_solver < 0 || c.sxz_solver > 3)
            return fail(ECSG_EINVAL, "sxz_solver must be 0 (Jacobi), 1 (Chebyshev), 2 (Jacobi, one workgroup) or 3 (direct)");
        if (c.sxz_solver == 2) {
            // the same sweeps, one workgroup per interface wherever the iterate fits in LDS
            cfg.sxz_solver = 0;
            lds_pref_maxw = JL_MAXW;
        }
        if (c.num_grids < 1 || c.num_grids > ECSG_MAX_GRIDS) return fail(ECSG_EINVAL, "num_grids out of range");
        if (c.radius < 1 || c.radius > RMAX) return fail(ECSG_EINVAL, "radius must be 1..6");
        if (c.radius > 1 && c.num_grids != 1)
            return fail(ECSG_EINVAL, "radius > 1 runs on a single grid (the interface formulas are radius 1)");
        if (c.imode != 0 && c.imode != 1) return fail(ECSG_EINVAL, "imode must be 0 (II) or 1 (ECI)");
        if (c.top_bc != 0 && c.top_bc != 1) return fail(ECSG_EINVAL, "top_bc must be 0 or 1");
        if (!(c.dt > 0)) return fail(ECSG_EINVAL, "dt must be positive");
        if (c.max_steps < 1) return fail(ECSG_EINVAL, "max_steps must be >= 1");
        if (sizeof(T) == 8) { tol_cap = 1e-12; tol_fac = 1e-5; } else { tol_cap = 1e-3; tol_fac = 1e-4; }
        if (const char *e = getenv("ECSG_TILE_BY")) tile_by = atoi(e);
        if (const char *e = getenv("ECSG_TILE_BYV")) tile_byv = atoi(e);
        if (const char *e = getenv("ECSG_TILE_ZC")) tile_zc = atoi(e);
        if (const char *e = getenv("ECSG_FUSED_FORM")) fused_form = atoi(e) == 8 ? 8 : atoi(e) == 2 ? 2 : 4;
        if (const char *e = getenv("ECSG_FUSED_ZC")) fused_zc = std::max(0, atoi(e));
        if (const char *e = getenv("ECSG_FUSED_LEAN")) fused_lean = atoi(e) != 0;
        if (tile_by != 1 && tile_by != 2 && tile_by != 4 && tile_by != 8) return fail(ECSG_EINVAL, "tile height must be 1, 2, 4 or 8");
        if (tile_byv != 0 && tile_byv != 1 && tile_byv != 2 && tile_byv != 4 && tile_byv != 8)
            return fail(ECSG_EINVAL, "velocity tile height must be 0, 1, 2, 4 or 8");
        if (tile_zc < 1) return fail(ECSG_EINVAL, "z chunk must be >= 1");
        HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        if (const char *e = getenv("ECSG_OVERLAP")) overlap_on = atoi(e) != 0 ? 1 : 0;
        G.resize(c.num_grids);
        for (int g = 0; g < c.num_grids; g++) {
            DevGrid<T> &d = G[g];
            for (int k = 0; k < 3; k++) {
                d.n[k] = c.n[g][k];
                d.h[k] = c.h[g][k];
                if (d.n[k] < 2 || d.n[k] < c.radius) return fail(ECSG_EINVAL, "each grid needs at least max(2, radius) cells per dimension");
                if (!(d.h[k] > 0)) return fail(ECSG_EINVAL, "grid steps must be positive");
            }
            // halo: indices [-H, n+H] in every dimension, H = radius
            const Layout lay = grid_layout(d.n, c.radius, (long)sizeof(T));
            d.P = lay.P;
            d.S2 = lay.S2;
            d.o = lay.o;
            d.astride = lay.astride;
            if (int rc = dalloc(&d.base, (size_t)d.astride * 19)) return rc;
            GridDev<T> &v = d.dev;
            for (int a = 0; a < 9; a++) v.f[a] = d.arr(a);
            for (int a = 0; a < 9; a++) v.c[a] = d.arr(9 + a);
            v.b = d.arr(18);
            v.nx = d.n[0]; v.ny = d.n[1]; v.nz = d.n[2];
            v.P = d.P; v.S2 = d.S2; v.o = d.o;
            // coeff = fd*dt/h with fd = 1 for radius 1 (init_del_fdcoeff.cpp:19, 44)
            v.cx = (T)((1.0 * c.dt) / d.h[0]);
            v.cy = (T)((1.0 * c.dt) / d.h[1]);
            v.cz = (T)((1.0 * c.dt) / d.h[2]);
            const int top = (g == 0), bottom = (g == c.num_grids - 1);
            v.top_eff = (top && c.top_bc != 1) ? 1 : 0;
            v.bottom = bottom;
            v.free_top = (top && c.top_bc) ? 1 : 0;
            v.iso = 0;
            // vy y-wall mirror planes: [0, nz-1] as the reference loops, or all
            v.zy = c.vy_mirror_all ? d.n[2] : d.n[2] - 1;
            set_tiling(d, tile_by, tile_zc);
            d.byv = tile_byv;
            if (c.radius > 1) {
                setup_radius(d, c, top, bottom);
                v.nblk = d.rblk;  // energy partials are per radius-R workgroup
            }
            // energy partials sized for the largest tiling (BY = 1, ZC = 1)
            const size_t maxblk = (size_t)((d.n[0] + BX) / BX) * (d.n[1] + 1) * (d.n[2] + 1);
            if (int rc = dalloc(&d.ke, maxblk)) return rc;
            if (int rc = dalloc(&d.se, maxblk)) return rc;
        }
        I.resize(c.num_grids - 1);
        for (int k = 0; k + 1 < c.num_grids; k++) {
            DevGrid<T> &F = G[k], &C = G[k + 1];
            const int r = F.n[0] / C.n[0];
            if (r < 2 || F.n[0] != r * C.n[0] || F.n[1] != r * C.n[1])
                return fail(ECSG_EINVAL, "interface " + std::to_string(k) + ": fine x/y cells must be an integer ratio r >= 2 of the coarse ones");
            if (c.imode == 0 && r != 2)
                return fail(ECSG_EINVAL, "interface " + std::to_string(k) + ": the II interpolation (imode 0) is defined for ratio 2 only");
            DevIface &f = I[k];
            f.r = r; f.nx2 = C.n[0]; f.ny2 = C.n[1]; f.nz1 = F.n[2];
            const long nl = (long)f.nx2 * (f.ny2 - 1);
            const long nn = (long)(f.nx2 - 1) * (f.ny2 - 1);
            if (int rc = dalloc(&f.B, nl)) return rc;
            if (int rc = dalloc(&f.M, nl)) return rc;
            if (int rc = dalloc(&f.Dp, nl)) return rc;
            if (int rc = dalloc(&f.Rp, nl)) return rc;
            if (int rc = dalloc(&f.U, nl)) return rc;
            if (int rc = dalloc(&f.Bs, nn)) return rc;
            if (int rc = dalloc(&f.A, 9 * nn)) return rc;
            f.nabs = (int)((nn + 255) / 256);
            if (int rc = dalloc(&f.absp, f.nabs)) return rc;
            const long nxw = (long)(f.nx2 + 1) * (f.ny2 + 1);
            if (int rc = dalloc(&f.X0, nxw)) return rc;
            if (int rc = dalloc(&f.X1, nxw)) return rc;
            int dev = 0, per = 0;
            HIPCHK(hipGetDevice(&dev));
            HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
            HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, (const void *)k_jacobi<T>, JT, 0));
            if (per < 1) return fail(ECSG_EHIP, "cooperative Jacobi kernel cannot be resident");
            // every unknown must be owned by one thread for the whole solve
            const long want = (nn + (long)JT * JPPT - 1) / ((long)JT * JPPT);
            max_coop = per * ncu;
            if (want > (long)max_coop)
                return fail(ECSG_EINVAL, "interface " + std::to_string(k) + " too large for the resident Jacobi solve");
            f.coop = (int)std::max(1L, want);
            int pert = 0;
            HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&pert, (const void *)k_jacobi_tiled<T, JTT>, JTT, 0));
            max_tiled = pert * ncu;
            HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&pert, (const void *)k_jacobi_tiled<T, JTT_FEW>, JTT_FEW, 0));
            max_tiled_few = pert * ncu;
            f.tiles = ((f.nx2 - 1 + TJ - 1) / TJ) * ((f.ny2 - 1 + TJ - 1) / TJ);
            const size_t np = std::max(4 * (size_t)f.coop + 1, (size_t)2 * f.tiles * 2 * SJ);
            if (int rc = dalloc(&f.jpart, np)) return rc;
            if (int rc = dalloc(&f.stats, 3)) return rc;
            if (cfg.sxz_solver == 3) {
                const long n0 = f.nx2 - 1, n1 = f.ny2 - 1;
                if (n0 > JD_MAXN0)
                    return fail(ECSG_EINVAL, "interface " + std::to_string(k) + ": coarse rows of more than " +
                                                 std::to_string(JD_MAXN0) + " unknowns for the direct vx-sxz solve");
                if ((double)n1 * n0 * n0 * sizeof(double) > DIRECT_MAX_BYTES)
                    return fail(ECSG_EINVAL, "interface " + std::to_string(k) + ": direct vx-sxz factors above " +
                                                 std::to_string(DIRECT_MAX_BYTES >> 20) + " MiB");
                if (int rc = dalloc(&f.W, (size_t)n1 * n0 * n0)) return rc;
            }
        }
        for (int k0 = 0; k0 < (int)I.size(); k0 += MAXI) {
            int tot = 0;
            int tott = 0;
            for (int k = k0; k < std::min((int)I.size(), k0 + MAXI); k++) { tot += I[k].coop; tott += I[k].tiles; }
            if (tot > max_coop || (cfg.sxz_solver == 0 && tott > max_tiled))
                return fail(ECSG_EINVAL, "interfaces too large for one resident Jacobi launch");
        }
        // interfaces whose iterate fits in LDS are solved by one workgroup
        // each (no cooperative launch: the step can be graph-captured)
        lds_maxw = 0;
        lds_ok = !I.empty();
        for (auto &f : I) {
            const long W = (long)(f.nx2 + 1) * (f.ny2 + 1);
            lds_maxw = std::max(lds_maxw, W);
            if (W > JL_MAXW) lds_ok = false;
        }
        if (lds_ok) {
            HIPCHK(hipFuncSetAttribute((const void *)k_jacobi_lds<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)(2 * JL_MAXW * sizeof(double))));
            HIPCHK(hipFuncSetAttribute((const void *)k_jacobi_lds_batch<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)(2 * JL_MAXW * sizeof(double))));
        }
        if (!I.empty()) {
            if (int rc = dalloc(&d_bar, 4)) return rc;
            if (int rc = set_barrier_timeout(bar_ms)) return rc;
        }
        if (int rc = dalloc(&d_step, 1)) return rc;
        if (int rc = dalloc(&d_energy, cfg.max_steps)) return rc;
        if (int rc = dalloc(&d_fix, 1)) return rc;
        if (int rc = dalloc(&d_src_old, 1)) return rc;
        if (int rc = dalloc(&d_src_loc, 3)) return rc;
        if (int rc = dalloc(&d_src_tr, 1)) return rc;
        if (int rc = dalloc(&d_rec, 4)) return rc;
        if (int rc = dalloc(&d_traces, 1)) return rc;
        return ECSG_OK;
    }

    // ----- radius > 1 (single grid) ----------------------------------------
    static void add_loop(MirrorSet &m, int f, int axis, int dbase, int dstep, int sbase, int sstep, int a0, int a1,
                         int b0, int b1, int nir, int sign)
    {
        if (nir < 1 || a1 < a0 || b1 < b0) return;
        MirrorLoop &L = m.m[m.n];
        L.f = f; L.axis = axis; L.dbase = dbase; L.dstep = dstep; L.sbase = sbase; L.sstep = sstep;
        L.a0 = a0; L.a1 = a1; L.b0 = b0; L.b1 = b1; L.nir = nir; L.sign = sign;
        m.start[m.n + 1] = m.start[m.n] + (long)(a1 - a0 + 1) * (b1 - b0 + 1) * nir;
        m.n++;
    }

    // Coefficients and the ghost-mirror loops of update_vel.cpp:73-246 and
    // update_stress.cpp:135-309, transcribed with their index ranges.  Axis
    // pairs (a, b): mirrored x -> (y, z), y -> (x, z), z -> (x, y).
    static void setup_radius(DevGrid<T> &d, const ecsg_config &c, int top, int bottom)
    {
        const int R = c.radius, nx = d.n[0], ny = d.n[1], nz = d.n[2];
        double fd[RMAX];
        fd_table(R, fd);
        for (int k = 0; k < 3; k++)
            for (int ir = 0; ir < R; ir++) d.rco.c[k][ir] = (T)(fd[ir] * c.dt / d.h[k]);
        const int vtop = (top && c.top_bc != 1) ? 1 : 0;  // update_vel.cpp:17-18
        const int zy = c.vy_mirror_all ? nz : nz - 1;
        MirrorSet &A = d.mva, &B = d.mvb, &S = d.mst;
        A = MirrorSet{}; B = MirrorSet{}; S = MirrorSet{};
        // velocities, z walls (update_vel.cpp:75-131) then x walls (:135-189)
        if (vtop) {
            add_loop(A, 2, 2, 0, -1, -1, 1, 0, nx - 1, 0, ny, R - c.top_bc, -1);
            add_loop(A, 0, 2, 0, -1, 0, 1, 0, nx - 1, 0, ny, R - 1, -1);
            add_loop(A, 1, 2, 0, -1, 0, 1, 0, nx - 1, 0, ny - 1, R - 1, -1);
        }
        if (bottom) {
            add_loop(A, 2, 2, nz - 1, 1, nz, -1, 0, nx - 1, 0, ny, R, -1);
            add_loop(A, 0, 2, nz, 1, nz, -1, 0, nx - 1, 0, ny, R - 1, -1);
            add_loop(A, 1, 2, nz, 1, nz, -1, 0, nx - 1, 0, ny - 1, R - 1, -1);
        }
        add_loop(A, 2, 0, 0, -1, -1, 1, 0, ny, 0, nz - 1, R, -1);
        add_loop(A, 0, 0, 0, -1, 0, 1, 0, ny, 0, nz, R - 1, -1);
        add_loop(A, 1, 0, 0, -1, -1, 1, 0, ny, 0, nz, R, -1);
        add_loop(A, 2, 0, nx - 1, 1, nx, -1, 0, ny, 0, nz - 1, R, -1);
        add_loop(A, 0, 0, nx, 1, nx, -1, 0, ny, 0, nz, R - 1, -1);
        add_loop(A, 1, 0, nx - 1, 1, nx, -1, 0, ny, 0, nz, R, -1);
        // velocities, y walls (update_vel.cpp:191-246)
        add_loop(B, 2, 1, 0, -1, 0, 1, 0, nx - 1, 0, nz - 1, R - 1, -1);
        add_loop(B, 0, 1, 0, -1, 0, 1, 0, nx, 0, nz - 1, R - 1, -1);
        add_loop(B, 1, 1, 0, -1, -1, 1, 0, nx - 1, 0, zy, R, -1);
        add_loop(B, 2, 1, ny, 1, ny, -1, 0, nx - 1, 0, nz - 1, R - 1, -1);
        add_loop(B, 0, 1, ny, 1, ny, -1, 0, nx, 0, nz - 1, R - 1, -1);
        add_loop(B, 1, 1, ny - 1, 1, ny, -1, 0, nx - 1, 0, zy, R, -1);
        // stresses (update_stress.cpp:135-309)
        const int flag = (top && c.top_bc) ? 1 : 0;
        add_loop(S, 3, 0, 0, -1, -1, 1, 0, ny, 0, nz, R - 1, 1);
        add_loop(S, 7, 0, 0, -1, 0, 1, 0, ny, 0, nz - 1, R - 1, 1);
        add_loop(S, 8, 0, 0, -1, 0, 1, 0, ny - 1, 0, nz, R - 1, 1);
        add_loop(S, 3, 0, nx - 1, 1, nx, -1, 0, ny, 0, nz, R - 1, 1);
        add_loop(S, 7, 0, nx, 1, nx, -1, 0, ny, 0, nz - 1, R - 1, 1);
        add_loop(S, 8, 0, nx, 1, nx, -1, 0, ny - 1, 0, nz, R - 1, 1);
        add_loop(S, 4, 1, 0, -1, 0, 1, 0, nx - 1, 0, nz, R - 1, 1);
        add_loop(S, 6, 1, 0, -1, -1, 1, 0, nx - 1, 0, nz - 1, R - 1, 1);
        add_loop(S, 8, 1, 0, -1, -1, 1, 0, nx, 0, nz, R - 1, 1);
        add_loop(S, 4, 1, ny, 1, ny, -1, 0, nx - 1, 0, nz, R - 1, 1);
        add_loop(S, 6, 1, ny - 1, 1, ny, -1, 0, nx - 1, 0, nz - 1, R - 1, 1);
        add_loop(S, 8, 1, ny - 1, 1, ny, -1, 0, nx, 0, nz, R - 1, 1);
        if (top) {
            add_loop(S, 5, 2, 0, -1, 0, 1, 0, nx - 1, 1, ny - 1, R - 1, 1);
            add_loop(S, 6, 2, 0, -1, -1, 1, 0, nx - 1, 0, ny - 1, R - 1 + flag, flag ? -1 : 1);
            add_loop(S, 7, 2, 0, -1, -1, 1, 1, nx - 1, 1, ny - 1, R - 1 + flag, flag ? -1 : 1);
        }
        if (bottom) {
            add_loop(S, 5, 2, nz, 1, nz, -1, 0, nx - 1, 1, ny - 1, R - 1, 1);
            add_loop(S, 6, 2, nz - 1, 1, nz, -1, 0, nx - 1, 0, ny - 1, R - 1, 1);
            add_loop(S, 7, 2, nz - 1, 1, nz, -1, 1, nx - 1, 1, ny - 1, R - 1, 1);
        }
        d.rblk = ((nx + BX) / BX) * ((ny + RBY) / RBY) * ((nz + RZC) / RZC);
    }

    void mirror_launch(DevGrid<T> &d, const MirrorSet &m)
    {
        const long n = m.start[m.n];
        if (n > 0) hipLaunchKernelGGL((k_mirror<T>), dim3((n + 255) / 256), dim3(256), 0, s, d.dev, m);
    }

    template <int R>
    void vel_r(DevGrid<T> &d, bool en, const int *srcp)
    {
        const dim3 blk(BX, RBY);
        if (en) hipLaunchKernelGGL((k_vel_r<T, R, true>), dim3(d.rblk), blk, 0, s, d.dev, d.rco, d.ke, srcp, d_src_old);
        else hipLaunchKernelGGL((k_vel_r<T, R, false>), dim3(d.rblk), blk, 0, s, d.dev, d.rco, d.ke, srcp, d_src_old);
    }
    template <int R>
    void stress_r(DevGrid<T> &d, bool en)
    {
        const dim3 blk(BX, RBY);
        if (en) hipLaunchKernelGGL((k_stress_r<T, R, true>), dim3(d.rblk), blk, 0, s, d.dev, d.rco, d.se);
        else hipLaunchKernelGGL((k_stress_r<T, R, false>), dim3(d.rblk), blk, 0, s, d.dev, d.rco, d.se);
    }
    // stencil updates only (no mirrors), radius > 1
    void vel_radius(DevGrid<T> &d, bool en, const int *srcp)
    {
        switch (cfg.radius) {
        case 2: vel_r<2>(d, en, srcp); break;
        case 3: vel_r<3>(d, en, srcp); break;
        case 4: vel_r<4>(d, en, srcp); break;
        case 5: vel_r<5>(d, en, srcp); break;
        default: vel_r<6>(d, en, srcp); break;
        }
    }
    void stress_radius(DevGrid<T> &d, bool en)
    {
        switch (cfg.radius) {
        case 2: stress_r<2>(d, en); break;
        case 3: stress_r<3>(d, en); break;
        case 4: stress_r<4>(d, en); break;
        case 5: stress_r<5>(d, en); break;
        default: stress_r<6>(d, en); break;
        }
    }

    // stencil tiling: BX x by threads per workgroup, zc planes per thread
    // tile_byv (ECSG_TILE_BYV): the overlapped step's velocity sweeps in
    // 64 x tile_byv workgroups (0: the stress sweeps' height).  On the
    // three-grid workload (1207 x 1207 x 81 finest grid, FP32) 64 x 2 takes
    // the step from 4.31 to 4.29 ms, but only by moving the Jacobi solve's
    // work from the velocity sweep to the stress sweep: beside two-wavefront
    // workgroups the grid-barrier launch becomes resident only as the
    // velocity sweep ends (velocity sweep 1.39 -> 1.26 ms, stress sweep
    // 2.17 -> 2.27 ms).  Off by default.
    int tile_by = 4, tile_zc = ZCHUNK_MODEL, tile_byv = 0;
    static void set_tiling(DevGrid<T> &d, int by, int zc)
    {
        GridDev<T> &v = d.dev;
        v.zc = zc;
        v.gx = (d.n[0] + 1 + BX - 1) / BX;
        v.gy = (d.n[1] + 1 + by - 1) / by;
        const int gz = (d.n[2] + 1 + zc - 1) / zc;
        v.nblk = v.gx * v.gy * gz;
        v.zb = 0;
        v.ze = d.n[2] + 1;
        d.by = by;
    }

    bool all_params() const
    {
        for (auto &g : G)
            if (!g.params) return false;
        return true;
    }

    // ReshapeField (field.cpp:428-476): 8-term sum in the reference order, /8
    static void reshape(const double *prim, const int n[3], const int pos[3], std::vector<double> &out)
    {
        const long px = n[0] + 1, py = n[1] + 1;
        const int e0 = pos[0] ? n[0] - 1 : n[0], e1 = pos[1] ? n[1] - 1 : n[1], e2 = pos[2] ? n[2] - 1 : n[2];
        out.assign((size_t)(e0 + 1) * (e1 + 1) * (e2 + 1), 0.0);
        auto PR = [&](int i, int j, int k) { return prim[((long)k * py + j) * px + i]; };
        long q = 0;
        for (int k = 0; k <= e2; k++)
            for (int j = 0; j <= e1; j++)
                for (int i = 0; i <= e0; i++, q++) {
                    if (!pos[0] && !pos[1] && !pos[2]) { out[q] = PR(i, j, k); continue; }
                    out[q] = (PR(i, j, k) + PR(i, j, k + pos[2]) + PR(i, j + pos[1], k) + PR(i, j + pos[1], k + pos[2])
                              + PR(i + pos[0], j, k) + PR(i + pos[0], j, k + pos[2]) + PR(i + pos[0], j + pos[1], k)
                              + PR(i + pos[0], j + pos[1], k + pos[2])) / 8;
                }
    }

    // copy a domain-only array (no halo) of position pos into device array a
    int upload_domain(int g, int a, const int pos[3], const std::vector<double> &v)
    {
        DevGrid<T> &d = G[g];
        const int e0 = pos[0] ? d.n[0] - 1 : d.n[0], e1 = pos[1] ? d.n[1] - 1 : d.n[1], e2 = pos[2] ? d.n[2] - 1 : d.n[2];
        std::vector<T> buf((size_t)d.astride, (T)0);
        long q = 0;
        for (int k = 0; k <= e2; k++)
            for (int j = 0; j <= e1; j++)
                for (int i = 0; i <= e0; i++, q++) buf[d.o + (long)k * d.S2 + (long)j * d.P + i] = (T)v[q];
        HIPCHK(hipMemcpy(d.arr(a), buf.data(), sizeof(T) * d.astride, hipMemcpyHostToDevice));
        return ECSG_OK;
    }

    int set_params(int g, const double *buoy, const double *const stiff[9], int iso) override
    {
        coef_ok = false;
        if (g < 0 || g >= (int)G.size()) return fail(ECSG_EINVAL, "grid index out of range");
        if (!buoy || !stiff) return fail(ECSG_EINVAL, "null parameter array");
        DevGrid<T> &d = G[g];
        const long np = (long)(d.n[0] + 1) * (d.n[1] + 1) * (d.n[2] + 1);
        for (long q = 0; q < np; q++)
            if (!(buoy[q] > 0) || !std::isfinite(buoy[q])) return fail(ECSG_EINVAL, "buoyancy must be positive and finite");
        std::vector<double> prim((size_t)np), out;
        static const int p0[3] = {0, 0, 0};
        for (int c = 0; c < 9; c++) {
            if (iso) {
                if (!stiff[0] || !stiff[4]) return fail(ECSG_EINVAL, "isotropic input needs stiff[0] (s11) and stiff[4] (s13)");
                for (long q = 0; q < np; q++) {
                    if (c >= 6) prim[q] = stiff[0][q] - stiff[4][q];
                    else if (c < 3) prim[q] = stiff[0][q];
                    else prim[q] = stiff[4][q];
                }
            } else {
                if (!stiff[c]) return fail(ECSG_EINVAL, "null stiffness array");
                for (long q = 0; q < np; q++) prim[q] = stiff[c][q];
            }
            for (long q = 0; q < np; q++)
                if (!std::isfinite(prim[q])) return fail(ECSG_EINVAL, "stiffness must be finite");
            reshape(prim.data(), d.n, ST_POS[c], out);
            if (c == 0 || c == 1 || c == 2 || c >= 6)
                for (double v : out)
                    if (!(v > 0)) return fail(ECSG_EINVAL, "c11, c22, c33, c44, c55, c66 must be positive");
            if (int rc = upload_stiff(g, c, out)) return rc;
        }
        d.staged = false;
        return finish_params(g, buoy, iso);
    }

    // stiffness c (staged values at ST_POS[c], domain extents) into the device
    int upload_stiff(int g, int c, const std::vector<double> &out)
    {
        DevGrid<T> &d = G[g];
        if (int rc = upload_domain(g, 9 + c, ST_POS[c], out)) return rc;
        if (c == 2) {
            const long pl = (long)d.n[0] * (d.n[1] + 1);
            d.c33_top.assign(out.begin(), out.begin() + pl);
            d.c33_bot.assign(out.end() - pl, out.end());
            for (auto &x : d.c33_top) x = (double)(T)x;
            for (auto &x : d.c33_bot) x = (double)(T)x;
        }
        return ECSG_OK;
    }

    // buoyancy (primary nodes) and everything that follows from a grid's parameters
    int finish_params(int g, const double *buoy, int iso)
    {
        DevGrid<T> &d = G[g];
        static const int p0[3] = {0, 0, 0};
        const long np = (long)(d.n[0] + 1) * (d.n[1] + 1) * (d.n[2] + 1);
        std::vector<double> b(buoy, buoy + np);
        if (int rc = upload_domain(g, 18, p0, b)) return rc;
        const long pl = (long)(d.n[0] + 1) * (d.n[1] + 1);
        d.b_top.assign(b.begin(), b.begin() + pl);
        d.b_bot.assign(b.end() - pl, b.end());
        for (auto &x : d.b_top) x = (double)(T)x;
        for (auto &x : d.b_bot) x = (double)(T)x;
        d.iso = iso ? 1 : 0;
        d.dev.iso = d.iso;
        d.params = true;
        grad_ready = false;
        for (int k = 0; k < (int)I.size(); k++)
            if ((k == g || k + 1 == g) && G[k].params && G[k + 1].params)
                if (int rc = build_iface(k)) return rc;
        invalidate_graph();
        return ECSG_OK;
    }

    // The parameters main.cpp holds after InitParField
    // (init_del_par_field.cpp:77-150): stiff[m] = stiffs_ptr[g][m], the Field
    // of stiffness m at its staggered position ST_POS[m] (halo 0, x fastest,
    // n or n+1 entries per dimension as Field::Allocate sizes them,
    // field.cpp:85-98), and buoys[g] at the primary nodes.  The values are
    // taken as they are (no averaging).  The isotropic compliance of the
    // energy (init_del_par_field.cpp:259-278) is used when stiff[0..2] are one
    // Field and stiff[3..5] another, as InitParField's pointers are.
    int set_params_staged(int g, const double *buoy, const double *const stiff[9]) override
    {
        coef_ok = false;
        if (g < 0 || g >= (int)G.size()) return fail(ECSG_EINVAL, "grid index out of range");
        if (!buoy || !stiff) return fail(ECSG_EINVAL, "null parameter array");
        DevGrid<T> &d = G[g];
        const long np = (long)(d.n[0] + 1) * (d.n[1] + 1) * (d.n[2] + 1);
        for (long q = 0; q < np; q++)
            if (!(buoy[q] > 0) || !std::isfinite(buoy[q])) return fail(ECSG_EINVAL, "buoyancy must be positive and finite");
        for (int c = 0; c < 9; c++)
            if (!stiff[c]) return fail(ECSG_EINVAL, "null stiffness array");
        std::vector<double> out;
        for (int c = 0; c < 9; c++) {
            const int *pos = ST_POS[c];
            const long len = (long)(pos[0] ? d.n[0] : d.n[0] + 1) * (pos[1] ? d.n[1] : d.n[1] + 1) *
                             (pos[2] ? d.n[2] : d.n[2] + 1);
            out.assign(stiff[c], stiff[c] + len);
            for (double v : out) {
                if (!std::isfinite(v)) return fail(ECSG_EINVAL, "stiffness must be finite");
                if ((c <= 2 || c >= 6) && !(v > 0)) return fail(ECSG_EINVAL, "c11, c22, c33, c44, c55, c66 must be positive");
            }
            if (int rc = upload_stiff(g, c, out)) return rc;
        }
        const int iso = stiff[0] == stiff[1] && stiff[1] == stiff[2] && stiff[3] == stiff[4] && stiff[4] == stiff[5];
        d.staged = true;
        return finish_params(g, buoy, iso);
    }

    static double hat(int a, int r) { return (double)(r - std::abs(a)) / (double)r; }

    // Time-invariant interface matrices (derived ratio-r forms of
    // update_ghost_vz.cpp:139-150 and update_ghost_sxz.cpp:96-133), built
    // from the stored (T-rounded) parameters, prefactored once.
    int build_iface(int k)
    {
        DevIface &f = I[k];
        DevGrid<T> &F = G[k], &C = G[k + 1];
        const int r = f.r, nx2 = f.nx2, ny2 = f.ny2;
        const double k1z = (double)F.dev.cz, k2z = (double)C.dev.cz;
        // vz-szz: tridiagonal per x-line; c33 at (1,0,0): plane index i + j*nx
        const int nl = ny2 - 1;
        std::vector<double> M((size_t)nl * nx2), Dp((size_t)nl * nx2), U((size_t)nl * nx2);
        const double r3 = (double)(r * r * r);
        const long nxF = F.n[0], nxC = C.n[0];
        for (int Ii = 0; Ii < nx2; Ii++) {
            std::vector<double> L(nl), D(nl), Uu(nl);
            for (int J = 1; J <= ny2 - 1; J++) {
                const int q = J - 1;
                double dd = 0, lo = 0, up = 0;
                for (int b = -(r - 1); b <= r - 1; b++) {
                    const int j = r * J + b;
                    const double wj = (double)(r - std::abs(b));
                    const double pj = hat(b, r);
                    for (int a = 0; a < r; a++) {
                        const int i = r * Ii + a;
                        const double cc = F.c33_bot[(long)j * nxF + i];
                        dd += wj * pj * cc;
                        if (b < 0) lo += wj * hat(b + r, r) * cc;
                        if (b > 0) up += wj * hat(b - r, r) * cc;
                    }
                }
                D[q] = r3 * k2z * C.c33_top[(long)J * nxC + Ii] + k1z * dd;
                L[q] = k1z * lo;
                Uu[q] = k1z * up;
            }
            // elimination factors (the reference recomputes them every step,
            // update_ghost_vz.cpp:211-216; they depend on time-invariant data)
            std::vector<double> Dd(D);
            std::vector<double> mm(nl, 0.0);
            for (int q = 1; q < nl; q++) {
                const double m = L[q] / Dd[q - 1];
                Dd[q] -= m * Uu[q - 1];
                mm[q] = m;
            }
            for (int q = 0; q < nl; q++) {
                M[(size_t)q * nx2 + Ii] = mm[q];
                Dp[(size_t)q * nx2 + Ii] = Dd[q];
                U[(size_t)q * nx2 + Ii] = Uu[q];
            }
        }
        HIPCHK(hipMemcpy(f.M, M.data(), sizeof(double) * M.size(), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(f.Dp, Dp.data(), sizeof(double) * Dp.size(), hipMemcpyHostToDevice));
        // the forward back substitution multiplies by the pivot reciprocals
        // (one dependent multiply per row instead of a division)
        std::vector<double> Rp(Dp.size());
        for (size_t q = 0; q < Dp.size(); q++) Rp[q] = 1.0 / Dp[q];
        HIPCHK(hipMemcpy(f.Rp, Rp.data(), sizeof(double) * Rp.size(), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(f.U, U.data(), sizeof(double) * U.size(), hipMemcpyHostToDevice));
        // vx-sxz: 9 bands on interior coarse nodes; buoyancy at primary nodes
        const int n0 = nx2 - 1, n1 = ny2 - 1;
        const long nn = (long)n0 * n1;
        std::vector<double> A((size_t)9 * nn, 0.0);
        const double r2 = (double)(r * r), r4 = r2 * r2;
        const long pxF = F.n[0] + 1, pxC = C.n[0] + 1;
        for (int J = 1; J <= ny2 - 1; J++)
            for (int Ii = 1; Ii <= nx2 - 1; Ii++) {
                const long q = (long)(J - 1) * n0 + (Ii - 1);
                double band[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
                for (int b = -(r - 1); b <= r - 1; b++)
                    for (int a = -(r - 1); a <= r - 1; a++) {
                        const int i = r * Ii + a, j = r * J + b;
                        const double w = (double)((r - std::abs(a)) * (r - std::abs(b)));
                        const double bb = F.b_bot[(long)j * pxF + i];
                        for (int dJ = -1; dJ <= 1; dJ++)
                            for (int dI = -1; dI <= 1; dI++) {
                                const int Ic = Ii + dI, Jc = J + dJ;
                                const int ai = i - r * Ic, bj = j - r * Jc;
                                if (std::abs(ai) >= r || std::abs(bj) >= r) continue;
                                if (Ic < 1 || Ic > nx2 - 1 || Jc < 1 || Jc > ny2 - 1) continue;
                                const double p = hat(ai, r) * hat(bj, r);
                                band[3 * (dI + 1) + (dJ + 1)] += w * p * bb;
                            }
                    }
                for (int m = 0; m < 9; m++) A[(size_t)m * nn + q] = k1z * band[m];
                A[(size_t)4 * nn + q] += r4 * C.b_top[(long)J * pxC + Ii] * k2z;
            }
        HIPCHK(hipMemcpy(f.A, A.data(), sizeof(double) * A.size(), hipMemcpyHostToDevice));
        // Gershgorin bound of the Jacobi iteration matrix D^-1 (A - D): the
        // Chebyshev recurrence needs an upper bound of its spectral radius
        double rho = 0;
        for (long q = 0; q < nn; q++) {
            double off = 0;
            for (int m = 0; m < 9; m++)
                if (m != 4) off += std::fabs(A[(size_t)m * nn + q]);
            rho = std::max(rho, off / A[(size_t)4 * nn + q]);
        }
        f.rho = rho;
        if (cfg.sxz_solver == 3) {
            std::vector<double> W;
            if (!sxz_schur_inverses(A, n0, n1, W))
                return fail(ECSG_EINVAL, "interface " + std::to_string(k) + ": singular vx-sxz block in the direct factorisation");
            HIPCHK(hipMemcpy(f.W, W.data(), sizeof(double) * W.size(), hipMemcpyHostToDevice));
        }
        f.ready = true;
        return ECSG_OK;
    }

    int set_source(int g, int ix, int iy, int iz, int type, const float *tr, int len) override
    {
        if (g < 0 || g >= (int)G.size()) return fail(ECSG_EINVAL, "source grid out of range");
        const DevGrid<T> &d = G[g];
        if (ix < 0 || iy < 0 || iz < 0 || ix > d.n[0] || iy > d.n[1] || iz > d.n[2])
            return fail(ECSG_EINVAL, "source location outside the grid");
        if (type != 0 && type != 1) return fail(ECSG_EINVAL, "source type must be 0 (vx force) or 1 (explosive)");
        if (len < 0 || (len > 0 && !tr)) return fail(ECSG_EINVAL, "bad source trace");
        hipFree(d_src_tr);
        d_src_tr = nullptr;
        if (int rc = dalloc(&d_src_tr, std::max(len, 1))) return rc;
        if (len) HIPCHK(hipMemcpy(d_src_tr, tr, sizeof(float) * len, hipMemcpyHostToDevice));
        src_g = g; src_loc[0] = ix; src_loc[1] = iy; src_loc[2] = iz; src_type = type; src_len = len;
        HIPCHK(hipMemcpy(d_src_loc, src_loc, sizeof(int) * 3, hipMemcpyHostToDevice));
        invalidate_graph();
        return ECSG_OK;
    }

    int set_receivers(int n, const int *loc) override
    {
        if (n < 0 || (n > 0 && !loc)) return fail(ECSG_EINVAL, "bad receiver list");
        for (int q = 0; q < n; q++) {
            const int g = loc[4 * q];
            if (g < 0 || g >= (int)G.size()) return fail(ECSG_EINVAL, "receiver grid out of range");
            const int *nn = G[g].n;
            for (int d = 0; d < 3; d++)
                if (loc[4 * q + 1 + d] < 0 || loc[4 * q + 1 + d] > nn[d]) return fail(ECSG_EINVAL, "receiver outside its grid");
        }
        hipFree(d_rec); hipFree(d_traces);
        d_rec = nullptr; d_traces = nullptr;
        if (int rc = dalloc(&d_rec, std::max(4 * n, 4))) return rc;
        if (n) HIPCHK(hipMemcpy(d_rec, loc, sizeof(int) * 4 * n, hipMemcpyHostToDevice));
        h_rec.assign(loc ? loc : nullptr, loc ? loc + 4 * n : nullptr);
        if (int rc = dalloc(&d_traces, (size_t)std::max(n, 1) * cfg.max_steps)) return rc;
        nrec = n;
        invalidate_graph();
        return ECSG_OK;
    }

    void invalidate_graph()
    {
        if (gexec) { hipGraphExecDestroy(gexec); gexec = nullptr; }
        if (graph) { hipGraphDestroy(graph); graph = nullptr; }
        if (gexec_f) { hipGraphExecDestroy(gexec_f); gexec_f = nullptr; }
        if (graph_f) { hipGraphDestroy(graph_f); graph_f = nullptr; }
    }

    // ----- fused step ------------------------------------------------------
    bool fused_ok() const { return use_fused && !cfg.check_energy; }
    int set_overlap(int on) override
    {
        HIPCHK(hipStreamSynchronize(s));
        overlap_on = on < 0 ? -1 : on ? 1 : 0;
        if (overlap_ok())
            if (int rc = ensure_s2()) return rc;
        invalidate_graph();
        return ECSG_OK;
    }
    int overlap_active() const override { return fused_ok() ? 0 : overlap_mode(); }

    // Longest wait at one grid barrier of the cooperative Jacobi kernels;
    // past it the kernel ends early and sync() reports ECSG_ETIMEOUT.
    int set_barrier_timeout(int ms) override
    {
        if (ms < 0) return fail(ECSG_EINVAL, "barrier timeout must be >= 0 ms");
        bar_ms = ms;
        int dev = 0, khz = 0;
        HIPCHK(hipGetDevice(&dev));
        if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess || khz <= 0)
            khz = 100000;  // the 100 MHz constant clock of CDNA devices
        bar_limit = (unsigned long long)ms * (unsigned long long)khz;
        return ECSG_OK;
    }
    // Exit tolerance of the forward Jacobi sweeps, tol = min(cap, fac *
    // AvgAbs(B)) as update_ghost_sxz.cpp:229-236 forms it; the defaults are
    // that line's constants of the storage precision (1e-3, 1e-4 single;
    // 1e-12, 1e-5 double), the solve itself runs in double either way
    int set_jacobi_tol(double cap, double fac) override
    {
        if (!(cap >= 0) || !(fac >= 0) || !std::isfinite(cap) || !std::isfinite(fac))
            return fail(ECSG_EINVAL, "Jacobi tolerance factors must be finite and >= 0");
        tol_cap = cap;
        tol_fac = fac;
        invalidate_graph();
        return ECSG_OK;
    }
    // A timed-out grid barrier: clear the state for the next launch and
    // report it (the solve it belonged to ended early, so the state is wrong)
    int check_barrier()
    {
        if (!d_bar) return ECSG_OK;
        unsigned h[3] = {0, 0, 0};
        HIPCHK(hipMemcpy(h, d_bar, sizeof(h), hipMemcpyDeviceToHost));
        if (!h[2]) return ECSG_OK;
        HIPCHK(hipMemset(d_bar, 0, 4 * sizeof(unsigned)));
        return fail(ECSG_ETIMEOUT, "a grid barrier of the cooperative Jacobi solve timed out; the wavefield is invalid");
    }
    // Each cooperative launch of the step has at most one workgroup per
    // compute unit.  Only then may it run beside the stencils on the other
    // stream: its workgroups are placed as the stencil workgroups leave, one
    // per CU, while a launch with several per CU (e.g. 757 tiles of 16 x 16
    // unknowns on 256 CUs) waits at its first barrier for room the stencils
    // hold; such launches get the GPU to themselves (serial step).
    bool coop_one_per_cu() const
    {
        if (no_grid_barrier()) return true;
        for (int k0 = 0; k0 < (int)I.size(); k0 += MAXI) {
            long tot = 0;
            for (int k = k0; k < std::min((int)I.size(), k0 + MAXI); k++)
                tot += cfg.sxz_solver == 0 ? I[k].tiles : I[k].coop;
            if (tot > ncu) return false;
        }
        return true;
    }

    int set_fused(int on) override
    {
        if (on && cfg.imode == 0 && !I.empty()) return fail(ECSG_EINVAL, "the fused pass runs the energy-conserving interface (imode 1) only");
        if (on && cfg.radius != 1) return fail(ECSG_EINVAL, "the fused pass is radius 1");
        HIPCHK(hipStreamSynchronize(s));
        if (on)
            for (auto &d : G)
                if (!d.alt) {
                    if (int rc = dalloc(&d.alt, (size_t)d.astride * 9)) return rc;
                    HIPCHK(hipMemcpy(d.alt, d.fields(), sizeof(T) * 9 * d.astride, hipMemcpyDeviceToDevice));
                }
        use_fused = on ? 1 : 0;
        invalidate_graph();
        return ECSG_OK;
    }

    FusedDev<T> fused_dev(int g)
    {
        DevGrid<T> &d = G[g];
        FusedDev<T> a{};
        a.g = d.dev;
        T *o = d.fset(d.cur ^ 1);
        for (int f = 0; f < 9; f++) a.o[f] = o + (long)f * d.astride;
        a.defer0 = g > 0;
        a.defern = g + 1 < (int)G.size();
        a.zc = fused_zc > 0 ? fused_zc : d.dev.zc;
        {
            const int nx = d.n[0], ny = d.n[1], nz = d.n[2];
            a.gx = (nx + 1 + 63) / 64;
            a.gy = (ny + 1 + 4 * fused_form - 1) / (4 * fused_form);
            // k_fused_lean's planes [1, nz-2), columns [64, 64 (gxl+1)) and
            // rows [4 FRL, 4 FRL (gyl+1)): every halo entry inside the walls
            a.zlo = nz >= 4 ? 1 : 0;
            a.zhi = nz >= 4 ? nz - 2 : 0;
            a.nzc = (a.zhi - a.zlo + a.zc - 1) / a.zc;
            a.gz = a.nzc + 2;
            const int gxl = nx >= 129 ? (nx - 65) / 64 : 0;
            const int gyl = std::max(0, (ny - 1) / (4 * FRL) - 1);
            const bool lean = fused_lean && gxl > 0 && gyl > 0 && a.nzc > 0 &&
                              (double)d.astride * sizeof(T) < 4294967296.0;
            a.cx0 = 64; a.cx1 = lean ? 64 * (gxl + 1) : 64;
            a.ry0 = 4 * FRL; a.ry1 = lean ? 4 * FRL * (gyl + 1) : 4 * FRL;
        }
        a.src_p = -1;
        if (g == src_g && src_len > 0) {
            a.src_p = d.o + (long)src_loc[2] * d.S2 + (long)src_loc[1] * d.P + src_loc[0];
            a.src_type = src_type;
            a.src_len = src_len;
            a.src_tr = d_src_tr;
        }
        a.step = d_step;
        return a;
    }

    int sync_sets()
    {
        for (auto &d : G)
            HIPCHK(hipMemcpyAsync(d.fset(d.cur ^ 1), d.fields(), sizeof(T) * 9 * d.astride, hipMemcpyDeviceToDevice, s));
        sets_dirty = false;
        return ECSG_OK;
    }

    // one time step with the fused pass (main.cpp:165-261, same results as
    // enqueue_step): fused velocity+stress into the other field set,
    // receivers/source, vz-szz ghosts, deferred interface-plane stresses,
    // then the new set becomes current for the sxz and syz ghosts
    int enqueue_fused_step()
    {
        const int ng = (int)G.size();
        for (int g = 0; g < ng; g++) {
            ev_begin();
            FusedDev<T> a = fused_dev(g);
            const unsigned nt = (unsigned)(a.gx * a.gy * a.gz);
            if (a.cx1 > a.cx0) {
                const DevGrid<T> &d = G[g];
                FusedLean<T> L{};
                L.A = a.g.f[0];
                L.B = a.o[0];
                L.Pm = d.base + 9 * d.astride;
                L.as = (unsigned)d.astride; L.P = (unsigned)d.P; L.S2 = (unsigned)d.S2; L.o = (unsigned)d.o;
                L.cx = d.dev.cx; L.cy = d.dev.cy; L.cz = d.dev.cz;
                L.bx0 = 1; L.by0 = 1;
                L.gxl = a.cx1 / 64 - 1; L.gyl = a.ry1 / (4 * FRL) - 1;
                L.zlo = a.zlo; L.zhi = a.zhi; L.zc = a.zc; L.nzc = a.nzc;
                L.src_p = a.src_p >= 0 ? (unsigned)a.src_p : 0xffffffffu;
                L.src_type = a.src_type; L.src_len = a.src_len; L.src_tr = a.src_tr; L.step = a.step;
                hipLaunchKernelGGL((k_fused_lean<T, FRL>), dim3((unsigned)(L.gxl * L.gyl * L.nzc)), dim3(256), 0, s, L);
            }
            if (fused_form == 2) hipLaunchKernelGGL((k_fused_w<T, 2>), dim3(nt), dim3(256), 0, s, a);
            else if (fused_form == 8) hipLaunchKernelGGL((k_fused_w<T, 8>), dim3(nt), dim3(256), 0, s, a);
            else hipLaunchKernelGGL((k_fused_w<T, 4>), dim3(nt), dim3(256), 0, s, a);
            ev_end(g == 0 ? 0 : 2);
        }
        HIPCHK(hipGetLastError());
        {
            PointArgs<T> pa = point_args();
            bool sdefer = false;
            if (src_g >= 0 && src_type == 1)
                sdefer = (src_g > 0 && src_loc[2] == 0) || (src_g + 1 < ng && src_loc[2] == G[src_g].n[2]);
            for (int g = 0; g < ng; g++) {
                pa.vx[g] = G[g].fset(G[g].cur ^ 1);
                pa.st[g][0] = pa.st[g][1] = pa.st[g][2] = pa.st[g][3] = nullptr;
            }
            pa.skip_s = sdefer ? 0 : 1;
            pa.energy = 0;
            ev_begin();
            hipLaunchKernelGGL((k_recsrc<T>), dim3(1), dim3(256), 0, s, pa);
            ev_end(5);
        }
        for (int k = 0; k + 1 < ng; k++) {
            DevIface &f = I[k];
            IfaceDev<T> a = iface_dev(k);
            for (int c = 0; c < 3; c++) {
                a.F.f[c] = G[k].fset(G[k].cur ^ 1) + (long)c * G[k].astride;
                a.C.f[c] = G[k + 1].fset(G[k + 1].cur ^ 1) + (long)c * G[k + 1].astride;
            }
            const long nl = (long)f.nx2 * (f.ny2 - 1);
            ev_begin();
            hipLaunchKernelGGL((k_gvz_rhs<T>), dim3((nl + 255) / 256), dim3(256), 0, s, a, f.B);
            launch_vz_solve(a, f);
            const long nf = (long)f.r * f.nx2 * (f.r * f.ny2 - 1);
            hipLaunchKernelGGL((k_gvz_fine<T>), dim3((nf + 255) / 256), dim3(256), 0, s, a);
            ev_end(4);
        }
        for (int g = 0; g < ng; g++) {
            DevGrid<T> &d = G[g];
            T *o = d.fset(d.cur ^ 1);
            const long as = d.astride;
            const int planes[2] = {g > 0 ? 0 : -1, g + 1 < ng ? d.n[2] : -1};
            for (int pl : planes)
                if (pl >= 0) {
                    const long nt = (long)d.n[0] * (d.n[1] + 1);
                    ev_begin();
                    hipLaunchKernelGGL((k_stress_defer<T>), dim3((nt + 255) / 256), dim3(256), 0, s, d.dev, o, o + as, o + 2 * as,
                                       o + 3 * as, o + 4 * as, o + 5 * as, pl);
                    ev_end(4);
                }
        }
        for (auto &d : G) {
            d.cur ^= 1;
            for (int f = 0; f < 9; f++) d.dev.f[f] = d.arr(f);
        }
        if (ng > 1) {
            ev_begin();
            for (int k = 0; k + 1 < ng; k++)
                if (int rc = launch_sxz_rhs(k)) return rc;
            if (int rc = launch_sxz_solves()) return rc;
            ev_end(4);
        }
        for (int k = 0; k + 1 < ng; k++)
            if (int rc = launch_ghost_syz(k)) return rc;
        HIPCHK(hipGetLastError());
        return ECSG_OK;
    }

    int step_fused(int n, int use_graph)
    {
        if (sets_dirty)
            if (int rc = sync_sets()) return rc;
        int left = n;
        const bool can_graph = use_graph && !graph_failed && !prof && (I.empty() || no_grid_barrier());
        if (can_graph) {
            if (G[0].cur == 1 && left > 0) {
                if (int rc = enqueue_fused_step()) return rc;
                left--;
            }
            if (left >= 2 && !gexec_f) {
                // a pair of steps returns the field sets to their places
                if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) == hipSuccess) {
                    int rc = enqueue_fused_step();
                    if (rc == ECSG_OK) rc = enqueue_fused_step();
                    hipGraph_t gr = nullptr;
                    const hipError_t e2 = hipStreamEndCapture(s, &gr);
                    if (rc == ECSG_OK && e2 == hipSuccess && gr && hipGraphInstantiate(&gexec_f, gr, nullptr, nullptr, 0) == hipSuccess) {
                        graph_f = gr;
                    } else {
                        if (gr) hipGraphDestroy(gr);
                        gexec_f = nullptr;
                        graph_failed = true;
                        hipGetLastError();
                        st = Status{};
                        // the capture toggled the host-side sets twice: unchanged
                    }
                } else {
                    graph_failed = true;
                    hipGetLastError();
                }
            }
            if (gexec_f)
                for (; left >= 2; left -= 2) HIPCHK(hipGraphLaunch(gexec_f, s));
        }
        for (; left > 0; left--)
            if (int rc = enqueue_fused_step()) return rc;
        host_step += n;
        return ECSG_OK;
    }

    // ----- launches --------------------------------------------------------
    // events come from a pool created before the measured steps, so the
    // instrumented launches cost one hipEventRecord each
    size_t ev_next = 0;
    hipEvent_t next_event()
    {
        if (ev_next < evs.size()) return evs[ev_next++];
        hipEvent_t e;
        hipEventCreate(&e);
        evs.push_back(e);
        ev_next++;
        return e;
    }
    // mark: time this launch also outside profile() (step_timed)
    void ev_begin(bool mark = false)
    {
        if (!prof && !mark) return;
        hipEventRecord(next_event(), s);
    }
    void ev_end(int kind, bool mark = false)
    {
        if (!prof && !mark) return;
        hipEventRecord(next_event(), s);
        ev_kind.push_back(kind);
    }

    template <int BYv>
    void vel_launch(const GridDev<T> &v, double *ke, bool en, const int *srcp)
    {
        const dim3 gr(v.nblk), bl(BX, BYv);
        if (fits32(v)) {
            if (en) hipLaunchKernelGGL((k_vel<T, true, BYv, unsigned>), gr, bl, 0, s, v, ke, srcp, d_src_old);
            else hipLaunchKernelGGL((k_vel<T, false, BYv, unsigned>), gr, bl, 0, s, v, ke, srcp, d_src_old);
        } else {
            if (en) hipLaunchKernelGGL((k_vel<T, true, BYv, long>), gr, bl, 0, s, v, ke, srcp, d_src_old);
            else hipLaunchKernelGGL((k_vel<T, false, BYv, long>), gr, bl, 0, s, v, ke, srcp, d_src_old);
        }
    }
    template <int BYv>
    void stress_launch(const GridDev<T> &v, double *se, bool en)
    {
        const dim3 gr(v.nblk), bl(BX, BYv);
        if (fits32(v)) {
            if (en) hipLaunchKernelGGL((k_stress<T, true, BYv, unsigned>), gr, bl, 0, s, v, se);
            else hipLaunchKernelGGL((k_stress<T, false, BYv, unsigned>), gr, bl, 0, s, v, se);
        } else {
            if (en) hipLaunchKernelGGL((k_stress<T, true, BYv, long>), gr, bl, 0, s, v, se);
            else hipLaunchKernelGGL((k_stress<T, false, BYv, long>), gr, bl, 0, s, v, se);
        }
    }
    // radius-1 stencil of grid g over the z planes [zb, ze) only (no energy)
    int stencil_part(int g, int zb, int ze, bool vel)
    {
        if (ze <= zb) return ECSG_OK;
        DevGrid<T> &d = G[g];
        GridDev<T> v = d.dev;
        v.zb = zb;
        v.ze = ze;
        // the velocity sweep's own workgroup height (byv), if set
        const int bv = vel && d.byv > 0 ? d.byv : d.by;
        v.gy = (d.n[1] + 1 + bv - 1) / bv;
        v.nblk = v.gx * v.gy * ((ze - zb + v.zc - 1) / v.zc);
        // kinds 0 / 1: the interior launch of grid 0, the long one
        const bool main_part = g == 0 && zb == 0;
        ev_begin(live && main_part);
        if (vel) {
            switch (bv) {
            case 1: vel_launch<1>(v, d.ke, false, nullptr); break;
            case 2: vel_launch<2>(v, d.ke, false, nullptr); break;
            case 8: vel_launch<8>(v, d.ke, false, nullptr); break;
            default: vel_launch<4>(v, d.ke, false, nullptr); break;
            }
        } else {
            switch (d.by) {
            case 1: stress_launch<1>(v, d.se, false); break;
            case 2: stress_launch<2>(v, d.se, false); break;
            case 8: stress_launch<8>(v, d.se, false); break;
            default: stress_launch<4>(v, d.se, false); break;
            }
        }
        ev_end(main_part ? (vel ? 0 : 1) : (vel ? 2 : 3), live && main_part);
        HIPCHK(hipGetLastError());
        return ECSG_OK;
    }

    int launch_vel(int g)
    {
        DevGrid<T> &d = G[g];
        const bool en = cfg.check_energy != 0;
        const int *srcp = (en && src_type == 0 && g == src_g && src_len > 0) ? d_src_loc : nullptr;
        const bool mark = live && g == 0;
        ev_begin(mark);
        if (cfg.radius > 1) {
            vel_radius(d, en, srcp);
            mirror_launch(d, d.mva);  // z and x walls
            mirror_launch(d, d.mvb);  // y walls, after the x mirrors read the old y-halo row
            ev_end(g == 0 ? 0 : 2, mark);
            HIPCHK(hipGetLastError());
            return ECSG_OK;
        }
        switch (d.by) {
        case 1: vel_launch<1>(d.dev, d.ke, en, srcp); break;
        case 2: vel_launch<2>(d.dev, d.ke, en, srcp); break;
        case 8: vel_launch<8>(d.dev, d.ke, en, srcp); break;
        default: vel_launch<4>(d.dev, d.ke, en, srcp); break;
        }
        ev_end(g == 0 ? 0 : 2, mark);
        HIPCHK(hipGetLastError());
        return ECSG_OK;
    }

    int launch_stress(int g)
    {
        DevGrid<T> &d = G[g];
        const bool mark = live && g == 0;
        ev_begin(mark);
        if (cfg.radius > 1) {
            stress_radius(d, cfg.check_energy != 0);
            mirror_launch(d, d.mst);
            ev_end(g == 0 ? 1 : 3, mark);
            HIPCHK(hipGetLastError());
            return ECSG_OK;
        }
        switch (d.by) {
        case 1: stress_launch<1>(d.dev, d.se, cfg.check_energy != 0); break;
        case 2: stress_launch<2>(d.dev, d.se, cfg.check_energy != 0); break;
        case 8: stress_launch<8>(d.dev, d.se, cfg.check_energy != 0); break;
        default: stress_launch<4>(d.dev, d.se, cfg.check_energy != 0); break;
        }
        ev_end(g == 0 ? 1 : 3, mark);
        HIPCHK(hipGetLastError());
        return ECSG_OK;
    }

    // the vz-szz tridiagonal solves: one wavefront per coarse x-line when the
    // line fits in LDS, else one thread per line
    void launch_vz_solve(const IfaceDev<T> &a, const DevIface &f)
    {
        const size_t lds = (size_t)32 * (f.ny2 - 1);
        if (lds <= (size_t)VZ_WAVE_LDS)
            hipLaunchKernelGGL((k_gvz_solve_wave<T>), dim3(f.nx2), dim3(64), lds, s, a, (const double *)f.B, (const double *)f.M,
                               (const double *)f.Rp, (const double *)f.U);
        else
            hipLaunchKernelGGL((k_gvz_solve<T>), dim3((f.nx2 + 63) / 64), dim3(64), 0, s, a, f.B, f.M, f.Rp, f.U);
    }

    IfaceDev<T> iface_dev(int k) const
    {
        IfaceDev<T> a{};
        a.F = G[k].dev;
        a.C = G[k + 1].dev;
        a.r = I[k].r; a.nx2 = I[k].nx2; a.ny2 = I[k].ny2; a.nz1 = I[k].nz1;
        a.k1x = (double)a.F.cx; a.k1y = (double)a.F.cy; a.k1z = (double)a.F.cz;
        a.k2x = (double)a.C.cx; a.k2y = (double)a.C.cy; a.k2z = (double)a.C.cz;
        return a;
    }

    int launch_ghost_vz(int k)
    {
        DevIface &f = I[k];
        IfaceDev<T> a = iface_dev(k);
        if (cfg.imode == 0) {
            const long nt = (long)f.nx2 * (f.ny2 - 1) + (long)(f.nx2 - 1) * f.ny2 + f.ny2;
            ev_begin();
            hipLaunchKernelGGL((k_ii_vz<T>), dim3((nt + 255) / 256), dim3(256), 0, s, a, cfg.ii_vz_xend ? 1 : 0);
            ev_end(4);
            HIPCHK(hipGetLastError());
            return ECSG_OK;
        }
        const long nl = (long)f.nx2 * (f.ny2 - 1);
        ev_begin();
        hipLaunchKernelGGL((k_gvz_rhs<T>), dim3((nl + 255) / 256), dim3(256), 0, s, a, f.B);
        launch_vz_solve(a, f);
        const long nf = (long)f.r * f.nx2 * (f.r * f.ny2 - 1);
        hipLaunchKernelGGL((k_gvz_fine<T>), dim3((nf + 255) / 256), dim3(256), 0, s, a);
        ev_end(4);
        HIPCHK(hipGetLastError());
        return ECSG_OK;
    }

    // right-hand side of the vx-sxz system of interface k
    int launch_sxz_rhs(int k)
    {
        DevIface &f = I[k];
        IfaceDev<T> a = iface_dev(k);
        hipLaunchKernelGGL((k_gsxz_rhs<T>), dim3(f.nabs), dim3(256), 0, s, a, f.Bs, f.absp);
        HIPCHK(hipGetLastError());
        return ECSG_OK;
    }

    JacLds<T> jac_lds(int k, bool adjoint) const
    {
        const DevIface &f = I[k];
        JacLds<T> J{};
        J.A = f.A;
        if (adjoint) {
            const AdjIface &a = AI[k];
            J.B = a.Ba; J.abs_part = a.absz; J.X0 = a.X0a; J.stats = a.stats;
        } else {
            J.B = f.Bs; J.abs_part = f.absp; J.X0 = f.X0; J.stats = f.stats;
            J.sxz1 = G[k].dev.f[7]; J.sxz2 = G[k + 1].dev.f[7];
            J.sxz1s = G[k].dev.st[7]; J.sxz2s = G[k + 1].dev.st[7];
        }
        J.o1 = G[k].o; J.P1 = G[k].P; J.S21 = G[k].S2;
        J.o2 = G[k + 1].o; J.P2 = G[k + 1].P; J.S22 = G[k + 1].S2;
        J.nx2 = f.nx2; J.ny2 = f.ny2; J.nz1 = f.nz1; J.r = f.r; J.nabs = f.nabs;
        return J;
    }

    int launch_jacobi_lds(bool adjoint)
    {
        JacLdsSet<T> set{};
        const int ni = (int)I.size();
        for (int k = 0; k < ni; k++) set.it[k] = jac_lds(k, adjoint);
        JacParams jp{};
        if (adjoint) {
            jp.tol_cap = 0; jp.tol_fac = 0; jp.rel_tol = adj_rel_tol; jp.solve_only = 1;
        } else {
            jp.tol_cap = tol_cap; jp.tol_fac = tol_fac; jp.rel_tol = 0; jp.solve_only = 0;
        }
        jp.floor = 1024 * 2.220446049250313e-16;
        jp.cap = JACOBI_CAP;
        hipLaunchKernelGGL((k_jacobi_lds<T>), dim3(ni), dim3(JL), (size_t)(2 * lds_maxw * sizeof(double)), s, set, jp);
        HIPCHK(hipGetLastError());
        return ECSG_OK;
    }

    // cooperative launch of the tiled Jacobi solve over blk tiles: JTT_FEW
    // threads per tile for launches of at most JTT_FEW_TILES tiles, else JTT
    int launch_tiled(JacBatch<T> &jb, int blk)
    {
        jb.bar = d_bar;
        jb.bar_limit = bar_limit;
        void *args[] = {(void *)&jb};
        if (blk <= JTT_FEW_TILES && blk <= max_tiled_few) {
            HIPCHK(launch_cooperative((const void *)k_jacobi_tiled<T, JTT_FEW>, dim3(blk), dim3(JTT_FEW), args, s));
            return ECSG_OK;
        }
        if (blk > max_tiled) return fail(ECSG_EINVAL, "interfaces too large for one resident Jacobi launch");
        HIPCHK(launch_cooperative((const void *)k_jacobi_tiled<T, JTT>, dim3(blk), dim3(JTT), args, s));
        return ECSG_OK;
    }

    // Jacobi solves + ghost scatter of interfaces [k0, k0+n) in one launch
    int launch_jacobi(int k0, int n)
    {
        JacBatch<T> jb{};
        int blk = 0;
        for (int q = 0; q < n; q++) {
            DevIface &f = I[k0 + q];
            JacItf<T> &J = jb.it[q];
            J.A = f.A; J.B = f.Bs; J.X0 = f.X0; J.X1 = f.X1; J.abs_part = f.absp; J.nabs = f.nabs;
            J.part = f.jpart; J.stats = f.stats;
            J.sxz1 = G[k0 + q].dev.f[7]; J.sxz2 = G[k0 + q + 1].dev.f[7];
            J.sxz1s = G[k0 + q].dev.st[7]; J.sxz2s = G[k0 + q + 1].dev.st[7];
            J.o1 = G[k0 + q].o; J.P1 = G[k0 + q].P; J.S21 = G[k0 + q].S2;
            J.o2 = G[k0 + q + 1].o; J.P2 = G[k0 + q + 1].P; J.S22 = G[k0 + q + 1].S2;
            J.nx2 = f.nx2; J.ny2 = f.ny2; J.nz1 = f.nz1; J.r = f.r;
            const bool tiled = cfg.sxz_solver == 0;
            J.blk0 = blk; J.nblk = tiled ? f.tiles : f.coop; J.rho2 = f.rho * f.rho;
            blk += J.nblk;
        }
        jb.n = n;
        jb.tol_cap = tol_cap; jb.tol_fac = tol_fac;
        jb.floor = 1024 * 2.220446049250313e-16;
        jb.cap = JACOBI_CAP;
        // the Chebyshev weights need a spectral bound below one
        bool cheb_ok = cfg.sxz_solver == 1;
        for (int q = 0; q < n; q++)
            if (!(I[k0 + q].rho < 1.0)) cheb_ok = false;
        jb.cheby = cheb_ok ? 1 : 0;
        void *args[] = {(void *)&jb};
        if (cfg.sxz_solver == 0) {
            // the reference's Jacobi sweeps, SJ per grid barrier in LDS tiles
            if (int rc = launch_tiled(jb, blk)) return rc;
        } else {
            if (blk > max_coop) return fail(ECSG_EINVAL, "interfaces too large for one resident Jacobi launch");
            jb.bar = d_bar;
            jb.bar_limit = bar_limit;
            HIPCHK(launch_cooperative((const void *)k_jacobi<T>, dim3(blk), dim3(JT), args, s));
        }
        return launch_sxz_fine(jb, k0, n);
    }

    // fine ghosts of interfaces [k0, k0+n) from the coarse ghosts just
    // written (n <= MAXI; jb's ghost views filled)
    int launch_sxz_fine(const JacBatch<T> &jb, int k0, int n)
    {
        long nfmax = 0;
        for (int q = 0; q < n; q++) {
            const long r = I[k0 + q].r;
            nfmax = std::max(nfmax, (r * I[k0 + q].nx2 - 1) * (r * I[k0 + q].ny2 - 1));
        }
        if (nfmax > (1L << 31) - 256) return fail(ECSG_EINVAL, "interface too large for the fine ghost kernel");
        hipLaunchKernelGGL((k_jacobi_fine<T>), dim3((unsigned)((nfmax + 255) / 256), n), dim3(256), 0, s, jb);
        HIPCHK(hipGetLastError());
        return ECSG_OK;
    }

    // Direct vx-sxz solves of interfaces [k0, k0+n) (one workgroup each),
    // then their fine ghosts
    int launch_sxz_direct(int k0, int n)
    {
        SxzDirectSet<T> set{};
        for (int q = 0; q < n; q++) {
            DevIface &f = I[k0 + q];
            const DevGrid<T> &C = G[k0 + q + 1];
            SxzDirect<T> &D = set.it[q];
            D.A = f.A; D.W = f.W; D.B = f.Bs; D.X0 = f.X0; D.Y = f.X1;
            D.sxz2 = C.dev.f[7]; D.sxz2s = C.dev.st[7];
            D.o2 = C.o; D.P2 = C.P; D.S22 = C.S2;
            D.nx2 = f.nx2; D.ny2 = f.ny2;
        }
        hipLaunchKernelGGL((k_sxz_direct<T>), dim3(n), dim3(JD), 0, s, set);
        HIPCHK(hipGetLastError());
        for (int q0 = 0; q0 < n; q0 += MAXI) {
            const int m = std::min(MAXI, n - q0);
            JacBatch<T> jb{};
            for (int q = 0; q < m; q++) {
                const int k = k0 + q0 + q;
                JacItf<T> &J = jb.it[q];
                J.sxz1 = G[k].dev.f[7]; J.sxz2 = G[k + 1].dev.f[7];
                J.sxz1s = G[k].dev.st[7]; J.sxz2s = G[k + 1].dev.st[7];
                J.o1 = G[k].o; J.P1 = G[k].P; J.S21 = G[k].S2;
                J.o2 = G[k + 1].o; J.P2 = G[k + 1].P; J.S22 = G[k + 1].S2;
                J.nx2 = I[k].nx2; J.ny2 = I[k].ny2; J.nz1 = I[k].nz1; J.r = I[k].r;
            }
            jb.n = m;
            if (int rc = launch_sxz_fine(jb, k0 + q0, m)) return rc;
        }
        return ECSG_OK;
    }

    // the vx-sxz solves of every interface, in the configured form
    int launch_sxz_solves()
    {
        const int ni = (int)I.size();
        if (cfg.sxz_solver == 3) return launch_sxz_direct(0, ni);
        if (use_lds() && cfg.sxz_solver == 0) return launch_jacobi_lds(false);
        for (int k0 = 0; k0 < ni; k0 += MAXI)
            if (int rc = launch_jacobi(k0, std::min(MAXI, ni - k0))) return rc;
        return ECSG_OK;
    }

    int launch_ghost_sxz(int k)
    {
        if (cfg.imode == 0) {
            const DevIface &f = I[k];
            const long nt = (long)(f.nx2 + 1) * (f.ny2 + 1);
            ev_begin();
            hipLaunchKernelGGL((k_ii_sxz<T>), dim3((nt + 255) / 256), dim3(256), 0, s, iface_dev(k));
            ev_end(4);
            HIPCHK(hipGetLastError());
            return ECSG_OK;
        }
        ev_begin();
        if (int rc = launch_sxz_rhs(k)) return rc;
        if (cfg.sxz_solver == 3) {
            if (int rc = launch_sxz_direct(k, 1)) return rc;
            ev_end(4);
            return ECSG_OK;
        }
        if (use_lds() && cfg.sxz_solver == 0) {
            JacLdsSet<T> set{};
            set.it[0] = jac_lds(k, false);
            JacParams jp{};
            jp.tol_cap = tol_cap; jp.tol_fac = tol_fac; jp.floor = 1024 * 2.220446049250313e-16; jp.cap = JACOBI_CAP;
            hipLaunchKernelGGL((k_jacobi_lds<T>), dim3(1), dim3(JL), (size_t)(2 * lds_maxw * sizeof(double)), s, set, jp);
            HIPCHK(hipGetLastError());
            ev_end(4);
            return ECSG_OK;
        }
        if (int rc = launch_jacobi(k, 1)) return rc;
        ev_end(4);
        return ECSG_OK;
    }

    int launch_ghost_syz(int k)
    {
        DevIface &f = I[k];
        IfaceDev<T> a = iface_dev(k);
        const long nc = (long)f.nx2 * f.ny2;
        ev_begin();
        if (cfg.imode == 0) {
            const long nt = nc + 2L * f.nx2 + (2L * f.ny2 - 2);
            hipLaunchKernelGGL((k_ii_syz<T>), dim3((nt + 255) / 256), dim3(256), 0, s, a);
        } else {
            hipLaunchKernelGGL((k_gsyz<T>), dim3((nc + 255) / 256), dim3(256), 0, s, a);
        }
        ev_end(4);
        HIPCHK(hipGetLastError());
        return ECSG_OK;
    }

    PointArgs<T> point_args()
    {
        PointArgs<T> a{};
        for (int g = 0; g < (int)G.size(); g++) {
            a.vx[g] = G[g].dev.f[0];
            a.sxx[g] = G[g].dev.f[3]; a.syy[g] = G[g].dev.f[4]; a.szz[g] = G[g].dev.f[5];
            a.st[g][0] = G[g].dev.st[0]; a.st[g][1] = G[g].dev.st[3]; a.st[g][2] = G[g].dev.st[4]; a.st[g][3] = G[g].dev.st[5];
            a.b[g] = G[g].dev.b;
            a.o[g] = G[g].o; a.P[g] = G[g].P; a.S2[g] = G[g].S2;
            for (int d = 0; d < 3; d++) a.n[g][d] = G[g].n[d];
        }
        a.nrec = nrec; a.rec = d_rec; a.traces = d_traces; a.max_steps = cfg.max_steps; a.step = d_step;
        a.src_g = src_g < 0 ? 0 : src_g; a.src_i = src_loc[0]; a.src_j = src_loc[1]; a.src_k = src_loc[2];
        a.src_type = src_type; a.src_len = src_g < 0 ? 0 : src_len; a.src_tr = d_src_tr;
        a.src_old = d_src_old; a.fix = d_fix; a.energy = cfg.check_energy;
        return a;
    }

    int launch_recsrc()
    {
        ev_begin();
        hipLaunchKernelGGL((k_recsrc<T>), dim3(1), dim3(256), 0, s, point_args());
        ev_end(5);
        HIPCHK(hipGetLastError());
        return ECSG_OK;
    }

    int launch_energy()
    {
        EnergyArgs e{};
        for (int g = 0; g < (int)G.size(); g++) {
            e.ke[g] = G[g].ke; e.se[g] = G[g].se; e.nblk[g] = G[g].dev.nblk;
            e.vol[g] = 0.5 * G[g].h[2] * G[g].h[1] * G[g].h[0];
        }
        e.ng = (int)G.size(); e.src_g = (src_type == 0 && src_g >= 0) ? src_g : -1; e.fix = d_fix;
        e.energy = d_energy; e.step = d_step; e.max_steps = cfg.max_steps;
        ev_begin();
        hipLaunchKernelGGL(k_energy, dim3(1), dim3(64), 0, s, e);
        ev_end(5);
        HIPCHK(hipGetLastError());
        return ECSG_OK;
    }

    int check_ready()
    {
        if (!all_params()) return fail(ECSG_ESTATE, "parameters not set for every grid");
        // a multi-step call that stopped on an error may have left the last
        // ghost update unjoined
        return ov_join();
    }

    // By default the overlap goes with the eager step of the tiled Jacobi
    // solve.  With the single-workgroup solve (graph-captured steps of small
    // interfaces) its extra launches and cross-stream waits cost more than
    // they hide: 0.58 against 0.25 ms per step at a 31 x 31 iterate, and 64
    // two-grid members stepped one at a time ran at half speed.
    bool overlap_ok() const
    {
        if (G.size() < 2 || cfg.radius != 1 || cfg.imode != 1 || cfg.check_energy) return false;
        if (!coop_one_per_cu()) return false;
        if (overlap_on >= 0) return overlap_on != 0;
        return !use_lds();
    }
    // the launches of a scope go to s2 (the interface stream) instead of s
    struct OnSecond {
        Engine *e;
        explicit OnSecond(Engine *x) : e(x) { std::swap(e->s, e->s2); }
        ~OnSecond() { std::swap(e->s, e->s2); }
    };
    // the interface stream and its events, created when first needed (an
    // idle second stream per handle would take a hardware queue from the
    // handles that step side by side)
    int ensure_s2()
    {
        if (s2) return ECSG_OK;
        HIPCHK(hipStreamCreateWithFlags(&s2, hipStreamNonBlocking));
        for (auto &e : ov_ev) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        return ECSG_OK;
    }
    // s waits for the sxz/syz ghosts of the last overlapped step
    int ov_join()
    {
        if (!ov_open) return ECSG_OK;
        ov_open = false;
        HIPCHK(hipStreamWaitEvent(s, ov_ev[3], 0));
        return ECSG_OK;
    }

    // planes [a, b) of grid g excluded: the planes of [0, nz] below a and
    // from b on (all of them when a >= b)
    int stencil_rest(int g, int a, int b, bool vel)
    {
        const int n1 = G[g].n[2] + 1;
        if (a >= b) return stencil_part(g, 0, n1, vel);
        if (int rc = stencil_part(g, 0, a, vel)) return rc;
        return stencil_part(g, b, n1, vel);
    }
    // Planes of grid g away from its interfaces.  Velocity interior [vlo,
    // vhi): every plane but z = 0 below an interface and z = nz above one.
    // Stress interior [slo, shi): also without z = 1 below and z = nz-1
    // above an interface (their stencils read the velocity edge planes).
    int vlo(int g) const { return g > 0 ? 1 : 0; }
    int vhi(int g) const { return g + 1 < (int)G.size() ? G[g].n[2] : G[g].n[2] + 1; }
    int slo(int g) const { return g > 0 ? 2 : 0; }
    int shi(int g) const { return g + 1 < (int)G.size() ? G[g].n[2] - 1 : G[g].n[2] + 1; }
    // the source and every receiver lie on stress-interior planes: the
    // receiver/source kernel may run before the velocity edge planes
    bool points_interior() const
    {
        if (src_len > 0 && src_g >= 0 && (src_loc[2] < slo(src_g) || src_loc[2] >= shi(src_g))) return false;
        for (size_t q = 0; q + 3 < h_rec.size(); q += 4)
            if (h_rec[q + 3] < slo(h_rec[q]) || h_rec[q + 3] >= shi(h_rec[q])) return false;
        return true;
    }
    int overlap_mode() const { return !overlap_ok() ? 0 : points_interior() ? 2 : 1; }

    // One time step (main.cpp:165-261) with the interface updates on s2.
    // An interface reads and writes, in its upper (fine) grid, only the
    // planes z = nz-1 .. nz (and the ghost plane below), in its lower
    // (coarse) grid only z = -1 .. 0; a stencil at plane z reads planes
    // z-1 .. z+1.  So the velocity update of every plane but z = nz of an
    // upper and z = 0 of a lower grid runs beside the previous step's
    // sxz/syz ghost solves, and so does the stress update of grid 0's planes
    // z = 0 .. nz-2 when the source and the receivers lie on such interior
    // planes (mode 2: their kernel then runs right after the interior
    // velocities); the other stress interiors run beside the vz ghost
    // solves, and the planes next to each interface after the ghosts they
    // read.  Every entry sees the same operations in the same order as in
    // the serial step.  With keep_open the step ends without s waiting for
    // its sxz/syz ghosts: the next step's interior velocities may start
    // first.
    int enqueue_step_overlap(bool keep_open)
    {
        if (int rc = ensure_s2()) return rc;
        const int ng = (int)G.size();
        const bool early = overlap_mode() == 2;
        // coarse grids first: their short launches run before the previous
        // step's Jacobi solve is resident, which slows them several times
        for (int g = ng - 1; g >= 0; g--)
            if (int rc = stencil_part(g, vlo(g), vhi(g), true)) return rc;
        if (early) {
            if (int rc = launch_recsrc()) return rc;
            if (int rc = stencil_part(0, slo(0), shi(0), false)) return rc;
        }
        if (int rc = ov_join()) return rc;
        for (int g = 0; g < ng; g++)
            if (int rc = stencil_rest(g, vlo(g), vhi(g), true)) return rc;
        if (!early)
            if (int rc = launch_recsrc()) return rc;
        HIPCHK(hipEventRecord(ov_ev[0], s));
        {
            OnSecond on2(this);
            HIPCHK(hipStreamWaitEvent(s, ov_ev[0], 0));
            for (int k = 0; k + 1 < ng; k++)
                if (int rc = launch_ghost_vz(k)) return rc;
            HIPCHK(hipEventRecord(ov_ev[1], s));
        }
        for (int g = early ? 1 : 0; g < ng; g++)
            if (int rc = stencil_part(g, slo(g), shi(g), false)) return rc;
        HIPCHK(hipStreamWaitEvent(s, ov_ev[1], 0));
        for (int g = 0; g < ng; g++)
            if (int rc = stencil_rest(g, slo(g), shi(g), false)) return rc;
        HIPCHK(hipEventRecord(ov_ev[2], s));
        {
            OnSecond on2(this);
            HIPCHK(hipStreamWaitEvent(s, ov_ev[2], 0));
            if (int rc = enqueue_ghost_stress()) return rc;
            HIPCHK(hipEventRecord(ov_ev[3], s));
        }
        ov_open = true;
        return keep_open ? ECSG_OK : ov_join();
    }

    // UpdateGhostSXZ of every interface (main.cpp:242-246): right-hand
    // sides, then the Jacobi solves of up to MAXI interfaces side by side;
    // then UpdateGhostSYZ
    int enqueue_ghost_stress()
    {
        const int ng = (int)G.size();
        if (ng > 1) {
            ev_begin();
            for (int k = 0; k + 1 < ng; k++)
                if (int rc = launch_sxz_rhs(k)) return rc;
            if (int rc = launch_sxz_solves()) return rc;
            ev_end(4);
        }
        for (int k = 0; k + 1 < ng; k++)
            if (int rc = launch_ghost_syz(k)) return rc;
        return ECSG_OK;
    }

    // one time step, main.cpp:165-261
    int enqueue_step(bool keep_open = false)
    {
        if (overlap_ok()) {
            const int rc = enqueue_step_overlap(keep_open);
            if (rc != ECSG_OK && s2) {
                // a failed launch: let the interface stream drain so that no
                // later call races its remaining work
                hipStreamSynchronize(s2);
                ov_open = false;
            }
            return rc;
        }
        const int ng = (int)G.size();
        for (int g = 0; g < ng; g++)
            if (int rc = launch_vel(g)) return rc;
        if (int rc = launch_recsrc()) return rc;
        for (int k = 0; k + 1 < ng; k++)
            if (int rc = launch_ghost_vz(k)) return rc;
        for (int g = 0; g < ng; g++)
            if (int rc = launch_stress(g)) return rc;
        if (cfg.check_energy)
            if (int rc = launch_energy()) return rc;
        if (cfg.imode == 0) {  // main.cpp:242-251 with the II ghosts
            for (int k = 0; k + 1 < ng; k++) {
                if (int rc = launch_ghost_sxz(k)) return rc;
                if (int rc = launch_ghost_syz(k)) return rc;
            }
            return ECSG_OK;
        }
        return enqueue_ghost_stress();
    }

    // ----- per-phase entry points (the host drives the order) ---------------
    int update_vel(int g) override
    {
        if (int rc = check_ready()) return rc;
        if (g < 0 || g >= (int)G.size()) return fail(ECSG_EINVAL, "grid index out of range");
        return launch_vel(g);
    }
    int record_receivers() override
    {
        if (int rc = check_ready()) return rc;
        // the receiver/source kernel does both; here only the receivers
        PointArgs<T> a = point_args();
        a.src_len = 0;
        a.energy = 1;  // do not advance the step counter
        hipLaunchKernelGGL((k_recsrc<T>), dim3(1), dim3(256), 0, s, a);
        HIPCHK(hipGetLastError());
        return ECSG_OK;
    }
    int add_source() override
    {
        if (int rc = check_ready()) return rc;
        PointArgs<T> a = point_args();
        a.nrec = 0;
        a.energy = 1;  // step counter advanced by end_step
        hipLaunchKernelGGL((k_recsrc<T>), dim3(1), dim3(256), 0, s, a);
        HIPCHK(hipGetLastError());
        return ECSG_OK;
    }
    int ghost_vz(int k) override
    {
        if (int rc = check_ready()) return rc;
        if (k < 0 || k >= (int)I.size()) return fail(ECSG_EINVAL, "interface index out of range");
        return launch_ghost_vz(k);
    }
    int ghost_sxz(int k) override
    {
        if (int rc = check_ready()) return rc;
        if (k < 0 || k >= (int)I.size()) return fail(ECSG_EINVAL, "interface index out of range");
        return launch_ghost_sxz(k);
    }
    int ghost_syz(int k) override
    {
        if (int rc = check_ready()) return rc;
        if (k < 0 || k >= (int)I.size()) return fail(ECSG_EINVAL, "interface index out of range");
        return launch_ghost_syz(k);
    }
    int update_stress(int g) override
    {
        if (int rc = check_ready()) return rc;
        if (g < 0 || g >= (int)G.size()) return fail(ECSG_EINVAL, "grid index out of range");
        return launch_stress(g);
    }
    // Energy of the current state: the kinetic partials were written by the
    // last velocity launches, the strain part needs the stresses before the
    // stress update, so this runs the stress reduction without updating.
    int energy(double *e) override
    {
        if (int rc = check_ready()) return rc;
        if (!cfg.check_energy) return fail(ECSG_ESTATE, "created without check_energy");
        // strain energy without modifying the fields: run the stress kernel on
        // a copy of the stress arrays would double the traffic; instead the
        // per-phase path computes it from a snapshot of the six stresses.
        size_t need = 0;
        for (const DevGrid<T> &d : G) need = std::max(need, (size_t)d.astride * 6);
        if (e_save_n < need) {
            HIPCHK(hipStreamSynchronize(s));
            hipFree(e_save);
            e_save = nullptr;
            e_save_n = 0;
            if (int rc = dalloc(&e_save, need)) return rc;
            e_save_n = need;
        }
        for (int g = 0; g < (int)G.size(); g++) {
            DevGrid<T> &d = G[g];
            T *save = e_save;
            HIPCHK(hipMemcpyAsync(save, d.arr(3), sizeof(T) * d.astride * 6, hipMemcpyDeviceToDevice, s));
            if (cfg.radius > 1) stress_radius(d, true);
            else switch (d.by) {
            case 1: stress_launch<1>(d.dev, d.se, true); break;
            case 2: stress_launch<2>(d.dev, d.se, true); break;
            case 8: stress_launch<8>(d.dev, d.se, true); break;
            default: stress_launch<4>(d.dev, d.se, true); break;
            }
            HIPCHK(hipGetLastError());
            HIPCHK(hipMemcpyAsync(d.arr(3), save, sizeof(T) * d.astride * 6, hipMemcpyDeviceToDevice, s));
        }
        // energy kernel writes energy[step] and advances the counter; undo
        int step = 0;
        HIPCHK(hipMemcpy(&step, d_step, sizeof(int), hipMemcpyDeviceToHost));
        if (int rc = launch_energy()) return rc;
        HIPCHK(hipMemcpyAsync(d_step, &step, sizeof(int), hipMemcpyHostToDevice, s));
        double v = 0;
        HIPCHK(hipStreamSynchronize(s));
        if (step < cfg.max_steps) HIPCHK(hipMemcpy(&v, d_energy + step, sizeof(double), hipMemcpyDeviceToHost));
        if (e) *e = v;
        return ECSG_OK;
    }
    int end_step() override
    {
        hipLaunchKernelGGL(k_end_step, dim3(1), dim3(64), 0, s, d_step);
        HIPCHK(hipGetLastError());
        host_step++;
        return ECSG_OK;
    }

    int residual(int which, double *out) override
    {
        if (which < 0 || which > 2) return fail(ECSG_EINVAL, "which must be 0, 1 or 2");
        const int ni = (int)I.size();
        if (ni == 0) return ECSG_OK;
        if (!d_diag)
            if (int rc = dalloc(&d_diag, DIAG_N)) return rc;
        for (int k = 0; k < ni; k++)
            hipLaunchKernelGGL((k_res<T>), dim3(1), dim3(256), 0, s, iface_dev(k), which, d_diag + k);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(out, d_diag, sizeof(double) * ni, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        return ECSG_OK;
    }

    // test_sbp.cpp:5-199 on grid g: seven residuals
    double *d_sbp = nullptr;
    int sbp(int g, double *out) override
    {
        if (int rc = check_ready()) return rc;
        if (g < 0 || g >= (int)G.size()) return fail(ECSG_EINVAL, "grid index out of range");
        DevGrid<T> &d = G[g];
        const int nblk = ((d.n[0] + BX) / BX) * ((d.n[1] + SBP_BY) / SBP_BY) * ((d.n[2] + SBP_ZC) / SBP_ZC);
        const size_t need = 7 * (size_t)nblk + 8;
        if (!d_sbp || sbp_n < need) {
            HIPCHK(hipStreamSynchronize(s));
            hipFree(d_sbp);
            d_sbp = nullptr;
            sbp_n = 0;
            if (int rc = dalloc(&d_sbp, need)) return rc;
            sbp_n = need;
        }
        hipLaunchKernelGGL((k_sbp_part<T>), dim3(nblk), dim3(BX, SBP_BY), 0, s, d.dev, cfg.radius, d_sbp, nblk);
        hipLaunchKernelGGL(k_sbp_final, dim3(7), dim3(64), 0, s, (const double *)d_sbp, nblk, d_sbp + 7 * (size_t)nblk);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(out, d_sbp + 7 * (size_t)nblk, 7 * sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        return ECSG_OK;
    }
    size_t sbp_n = 0;

    static constexpr int DIAG_PARTS = 1024;
    static constexpr int DIAG_N = DIAG_PARTS + 8;

    // Field::AvgAbs of wavefield sel on grid g (over the field's own range)
    int avg_abs(int g, int sel, double *out) override
    {
        if (int rc = check_ready()) return rc;
        if (g < 0 || g >= (int)G.size() || sel < 0 || sel > 8) return fail(ECSG_EINVAL, "bad grid or wavefield selector");
        if (!d_diag)
            if (int rc = dalloc(&d_diag, DIAG_N)) return rc;
        const DevGrid<T> &d = G[g];
        int len[3];
        for (int k = 0; k < 3; k++) len[k] = SIM_POS[sel][k] ? d.n[k] : d.n[k] + 1;
        const long rows = (long)len[1] * len[2];
        const int nb = (int)std::min<long>(rows, DIAG_PARTS);
        hipLaunchKernelGGL((k_avgabs_part<T>), dim3(nb), dim3(256), 0, s, (const T *)d.arr(sel), d.o, d.P, d.S2,
                           len[0], len[1], len[2], d_diag);
        hipLaunchKernelGGL(k_avgabs_final, dim3(1), dim3(64), 0, s, (const double *)d_diag, nb, (double)len[2],
                           (double)len[1], (double)len[0], d_diag + DIAG_PARTS);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(out, d_diag + DIAG_PARTS, sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        return ECSG_OK;
    }

    int step_n(int n, int use_graph) override
    {
        if (int rc = check_ready()) return rc;
        if (n < 0) return fail(ECSG_EINVAL, "negative step count");
        if (host_step + n > cfg.max_steps) return fail(ECSG_EINVAL, "max_steps exceeded");
        if (fused_ok()) return step_fused(n, use_graph);
        if (overlap_ok())  // before any capture
            if (int rc = ensure_s2()) return rc;
        // The tiled interface Jacobi is a grid-barrier launch ordered behind
        // every other handle's through a process-wide event, which a captured
        // graph cannot hold, so composite models with interfaces too large
        // for the LDS solver run eagerly.
        if (use_graph && !graph_failed && !prof && (I.empty() || cfg.imode == 0 || no_grid_barrier())) {
            if (!gexec) {
                if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) == hipSuccess) {
                    int rc = enqueue_step();
                    hipGraph_t gr = nullptr;
                    hipError_t e2 = hipStreamEndCapture(s, &gr);
                    if (rc == ECSG_OK && e2 == hipSuccess && gr &&
                        hipGraphInstantiate(&gexec, gr, nullptr, nullptr, 0) == hipSuccess) {
                        graph = gr;
                    } else {
                        if (gr) hipGraphDestroy(gr);
                        gexec = nullptr;
                        graph_failed = true;
                        hipGetLastError();
                        st = Status{};
                    }
                } else {
                    graph_failed = true;
                    hipGetLastError();
                }
            }
            if (gexec) {
                for (int i = 0; i < n; i++) HIPCHK(hipGraphLaunch(gexec, s));
                host_step += n;
                return ECSG_OK;
            }
        }
        for (int i = 0; i < n; i++)
            if (int rc = enqueue_step(i + 1 < n)) return rc;
        host_step += n;
        return ECSG_OK;
    }

    int steps_done() const override { return host_step; }

    int sync() override
    {
        if (int rc = ov_join()) return rc;
        HIPCHK(hipStreamSynchronize(s));
        if (int rc = check_barrier()) return rc;
        for (int k = 0; k < (int)I.size(); k++) {
            long long st3[3];
            HIPCHK(hipMemcpy(st3, I[k].stats, sizeof(st3), hipMemcpyDeviceToHost));
            if (st3[2]) return fail(ECSG_ENOCONV, "Jacobi iteration cap reached on interface " + std::to_string(k));
        }
        return ECSG_OK;
    }

    // ----- transfers --------------------------------------------------------
    int dims(int g, int sel, int len[3]) const override
    {
        if (g < 0 || g >= (int)G.size() || sel < 0 || sel > 18) return -1;
        const DevGrid<T> &d = G[g];
        static const int p0[3] = {0, 0, 0};
        const int *pos = sel < 9 ? SIM_POS[sel] : (sel == 9 ? p0 : ST_POS[sel - 10]);
        const int r = sel < 9 ? cfg.radius : 0;
        for (int k = 0; k < 3; k++) len[k] = pos[k] ? d.n[k] + 2 * r : d.n[k] + 1 + 2 * r;
        return r;
    }

    int device_array(int sel) const
    {
        if (sel < 9) return sel;
        if (sel == 9) return 18;
        return 9 + (sel - 10);
    }

    int get_field(int g, int sel, double *out) override
    {
        int len[3];
        const int r = dims(g, sel, len);
        if (r < 0) return fail(ECSG_EINVAL, "bad grid or field selector");
        DevGrid<T> &d = G[g];
        HIPCHK(hipStreamSynchronize(s));
        std::vector<T> buf((size_t)d.astride);
        HIPCHK(hipMemcpy(buf.data(), d.arr(device_array(sel)), sizeof(T) * d.astride, hipMemcpyDeviceToHost));
        long q = 0;
        for (int k = -r; k < len[2] - r; k++)
            for (int j = -r; j < len[1] - r; j++)
                for (int i = -r; i < len[0] - r; i++, q++) out[q] = (double)buf[d.o + (long)k * d.S2 + (long)j * d.P + i];
        return ECSG_OK;
    }

    int set_field(int g, int sel, const double *in) override
    {
        int len[3];
        const int r = dims(g, sel, len);
        if (r < 0 || sel >= 9) return fail(ECSG_EINVAL, "only wavefields (0..8) can be set");
        DevGrid<T> &d = G[g];
        HIPCHK(hipStreamSynchronize(s));
        std::vector<T> buf((size_t)d.astride);
        HIPCHK(hipMemcpy(buf.data(), d.arr(sel), sizeof(T) * d.astride, hipMemcpyDeviceToHost));
        long q = 0;
        for (int k = -r; k < len[2] - r; k++)
            for (int j = -r; j < len[1] - r; j++)
                for (int i = -r; i < len[0] - r; i++, q++) buf[d.o + (long)k * d.S2 + (long)j * d.P + i] = (T)in[q];
        // both field sets: entries no kernel writes must agree between them
        HIPCHK(hipMemcpy(d.fset(0) + (long)sel * d.astride, buf.data(), sizeof(T) * d.astride, hipMemcpyHostToDevice));
        if (d.alt) HIPCHK(hipMemcpy(d.alt + (long)sel * d.astride, buf.data(), sizeof(T) * d.astride, hipMemcpyHostToDevice));
        return ECSG_OK;
    }

    int get_traces(double *out) override
    {
        HIPCHK(hipStreamSynchronize(s));
        const int n = std::min(host_step, cfg.max_steps);
        std::vector<T> buf((size_t)std::max(nrec, 1) * cfg.max_steps);
        if (nrec) HIPCHK(hipMemcpy(buf.data(), d_traces, sizeof(T) * nrec * cfg.max_steps, hipMemcpyDeviceToHost));
        for (int q = 0; q < nrec; q++)
            for (int t = 0; t < n; t++) out[(long)q * n + t] = (double)buf[(long)q * cfg.max_steps + t];
        return ECSG_OK;
    }

    int get_energy(double *out) override
    {
        HIPCHK(hipStreamSynchronize(s));
        const int n = std::min(host_step, cfg.max_steps);
        if (n) HIPCHK(hipMemcpy(out, d_energy, sizeof(double) * n, hipMemcpyDeviceToHost));
        return ECSG_OK;
    }

    int jacobi_stats(int k, long long *tot, int *last, int *capped) override
    {
        if (k < 0 || k >= (int)I.size()) return fail(ECSG_EINVAL, "interface index out of range");
        HIPCHK(hipStreamSynchronize(s));
        long long st3[3];
        HIPCHK(hipMemcpy(st3, I[k].stats, sizeof(st3), hipMemcpyDeviceToHost));
        if (tot) *tot = st3[0];
        if (last) *last = (int)st3[1];
        if (capped) *capped = (int)st3[2];
        return ECSG_OK;
    }

    int profile(int nsteps, double ms[6], int cnt[6]) override
    {
        if (int rc = check_ready()) return rc;
        if (host_step + nsteps > cfg.max_steps) return fail(ECSG_EINVAL, "max_steps exceeded");
        HIPCHK(hipStreamSynchronize(s));
        // pre-create the events: launches per step times two
        size_t per_step = 2 * (6 * G.size() + 2 + 3 * I.size() + (cfg.check_energy ? 1 : 0));
        while (evs.size() < per_step * (size_t)nsteps) {
            hipEvent_t e;
            HIPCHK(hipEventCreate(&e));
            evs.push_back(e);
        }
        ev_next = 0;
        ev_kind.clear();
        prof = true;
        if (fused_ok() && sets_dirty) sync_sets();
        for (int i = 0; i < nsteps; i++)
            if (int rc = fused_ok() ? enqueue_fused_step() : enqueue_step(i + 1 < nsteps)) { prof = false; return rc; }
        host_step += nsteps;
        prof = false;
        HIPCHK(hipStreamSynchronize(s));
        for (int k = 0; k < 6; k++) { ms[k] = 0; cnt[k] = 0; }
        for (size_t q = 0; q < ev_kind.size(); q++) {
            float t = 0;
            hipEventElapsedTime(&t, evs[2 * q], evs[2 * q + 1]);
            const int kind = ev_kind[q];
            ms[kind] += t; cnt[kind]++;
            if (kind == 0) { ms[2] += t; cnt[2]++; }
            if (kind == 1) { ms[3] += t; cnt[3]++; }
        }
        for (int k = 0; k < 6; k++) ms[k] = cnt[k] ? ms[k] / cnt[k] : 0;
        ev_kind.clear();
        ev_next = 0;
        return ECSG_OK;
    }

    // nsteps steps as step_n without a graph, with HIP events (on the
    // launch stream) around the grid-0 velocity and stress stencil launches
    // only; ms: their average durations.  Blocks until the steps are done.
    int step_timed(int nsteps, double ms[2]) override
    {
        if (int rc = check_ready()) return rc;
        if (nsteps < 0) return fail(ECSG_EINVAL, "negative step count");
        if (host_step + nsteps > cfg.max_steps) return fail(ECSG_EINVAL, "max_steps exceeded");
        if (fused_ok()) return fail(ECSG_EINVAL, "step_timed runs the separate stencil passes (fused pass on)");
        HIPCHK(hipStreamSynchronize(s));
        while (evs.size() < (size_t)4 * nsteps) {
            hipEvent_t e;
            HIPCHK(hipEventCreate(&e));
            evs.push_back(e);
        }
        ev_next = 0;
        ev_kind.clear();
        live = true;
        for (int i = 0; i < nsteps; i++)
            if (int rc = enqueue_step(i + 1 < nsteps)) { live = false; return rc; }
        live = false;
        host_step += nsteps;
        if (int rc = sync()) return rc;
        double sum[2] = {0, 0};
        int cnt[2] = {0, 0};
        for (size_t q = 0; q < ev_kind.size(); q++) {
            float t = 0;
            HIPCHK(hipEventElapsedTime(&t, evs[2 * q], evs[2 * q + 1]));
            if (ev_kind[q] == 0 || ev_kind[q] == 1) { sum[ev_kind[q]] += t; cnt[ev_kind[q]]++; }
        }
        for (int k = 0; k < 2; k++) ms[k] = cnt[k] ? sum[k] / cnt[k] : 0;
        ev_kind.clear();
        ev_next = 0;
        return ECSG_OK;
    }

    // ======================================================================
    // Adjoint-state gradient of the receiver misfit
    //   J = 1/2 sum_n sum_r (vx_r(n) - obs_r(n))^2
    // with respect to the parameters given to set_params.  Forward pass with
    // checkpoints every ck_every steps; then, segment by segment from the
    // last, the forward steps are recomputed from the checkpoint into a ring
    // of saved states and the transposed steps run backwards over them.
    // ======================================================================
    struct AdjIface {
        double *mu_vz = nullptr, *mu_syz = nullptr, *den = nullptr;
        double *Ba = nullptr, *X0a = nullptr, *X1a = nullptr, *jpart = nullptr, *absz = nullptr;
        long long *stats = nullptr;
    };
    struct MirrorDev {
        long *src = nullptr;
        int *start = nullptr;
        long *halo = nullptr;
        int n = 0, nh = 0;
    };
    struct AdjGrid {
        T *lam = nullptr;       // 9 adjoint wavefields
        T *grad = nullptr;      // 12 gradient accumulators (9 staged stiffness, b node, b xy mean, b xz mean)
        T *coef = nullptr;      // 12 masked coefficients of the transposed stencils
        MirrorDev vA, vB, sm;   // transposed ghost mirrors: vy y walls, other velocity mirrors, free surface
    };
    std::vector<AdjIface> AI;
    std::vector<AdjGrid> AG;
    bool adj_ready = false;
    long *d_rloc = nullptr;
    int *d_rstart = nullptr, *d_rrec = nullptr;
    int nrloc = 0;
    T *d_resid = nullptr;
    int ck_every = 50, ck_fmt = 0;
    int ck_alloc = 0, ck_alloc_fmt = -1, ring_alloc = 0;
    std::vector<T *> ck_nat;              // [ck][grid]
    std::vector<unsigned short *> ck_red; // [ck][grid]
    std::vector<double *> ck_x0;          // [ck][iface]
    float *ck_scale = nullptr;            // [ck][grid][9]
    float *ck_part = nullptr;
    std::vector<T *> ring;                // [slot][grid]
    std::vector<std::vector<double>> gprim;  // [grid*10 + a]: 0 buoyancy, 1..9 stiffness inputs
    double adj_rel_tol = 1e-11;
    static constexpr int CKP = 512;       // partial maxima per array

    template <typename P>
    static void dfree(P *&p) { if (p) hipFree((void *)p); p = nullptr; }
    void ck_free()
    {
        for (auto p : ck_nat) dfree(p);
        for (auto p : ck_red) dfree(p);
        for (auto p : ck_x0) dfree(p);
        for (auto p : ring) dfree(p);
        ck_nat.clear(); ck_red.clear(); ck_x0.clear(); ring.clear();
        dfree(ck_scale); dfree(ck_part);
        ck_alloc = 0; ck_alloc_fmt = -1; ring_alloc = 0;
    }
    void adj_free()
    {
        for (auto &a : AI) {
            dfree(a.mu_vz); dfree(a.mu_syz); dfree(a.den); dfree(a.Ba); dfree(a.X0a); dfree(a.X1a); dfree(a.jpart); dfree(a.absz);
            dfree(a.stats);
        }
        for (auto &g : AG) {
            dfree(g.lam); dfree(g.grad); dfree(g.coef);
            for (MirrorDev *m : {&g.vA, &g.vB, &g.sm}) { dfree(m->src); dfree(m->start); dfree(m->halo); }
        }
        AI.clear(); AG.clear();
        dfree(d_rloc); dfree(d_rstart); dfree(d_rrec); dfree(d_resid);
        ck_free();
        adj_ready = false;
        coef_ok = false;
    }

    template <typename P>
    int upload(P **dst, const std::vector<P> &v)
    {
        if (int rc = dalloc(dst, std::max<size_t>(v.size(), 1))) return rc;
        if (!v.empty()) HIPCHK(hipMemcpy(*dst, v.data(), sizeof(P) * v.size(), hipMemcpyHostToDevice));
        return ECSG_OK;
    }

    // work buffers of the transposed interface solves of interface k
    int build_adj_iface(int k)
    {
        const DevIface &f = I[k];
        AdjIface &ai = AI[k];
        const int nx2 = f.nx2, ny2 = f.ny2;
        const long nl = (long)nx2 * (ny2 - 1), nn = (long)(nx2 - 1) * (ny2 - 1);
        const long nxw = (long)(nx2 + 1) * (ny2 + 1);
        if (int rc = dalloc(&ai.mu_vz, nl)) return rc;
        if (int rc = dalloc(&ai.mu_syz, (long)nx2 * ny2)) return rc;
        if (int rc = dalloc(&ai.den, (long)nx2 * ny2)) return rc;
        if (int rc = dalloc(&ai.Ba, nn)) return rc;
        if (int rc = dalloc(&ai.X0a, nxw)) return rc;
        if (int rc = dalloc(&ai.X1a, nxw)) return rc;
        if (int rc = dalloc(&ai.absz, f.nabs)) return rc;
        if (int rc = dalloc(&ai.jpart, (size_t)2 * f.tiles * 2 * SJ + 1)) return rc;
        if (int rc = dalloc(&ai.stats, 3)) return rc;
        return ECSG_OK;
    }

    // transposed ghost mirrors of grid g (the mirror loops of k_vel and k_stress)
    int upload_mirror(std::vector<std::pair<long, long>> &v, MirrorDev &m)
    {
        std::stable_sort(v.begin(), v.end(), [](const std::pair<long, long> &a, const std::pair<long, long> &b) { return a.first < b.first; });
        std::vector<long> src, halo;
        std::vector<int> st;
        for (size_t i = 0; i < v.size(); i++) {
            if (i == 0 || v[i].first != v[i - 1].first) { src.push_back(v[i].first); st.push_back((int)i); }
            halo.push_back(v[i].second);
        }
        st.push_back((int)v.size());
        m.n = (int)src.size();
        m.nh = (int)halo.size();
        if (int rc = upload(&m.src, src)) return rc;
        if (int rc = upload(&m.start, st)) return rc;
        if (int rc = upload(&m.halo, halo)) return rc;
        return ECSG_OK;
    }

    int build_adj_grid(int g)
    {
        DevGrid<T> &d = G[g];
        AdjGrid &a = AG[g];
        const int nx = d.n[0], ny = d.n[1], nz = d.n[2];
        const long as = d.astride;
        auto ix = [&](int f, int i, int j, int k) { return (long)f * as + d.o + (long)k * d.S2 + (long)j * d.P + i; };
        enum { VY = 1, VZ = 2, SYZ = 6, SXZ = 7 };
        std::vector<std::pair<long, long>> A, B, S;  // (source, halo)
        for (int k = 0; k <= d.dev.zy; k++)
            for (int i = 0; i <= nx - 1; i++) {
                A.push_back({ix(VY, i, 0, k), ix(VY, i, -1, k)});
                A.push_back({ix(VY, i, ny - 1, k), ix(VY, i, ny, k)});
            }
        if (d.dev.top_eff)
            for (int j = 0; j <= ny; j++)
                for (int i = 0; i <= nx - 1; i++) B.push_back({ix(VZ, i, j, 0), ix(VZ, i, j, -1)});
        if (d.dev.bottom)
            for (int j = 0; j <= ny; j++)
                for (int i = 0; i <= nx - 1; i++) B.push_back({ix(VZ, i, j, nz - 1), ix(VZ, i, j, nz)});
        for (int k = 0; k <= nz - 1; k++)
            for (int j = 0; j <= ny; j++) {
                B.push_back({ix(VZ, 0, j, k), ix(VZ, -1, j, k)});
                B.push_back({ix(VZ, nx - 1, j, k), ix(VZ, nx, j, k)});
            }
        for (int k = 0; k <= nz; k++)
            for (int j = 0; j <= ny; j++) {
                B.push_back({ix(VY, 0, j, k), ix(VY, -1, j, k)});
                B.push_back({ix(VY, nx - 1, j, k), ix(VY, nx, j, k)});
            }
        if (d.dev.free_top) {
            for (int j = 0; j <= ny - 1; j++)
                for (int i = 0; i <= nx - 1; i++) S.push_back({ix(SYZ, i, j, 0), ix(SYZ, i, j, -1)});
            for (int j = 1; j <= ny - 1; j++)
                for (int i = 1; i <= nx - 1; i++) S.push_back({ix(SXZ, i, j, 0), ix(SXZ, i, j, -1)});
        }
        if (int rc = upload_mirror(A, a.vA)) return rc;
        if (int rc = upload_mirror(B, a.vB)) return rc;
        if (int rc = upload_mirror(S, a.sm)) return rc;
        if (int rc = dalloc(&a.lam, (size_t)9 * as)) return rc;
        if (int rc = dalloc(&a.grad, (size_t)12 * as)) return rc;
        if (int rc = dalloc(&a.coef, (size_t)12 * as)) return rc;
        return ECSG_OK;
    }

    int build_adj_receivers()
    {
        dfree(d_rloc); dfree(d_rstart); dfree(d_rrec);
        std::vector<int> loc4((size_t)4 * std::max(nrec, 1));
        if (nrec) HIPCHK(hipMemcpy(loc4.data(), d_rec, sizeof(int) * 4 * nrec, hipMemcpyDeviceToHost));
        std::vector<std::pair<long, int>> v;
        for (int q = 0; q < nrec; q++) {
            const int g = loc4[4 * q];
            const DevGrid<T> &d = G[g];
            const long idx = d.o + (long)loc4[4 * q + 3] * d.S2 + (long)loc4[4 * q + 2] * d.P + loc4[4 * q + 1];
            v.push_back({((long)g << 40) | idx, q});
        }
        std::stable_sort(v.begin(), v.end(), [](const std::pair<long, int> &a, const std::pair<long, int> &b) { return a.first < b.first; });
        std::vector<long> uk;
        std::vector<int> st, rc2;
        for (size_t i = 0; i < v.size(); i++) {
            if (i == 0 || v[i].first != v[i - 1].first) { uk.push_back(v[i].first); st.push_back((int)i); }
            rc2.push_back(v[i].second);
        }
        st.push_back((int)v.size());
        nrloc = (int)uk.size();
        if (int rc = upload(&d_rloc, uk)) return rc;
        if (int rc = upload(&d_rstart, st)) return rc;
        if (int rc = upload(&d_rrec, rc2)) return rc;
        return ECSG_OK;
    }

    int adj_build()
    {
        if (adj_ready) return ECSG_OK;
        if (int rc = check_ready()) return rc;
        if (cfg.imode == 0 && !I.empty()) return fail(ECSG_EINVAL, "the adjoint is implemented for the energy-conserving interface (imode 1)");
        if (cfg.radius != 1) return fail(ECSG_EINVAL, "the adjoint is implemented for radius 1");
        adj_free();
        AG.resize(G.size());
        AI.resize(I.size());
        for (int g = 0; g < (int)G.size(); g++)
            if (int rc = build_adj_grid(g)) return rc;
        for (int k = 0; k < (int)I.size(); k++)
            if (int rc = build_adj_iface(k)) return rc;
        adj_ready = true;
        return ECSG_OK;
    }

    const T *fview(int slot, int g, int f) const
    {
        return slot < 0 ? (const T *)G[g].arr(f) : ring[(size_t)slot * G.size() + g] + (long)f * G[g].astride;
    }

    IfAdj<T> if_adj(int k, int slot_v, int slot_s)
    {
        IfAdj<T> a{};
        const DevGrid<T> &F = G[k], &C = G[k + 1];
        for (int f = 0; f < 9; f++) {
            a.lf[f] = AG[k].lam + (long)f * F.astride;
            a.lc[f] = AG[k + 1].lam + (long)f * C.astride;
            a.ff[f] = fview(f < 3 ? slot_v : slot_s, k, f);
            a.fc[f] = fview(f < 3 ? slot_v : slot_s, k + 1, f);
            a.cf[f] = F.dev.c[f];
            a.cc[f] = C.dev.c[f];
        }
        for (int m = 0; m < 12; m++) {
            a.gf[m] = AG[k].grad + (long)m * F.astride;
            a.gc[m] = AG[k + 1].grad + (long)m * C.astride;
        }
        a.bf = F.dev.b; a.bc = C.dev.b;
        a.oF = F.o; a.PF = F.P; a.SF = F.S2; a.oC = C.o; a.PC = C.P; a.SC = C.S2;
        a.r = I[k].r; a.nx2 = I[k].nx2; a.ny2 = I[k].ny2; a.nz1 = I[k].nz1;
        a.k1x = (double)F.dev.cx; a.k1y = (double)F.dev.cy; a.k1z = (double)F.dev.cz;
        a.k2x = (double)C.dev.cx; a.k2y = (double)C.dev.cy; a.k2z = (double)C.dev.cz;
        return a;
    }

    AdjGridDev<T> adj_dev(int g, int slot_v, int slot_s)
    {
        const DevGrid<T> &d = G[g];
        AdjGridDev<T> a{};
        for (int f = 0; f < 9; f++) a.lam[f] = AG[g].lam + (long)f * d.astride;
        for (int f = 0; f < 3; f++) a.fv[f] = fview(slot_v, g, f);
        for (int f = 0; f < 6; f++) a.fs[f] = fview(slot_s, g, 3 + f);
        for (int m = 0; m < 9; m++) a.gc[m] = AG[g].grad + (long)m * d.astride;
        for (int m = 0; m < 3; m++) a.gb[m] = AG[g].grad + (long)(9 + m) * d.astride;
        for (int m = 0; m < 9; m++) a.c[m] = d.dev.c[m];
        a.b = d.dev.b;
        a.nx = d.n[0]; a.ny = d.n[1]; a.nz = d.n[2];
        a.P = d.P; a.S2 = d.S2; a.o = d.o;
        a.cx = d.dev.cx; a.cy = d.dev.cy; a.cz = d.dev.cz;
        a.top_eff = d.dev.top_eff; a.bottom = d.dev.bottom; a.free_top = d.dev.free_top;
        // output boxes: [-1, n] x [-1, n] (stress transpose) or [0, n] x [0, n]
        // (velocity transpose) in x, y; [-1, n] in z
        for (int m = 0; m < 12; m++) a.ac[m] = AG[g].coef + (long)m * d.astride;
        a.gx = (d.n[0] + 2 + AXO - 1) / AXO;
        a.gy = (d.n[1] + 2 + AWY - 1) / AWY;
        a.gz = (d.n[2] + 2 + AZ2 - 1) / AZ2;
        return a;
    }

    static int nb(long n) { return (int)((n + 255) / 256); }

    void launch_mirror(int g, const MirrorDev &m)
    {
        if (!m.n) return;
        hipLaunchKernelGGL((k_mirror_T<T>), dim3(nb(m.n)), dim3(256), 0, s, AG[g].lam, m.src, m.start, m.halo, m.n);
        hipLaunchKernelGGL((k_zero_list<T>), dim3(nb(m.nh)), dim3(256), 0, s, AG[g].lam, m.halo, m.nh);
    }

    int launch_jacobi_adj(int k0, int n)
    {
        JacBatch<T> jb{};
        int blk = 0;
        for (int q = 0; q < n; q++) {
            DevIface &f = I[k0 + q];
            AdjIface &a = AI[k0 + q];
            JacItf<T> &J = jb.it[q];
            J.A = f.A; J.B = a.Ba; J.X0 = a.X0a; J.X1 = a.X1a; J.abs_part = a.absz; J.nabs = f.nabs;
            J.part = a.jpart; J.stats = a.stats;
            J.sxz1 = nullptr; J.sxz2 = nullptr; J.sxz1s = nullptr; J.sxz2s = nullptr;
            J.nx2 = f.nx2; J.ny2 = f.ny2; J.nz1 = f.nz1; J.r = f.r;
            J.blk0 = blk; J.nblk = f.tiles; J.rho2 = 0;
            blk += J.nblk;
        }
        jb.n = n;
        jb.tol_cap = 0; jb.tol_fac = 0;
        jb.floor = 1024 * 2.220446049250313e-16;
        jb.rel_tol = adj_rel_tol;
        jb.cap = JACOBI_CAP;
        jb.cheby = 0;
        jb.solve_only = 1;
        return launch_tiled(jb, blk);
    }

    // One transposed step: forward step n read from ring slot slot_f (state
    // after the step) and slot_p (state before it); -1 = the live state.
    int adj_step_impl(int n, int slot_f, int slot_p, bool inject)
    {
        const int ng = (int)G.size(), ni = (int)I.size();
        // vy-syz (pointwise)
        for (int k = ni - 1; k >= 0; k--) {
            IfAdj<T> a = if_adj(k, slot_f, slot_f);
            const int nx1 = a.r * a.nx2, ny1 = a.r * a.ny2;
            hipLaunchKernelGGL((k_asyz_ghost<T>), dim3(nb((long)a.nx2 * a.ny2)), dim3(256), 0, s, a, (const double *)AI[k].den, AI[k].mu_syz);
            hipLaunchKernelGGL((k_asyz_fine0<T>), dim3(nb((long)nx1 * ny1)), dim3(256), 0, s, a);
            hipLaunchKernelGGL((k_asyz_fine<T>), dim3(nb((long)(nx1 + 1) * (ny1 + 1))), dim3(256), 0, s, a, (const double *)AI[k].mu_syz);
            hipLaunchKernelGGL((k_asyz_coarse<T>), dim3(nb((long)(a.nx2 + 1) * (a.ny2 + 1))), dim3(256), 0, s, a,
                               (const double *)AI[k].mu_syz);
        }
        // vx-sxz (Jacobi)
        if (ni) {
            for (int k = ni - 1; k >= 0; k--) {
                IfAdj<T> a = if_adj(k, slot_f, slot_f);
                const int nx1 = a.r * a.nx2, ny1 = a.r * a.ny2;
                hipLaunchKernelGGL((k_asxz_ghost<T>), dim3(nb((long)(a.nx2 - 1) * (a.ny2 - 1))), dim3(256), 0, s, a, AI[k].Ba);
                hipLaunchKernelGGL((k_asxz_fine0<T>), dim3(nb((long)(nx1 - 1) * (ny1 - 1))), dim3(256), 0, s, a);
            }
            if (use_lds()) {
                if (int rc = launch_jacobi_lds(true)) return rc;
            } else {
                for (int k0 = 0; k0 < ni; k0 += MAXI)
                    if (int rc = launch_jacobi_adj(k0, std::min(MAXI, ni - k0))) return rc;
            }
            for (int k = ni - 1; k >= 0; k--) {
                IfAdj<T> a = if_adj(k, slot_f, slot_f);
                const int nx1 = a.r * a.nx2, ny1 = a.r * a.ny2;
                hipLaunchKernelGGL((k_asxz_fine<T>), dim3(nb((long)(nx1 + 1) * (ny1 + 1))), dim3(256), 0, s, a, (const double *)AI[k].X0a);
                hipLaunchKernelGGL((k_asxz_coarse<T>), dim3(nb((long)(a.nx2 + 1) * (a.ny2 + 1))), dim3(256), 0, s, a,
                                   (const double *)AI[k].X0a);
            }
        }
        // stress stencils
        for (int g = ng - 1; g >= 0; g--) {
            launch_mirror(g, AG[g].sm);
            AdjGridDev<T> a = adj_dev(g, slot_f, slot_p);
            const int gb = ((a.nx + 64) / 64) * ((a.ny + GBY) / GBY) * ((a.nz + GZC) / GZC);
            if (fits32(a)) {
                hipLaunchKernelGGL((k_adj_stress<T, unsigned>), dim3(a.gx * a.gy * a.gz), dim3(64, AWY), 0, s, a);
                hipLaunchKernelGGL((k_adj_grad<T, unsigned>), dim3(gb), dim3(64, GBY), 0, s, a);
            } else {
                hipLaunchKernelGGL((k_adj_stress<T, long>), dim3(a.gx * a.gy * a.gz), dim3(64, AWY), 0, s, a);
                hipLaunchKernelGGL((k_adj_grad<T, long>), dim3(gb), dim3(64, GBY), 0, s, a);
            }
        }
        // vz-szz (tridiagonal)
        for (int k = ni - 1; k >= 0; k--) {
            IfAdj<T> a = if_adj(k, slot_f, slot_p);
            const int nx1 = a.r * a.nx2, ny1 = a.r * a.ny2;
            hipLaunchKernelGGL((k_avz_ghost<T>), dim3(nb((long)a.nx2 * (a.ny2 + 1))), dim3(256), 0, s, a, AI[k].mu_vz);
            hipLaunchKernelGGL((k_avz_fine0<T>), dim3(nb((long)nx1 * (ny1 - 1))), dim3(256), 0, s, a);
            if ((size_t)32 * (I[k].ny2 - 1) <= (size_t)VZ_WAVE_LDS)
                hipLaunchKernelGGL(k_tri_solve_wave, dim3(I[k].nx2), dim3(64), (size_t)32 * (I[k].ny2 - 1), s, AI[k].mu_vz,
                                   (const double *)I[k].M, (const double *)I[k].Dp, (const double *)I[k].U, I[k].nx2, I[k].ny2 - 1);
            else
                hipLaunchKernelGGL(k_tri_solve, dim3((I[k].nx2 + 63) / 64), dim3(64), 0, s, AI[k].mu_vz, I[k].M, I[k].Dp, I[k].U,
                                   I[k].nx2, I[k].ny2 - 1);
            hipLaunchKernelGGL((k_avz_fine<T>), dim3(nb((long)(nx1 + 1) * (ny1 + 1))), dim3(256), 0, s, a, (const double *)AI[k].mu_vz);
            hipLaunchKernelGGL((k_avz_coarse<T>), dim3(nb((long)(a.nx2 + 1) * (a.ny2 + 1))), dim3(256), 0, s, a,
                               (const double *)AI[k].mu_vz);
        }
        if (inject && nrloc) {
            RecAdj<T> ra{};
            for (int g = 0; g < ng; g++) ra.lamvx[g] = AG[g].lam;
            ra.loc = d_rloc; ra.start = d_rstart; ra.rec = d_rrec; ra.resid = d_resid; ra.nloc = nrloc; ra.nsteps = resid_steps;
            hipLaunchKernelGGL((k_adj_rec<T>), dim3(nb(nrloc)), dim3(256), 0, s, ra, n);
        }
        for (int g = ng - 1; g >= 0; g--) {
            launch_mirror(g, AG[g].vA);
            launch_mirror(g, AG[g].vB);
            AdjGridDev<T> a = adj_dev(g, slot_f, slot_p);
            const int gb = ((a.nx + 64) / 64) * ((a.ny + GBY) / GBY) * ((a.nz + GZC) / GZC);
            if (fits32(a)) {
                hipLaunchKernelGGL((k_adj_vel<T, unsigned>), dim3(a.gx * a.gy * a.gz), dim3(64, AWY), 0, s, a);
                hipLaunchKernelGGL((k_adj_gradb<T, unsigned>), dim3(gb), dim3(64, GBY), 0, s, a);
            } else {
                hipLaunchKernelGGL((k_adj_vel<T, long>), dim3(a.gx * a.gy * a.gz), dim3(64, AWY), 0, s, a);
                hipLaunchKernelGGL((k_adj_gradb<T, long>), dim3(gb), dim3(64, GBY), 0, s, a);
            }
        }
        HIPCHK(hipGetLastError());
        return ECSG_OK;
    }
    int resid_steps = 0;

    // masked coefficients: rebuilt only after the parameters change
    bool coef_ok = false;
    int adj_coef()
    {
        if (coef_ok) return ECSG_OK;
        coef_ok = true;
        for (int g = 0; g < (int)G.size(); g++) {
            const DevGrid<T> &d = G[g];
            const long np = (long)(d.n[0] + 1) * (d.n[1] + 1) * (d.n[2] + 1);
            HIPCHK(hipMemsetAsync(AG[g].coef, 0, sizeof(T) * 12 * d.astride, s));
            hipLaunchKernelGGL((k_adj_coef<T>), dim3(nb(np)), dim3(256), 0, s, d.dev, AG[g].coef, d.astride);
        }
        HIPCHK(hipGetLastError());
        return ECSG_OK;
    }

    int adj_zero()
    {
        if (int rc = adj_coef()) return rc;
        for (int g = 0; g < (int)G.size(); g++) {
            HIPCHK(hipMemsetAsync(AG[g].lam, 0, sizeof(T) * 9 * G[g].astride, s));
            HIPCHK(hipMemsetAsync(AG[g].grad, 0, sizeof(T) * 12 * G[g].astride, s));
        }
        for (int k = 0; k < (int)I.size(); k++) {
            const long nxw = (long)(I[k].nx2 + 1) * (I[k].ny2 + 1);
            HIPCHK(hipMemsetAsync(AI[k].X0a, 0, sizeof(double) * nxw, s));
            HIPCHK(hipMemsetAsync(AI[k].X1a, 0, sizeof(double) * nxw, s));
            HIPCHK(hipMemsetAsync(AI[k].stats, 0, sizeof(long long) * 3, s));
            const long nc = (long)I[k].nx2 * I[k].ny2;
            hipLaunchKernelGGL((k_gsyz_den<T>), dim3(nb(nc)), dim3(256), 0, s, iface_dev(k), AI[k].den);
        }
        HIPCHK(hipGetLastError());
        return ECSG_OK;
    }

    int reset() override
    {
        for (auto &d : G) {
            HIPCHK(hipMemsetAsync(d.fset(0), 0, sizeof(T) * 9 * d.astride, s));
            if (d.alt) HIPCHK(hipMemsetAsync(d.alt, 0, sizeof(T) * 9 * d.astride, s));
        }
        for (auto &f : I) {
            const long nxw = (long)(f.nx2 + 1) * (f.ny2 + 1);
            HIPCHK(hipMemsetAsync(f.X0, 0, sizeof(double) * nxw, s));
            HIPCHK(hipMemsetAsync(f.X1, 0, sizeof(double) * nxw, s));
            HIPCHK(hipMemsetAsync(f.stats, 0, sizeof(long long) * 3, s));
        }
        if (d_traces) HIPCHK(hipMemsetAsync(d_traces, 0, sizeof(T) * std::max(nrec, 1) * cfg.max_steps, s));
        HIPCHK(hipMemsetAsync(d_energy, 0, sizeof(double) * cfg.max_steps, s));
        hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, s, d_step, 0);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(s));
        host_step = 0;
        return ECSG_OK;
    }

    int set_checkpointing(int every, int format) override
    {
        if (every < 0) return fail(ECSG_EINVAL, "checkpoint interval must be >= 0 (0: sized from free device memory)");
        if (format < 0 || format > 2) return fail(ECSG_EINVAL, "checkpoint format must be 0 (native), 1 (fp16) or 2 (bf16)");
        ck_every = every;
        ck_fmt = format;
        return ECSG_OK;
    }

    // Largest ring (saved states for the transposed steps) a gradient may
    // hold; 0: sized from free device memory.  An interval longer than the
    // ring is processed in two levels (gradient).
    int ring_steps = 0;
    int set_ring_steps(int k) override
    {
        if (k < 0) return fail(ECSG_EINVAL, "ring steps must be >= 0 (0: sized from free device memory)");
        ring_steps = k;
        return ECSG_OK;
    }

    // Recomputed forward steps of a gradient over N steps with checkpoint
    // interval K1 and ring K2 (<= K1): every segment but the last (one
    // level), or every segment, then every sub-segment but each segment's
    // last (two levels).  The first segment (sub-segment) takes the
    // remainder.
    static long long recomputed(int N, int K1, int K2)
    {
        const int nseg = (N + K1 - 1) / K1;
        if (K2 >= K1) return (long long)N - std::min(K1, N);
        long long r = 0;
        for (int j = 0; j < nseg; j++) {
            const int L = j == 0 ? N - (nseg - 1) * K1 : K1;
            r += L + (L > K2 ? L - K2 : 0);
        }
        return r;
    }
    // checkpoints a gradient holds: one per segment but the last (one
    // level); one per segment plus one per sub-segment but the last of the
    // longest segment (two levels)
    static int ck_count(int N, int K1, int K2)
    {
        const int nseg = (N + K1 - 1) / K1;
        if (K2 >= K1) return std::max(nseg - 1, 1);
        return nseg + (std::min(K1, N) + K2 - 1) / K2 - 1;
    }

    // Checkpoint interval K1 and ring K2 from the free device memory, less a
    // reserve: one level (ring = interval, the largest that fits) when that
    // fits, else the two-level pair with the fewest recomputed steps among
    // those that fit.  k1_fixed > 0: that interval, the ring sized to it.
    int auto_interval(int N, int k1_fixed, int *K1_out, int *K2_out)
    {
        size_t fr = 0, tot = 0;
        HIPCHK(hipMemGetInfo(&fr, &tot));
        double state = 0, ckb = 0;
        for (const DevGrid<T> &d : G) {
            state += 9.0 * (double)d.astride * sizeof(T);
            ckb += 9.0 * (double)d.astride * (ck_fmt == 0 ? sizeof(T) : sizeof(unsigned short));
        }
        for (const DevIface &f : I) ckb += 8.0 * (f.nx2 + 1) * (f.ny2 + 1);
        // what this handle already holds for the ring and checkpoints is reusable
        double avail = (double)fr + (double)ring_alloc * state + (double)ck_alloc * ckb;
        avail -= std::max(4.0e9, 0.05 * (double)tot);
        auto fits = [&](int k1, int k2) { return (k2 + 1) * state + ck_count(N, k1, k2) * ckb <= avail; };
        const int kmax = ring_steps > 0 ? std::min(ring_steps, N) : N;
        if (k1_fixed > 0) {
            const int k1 = std::min(k1_fixed, N);
            int k2 = std::min(k1, kmax);
            while (k2 > 1 && !fits(k1, k2)) k2--;
            if (!fits(k1, k2))
                return fail(ECSG_ENOMEM, "gradient: the checkpoints of interval " + std::to_string(k1) +
                                             " and a ring of one state do not fit in device memory");
            *K1_out = k1;
            *K2_out = k2;
            return ECSG_OK;
        }
        for (int k = kmax; k >= 1; k--)
            if (fits(k, k)) {
                *K1_out = *K2_out = k;
                return ECSG_OK;
            }
        long long best = -1;
        int b1 = 0, b2 = 0;
        for (int k2 = std::min(kmax, 4096); k2 >= 1; k2--) {
            if ((k2 + 1) * state + 2 * ckb > avail) continue;
            for (int m = 2; (long long)m * k2 <= (long long)2 * N; m++) {
                const int k1 = std::min(m * k2, N);
                if (k1 <= k2 || !fits(k1, k2)) continue;
                const long long r = recomputed(N, k1, k2);
                if (best < 0 || r < best) { best = r; b1 = k1; b2 = k2; }
            }
        }
        if (best < 0) return fail(ECSG_ENOMEM, "gradient: no checkpoint interval fits in device memory");
        *K1_out = b1;
        *K2_out = b2;
        return ECSG_OK;
    }
    int last_K = 0, last_K2 = 0;
    long long last_recomputed = 0;
    int checkpoint_interval() const override { return last_K; }
    void checkpoint_levels(int *k1, int *k2, long long *rec) const override
    {
        if (k1) *k1 = last_K;
        if (k2) *k2 = last_K2;
        if (rec) *rec = last_recomputed;
    }

    int ck_prepare(int nck, int K)
    {
        const int ng = (int)G.size(), ni = (int)I.size();
        if (K + 1 > ring_alloc) {
            for (auto p : ring) dfree(p);
            ring.clear();
            ring_alloc = 0;
            for (int q = 0; q <= K; q++)
                for (int g = 0; g < ng; g++) {
                    T *p = nullptr;
                    if (int rc = dalloc(&p, (size_t)9 * G[g].astride)) return rc;
                    ring.push_back(p);
                }
            ring_alloc = K + 1;
        }
        if (nck > ck_alloc || ck_fmt != ck_alloc_fmt) {
            for (auto p : ck_nat) dfree(p);
            for (auto p : ck_red) dfree(p);
            for (auto p : ck_x0) dfree(p);
            ck_nat.clear(); ck_red.clear(); ck_x0.clear();
            dfree(ck_scale); dfree(ck_part);
            for (int j = 0; j < nck; j++) {
                for (int g = 0; g < ng; g++) {
                    if (ck_fmt == 0) {
                        T *p = nullptr;
                        if (int rc = dalloc(&p, (size_t)9 * G[g].astride)) return rc;
                        ck_nat.push_back(p);
                    } else {
                        unsigned short *p = nullptr;
                        if (int rc = dalloc(&p, (size_t)9 * G[g].astride)) return rc;
                        ck_red.push_back(p);
                    }
                }
                for (int k = 0; k < ni; k++) {
                    double *p = nullptr;
                    if (int rc = dalloc(&p, (size_t)(I[k].nx2 + 1) * (I[k].ny2 + 1))) return rc;
                    ck_x0.push_back(p);
                }
            }
            if (int rc = dalloc(&ck_scale, (size_t)nck * ng * 9)) return rc;
            if (int rc = dalloc(&ck_part, (size_t)CKP)) return rc;
            ck_alloc = nck;
            ck_alloc_fmt = ck_fmt;
        }
        return ECSG_OK;
    }

    int ck_save(int j)
    {
        const int ng = (int)G.size(), ni = (int)I.size();
        for (int g = 0; g < ng; g++) {
            const long as = G[g].astride;
            if (ck_fmt == 0) {
                HIPCHK(hipMemcpyAsync(ck_nat[(size_t)j * ng + g], G[g].fields(), sizeof(T) * 9 * as, hipMemcpyDeviceToDevice, s));
            } else {
                for (int f = 0; f < 9; f++) {
                    const T *x = G[g].arr(f);
                    float *sc = ck_scale + ((size_t)j * ng + g) * 9 + f;
                    hipLaunchKernelGGL((k_maxabs<T>), dim3(CKP), dim3(256), 0, s, x, as, ck_part);
                    hipLaunchKernelGGL(k_ck_scale, dim3(1), dim3(64), 0, s, (const float *)ck_part, CKP, sc);
                    hipLaunchKernelGGL((k_ck_pack<T>), dim3(1024), dim3(256), 0, s, x, as, ck_red[(size_t)j * ng + g] + (long)f * as,
                                       (const float *)sc, ck_fmt);
                }
            }
        }
        for (int k = 0; k < ni; k++)
            HIPCHK(hipMemcpyAsync(ck_x0[(size_t)j * ni + k], I[k].X0, sizeof(double) * (I[k].nx2 + 1) * (I[k].ny2 + 1),
                                  hipMemcpyDeviceToDevice, s));
        HIPCHK(hipGetLastError());
        return ECSG_OK;
    }

    int ck_restore(int j, int step)
    {
        const int ng = (int)G.size(), ni = (int)I.size();
        for (int g = 0; g < ng; g++) {
            const long as = G[g].astride;
            if (ck_fmt == 0) {
                HIPCHK(hipMemcpyAsync(G[g].fields(), ck_nat[(size_t)j * ng + g], sizeof(T) * 9 * as, hipMemcpyDeviceToDevice, s));
            } else {
                for (int f = 0; f < 9; f++)
                    hipLaunchKernelGGL((k_ck_unpack<T>), dim3(1024), dim3(256), 0, s, G[g].arr(f), as,
                                       (const unsigned short *)(ck_red[(size_t)j * ng + g] + (long)f * as),
                                       (const float *)(ck_scale + ((size_t)j * ng + g) * 9 + f), ck_fmt);
            }
        }
        for (int k = 0; k < ni; k++)
            HIPCHK(hipMemcpyAsync(I[k].X0, ck_x0[(size_t)j * ni + k], sizeof(double) * (I[k].nx2 + 1) * (I[k].ny2 + 1),
                                  hipMemcpyDeviceToDevice, s));
        hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, s, d_step, step);
        sets_dirty = true;
        HIPCHK(hipGetLastError());
        return ECSG_OK;
    }

    // direct the forward kernels' second writes to ring slot `slot` (-1: off)
    void set_store(int slot)
    {
        const int ng = (int)G.size();
        for (int g = 0; g < ng; g++)
            for (int f = 0; f < 9; f++)
                G[g].dev.st[f] = slot < 0 ? nullptr : ring[(size_t)slot * ng + g] + (long)f * G[g].astride;
    }

    int ring_save(int slot)
    {
        const int ng = (int)G.size();
        for (int g = 0; g < ng; g++)
            HIPCHK(hipMemcpyAsync(ring[(size_t)slot * ng + g], G[g].fields(), sizeof(T) * 9 * G[g].astride, hipMemcpyDeviceToDevice, s));
        return ECSG_OK;
    }

    int gradient(int N, const double *obs, double *misfit) override
    {
        if (int rc = check_ready()) return rc;
        if (N < 1 || N > cfg.max_steps) return fail(ECSG_EINVAL, "gradient: step count out of range");
        if (nrec > 0 && !obs) return fail(ECSG_EINVAL, "gradient: null observed traces");
        if (int rc = adj_build()) return rc;
        if (int rc = build_adj_receivers()) return rc;
        int K = 0, K2 = 0;
        if (int rc = auto_interval(N, ck_every, &K, &K2)) return rc;
        K = std::min(K, N);
        K2 = std::min(K2, K);
        last_K = K;
        last_K2 = K2;
        last_recomputed = recomputed(N, K, K2);
        const bool two = K2 < K;
        const int nseg = (N + K - 1) / K;
        // segment j covers steps [seg0(j), seg0(j) + seglen(j)); the first
        // takes the remainder so that the last one, which is never
        // recomputed in one-level mode, is a full K steps: N - K steps are
        // recomputed
        const int r0 = N - (nseg - 1) * K;
        auto seg0 = [&](int j) { return j == 0 ? 0 : r0 + (j - 1) * K; };
        auto seglen = [&](int j) { return j == 0 ? r0 : K; };
        if (int rc = ck_prepare(ck_count(N, K, K2), K2)) return rc;
        if (int rc = reset()) return rc;
        // forward pass with checkpoints; in one-level mode the last segment
        // fills the ring directly (the states the backward pass needs), so
        // it is never recomputed; in two-level mode every segment has a
        // checkpoint
        for (int j = 0; j < nseg; j++) {
            const bool last = j == nseg - 1 && !two;
            const int len = seglen(j);
            if (!last) {
                if (int rc = ck_save(j)) return rc;
                for (int t = 0; t < len; t++)
                    if (int rc = enqueue_step(t + 1 < len)) return rc;
            } else {
                if (int rc = ring_save(0)) return rc;
                for (int t = 0; t < len; t++) {
                    set_store(t + 1);
                    const int rc = enqueue_step(t + 1 < len);
                    set_store(-1);
                    if (rc) return rc;
                }
            }
        }
        HIPCHK(hipStreamSynchronize(s));
        if (int rc = check_barrier()) return rc;
        // residuals and misfit
        std::vector<T> tr((size_t)std::max(nrec, 1) * cfg.max_steps);
        if (nrec) HIPCHK(hipMemcpy(tr.data(), d_traces, sizeof(T) * nrec * cfg.max_steps, hipMemcpyDeviceToHost));
        std::vector<T> res((size_t)std::max(nrec, 1) * N, (T)0);
        double J = 0;
        for (int q = 0; q < nrec; q++)
            for (int t = 0; t < N; t++) {
                const double rr = (double)tr[(size_t)q * cfg.max_steps + t] - obs[(size_t)q * N + t];
                res[(size_t)q * N + t] = (T)rr;
                J += 0.5 * rr * rr;
            }
        if (misfit) *misfit = J;
        dfree(d_resid);
        if (int rc = upload(&d_resid, res)) return rc;
        resid_steps = N;
        if (int rc = adj_zero()) return rc;
        // recompute `len` steps from the current state into ring slots
        // 1..len (slot 0: the state before them), then the transposed steps
        // over them, the last first
        auto ring_pass = [&](int len) -> int {
            if (int rc = ring_save(0)) return rc;
            for (int t = 0; t < len; t++) {
                // the step's kernels also write every value into ring slot t+1
                set_store(t + 1);
                const int rc = enqueue_step(t + 1 < len);
                set_store(-1);
                if (rc) return rc;
            }
            return ECSG_OK;
        };
        auto adj_pass = [&](int step0, int len) -> int {
            for (int t = len - 1; t >= 0; t--)
                if (int rc = adj_step_impl(step0 + t, t + 1, t, true)) return rc;
            return ECSG_OK;
        };
        // backward over the segments
        int left_at = N;  // the step the last recomputation ended at
        for (int j = nseg - 1; j >= 0; j--) {
            const int len = seglen(j), a = seg0(j);
            if (!two) {
                if (j < nseg - 1) {
                    if (int rc = ck_restore(j, a)) return rc;
                    if (int rc = ring_pass(len)) return rc;
                    left_at = a + len;
                }
                if (int rc = adj_pass(a, len)) return rc;
                continue;
            }
            // two levels: sub-segment i covers [a + sub0(i), + sublen(i)),
            // the first taking the remainder; one pass over the segment
            // leaves a checkpoint at every sub-segment but the last, which
            // goes into the ring; each other sub-segment is recomputed from
            // its checkpoint into the ring before its transposed steps
            const int nsub = (len + K2 - 1) / K2, rs = len - (nsub - 1) * K2;
            auto sub0 = [&](int i) { return i == 0 ? 0 : rs + (i - 1) * K2; };
            auto sublen = [&](int i) { return i == 0 ? rs : K2; };
            if (int rc = ck_restore(j, a)) return rc;
            for (int i = 0; i + 1 < nsub; i++) {
                if (int rc = ck_save(nseg + i)) return rc;
                const int sl = sublen(i);
                for (int t = 0; t < sl; t++)
                    if (int rc = enqueue_step(t + 1 < sl)) return rc;
            }
            if (int rc = ring_pass(sublen(nsub - 1))) return rc;
            if (int rc = adj_pass(a + sub0(nsub - 1), sublen(nsub - 1))) return rc;
            left_at = a + len;
            for (int i = nsub - 2; i >= 0; i--) {
                if (int rc = ck_restore(nseg + i, a + sub0(i))) return rc;
                if (int rc = ring_pass(sublen(i))) return rc;
                if (int rc = adj_pass(a + sub0(i), sublen(i))) return rc;
                left_at = a + sub0(i) + sublen(i);
            }
        }
        HIPCHK(hipStreamSynchronize(s));
        if (int rc = check_barrier()) return rc;
        host_step = left_at;  // the state the last recomputation left
        for (int k = 0; k < (int)I.size(); k++) {
            long long st3[3];
            HIPCHK(hipMemcpy(st3, AI[k].stats, sizeof(st3), hipMemcpyDeviceToHost));
            if (st3[2]) return fail(ECSG_ENOCONV, "adjoint Jacobi iteration cap reached on interface " + std::to_string(k));
        }
        gprim.clear();
        grad_ready = true;
        return ECSG_OK;
    }
    bool grad_ready = false;

    // transposed parameter staging: gradients at the staged positions ->
    // gradients w.r.t. the primary-node inputs of set_params
    int finalize_gradient()
    {
        const int ng = (int)G.size();
        gprim.assign((size_t)ng * 10, std::vector<double>());
        for (int g = 0; g < ng; g++) {
            const DevGrid<T> &d = G[g];
            const long as = d.astride;
            std::vector<T> hg((size_t)12 * as);
            HIPCHK(hipMemcpy(hg.data(), AG[g].grad, sizeof(T) * 12 * as, hipMemcpyDeviceToHost));
            const int n0 = d.n[0], n1 = d.n[1], n2 = d.n[2];
            const long px = n0 + 1, py = n1 + 1;
            const long np = px * py * (n2 + 1);
            auto PI = [&](int i, int j, int k) { return ((long)k * py + j) * px + i; };
            auto DI = [&](int i, int j, int k) { return d.o + (long)k * d.S2 + (long)j * d.P + i; };
            std::vector<std::vector<double>> gm(9, std::vector<double>((size_t)np, 0.0));
            for (int m = 0; m < 9; m++) {
                const int *pos = ST_POS[m];
                const int e0 = pos[0] ? n0 - 1 : n0, e1 = pos[1] ? n1 - 1 : n1, e2 = pos[2] ? n2 - 1 : n2;
                std::vector<double> &o = gm[m];
                if (d.staged) {
                    // w.r.t. the staged Field itself, in its own layout
                    o.assign((size_t)(e0 + 1) * (e1 + 1) * (e2 + 1), 0.0);
                    long q = 0;
                    for (int k = 0; k <= e2; k++)
                        for (int j = 0; j <= e1; j++)
                            for (int i = 0; i <= e0; i++, q++) o[q] = (double)hg[(size_t)m * as + DI(i, j, k)];
                    continue;
                }
                for (int k = 0; k <= e2; k++)
                    for (int j = 0; j <= e1; j++)
                        for (int i = 0; i <= e0; i++) {
                            const double v = (double)hg[(size_t)m * as + DI(i, j, k)] / 8;
                            if (v == 0) continue;
                            for (int c = 0; c < 8; c++)
                                o[PI(i + ((c >> 2) & 1) * pos[0], j + ((c >> 1) & 1) * pos[1], k + (c & 1) * pos[2])] += v;
                        }
            }
            if (d.iso && !d.staged) {
                std::vector<double> s11((size_t)np, 0.0), s13((size_t)np, 0.0);
                for (long q = 0; q < np; q++) {
                    s11[q] = gm[0][q] + gm[1][q] + gm[2][q] + gm[6][q] + gm[7][q] + gm[8][q];
                    s13[q] = gm[3][q] + gm[4][q] + gm[5][q] - gm[6][q] - gm[7][q] - gm[8][q];
                }
                for (int m = 0; m < 9; m++) gprim[(size_t)g * 10 + 1 + m].assign((size_t)np, 0.0);
                gprim[(size_t)g * 10 + 1 + 0] = s11;
                gprim[(size_t)g * 10 + 1 + 4] = s13;
            } else {
                for (int m = 0; m < 9; m++) gprim[(size_t)g * 10 + 1 + m] = gm[m];
            }
            std::vector<double> gb((size_t)np, 0.0);
            for (int k = 0; k <= n2; k++)
                for (int j = 0; j <= n1; j++)
                    for (int i = 0; i <= n0; i++) gb[PI(i, j, k)] += (double)hg[(size_t)9 * as + DI(i, j, k)];
            for (int k = 0; k <= n2; k++)
                for (int j = 0; j <= n1 - 1; j++)
                    for (int i = 0; i <= n0 - 1; i++) {
                        const double v = 0.25 * (double)hg[(size_t)10 * as + DI(i, j, k)];
                        if (v == 0) continue;
                        gb[PI(i, j, k)] += v; gb[PI(i + 1, j, k)] += v; gb[PI(i, j + 1, k)] += v; gb[PI(i + 1, j + 1, k)] += v;
                    }
            for (int k = 0; k <= n2 - 1; k++)
                for (int j = 0; j <= n1; j++)
                    for (int i = 0; i <= n0 - 1; i++) {
                        const double v = 0.25 * (double)hg[(size_t)11 * as + DI(i, j, k)];
                        if (v == 0) continue;
                        gb[PI(i, j, k)] += v; gb[PI(i, j, k + 1)] += v; gb[PI(i + 1, j, k)] += v; gb[PI(i + 1, j, k + 1)] += v;
                    }
            gprim[(size_t)g * 10] = gb;
        }
        return ECSG_OK;
    }

    int get_gradient(int g, int sel, double *out) override
    {
        if (g < 0 || g >= (int)G.size()) return fail(ECSG_EINVAL, "grid index out of range");
        if (sel < ECSG_BUOY || sel > ECSG_C11 + 8) return fail(ECSG_EINVAL, "selector must be ECSG_BUOY or ECSG_C11..ECSG_C11+8");
        if (!grad_ready) return fail(ECSG_ESTATE, "no gradient computed yet");
        if (gprim.size() != G.size() * 10)
            if (int rc = finalize_gradient()) return rc;
        const std::vector<double> &v = gprim[(size_t)g * 10 + (sel == ECSG_BUOY ? 0 : 1 + sel - ECSG_C11)];
        std::copy(v.begin(), v.end(), out);
        return ECSG_OK;
    }

    int adjoint_step() override
    {
        if (int rc = adj_build()) return rc;
        if (int rc = adj_coef()) return rc;
        for (int k = 0; k < (int)I.size(); k++) {
            const long nc = (long)I[k].nx2 * I[k].ny2;
            hipLaunchKernelGGL((k_gsyz_den<T>), dim3(nb(nc)), dim3(256), 0, s, iface_dev(k), AI[k].den);
        }
        if (int rc = adj_step_impl(0, -1, -1, false)) return rc;
        HIPCHK(hipStreamSynchronize(s));
        if (int rc = check_barrier()) return rc;
        for (int k = 0; k < (int)I.size(); k++) {
            long long st3[3];
            HIPCHK(hipMemcpy(st3, AI[k].stats, sizeof(st3), hipMemcpyDeviceToHost));
            if (st3[2]) return fail(ECSG_ENOCONV, "adjoint Jacobi iteration cap reached on interface " + std::to_string(k));
        }
        return ECSG_OK;
    }

    int adj_field(int g, int sel, double *out, const double *in) override
    {
        if (int rc = adj_build()) return rc;
        int len[3];
        const int r = dims(g, sel, len);
        if (r < 0 || sel >= 9) return fail(ECSG_EINVAL, "adjoint fields are wavefields (0..8)");
        DevGrid<T> &d = G[g];
        HIPCHK(hipStreamSynchronize(s));
        std::vector<T> buf((size_t)d.astride);
        T *dev = AG[g].lam + (long)sel * d.astride;
        HIPCHK(hipMemcpy(buf.data(), dev, sizeof(T) * d.astride, hipMemcpyDeviceToHost));
        long q = 0;
        for (int k = -r; k < len[2] - r; k++)
            for (int j = -r; j < len[1] - r; j++)
                for (int i = -r; i < len[0] - r; i++, q++) {
                    T &x = buf[d.o + (long)k * d.S2 + (long)j * d.P + i];
                    if (in) x = (T)in[q];
                    if (out) out[q] = (double)x;
                }
        if (in) HIPCHK(hipMemcpy(dev, buf.data(), sizeof(T) * d.astride, hipMemcpyHostToDevice));
        return ECSG_OK;
    }

    // per grid: the stencil tiling (workgroup height, z chunk) with the
    // shortest velocity + stress launches, timed on a saved copy of the state
    int autotune(int reps, int *by_out, int *zc_out) override
    {
        if (int rc = check_ready()) return rc;
        if (reps < 1) return fail(ECSG_EINVAL, "reps must be >= 1");
        if (cfg.radius != 1) return fail(ECSG_EINVAL, "the tiling autotuner is for the radius-1 stencils");
        HIPCHK(hipStreamSynchronize(s));
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        const int bys[4] = {1, 2, 4, 8}, zcs[5] = {1, 2, 4, 8, 16};
        for (int g = 0; g < (int)G.size(); g++) {
            DevGrid<T> &d = G[g];
            T *save = nullptr;
            HIPCHK(hipMalloc((void **)&save, sizeof(T) * 9 * d.astride));
            HIPCHK(hipMemcpy(save, d.fields(), sizeof(T) * 9 * d.astride, hipMemcpyDeviceToDevice));
            float best = 1e30f;
            int bb = d.by, bz = d.dev.zc;
            for (int b : bys)
                for (int z : zcs) {
                    set_tiling(d, b, z);
                    launch_vel(g);
                    launch_stress(g);
                    HIPCHK(hipEventRecord(e0, s));
                    for (int r = 0; r < reps; r++) {
                        launch_vel(g);
                        launch_stress(g);
                    }
                    HIPCHK(hipEventRecord(e1, s));
                    HIPCHK(hipEventSynchronize(e1));
                    float ms = 0;
                    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
                    if (ms < best) { best = ms; bb = b; bz = z; }
                }
            set_tiling(d, bb, bz);
            HIPCHK(hipMemcpy(d.fields(), save, sizeof(T) * 9 * d.astride, hipMemcpyDeviceToDevice));
            hipFree(save);
            if (by_out) by_out[g] = bb;
            if (zc_out) zc_out[g] = bz;
        }
        hipEventDestroy(e0);
        hipEventDestroy(e1);
        invalidate_graph();
        return ECSG_OK;
    }

    int adjoint_stats(long long *sweeps, double *unused) override
    {
        long long tot = 0;
        for (int k = 0; k < (int)AI.size(); k++) {
            long long st3[3];
            HIPCHK(hipMemcpy(st3, AI[k].stats, sizeof(st3), hipMemcpyDeviceToHost));
            tot += st3[0];
        }
        if (sweeps) *sweeps = tot;
        if (unused) *unused = 0;
        return ECSG_OK;
    }
};


// ---------------------------------------------------------------------------
// Batch: many models advanced together, one launch per phase of the step.
class BatchBase {
public:
    virtual ~BatchBase() {}
    Status st;
    int fail(int code, const std::string &m)
    {
        st.code = code;
        st.msg = m;
        return code;
    }
    virtual int step_n(int n, int use_graph) = 0;
    virtual int sync() = 0;
    virtual int set_tiling(int by, int zc) = 0;
    virtual int autotune(int reps, int *by, int *zc) = 0;
    virtual int gradient(int nsteps, const double *const *obs, double *misfit, const int *group) = 0;
    virtual int last_interval() const = 0;
};

template <typename T>
class Batch : public BatchBase {
public:
    std::vector<Engine<T> *> M;
    hipStream_t s = nullptr;
    hipEvent_t done = nullptr;
    int by = 4, zc = ZCHUNK;
    GridDev<T> *d_g = nullptr;
    int *d_gs = nullptr;
    int ng = 0, gtot = 0;
    bool all32 = true;  // every member grid takes the 32-bit-offset stencil sweeps
    IfWork<T> *d_if = nullptr;
    int *d_is[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    int ni = 0, itot[5] = {0, 0, 0, 0, 0};
    JacLds<T> *d_jl = nullptr;
    long maxw = 0;
    long maxn_vz = 0;  // longest vz-szz line (the wavefront solves' LDS)
    PointArgs<T> *d_pa = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    // Interface overlap of the batched step (as Engine::enqueue_step_overlap,
    // schedule 1): the stencils over four plane sets -- velocity interiors
    // and edge planes, stress interiors and edge planes of every member grid
    // (tables d_pg[q], tile offsets d_ps[q]) -- on s, the vz ghosts and the
    // sxz/syz ghosts of every interface on s2.  A captured graph holds
    // OV_STEPS steps, each step's sxz/syz ghosts beside the next one's
    // velocity interiors.
    static constexpr int OV_STEPS = 8;
    bool ov = false;                   // ECSG_BATCH_OVERLAP=0: the serial step
    hipStream_t s2 = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    GridDev<T> *d_pg[4] = {nullptr, nullptr, nullptr, nullptr};
    int *d_ps[4] = {nullptr, nullptr, nullptr, nullptr};
    int pn[4] = {0, 0, 0, 0}, ptot[4] = {0, 0, 0, 0};
    hipGraph_t graph_ov[2] = {nullptr, nullptr};        // OV_STEPS steps, one step
    hipGraphExec_t gexec_ov[2] = {nullptr, nullptr};

    ~Batch() override
    {
        // let launched steps finish before their graphs and tables go
        if (s) hipStreamSynchronize(s);
        if (s2) hipStreamSynchronize(s2);
        drop_graph();
        hipFree(d_g); hipFree(d_gs); hipFree(d_if); hipFree(d_jl); hipFree(d_pa);
        for (auto p : d_is) hipFree(p);
        for (auto p : d_pg) hipFree(p);
        for (auto p : d_ps) hipFree(p);
        free_grad_tables();
        if (done) hipEventDestroy(done);
        for (auto e : ev)
            if (e) hipEventDestroy(e);
        if (s2) hipStreamDestroy(s2);
        if (s) hipStreamDestroy(s);
    }
    void drop_graph()
    {
        if (gexec) { hipGraphExecDestroy(gexec); gexec = nullptr; }
        if (graph) { hipGraphDestroy(graph); graph = nullptr; }
        for (int q = 0; q < 2; q++) {
            if (gexec_ov[q]) { hipGraphExecDestroy(gexec_ov[q]); gexec_ov[q] = nullptr; }
            if (graph_ov[q]) { hipGraphDestroy(graph_ov[q]); graph_ov[q] = nullptr; }
        }
    }
    template <typename P>
    int upload(P **dst, const std::vector<P> &v)
    {
        hipFree(*dst);
        *dst = nullptr;
        HIPCHK(hipMalloc((void **)dst, sizeof(P) * std::max<size_t>(v.size(), 1)));
        if (!v.empty()) HIPCHK(hipMemcpy(*dst, v.data(), sizeof(P) * v.size(), hipMemcpyHostToDevice));
        return ECSG_OK;
    }

    int init(const std::vector<Engine<T> *> &members)
    {
        M = members;
        for (size_t m = 0; m < M.size(); m++) {
            Engine<T> *e = M[m];
            const std::string id = "member " + std::to_string(m) + ": ";
            if (!e->all_params()) return fail(ECSG_ESTATE, id + "parameters not set for every grid");
            if (e->cfg.check_energy) return fail(ECSG_EINVAL, id + "batched models run without the energy check");
            if (e->cfg.sxz_solver != 0) return fail(ECSG_EINVAL, id + "batched models use the Jacobi sweeps (sxz_solver 0)");
            if (e->tol_cap != M[0]->tol_cap || e->tol_fac != M[0]->tol_fac)
                return fail(ECSG_EINVAL, id + "members must share one Jacobi tolerance");
            if (e->cfg.imode != 1 && !e->I.empty()) return fail(ECSG_EINVAL, id + "batched composite models use the energy-conserving interface (imode 1)");
            if (e->cfg.radius != 1) return fail(ECSG_EINVAL, id + "batched models are radius 1");
            if (!e->I.empty() && !e->lds_ok) return fail(ECSG_EINVAL, id + "interface too large for the LDS Jacobi of a batch");
            for (const DevIface &f : e->I)
                if ((size_t)32 * (f.ny2 - 1) > (size_t)VZ_WAVE_LDS) return fail(ECSG_EINVAL, id + "vz-szz line too long for a batch");
        }
        HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        HIPCHK(hipEventCreateWithFlags(&done, hipEventDisableTiming));
        if (const char *v = getenv("ECSG_BATCH_PERSIST")) persist = atoi(v);
        {
            int dev = 0;
            HIPCHK(hipGetDevice(&dev));
            HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
        }
        bool any_iface = false;
        for (Engine<T> *e : M) any_iface = any_iface || !e->I.empty();
        ov = any_iface;
        if (const char *v = getenv("ECSG_BATCH_OVERLAP")) ov = ov && atoi(v) != 0;
        if (ov) {
            HIPCHK(hipStreamCreateWithFlags(&s2, hipStreamNonBlocking));
            for (auto &e : ev) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        }
        return build_tables();
    }

    // descriptor tables of the members' current field sets
    std::vector<int> built_cur;
    int build_tables()
    {
        built_cur.clear();
        for (Engine<T> *e : M) built_cur.push_back(e->G[0].cur);
        maxw = 0;
        maxn_vz = 0;
        if (int rc = build_grids()) return rc;
        // interfaces, Jacobi descriptors and point kernels
        std::vector<IfWork<T>> ifs;
        std::vector<int> st[5];
        std::vector<JacLds<T>> jl;
        std::vector<PointArgs<T>> pa;
        int acc[5] = {0, 0, 0, 0, 0};
        for (Engine<T> *e : M) {
            pa.push_back(e->point_args());
            for (int k = 0; k < (int)e->I.size(); k++) {
                DevIface &f = e->I[k];
                IfWork<T> w{};
                w.a = e->iface_dev(k);
                w.B = f.B; w.M = f.M; w.Rp = f.Rp; w.U = f.U; w.Bs = f.Bs; w.absp = f.absp;
                ifs.push_back(w);
                const long r = f.r;
                const long nb[5] = {((long)f.nx2 * (f.ny2 - 1) + 255) / 256, (long)f.nx2,
                                    (r * f.nx2 * (r * f.ny2 - 1) + 255) / 256, (long)f.nabs, ((long)f.nx2 * f.ny2 + 255) / 256};
                for (int q = 0; q < 5; q++) { st[q].push_back(acc[q]); acc[q] += (int)nb[q]; }
                jl.push_back(e->jac_lds(k, false));
                maxw = std::max(maxw, (long)(f.nx2 + 1) * (f.ny2 + 1));
                maxn_vz = std::max(maxn_vz, (long)(f.ny2 - 1));
            }
        }
        ni = (int)ifs.size();
        for (int q = 0; q < 5; q++) itot[q] = acc[q];
        if (int rc = upload(&d_if, ifs)) return rc;
        for (int q = 0; q < 5; q++)
            if (int rc = upload(&d_is[q], st[q])) return rc;
        if (int rc = upload(&d_jl, jl)) return rc;
        if (int rc = upload(&d_pa, pa)) return rc;
        if (ni) HIPCHK(hipFuncSetAttribute((const void *)k_jacobi_lds_batch<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)(2 * JL_MAXW * sizeof(double))));
        return ECSG_OK;
    }

    // flat tile list over every grid of every member, with the batch tiling
    int build_grids()
    {
        std::vector<GridDev<T>> gs;
        std::vector<int> st;
        int acc = 0;
        for (Engine<T> *e : M)
            for (auto &d : e->G) {
                GridDev<T> v = d.dev;
                for (int f = 0; f < 9; f++) v.st[f] = nullptr;
                v.zc = zc;
                v.gx = (d.n[0] + 1 + BX - 1) / BX;
                v.gy = (d.n[1] + 1 + by - 1) / by;
                v.nblk = v.gx * v.gy * ((d.n[2] + 1 + zc - 1) / zc);
                v.zb = 0;
                v.ze = d.n[2] + 1;
                gs.push_back(v);
                st.push_back(acc);
                acc += v.nblk;
            }
        ng = (int)gs.size();
        gtot = acc;
        all32 = true;
        for (const auto &v : gs) all32 = all32 && fits32(v);
        if (int rc = upload(&d_g, gs)) return rc;
        if (int rc = upload(&d_gs, st)) return rc;
        // the plane sets of the overlapped step (Engine::vlo .. shi)
        for (int q = 0; q < 4; q++) {
            std::vector<GridDev<T>> pg;
            std::vector<int> ps;
            ptot[q] = plane_set(q, gs.data(), pg, ps);
            pn[q] = (int)pg.size();
            if (int rc = upload(&d_pg[q], pg)) return rc;
            if (int rc = upload(&d_ps[q], ps)) return rc;
        }
        drop_graph();
        return ECSG_OK;
    }
    // Plane set q of the overlapped step (0: velocity interiors, 1: velocity
    // edge planes, 2: stress interiors, 3: stress edge planes) of the member
    // grids gs[0..] in member order: descriptors appended to pg, their first
    // tiles to ps; returns the set's tile count.  The same for every table
    // of one tiling, so one ps serves the gradient's ring-slot tables too.
    int plane_set(int q, const GridDev<T> *gs, std::vector<GridDev<T>> &pg, std::vector<int> &ps) const
    {
        int pacc = 0;
        int k = 0;
        for (Engine<T> *e : M)
            for (int g = 0; g < (int)e->G.size(); g++, k++) {
                const int n1 = e->G[g].n[2] + 1;
                const int lo = q < 2 ? e->vlo(g) : e->slo(g), hi = q < 2 ? e->vhi(g) : e->shi(g);
                int r[2][2] = {{lo, hi}, {0, 0}};
                if (q == 1 || q == 3) {
                    if (lo >= hi) { r[0][0] = 0; r[0][1] = n1; }
                    else { r[0][0] = 0; r[0][1] = lo; r[1][0] = hi; r[1][1] = n1; }
                }
                for (auto &ab : r) {
                    if (ab[1] <= ab[0]) continue;
                    GridDev<T> v = gs[k];
                    v.zb = ab[0];
                    v.ze = ab[1];
                    v.nblk = v.gx * v.gy * ((ab[1] - ab[0] + zc - 1) / zc);
                    pg.push_back(v);
                    ps.push_back(pacc);
                    pacc += v.nblk;
                }
            }
        return pacc;
    }

    // workgroups of a batched stencil launch over `tiles` tiles: one per
    // tile, or (ECSG_BATCH_PERSIST = k > 0) k per CU, a multiple of 8, each
    // looping over tiles
    int persist = 0, ncu = 0;
    int wgs(int tiles) const
    {
        if (persist <= 0 || ncu <= 0) return tiles;
        const int w = (persist * ncu + 7) / 8 * 8;
        return std::min(tiles, w);
    }
    template <int BYv>
    void stencils(bool vel, const GridDev<T> *g)
    {
        const dim3 gr(wgs(gtot)), bl(BX, BYv);
        if (all32) {
            if (vel) hipLaunchKernelGGL((k_vel_batch<T, BYv, unsigned>), gr, bl, 0, s, g, (const int *)d_gs, ng, gtot);
            else hipLaunchKernelGGL((k_stress_batch<T, BYv, unsigned>), gr, bl, 0, s, g, (const int *)d_gs, ng, gtot);
        } else {
            if (vel) hipLaunchKernelGGL((k_vel_batch<T, BYv, long>), gr, bl, 0, s, g, (const int *)d_gs, ng, gtot);
            else hipLaunchKernelGGL((k_stress_batch<T, BYv, long>), gr, bl, 0, s, g, (const int *)d_gs, ng, gtot);
        }
    }
    void launch_stencils(bool vel, const GridDev<T> *g = nullptr)
    {
        if (!g) g = d_g;
        switch (by) {
        case 1: stencils<1>(vel, g); break;
        case 2: stencils<2>(vel, g); break;
        case 8: stencils<8>(vel, g); break;
        default: stencils<4>(vel, g); break;
        }
    }

    template <int BYv>
    void stencils_set(bool vel, int q, const GridDev<T> *g)
    {
        if (ptot[q] == 0) return;
        const dim3 gr(wgs(ptot[q])), bl(BX, BYv);
        const int *st = d_ps[q];
        if (all32) {
            if (vel) hipLaunchKernelGGL((k_vel_batch<T, BYv, unsigned>), gr, bl, 0, s, g, st, pn[q], ptot[q]);
            else hipLaunchKernelGGL((k_stress_batch<T, BYv, unsigned>), gr, bl, 0, s, g, st, pn[q], ptot[q]);
        } else {
            if (vel) hipLaunchKernelGGL((k_vel_batch<T, BYv, long>), gr, bl, 0, s, g, st, pn[q], ptot[q]);
            else hipLaunchKernelGGL((k_stress_batch<T, BYv, long>), gr, bl, 0, s, g, st, pn[q], ptot[q]);
        }
    }
    void launch_set(int q, const GridDev<T> *g)
    {
        switch (by) {
        case 1: stencils_set<1>(q < 2, q, g); break;
        case 2: stencils_set<2>(q < 2, q, g); break;
        case 8: stencils_set<8>(q < 2, q, g); break;
        default: stencils_set<4>(q < 2, q, g); break;
        }
    }
    void launch_set(int q) { launch_set(q, d_pg[q]); }

    // nsteps batched time steps with the interface overlap: every member
    // grid's velocity interior beside the previous step's sxz/syz ghosts,
    // then (after them) the velocity edge planes and the receivers/sources;
    // the vz ghosts beside the stress interiors, the stress edge planes after
    // them; the last step's sxz/syz ghosts are joined at the end.  Every
    // entry sees the operations of the serial step in the same order.
    // tables of one overlapped step: the plane sets (as d_pg), interface
    // work, LDS Jacobi and point descriptors
    struct OvTables {
        const GridDev<T> *pg[4];
        const IfWork<T> *w;
        const JacLds<T> *jl;
        const PointArgs<T> *pa;
    };
    OvTables ov_tables() const
    {
        return OvTables{{d_pg[0], d_pg[1], d_pg[2], d_pg[3]}, d_if, d_jl, d_pa};
    }
    int enqueue_ov(int nsteps)
    {
        return enqueue_ov_steps(nsteps, [this](int) { return ov_tables(); });
    }
    template <typename F>
    int enqueue_ov_steps(int nsteps, F tables)
    {
        for (int i = 0; i < nsteps; i++) {
            const OvTables tb = tables(i);
            launch_set(0, tb.pg[0]);
            if (i > 0) HIPCHK(hipStreamWaitEvent(s, ev[3], 0));
            launch_set(1, tb.pg[1]);
            hipLaunchKernelGGL((k_recsrc_batch<T>), dim3((int)M.size()), dim3(256), 0, s, tb.pa);
            HIPCHK(hipEventRecord(ev[0], s));
            HIPCHK(hipStreamWaitEvent(s2, ev[0], 0));
            hipLaunchKernelGGL((k_gvz_rhs_batch<T>), dim3(itot[0]), dim3(256), 0, s2, tb.w, (const int *)d_is[0], ni);
            hipLaunchKernelGGL((k_gvz_solve_wave_batch<T>), dim3(itot[1]), dim3(64), (size_t)(32 * maxn_vz), s2, tb.w,
                               (const int *)d_is[1], ni);
            hipLaunchKernelGGL((k_gvz_fine_batch<T>), dim3(itot[2]), dim3(256), 0, s2, tb.w, (const int *)d_is[2], ni);
            HIPCHK(hipEventRecord(ev[1], s2));
            launch_set(2, tb.pg[2]);
            HIPCHK(hipStreamWaitEvent(s, ev[1], 0));
            launch_set(3, tb.pg[3]);
            HIPCHK(hipEventRecord(ev[2], s));
            HIPCHK(hipStreamWaitEvent(s2, ev[2], 0));
            hipLaunchKernelGGL((k_gsxz_rhs_batch<T>), dim3(itot[3]), dim3(256), 0, s2, tb.w, (const int *)d_is[3], ni);
            JacParams jp{};
            jp.tol_cap = M[0]->tol_cap; jp.tol_fac = M[0]->tol_fac;
            jp.floor = 1024 * 2.220446049250313e-16;
            jp.cap = JACOBI_CAP;
            hipLaunchKernelGGL((k_jacobi_lds_batch<T>), dim3(ni), dim3(JL), (size_t)(2 * maxw * sizeof(double)), s2, tb.jl, jp);
            hipLaunchKernelGGL((k_gsyz_batch<T>), dim3(itot[4]), dim3(256), 0, s2, tb.w, (const int *)d_is[4], ni);
            HIPCHK(hipEventRecord(ev[3], s2));
        }
        if (nsteps > 0) HIPCHK(hipStreamWaitEvent(s, ev[3], 0));
        HIPCHK(hipGetLastError());
        return ECSG_OK;
    }
    int capture_ov(int q, int nsteps)
    {
        if (gexec_ov[q]) return ECSG_OK;
        HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        const int rc = enqueue_ov(nsteps);
        hipGraph_t gr = nullptr;
        const hipError_t e2 = hipStreamEndCapture(s, &gr);
        if (rc != ECSG_OK || e2 != hipSuccess || !gr) {
            if (gr) hipGraphDestroy(gr);
            hipGetLastError();
            return fail(ECSG_EHIP, "graph capture of the overlapped batched step failed");
        }
        graph_ov[q] = gr;
        HIPCHK(hipGraphInstantiate(&gexec_ov[q], gr, nullptr, nullptr, 0));
        return ECSG_OK;
    }

    // one batched time step (main.cpp:165-261 for every member)
    int enqueue() { return enqueue_tables(d_g, d_if, d_jl, d_pa); }

    // the same over other descriptor tables (those of a ring-store slot)
    int enqueue_tables(const GridDev<T> *g, const IfWork<T> *w, const JacLds<T> *jl, const PointArgs<T> *pa)
    {
        launch_stencils(true, g);
        hipLaunchKernelGGL((k_recsrc_batch<T>), dim3((int)M.size()), dim3(256), 0, s, pa);
        if (ni) {
            hipLaunchKernelGGL((k_gvz_rhs_batch<T>), dim3(itot[0]), dim3(256), 0, s, w, (const int *)d_is[0], ni);
            hipLaunchKernelGGL((k_gvz_solve_wave_batch<T>), dim3(itot[1]), dim3(64), (size_t)(32 * maxn_vz), s, w, (const int *)d_is[1], ni);
            hipLaunchKernelGGL((k_gvz_fine_batch<T>), dim3(itot[2]), dim3(256), 0, s, w, (const int *)d_is[2], ni);
        }
        launch_stencils(false, g);
        if (ni) {
            hipLaunchKernelGGL((k_gsxz_rhs_batch<T>), dim3(itot[3]), dim3(256), 0, s, w, (const int *)d_is[3], ni);
            JacParams jp{};
            jp.tol_cap = M[0]->tol_cap; jp.tol_fac = M[0]->tol_fac;
            jp.floor = 1024 * 2.220446049250313e-16;
            jp.cap = JACOBI_CAP;
            hipLaunchKernelGGL((k_jacobi_lds_batch<T>), dim3(ni), dim3(JL), (size_t)(2 * maxw * sizeof(double)), s, jl, jp);
            hipLaunchKernelGGL((k_gsyz_batch<T>), dim3(itot[4]), dim3(256), 0, s, w, (const int *)d_is[4], ni);
        }
        HIPCHK(hipGetLastError());
        return ECSG_OK;
    }

    int step_n(int n, int use_graph) override
    {
        if (n < 0) return fail(ECSG_EINVAL, "negative step count");
        for (Engine<T> *e : M) {
            if (e->host_step + n > e->cfg.max_steps) return fail(ECSG_EINVAL, "max_steps exceeded by a member");
            HIPCHK(hipStreamSynchronize(e->s));
        }
        // a member stepped on its own since (fused steps swap its field sets)
        for (size_t m = 0; m < M.size(); m++)
            if (M[m]->G[0].cur != built_cur[m]) {
                if (int rc = build_tables()) return rc;
                break;
            }
        if (ov) {
            if (use_graph) {
                int left = n;
                if (left >= OV_STEPS) {
                    if (int rc = capture_ov(0, OV_STEPS)) return rc;
                    for (; left >= OV_STEPS; left -= OV_STEPS) HIPCHK(hipGraphLaunch(gexec_ov[0], s));
                }
                if (left > 0) {
                    if (int rc = capture_ov(1, 1)) return rc;
                    for (; left > 0; left--) HIPCHK(hipGraphLaunch(gexec_ov[1], s));
                }
            } else if (int rc = enqueue_ov(n)) {
                return rc;
            }
        } else {
            if (use_graph && !gexec) {
                HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                const int rc = enqueue();
                hipGraph_t gr = nullptr;
                const hipError_t e2 = hipStreamEndCapture(s, &gr);
                if (rc != ECSG_OK || e2 != hipSuccess || !gr) {
                    if (gr) hipGraphDestroy(gr);
                    return fail(ECSG_EHIP, "graph capture of the batched step failed");
                }
                graph = gr;
                HIPCHK(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
            }
            for (int i = 0; i < n; i++) {
                if (use_graph) HIPCHK(hipGraphLaunch(gexec, s));
                else if (int rc = enqueue()) return rc;
            }
        }
        // members' own calls (traces, fields) are ordered after the batch
        HIPCHK(hipEventRecord(done, s));
        for (Engine<T> *e : M) {
            HIPCHK(hipStreamWaitEvent(e->s, done, 0));
            e->host_step += n;
        }
        return ECSG_OK;
    }

    int sync() override
    {
        HIPCHK(hipStreamSynchronize(s));
        for (size_t m = 0; m < M.size(); m++)
            if (int rc = M[m]->sync()) return fail(rc, "member " + std::to_string(m) + ": " + M[m]->st.msg);
        return ECSG_OK;
    }

    int set_tiling(int b, int z) override
    {
        if (b != 1 && b != 2 && b != 4 && b != 8) return fail(ECSG_EINVAL, "tile height must be 1, 2, 4 or 8");
        if (z < 1) return fail(ECSG_EINVAL, "z chunk must be >= 1");
        by = b;
        zc = z;
        return build_grids();
    }

    // times the batched stencil launches for each tiling candidate on a saved
    // copy of the state, keeps the fastest, restores the state
    int autotune(int reps, int *by_out, int *zc_out) override
    {
        if (reps < 1) return fail(ECSG_EINVAL, "reps must be >= 1");
        for (Engine<T> *e : M) HIPCHK(hipStreamSynchronize(e->s));
        HIPCHK(hipStreamSynchronize(s));
        std::vector<T *> save;
        for (Engine<T> *e : M)
            for (auto &d : e->G) {
                T *p = nullptr;
                HIPCHK(hipMalloc((void **)&p, sizeof(T) * 9 * d.astride));
                HIPCHK(hipMemcpy(p, d.fields(), sizeof(T) * 9 * d.astride, hipMemcpyDeviceToDevice));
                save.push_back(p);
            }
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        const int bys[4] = {1, 2, 4, 8}, zcs[4] = {2, 4, 8, 16};
        // the stencil launches of the step that runs: with the interface
        // overlap, its four plane sets in their order
        auto sweeps = [&]() {
            if (ov) {
                for (int q = 0; q < 4; q++) launch_set(q);
            } else {
                launch_stencils(true);
                launch_stencils(false);
            }
        };
        // two passes over the candidates, each candidate's faster one counts
        // (one pass picked tilings whose steps differed by 8 % from run to run)
        float tbest[4][4];
        for (auto &row : tbest)
            for (auto &t : row) t = 1e30f;
        for (int pass = 0; pass < 2; pass++)
            for (int i = 0; i < 4; i++)
                for (int j = 0; j < 4; j++) {
                    if (int rc = set_tiling(bys[i], zcs[j])) return rc;
                    sweeps();
                    HIPCHK(hipEventRecord(e0, s));
                    for (int r = 0; r < reps; r++) sweeps();
                    HIPCHK(hipEventRecord(e1, s));
                    HIPCHK(hipEventSynchronize(e1));
                    float ms = 0;
                    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
                    tbest[i][j] = std::min(tbest[i][j], ms);
                }
        float best = 1e30f;
        int bb = by, bz = zc;
        for (int i = 0; i < 4; i++)
            for (int j = 0; j < 4; j++)
                if (tbest[i][j] < best) { best = tbest[i][j]; bb = bys[i]; bz = zcs[j]; }
        hipEventDestroy(e0);
        hipEventDestroy(e1);
        size_t q = 0;
        for (Engine<T> *e : M)
            for (auto &d : e->G) {
                HIPCHK(hipMemcpy(d.fields(), save[q], sizeof(T) * 9 * d.astride, hipMemcpyDeviceToDevice));
                hipFree(save[q]);
                q++;
            }
        if (int rc = set_tiling(bb, bz)) return rc;
        if (by_out) *by_out = bb;
        if (zc_out) *zc_out = bz;
        return ECSG_OK;
    }
    // ======================================================================
    // Gradients of every member (each member's ecsg_gradient, bit for bit)
    // with one launch per phase across members.  Every member keeps its own
    // checkpoints, ring of saved states, residuals and accumulators; the
    // forward steps are batched steps over descriptor tables whose second
    // destinations point into each member's ring slot, the transposed steps
    // batched launches of the adjoint kernel bodies.
    // ======================================================================
    int K_used = 0;
    int last_interval() const override { return K_used; }
    GridDev<T> *d_gF = nullptr;        // [slot][grid]: forward stencils, ring slot `slot` (0: none)
    GridDev<T> *d_pgF[4] = {nullptr, nullptr, nullptr, nullptr};  // [slot][pn[q]]: plane sets of the overlapped step
    IfWork<T> *d_ifF = nullptr;        // [slot][interface]
    JacLds<T> *d_jlF = nullptr;        // [slot][interface]
    PointArgs<T> *d_paF = nullptr;     // [slot][member]
    AdjGridDev<T> *d_ag = nullptr;     // [t][grid]: transposed step over ring slots t+1, t
    bool adj32 = false;                // every grid's adjoint sweeps index in 32 bits
    AdjIfWork<T> *d_aw = nullptr;      // [t][interface, by level]
    MirrorWork<T> *d_mw = nullptr;     // [set][grid], sets: free surface, vy y walls, other velocity mirrors
    JacLds<T> *d_jla = nullptr;        // transposed vx-sxz systems
    RecAdj<T> *d_ra = nullptr;
    int *d_gt = nullptr, *d_gtg = nullptr;   // tile starts: adjoint stencils, gradient sweeps
    int gt_tot = 0, gtg_tot = 0;
    int *d_ms[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    int ms_tot[6] = {0, 0, 0, 0, 0, 0};
    int *d_ras = nullptr;
    int ra_tot = 0;
    int *d_aps = nullptr;              // block starts, [level][phase][interfaces of the level]
    std::vector<int> lev_off, lev_n, aps_off, aps_tot;
    int naw = 0, nlev = 0;

    void free_grad_tables()
    {
        for (void *p : {(void *)d_gF, (void *)d_ifF, (void *)d_jlF, (void *)d_paF, (void *)d_ag, (void *)d_aw, (void *)d_mw,
                        (void *)d_jla, (void *)d_ra, (void *)d_gt, (void *)d_gtg, (void *)d_ras, (void *)d_aps})
            hipFree(p);
        for (auto &p : d_ms) { hipFree(p); p = nullptr; }
        for (auto &p : d_pgF) { hipFree(p); p = nullptr; }
        d_gF = nullptr; d_ifF = nullptr; d_jlF = nullptr; d_paF = nullptr; d_ag = nullptr; d_aw = nullptr; d_mw = nullptr;
        d_jla = nullptr; d_ra = nullptr; d_gt = nullptr; d_gtg = nullptr; d_ras = nullptr; d_aps = nullptr;
    }

    // members' calls issued on the batch stream for the duration of a scope
    struct MembersOnBatch {
        Batch *b;
        std::vector<hipStream_t> saved;
        explicit MembersOnBatch(Batch *x) : b(x)
        {
            for (Engine<T> *e : b->M) { saved.push_back(e->s); e->s = b->s; }
        }
        ~MembersOnBatch()
        {
            for (size_t m = 0; m < b->M.size(); m++) b->M[m]->s = saved[m];
        }
    };

    // the largest interval whose rings and checkpoints of all members fit in
    // free device memory (with a reserve)
    int batch_interval(int N, int *K_out)
    {
        size_t fr = 0, tot = 0;
        HIPCHK(hipMemGetInfo(&fr, &tot));
        double state = 0, ckb = 0, held = 0;
        for (Engine<T> *e : M) {
            double st1 = 0, ck1 = 0;
            for (const DevGrid<T> &d : e->G) {
                st1 += 9.0 * (double)d.astride * sizeof(T);
                ck1 += 9.0 * (double)d.astride * (e->ck_fmt == 0 ? sizeof(T) : sizeof(unsigned short));
            }
            for (const DevIface &f : e->I) ck1 += 8.0 * (f.nx2 + 1) * (f.ny2 + 1);
            state += st1;
            ckb += ck1;
            held += (double)e->ring_alloc * st1 + (double)e->ck_alloc * ck1;
        }
        const double avail = (double)fr + held - std::max(4.0e9, 0.05 * (double)tot);
        int K = 1;
        for (int k = N; k >= 1; k--) {
            const int nseg = (N + k - 1) / k;
            if ((k + 1) * state + std::max(nseg - 1, 1) * ckb <= avail) { K = k; break; }
        }
        *K_out = K;
        return ECSG_OK;
    }

    template <typename P>
    int upload_vec(P **dst, const std::vector<P> &v) { return upload(dst, v); }

    // descriptor tables of one gradient call (K: ring slots 1..K)
    int build_grad_tables(int K)
    {
        free_grad_tables();
        std::vector<GridDev<T>> gF;
        std::vector<IfWork<T>> ifF;
        std::vector<JacLds<T>> jlF;
        std::vector<PointArgs<T>> paF;
        for (int sl = 0; sl <= K; sl++) {
            for (Engine<T> *e : M) e->set_store(sl == 0 ? -1 : sl);
            for (Engine<T> *e : M) {
                for (auto &d : e->G) {
                    GridDev<T> v = d.dev;
                    v.zc = zc;
                    v.gx = (d.n[0] + 1 + BX - 1) / BX;
                    v.gy = (d.n[1] + 1 + by - 1) / by;
                    v.nblk = v.gx * v.gy * ((d.n[2] + 1 + zc - 1) / zc);
                    v.zb = 0;
                    v.ze = d.n[2] + 1;
                    gF.push_back(v);
                }
                for (int k = 0; k < (int)e->I.size(); k++) {
                    DevIface &f = e->I[k];
                    IfWork<T> w{};
                    w.a = e->iface_dev(k);
                    w.B = f.B; w.M = f.M; w.Rp = f.Rp; w.U = f.U; w.Bs = f.Bs; w.absp = f.absp;
                    ifF.push_back(w);
                    jlF.push_back(e->jac_lds(k, false));
                }
                paF.push_back(e->point_args());
            }
            for (Engine<T> *e : M) e->set_store(-1);
        }
        if (int rc = upload(&d_gF, gF)) return rc;
        if (ov) {
            // the overlapped step's plane sets of every slot ([slot][pn[q]])
            for (int q = 0; q < 4; q++) {
                std::vector<GridDev<T>> pg;
                std::vector<int> ps;
                for (int sl = 0; sl <= K; sl++) {
                    ps.clear();
                    plane_set(q, gF.data() + (size_t)sl * ng, pg, ps);
                }
                if (int rc = upload(&d_pgF[q], pg)) return rc;
            }
        }
        if (int rc = upload(&d_ifF, ifF)) return rc;
        if (int rc = upload(&d_jlF, jlF)) return rc;
        if (int rc = upload(&d_paF, paF)) return rc;
        // transposed steps: grid descriptors and their tile starts
        std::vector<AdjGridDev<T>> ag;
        std::vector<int> gt, gtg;
        gt_tot = gtg_tot = 0;
        adj32 = true;
        for (int t = 0; t < K; t++)
            for (Engine<T> *e : M)
                for (int g = 0; g < (int)e->G.size(); g++) {
                    const AdjGridDev<T> a = e->adj_dev(g, t + 1, t);
                    adj32 = adj32 && fits32(a);
                    ag.push_back(a);
                    if (t == 0) {
                        gt.push_back(gt_tot); gt_tot += a.gx * a.gy * a.gz;
                        gtg.push_back(gtg_tot); gtg_tot += adj_grad_tiles(a);
                    }
                }
        if (int rc = upload(&d_ag, ag)) return rc;
        if (int rc = upload(&d_gt, gt)) return rc;
        if (int rc = upload(&d_gtg, gtg)) return rc;
        // interfaces by level (interface index k), members in order
        nlev = 0;
        for (Engine<T> *e : M) nlev = std::max(nlev, (int)e->I.size());
        lev_off.assign(nlev, 0);
        lev_n.assign(nlev, 0);
        std::vector<std::pair<Engine<T> *, int>> order;
        for (int L = 0; L < nlev; L++) {
            lev_off[L] = (int)order.size();
            for (Engine<T> *e : M)
                if (L < (int)e->I.size()) order.push_back({e, L});
            lev_n[L] = (int)order.size() - lev_off[L];
        }
        naw = (int)order.size();
        std::vector<AdjIfWork<T>> aw;
        for (int t = 0; t < K; t++)
            for (auto &ek : order) {
                Engine<T> *e = ek.first;
                const int k = ek.second;
                AdjIfWork<T> w{};
                w.a = e->if_adj(k, t + 1, t + 1);
                w.ap = e->if_adj(k, t + 1, t);
                w.Ba = e->AI[k].Ba; w.X0a = e->AI[k].X0a; w.mu_vz = e->AI[k].mu_vz; w.mu_syz = e->AI[k].mu_syz;
                w.den = e->AI[k].den;
                w.M = e->I[k].M; w.Dp = e->I[k].Dp; w.U = e->I[k].U;
                aw.push_back(w);
            }
        if (int rc = upload(&d_aw, aw)) return rc;
        std::vector<int> aps;
        aps_off.assign((size_t)nlev * AB_NPHASE, 0);
        aps_tot.assign((size_t)nlev * AB_NPHASE, 0);
        for (int L = 0; L < nlev; L++)
            for (int ph = 0; ph < AB_NPHASE; ph++) {
                aps_off[L * AB_NPHASE + ph] = (int)aps.size();
                int acc = 0;
                for (int q = 0; q < lev_n[L]; q++) {
                    const auto &ek = order[lev_off[L] + q];
                    const DevIface &f = ek.first->I[ek.second];
                    const long nt = adj_if_threads(ph, f.r, f.nx2, f.ny2);
                    aps.push_back(acc);
                    acc += (int)(ph == AB_VZ_TRI ? nt : (nt + 255) / 256);
                }
                aps_tot[L * AB_NPHASE + ph] = acc;
            }
        if (int rc = upload(&d_aps, aps)) return rc;
        // transposed ghost mirrors and Jacobi descriptors
        std::vector<MirrorWork<T>> mw;
        std::vector<int> ms[6];
        for (int q = 0; q < 6; q++) ms_tot[q] = 0;
        for (int set = 0; set < 3; set++)
            for (Engine<T> *e : M)
                for (int g = 0; g < (int)e->G.size(); g++) {
                    const auto &md = set == 0 ? e->AG[g].sm : set == 1 ? e->AG[g].vA : e->AG[g].vB;
                    MirrorWork<T> w{};
                    w.lam = e->AG[g].lam; w.src = md.src; w.start = md.start; w.halo = md.halo; w.n = md.n; w.nh = md.nh;
                    mw.push_back(w);
                    ms[2 * set].push_back(ms_tot[2 * set]); ms_tot[2 * set] += (md.n + 255) / 256;
                    ms[2 * set + 1].push_back(ms_tot[2 * set + 1]); ms_tot[2 * set + 1] += (md.nh + 255) / 256;
                }
        if (int rc = upload(&d_mw, mw)) return rc;
        for (int q = 0; q < 6; q++)
            if (int rc = upload(&d_ms[q], ms[q])) return rc;
        std::vector<JacLds<T>> jla;
        for (Engine<T> *e : M)
            for (int k = 0; k < (int)e->I.size(); k++) jla.push_back(e->jac_lds(k, true));
        if (int rc = upload(&d_jla, jla)) return rc;
        return ECSG_OK;
    }

    // misfit injection descriptors (after the residuals are on the device)
    int build_rec_table()
    {
        std::vector<RecAdj<T>> ra;
        std::vector<int> st;
        ra_tot = 0;
        for (Engine<T> *e : M) {
            RecAdj<T> r{};
            for (int g = 0; g < (int)e->G.size(); g++) r.lamvx[g] = e->AG[g].lam;
            r.loc = e->d_rloc; r.start = e->d_rstart; r.rec = e->d_rrec; r.resid = e->d_resid;
            r.nloc = e->nrloc; r.nsteps = e->resid_steps;
            ra.push_back(r);
            st.push_back(ra_tot);
            ra_tot += (e->nrloc + 255) / 256;
        }
        hipFree(d_ra); d_ra = nullptr;
        hipFree(d_ras); d_ras = nullptr;
        if (int rc = upload(&d_ra, ra)) return rc;
        return upload(&d_ras, st);
    }

    template <int PH>
    void if_phase(int t, int L)
    {
        const int tot = aps_tot[L * AB_NPHASE + PH];
        if (!tot) return;
        hipLaunchKernelGGL((k_adj_if_batch<T, PH>), dim3(tot), dim3(256), 0, s,
                           (const AdjIfWork<T> *)(d_aw + (long)t * naw + lev_off[L]),
                           (const int *)(d_aps + aps_off[L * AB_NPHASE + PH]), lev_n[L]);
    }
    void mirror_phase(int set)
    {
        const int ngr = ng;  // grids of all members
        const MirrorWork<T> *w = d_mw + (long)set * ngr;
        if (ms_tot[2 * set])
            hipLaunchKernelGGL((k_mirror_batch<T, false>), dim3(ms_tot[2 * set]), dim3(256), 0, s, w, (const int *)d_ms[2 * set], ngr);
        if (ms_tot[2 * set + 1])
            hipLaunchKernelGGL((k_mirror_batch<T, true>), dim3(ms_tot[2 * set + 1]), dim3(256), 0, s, w, (const int *)d_ms[2 * set + 1],
                               ngr);
    }
    template <int PH>
    void grid_phase(int t)
    {
        const bool grad = PH == AG_GRAD || PH == AG_GRADB;
        const int tot = grad ? gtg_tot : gt_tot;
        if (adj32)
            hipLaunchKernelGGL((k_adj_grid_batch<T, PH, unsigned>), dim3(tot), dim3(64, grad ? GBY : AWY), 0, s,
                               (const AdjGridDev<T> *)(d_ag + (long)t * ng), (const int *)(grad ? d_gtg : d_gt), ng, tot);
        else
            hipLaunchKernelGGL((k_adj_grid_batch<T, PH, long>), dim3(tot), dim3(64, grad ? GBY : AWY), 0, s,
                               (const AdjGridDev<T> *)(d_ag + (long)t * ng), (const int *)(grad ? d_gtg : d_gt), ng, tot);
    }

    // one transposed step of every member (adj_step_impl's phases and order):
    // forward step n, ring slots t+1 (after) and t (before)
    int adj_step(int n, int t, bool inject)
    {
        for (int L = nlev - 1; L >= 0; L--) {
            if_phase<AB_SYZ_GHOST>(t, L);
            if_phase<AB_SYZ_FINE0>(t, L);
            if_phase<AB_SYZ_FINE>(t, L);
            if_phase<AB_SYZ_COARSE>(t, L);
        }
        if (nlev) {
            for (int L = nlev - 1; L >= 0; L--) {
                if_phase<AB_SXZ_GHOST>(t, L);
                if_phase<AB_SXZ_FINE0>(t, L);
            }
            JacParams jp{};
            jp.tol_cap = 0; jp.tol_fac = 0; jp.rel_tol = M[0]->adj_rel_tol; jp.solve_only = 1;
            jp.floor = 1024 * 2.220446049250313e-16;
            jp.cap = JACOBI_CAP;
            hipLaunchKernelGGL((k_jacobi_lds_batch<T>), dim3(naw), dim3(JL), (size_t)(2 * maxw * sizeof(double)), s,
                               (const JacLds<T> *)d_jla, jp);
            for (int L = nlev - 1; L >= 0; L--) {
                if_phase<AB_SXZ_FINE>(t, L);
                if_phase<AB_SXZ_COARSE>(t, L);
            }
        }
        mirror_phase(0);
        grid_phase<AG_STRESS>(t);
        grid_phase<AG_GRAD>(t);
        for (int L = nlev - 1; L >= 0; L--) {
            if_phase<AB_VZ_GHOST>(t, L);
            if_phase<AB_VZ_FINE0>(t, L);
            const int tot = aps_tot[L * AB_NPHASE + AB_VZ_TRI];
            if (tot)
                hipLaunchKernelGGL((k_adj_tri_batch<T>), dim3(tot), dim3(64), (size_t)(32 * maxn_vz), s,
                                   (const AdjIfWork<T> *)(d_aw + (long)t * naw + lev_off[L]),
                                   (const int *)(d_aps + aps_off[L * AB_NPHASE + AB_VZ_TRI]), lev_n[L]);
            if_phase<AB_VZ_FINE>(t, L);
            if_phase<AB_VZ_COARSE>(t, L);
        }
        if (inject && ra_tot)
            hipLaunchKernelGGL((k_adj_rec_batch<T>), dim3(ra_tot), dim3(256), 0, s, (const RecAdj<T> *)d_ra, (const int *)d_ras,
                               (int)M.size(), n);
        mirror_phase(1);
        mirror_phase(2);
        grid_phase<AG_VEL>(t);
        grid_phase<AG_GRADB>(t);
        HIPCHK(hipGetLastError());
        return ECSG_OK;
    }

    int fwd_step(int slot)
    {
        const long nif = ni, nm = (long)M.size();
        return enqueue_tables(d_gF + (long)slot * ng, d_ifF + (long)slot * nif, d_jlF + (long)slot * nif, d_paF + (long)slot * nm);
    }
    // n forward steps, step i into ring slot `ring ? slot0 + i : slot0`;
    // with the interface overlap the steps are enqueued as one overlapped
    // sequence (each step's sxz/syz ghosts beside the next one's velocity
    // interiors)
    int fwd_steps(int slot0, int n, bool ring)
    {
        if (!ov) {
            for (int i = 0; i < n; i++)
                if (int rc = fwd_step(ring ? slot0 + i : slot0)) return rc;
            return ECSG_OK;
        }
        const long nif = ni, nm = (long)M.size();
        return enqueue_ov_steps(n, [&](int i) {
            const long sl = ring ? slot0 + i : slot0;
            OvTables tb;
            for (int q = 0; q < 4; q++) tb.pg[q] = d_pgF[q] + sl * pn[q];
            tb.w = d_ifF + sl * nif;
            tb.jl = d_jlF + sl * nif;
            tb.pa = d_paF + sl * nm;
            return tb;
        });
    }

    int gradient(int N, const double *const *obs, double *misfit, const int *group) override
    {
        if (N < 1) return fail(ECSG_EINVAL, "gradient: step count out of range");
        const int nm = (int)M.size();
        for (int m = 0; m < nm; m++) {
            Engine<T> *e = M[m];
            const std::string id = "member " + std::to_string(m) + ": ";
            if (N > e->cfg.max_steps) return fail(ECSG_EINVAL, id + "gradient: step count above max_steps");
            if (e->nrec > 0 && (!obs || !obs[m])) return fail(ECSG_EINVAL, id + "gradient: null observed traces");
            if (e->ck_every != M[0]->ck_every || e->ck_fmt != M[0]->ck_fmt)
                return fail(ECSG_EINVAL, "gradient: members must share one checkpoint interval and format");
            if (e->adj_rel_tol != M[0]->adj_rel_tol) return fail(ECSG_EINVAL, "gradient: members must share one adjoint tolerance");
            HIPCHK(hipStreamSynchronize(e->s));
            if (int rc = e->adj_build()) return fail(rc, id + e->st.msg);
            if (int rc = e->build_adj_receivers()) return fail(rc, id + e->st.msg);
        }
        if (group)
            for (int m = 0; m < nm; m++)
                for (int q = 0; q < m; q++)
                    if (group[q] == group[m]) {
                        if (M[q]->G.size() != M[m]->G.size()) return fail(ECSG_EINVAL, "gradient: a group's members differ in shape");
                        for (size_t g = 0; g < M[m]->G.size(); g++)
                            for (int d = 0; d < 3; d++)
                                if (M[q]->G[g].n[d] != M[m]->G[g].n[d])
                                    return fail(ECSG_EINVAL, "gradient: a group's members differ in shape");
                        break;
                    }
        HIPCHK(hipStreamSynchronize(s));
        for (size_t m = 0; m < M.size(); m++)
            if (M[m]->G[0].cur != built_cur[m])
                if (int rc = build_tables()) return rc;
        int K = M[0]->ck_every;
        if (K == 0)
            if (int rc = batch_interval(N, &K)) return rc;
        K = std::min(K, N);
        K_used = K;
        const int nseg = (N + K - 1) / K;
        const int r0 = N - (nseg - 1) * K;
        auto seg0 = [&](int j) { return j == 0 ? 0 : r0 + (j - 1) * K; };
        auto seglen = [&](int j) { return j == 0 ? r0 : K; };
        MembersOnBatch on(this);
        for (int m = 0; m < nm; m++) {
            Engine<T> *e = M[m];
            if (int rc = e->ck_prepare(std::max(nseg - 1, 1), K)) return fail(rc, "member " + std::to_string(m) + ": " + e->st.msg);
            e->last_K = K;
            if (int rc = e->reset()) return fail(rc, "member " + std::to_string(m) + ": " + e->st.msg);
        }
        if (int rc = build_grad_tables(K)) return rc;
        // forward pass with checkpoints; the last segment fills the rings
        for (int j = 0; j < nseg; j++) {
            const int len = seglen(j);
            if (j < nseg - 1) {
                for (Engine<T> *e : M)
                    if (int rc = e->ck_save(j)) return fail(rc, e->st.msg);
                if (int rc = fwd_steps(0, len, false)) return rc;
            } else {
                for (Engine<T> *e : M)
                    if (int rc = e->ring_save(0)) return fail(rc, e->st.msg);
                if (int rc = fwd_steps(1, len, true)) return rc;
            }
        }
        HIPCHK(hipStreamSynchronize(s));
        // residuals and misfits, member by member
        for (int m = 0; m < nm; m++) {
            Engine<T> *e = M[m];
            const int nrec = e->nrec, ms = e->cfg.max_steps;
            std::vector<T> tr((size_t)std::max(nrec, 1) * ms);
            if (nrec) HIPCHK(hipMemcpy(tr.data(), e->d_traces, sizeof(T) * nrec * ms, hipMemcpyDeviceToHost));
            std::vector<T> res((size_t)std::max(nrec, 1) * N, (T)0);
            double J = 0;
            for (int q = 0; q < nrec; q++)
                for (int t = 0; t < N; t++) {
                    const double rr = (double)tr[(size_t)q * ms + t] - obs[m][(size_t)q * N + t];
                    res[(size_t)q * N + t] = (T)rr;
                    J += 0.5 * rr * rr;
                }
            if (misfit) misfit[m] = J;
            Engine<T>::dfree(e->d_resid);
            if (int rc = e->upload(&e->d_resid, res)) return fail(rc, e->st.msg);
            e->resid_steps = N;
            if (int rc = e->adj_zero()) return fail(rc, e->st.msg);
        }
        if (int rc = build_rec_table()) return rc;
        // backward over the segments
        for (int j = nseg - 1; j >= 0; j--) {
            const int len = seglen(j);
            if (j < nseg - 1) {
                for (Engine<T> *e : M) {
                    if (int rc = e->ck_restore(j, seg0(j))) return fail(rc, e->st.msg);
                    if (int rc = e->ring_save(0)) return fail(rc, e->st.msg);
                }
                if (int rc = fwd_steps(1, len, true)) return rc;
            }
            for (int t = len - 1; t >= 0; t--)
                if (int rc = adj_step(seg0(j) + t, t, true)) return rc;
        }
        // a model's gradient: its members' accumulators added, in member
        // order, into the first member of its group
        if (group)
            for (int m = 0; m < nm; m++) {
                int lead = m;
                for (int q = 0; q < m; q++)
                    if (group[q] == group[m]) { lead = q; break; }
                if (lead == m) continue;
                for (size_t g = 0; g < M[m]->G.size(); g++) {
                    const long n12 = 12 * M[m]->G[g].astride;
                    hipLaunchKernelGGL((k_add_into<T>), dim3((unsigned)std::min<long>((n12 + 255) / 256, 4096)), dim3(256), 0, s,
                                       M[lead]->AG[g].grad, (const T *)M[m]->AG[g].grad, n12);
                }
            }
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(s));
        for (int m = 0; m < nm; m++) {
            Engine<T> *e = M[m];
            for (int k = 0; k < (int)e->I.size(); k++) {
                long long st3[3];
                HIPCHK(hipMemcpy(st3, e->AI[k].stats, sizeof(st3), hipMemcpyDeviceToHost));
                if (st3[2]) return fail(ECSG_ENOCONV, "member " + std::to_string(m) + ": adjoint Jacobi iteration cap reached");
            }
            e->host_step = nseg > 1 ? r0 : N;
            e->gprim.clear();
            e->grad_ready = true;
        }
        return ECSG_OK;
    }
};

}  // namespace

struct ecsg_batch {
    std::unique_ptr<BatchBase> b;
    std::string err;
};

struct ecsg_ctx {
    std::unique_ptr<EngineBase> e;
    std::string err;
};

static int ret(ecsg_ctx *c, int rc)
{
    if (c && rc != ECSG_OK) c->err = c->e->st.msg;
    return rc;
}

// Copy of known size for scaling the HBM byte counters; H = 3 uses the
// stencils' non-temporal loads and stores, H = 0 plain ones.  With
// InPlace (src == dst) every value is read and written back by one lane, the
// access pattern of the stencils' updated fields.
template <int H, bool InPlace>
__global__ __launch_bounds__(256) void k_copy_probe(float *dst, const float *src, uint64_t n)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        if constexpr (H == 3) __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
        else dst[i] = src[i];
    }
}

// A stencil sweep's memory traffic without its arithmetic: NR FP32 arrays
// read and the first NW written back in place (a[k] + a[k+NW] + a[k+2NW] +
// ... over all NR arrays, so that every load is used), one node per lane, with the sweeps' cache hints, over the nodes
// of a radius-1 grid in the device layout -- in the stencil tiling (64 x BYv
// rows x zc planes, workgroups in XCD order, z marched per lane), or as one
// flat pass over the arrays' planes 0..nz (halo rows and padding included).
// NR = 18, NW = 6: the stress sweep (9 stiffnesses, 3 velocities and the 6
// stresses it updates); NR = 10, NW = 3: the velocity sweep (6 stresses, the
// buoyancy and the 3 velocities it updates).  Its time is the streaming bound of
// the traversal on this GPU.
struct StreamProbe {
    float *a[18];
    int nx, ny, nz, gx, gy, zc, nblk;
    long P, S2, o;
};

template <int NR, int NW>
__device__ __forceinline__ void stream_node(const StreamProbe &g, long p)
{
    float v[NR];
#pragma unroll
    for (int k = 0; k < NR; k++) v[k] = ldc(g.a[k], p);
#pragma unroll
    for (int k = 0; k < NW; k++) {
        float u = v[k];
#pragma unroll
        for (int j = NW + k; j < NR; j += NW) u = u + v[j];
        stc(g.a[k], p, u);
    }
}

template <int BYv, int NR, int NW>
__global__ __launch_bounds__(64 * BYv) void k_stream_tiles(StreamProbe g)
{
    const int tile = sweep_tile(blockIdx.x, g.nblk, g.gx);
    const int bx = tile % g.gx, rest = tile / g.gx, by = rest % g.gy, bz = rest / g.gy;
    const int ix = bx * 64 + threadIdx.x, iy = by * BYv + threadIdx.y;
    if (ix > g.nx || iy > g.ny) return;
    const int z0 = bz * g.zc, z1 = min(z0 + g.zc, g.nz + 1);
    long p = g.o + (long)z0 * g.S2 + (long)iy * g.P + ix;
    for (int iz = z0; iz < z1; ++iz, p += g.S2) stream_node<NR, NW>(g, p);
}

template <int NR, int NW>
__global__ __launch_bounds__(256) void k_stream_flat(StreamProbe g)
{
    const long n = (long)(g.nz + 1) * g.S2;
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    stream_node<NR, NW>(g, g.S2 + i);  // planes 0 .. nz
}

template <int NR, int NW>
static void launch_stream(const StreamProbe &g, int by, long cnt, hipStream_t st)
{
    if (g.zc == 0) k_stream_flat<NR, NW><<<(unsigned)((cnt + 255) / 256), 256, 0, st>>>(g);
    else if (by == 1) k_stream_tiles<1, NR, NW><<<g.nblk, dim3(64, 1), 0, st>>>(g);
    else if (by == 2) k_stream_tiles<2, NR, NW><<<g.nblk, dim3(64, 2), 0, st>>>(g);
    else if (by == 4) k_stream_tiles<4, NR, NW><<<g.nblk, dim3(64, 4), 0, st>>>(g);
    else k_stream_tiles<8, NR, NW><<<g.nblk, dim3(64, 8), 0, st>>>(g);
}

extern "C" {

const char *ecsg_version(void) { return "ecsg_hip 0.1 (gfx950)"; }

int ecsg_create(const ecsg_config *cfg, ecsg_ctx **out)
{
    if (!cfg || !out) return ECSG_EINVAL;
    *out = nullptr;
    ecsg_ctx *c = new ecsg_ctx();
    if (cfg->precision == ECSG_F64) c->e.reset(new Engine<double>());
    else if (cfg->precision == ECSG_F32) c->e.reset(new Engine<float>());
    else { delete c; return ECSG_EINVAL; }
    int rc = c->e->init(*cfg);
    if (rc != ECSG_OK) c->err = c->e->st.msg;
    *out = c;  // returned even on failure so ecsg_last_error can report
    return rc;
}

void ecsg_destroy(ecsg_ctx *ctx) { delete ctx; }
const char *ecsg_last_error(const ecsg_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

#define CTX_CHECK(ctx) \
    if (!(ctx) || !(ctx)->e) return ECSG_EINVAL

int ecsg_set_params(ecsg_ctx *ctx, int g, const double *buoy, const double *const stiff[9], int iso)
{
    CTX_CHECK(ctx);
    return ret(ctx, ctx->e->set_params(g, buoy, stiff, iso));
}
int ecsg_set_params_staged(ecsg_ctx *ctx, int g, const double *buoy, const double *const stiff[9])
{
    CTX_CHECK(ctx);
    return ret(ctx, ctx->e->set_params_staged(g, buoy, stiff));
}
int ecsg_set_source(ecsg_ctx *ctx, int g, int ix, int iy, int iz, int type, const float *trace, int len)
{
    CTX_CHECK(ctx);
    return ret(ctx, ctx->e->set_source(g, ix, iy, iz, type, trace, len));
}
int ecsg_set_receivers(ecsg_ctx *ctx, int n, const int *loc)
{
    CTX_CHECK(ctx);
    return ret(ctx, ctx->e->set_receivers(n, loc));
}
int ecsg_update_vel(ecsg_ctx *ctx, int g) { CTX_CHECK(ctx); return ret(ctx, ctx->e->update_vel(g)); }
int ecsg_record_receivers(ecsg_ctx *ctx) { CTX_CHECK(ctx); return ret(ctx, ctx->e->record_receivers()); }
int ecsg_add_source(ecsg_ctx *ctx) { CTX_CHECK(ctx); return ret(ctx, ctx->e->add_source()); }
int ecsg_update_ghost_vz(ecsg_ctx *ctx, int k) { CTX_CHECK(ctx); return ret(ctx, ctx->e->ghost_vz(k)); }
int ecsg_compute_energy(ecsg_ctx *ctx, double *e) { CTX_CHECK(ctx); return ret(ctx, ctx->e->energy(e)); }
int ecsg_update_stress(ecsg_ctx *ctx, int g) { CTX_CHECK(ctx); return ret(ctx, ctx->e->update_stress(g)); }
int ecsg_update_ghost_sxz(ecsg_ctx *ctx, int k) { CTX_CHECK(ctx); return ret(ctx, ctx->e->ghost_sxz(k)); }
int ecsg_update_ghost_syz(ecsg_ctx *ctx, int k) { CTX_CHECK(ctx); return ret(ctx, ctx->e->ghost_syz(k)); }
int ecsg_end_step(ecsg_ctx *ctx) { CTX_CHECK(ctx); return ret(ctx, ctx->e->end_step()); }
int ecsg_compute_res(ecsg_ctx *ctx, int which, double *out) { CTX_CHECK(ctx); return ret(ctx, ctx->e->residual(which, out)); }
int ecsg_field_avgabs(ecsg_ctx *ctx, int g, int sel, double *out) { CTX_CHECK(ctx); return ret(ctx, ctx->e->avg_abs(g, sel, out)); }
int ecsg_sbp_residuals(ecsg_ctx *ctx, int g, double out[7]) { CTX_CHECK(ctx); return ret(ctx, ctx->e->sbp(g, out)); }
int ecsg_copy_probe(float *dst, const float *src, uint64_t n, int hints, void *stream)
{
    if (!dst || !src || (hints != 0 && hints != 3)) return ECSG_EINVAL;
    if (n == 0) return ECSG_OK;
    const unsigned blocks = (unsigned)std::min<uint64_t>((n + 255) / 256, 256 * 64);
    hipStream_t st = (hipStream_t)stream;
    if (src == dst) {
        if (hints == 3) k_copy_probe<3, true><<<blocks, 256, 0, st>>>(dst, src, n);
        else k_copy_probe<0, true><<<blocks, 256, 0, st>>>(dst, src, n);
    } else {
        if (hints == 3) k_copy_probe<3, false><<<blocks, 256, 0, st>>>(dst, src, n);
        else k_copy_probe<0, false><<<blocks, 256, 0, st>>>(dst, src, n);
    }
    return hipGetLastError() == hipSuccess ? ECSG_OK : ECSG_EHIP;
}
int ecsg_cfl_dt(const ecsg_config *cfg, const double *const buoy[], const double *const stiff[][9], int iso,
                double frac, double *dt_out)
{
    if (!cfg || !buoy || !stiff || !dt_out || !(frac > 0) || cfg->num_grids < 1 || cfg->num_grids > ECSG_MAX_GRIDS)
        return ECSG_EINVAL;
    double best = INFINITY;
    for (int g = 0; g < cfg->num_grids; g++) {
        const int *n = cfg->n[g];
        const double *h = cfg->h[g];
        if (n[0] < 1 || n[1] < 1 || n[2] < 1 || !(h[0] > 0) || !(h[1] > 0) || !(h[2] > 0)) return ECSG_EINVAL;
        const double *src[4] = {buoy[g], stiff[g][0], iso ? nullptr : stiff[g][1], iso ? nullptr : stiff[g][2]};
        const int na = iso ? 2 : 4;
        for (int a = 0; a < na; a++)
            if (!src[a]) return ECSG_EINVAL;
        const long np = (long)(n[0] + 1) * (n[1] + 1) * (n[2] + 1);
        for (int a = 0; a < na; a++)
            for (long q = 0; q < np; q++)
                if (!(src[a][q] > 0) || !std::isfinite(src[a][q])) return ECSG_EINVAL;
        const int nblk = (int)std::min<long>((np + 255) / 256, 2048);
        double *d = nullptr;
        if (hipMalloc((void **)&d, sizeof(double) * ((size_t)na * np + nblk + 1)) != hipSuccess) return ECSG_ENOMEM;
        double *dv[4] = {nullptr, nullptr, nullptr, nullptr};
        int rc = ECSG_OK;
        for (int a = 0; a < na && rc == ECSG_OK; a++) {
            dv[a] = d + (size_t)a * np;
            if (hipMemcpy(dv[a], src[a], sizeof(double) * np, hipMemcpyHostToDevice) != hipSuccess) rc = ECSG_EHIP;
        }
        double *part = d + (size_t)na * np, *res = part + nblk;
        double m = 0.0;
        if (rc == ECSG_OK) {
            hipLaunchKernelGGL(k_cfl_part, dim3(nblk), dim3(256), 0, 0, dv[0], dv[1], dv[2], dv[3], np, part);
            hipLaunchKernelGGL(k_cfl_final, dim3(1), dim3(64), 0, 0, part, nblk, res);
            if (hipGetLastError() != hipSuccess || hipMemcpy(&m, res, sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
                rc = ECSG_EHIP;
        }
        hipFree(d);
        if (rc != ECSG_OK) return rc;
        // h_min / (vp sqrt 3), vp = sqrt(max c b)
        const double vp = std::sqrt(m);
        const double hmin = std::min(std::min(h[0], h[1]), h[2]);
        best = std::min(best, hmin / (vp * std::sqrt(3.0)));
    }
    *dt_out = frac * best;
    return ECSG_OK;
}
// Scratch of ecsg_cfl_dt_noalloc: per grid, CFL_NB block maxima and the
// grid's maximum (a stated, fixed size: 8 grids x 1025 doubles = 65.6 KB,
// held by every GPU the library runs on until the process ends).  Every
// block writes its entry before the final pass reads it, so nothing needs
// clearing between calls; calls that overlap on one GPU race.  In host code
// the name is not a device address: kernels get cfl_scratch_ptr[device].
constexpr int CFL_NB = 1024;
static_assert(ECSG_MAX_GRIDS <= 8, "CflDtArgs holds 8 grid steps");
static __device__ __align__(256) double scratch[ECSG_MAX_GRIDS * (CFL_NB + 1)];
static std::mutex scratch_mu;
static double *cfl_scratch_ptr[64];

int ecsg_cfl_dt_noalloc(const ecsg_config *cfg, const double *const buoy[], const double *const stiff[][9], int iso,
                        double frac, double *dt_dev, void *stream)
{
    if (!cfg || !buoy || !stiff || !dt_dev || !(frac > 0)) return ECSG_EINVAL;
    if (cfg->num_grids < 1 || cfg->num_grids > ECSG_MAX_GRIDS) return ECSG_EINVAL;  // beyond the scratch
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return ECSG_EHIP;
    double *sp = nullptr;
    {
        std::lock_guard<std::mutex> lock(scratch_mu);
        if (!cfl_scratch_ptr[dev]) {
            void *p = nullptr;
            if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return ECSG_EHIP;
            cfl_scratch_ptr[dev] = (double *)p;
        }
        sp = cfl_scratch_ptr[dev];
    }
    hipStream_t st = (hipStream_t)stream;
    CflDtArgs a{};
    a.ng = cfg->num_grids;
    a.stride = CFL_NB + 1;
    a.frac = frac;
    for (int g = 0; g < cfg->num_grids; g++) {
        const int *n = cfg->n[g];
        const double *h = cfg->h[g];
        if (n[0] < 1 || n[1] < 1 || n[2] < 1 || !(h[0] > 0) || !(h[1] > 0) || !(h[2] > 0)) return ECSG_EINVAL;
        if (!buoy[g] || !stiff[g][0] || (!iso && (!stiff[g][1] || !stiff[g][2]))) return ECSG_EINVAL;
        a.hmin[g] = std::min(std::min(h[0], h[1]), h[2]);
        const long np = (long)(n[0] + 1) * (n[1] + 1) * (n[2] + 1);
        const int nblk = (int)std::min<long>((np + 255) / 256, CFL_NB);
        double *part = sp + (long)g * (CFL_NB + 1);
        hipLaunchKernelGGL(k_cfl_part, dim3(nblk), dim3(256), 0, st, buoy[g], stiff[g][0], iso ? nullptr : stiff[g][1],
                           iso ? nullptr : stiff[g][2], np, part);
        hipLaunchKernelGGL(k_cfl_final, dim3(1), dim3(64), 0, st, part, nblk, part + CFL_NB);
    }
    hipLaunchKernelGGL(k_cfl_dt, dim3(1), dim3(64), 0, st, sp + CFL_NB, a, dt_dev);
    return hipGetLastError() == hipSuccess ? ECSG_OK : ECSG_EHIP;
}
uint64_t ecsg_stream_probe_floats(int nx, int ny, int nz)
{
    if (nx < 1 || ny < 1 || nz < 1) return 0;
    const int n[3] = {nx, ny, nz};
    return 18 * (uint64_t)grid_layout(n, 1, (long)sizeof(float)).astride;
}
int ecsg_stream_probe(float *base, int nx, int ny, int nz, int by, int zc, int pattern, void *stream)
{
    if (!base || nx < 1 || ny < 1 || nz < 1 || zc < 0 || (pattern != 0 && pattern != 1)) return ECSG_EINVAL;
    if (by != 1 && by != 2 && by != 4 && by != 8) return ECSG_EINVAL;
    const int n[3] = {nx, ny, nz};
    const Layout lay = grid_layout(n, 1, (long)sizeof(float));
    StreamProbe g;
    for (int k = 0; k < 18; k++) g.a[k] = base + k * lay.astride;
    g.nx = nx; g.ny = ny; g.nz = nz;
    g.P = lay.P; g.S2 = lay.S2; g.o = lay.o;
    hipStream_t st = (hipStream_t)stream;
    g.gx = g.gy = g.zc = g.nblk = 0;
    if (zc > 0) {
        g.zc = zc;
        g.gx = (nx + 1 + 63) / 64;
        g.gy = (ny + 1 + by - 1) / by;
        const long nb = (long)g.gx * g.gy * ((nz + 1 + zc - 1) / zc);
        if (nb > (1L << 30)) return ECSG_EINVAL;
        g.nblk = (int)nb;
    }
    const long cnt = (long)(nz + 1) * lay.S2;
    if (pattern == 0) launch_stream<18, 6>(g, by, cnt, st);
    else launch_stream<10, 3>(g, by, cnt, st);
    return hipGetLastError() == hipSuccess ? ECSG_OK : ECSG_EHIP;
}
int ecsg_step_n(ecsg_ctx *ctx, int nsteps, int use_graph) { CTX_CHECK(ctx); return ret(ctx, ctx->e->step_n(nsteps, use_graph)); }
int ecsg_set_overlap(ecsg_ctx *ctx, int on) { CTX_CHECK(ctx); return ret(ctx, ctx->e->set_overlap(on)); }
int ecsg_step_n_timed(ecsg_ctx *ctx, int nsteps, double ms_out[2])
{
    CTX_CHECK(ctx);
    if (!ms_out) return ECSG_EINVAL;
    return ret(ctx, ctx->e->step_timed(nsteps, ms_out));
}
int ecsg_overlap_active(const ecsg_ctx *ctx) { return ctx && ctx->e ? ctx->e->overlap_active() : 0; }
int ecsg_set_barrier_timeout(ecsg_ctx *ctx, int ms) { CTX_CHECK(ctx); return ret(ctx, ctx->e->set_barrier_timeout(ms)); }
int ecsg_set_jacobi_tol(ecsg_ctx *ctx, double tol_cap, double tol_fac)
{
    CTX_CHECK(ctx);
    return ret(ctx, ctx->e->set_jacobi_tol(tol_cap, tol_fac));
}
int ecsg_steps_done(const ecsg_ctx *ctx) { return (ctx && ctx->e) ? ctx->e->steps_done() : -1; }
int ecsg_synchronize(ecsg_ctx *ctx) { CTX_CHECK(ctx); return ret(ctx, ctx->e->sync()); }
int ecsg_field_dims(const ecsg_ctx *ctx, int g, int sel, int len[3])
{
    if (!ctx || !ctx->e) return ECSG_EINVAL;
    return ctx->e->dims(g, sel, len) < 0 ? ECSG_EINVAL : ECSG_OK;
}
int ecsg_get_field(ecsg_ctx *ctx, int g, int sel, double *out) { CTX_CHECK(ctx); return ret(ctx, ctx->e->get_field(g, sel, out)); }
int ecsg_set_field(ecsg_ctx *ctx, int g, int sel, const double *in) { CTX_CHECK(ctx); return ret(ctx, ctx->e->set_field(g, sel, in)); }
int ecsg_get_traces(ecsg_ctx *ctx, double *out) { CTX_CHECK(ctx); return ret(ctx, ctx->e->get_traces(out)); }
int ecsg_get_energy(ecsg_ctx *ctx, double *out) { CTX_CHECK(ctx); return ret(ctx, ctx->e->get_energy(out)); }
int ecsg_jacobi_stats(ecsg_ctx *ctx, int k, long long *total, int *last, int *capped)
{
    CTX_CHECK(ctx);
    return ret(ctx, ctx->e->jacobi_stats(k, total, last, capped));
}
int ecsg_profile_kernels(ecsg_ctx *ctx, int nsteps, double ms_out[6], int count_out[6])
{
    CTX_CHECK(ctx);
    return ret(ctx, ctx->e->profile(nsteps, ms_out, count_out));
}

int ecsg_reset(ecsg_ctx *ctx) { CTX_CHECK(ctx); return ret(ctx, ctx->e->reset()); }
int ecsg_set_checkpointing(ecsg_ctx *ctx, int every, int format)
{
    CTX_CHECK(ctx);
    return ret(ctx, ctx->e->set_checkpointing(every, format));
}
int ecsg_checkpoint_interval(const ecsg_ctx *ctx) { return (ctx && ctx->e) ? ctx->e->checkpoint_interval() : 0; }
int ecsg_set_ring_steps(ecsg_ctx *ctx, int steps) { CTX_CHECK(ctx); return ret(ctx, ctx->e->set_ring_steps(steps)); }
int ecsg_checkpoint_levels(const ecsg_ctx *ctx, int *interval, int *ring, long long *recomputed)
{
    if (!ctx || !ctx->e) return ECSG_EINVAL;
    ctx->e->checkpoint_levels(interval, ring, recomputed);
    return ECSG_OK;
}
int ecsg_gradient(ecsg_ctx *ctx, int nsteps, const double *observed, double *misfit)
{
    CTX_CHECK(ctx);
    return ret(ctx, ctx->e->gradient(nsteps, observed, misfit));
}
int ecsg_get_gradient(ecsg_ctx *ctx, int g, int sel, double *out)
{
    CTX_CHECK(ctx);
    if (!out) return ECSG_EINVAL;
    return ret(ctx, ctx->e->get_gradient(g, sel, out));
}
int ecsg_adjoint_step(ecsg_ctx *ctx) { CTX_CHECK(ctx); return ret(ctx, ctx->e->adjoint_step()); }
int ecsg_get_adjoint_field(ecsg_ctx *ctx, int g, int sel, double *out)
{
    CTX_CHECK(ctx);
    if (!out) return ECSG_EINVAL;
    return ret(ctx, ctx->e->adj_field(g, sel, out, nullptr));
}
int ecsg_set_adjoint_field(ecsg_ctx *ctx, int g, int sel, const double *in)
{
    CTX_CHECK(ctx);
    if (!in) return ECSG_EINVAL;
    return ret(ctx, ctx->e->adj_field(g, sel, nullptr, in));
}
int ecsg_adjoint_stats(ecsg_ctx *ctx, long long *sweeps)
{
    CTX_CHECK(ctx);
    return ret(ctx, ctx->e->adjoint_stats(sweeps, nullptr));
}

int ecsg_set_fused(ecsg_ctx *ctx, int on) { CTX_CHECK(ctx); return ret(ctx, ctx->e->set_fused(on)); }

int ecsg_autotune(ecsg_ctx *ctx, int reps, int *by_out, int *zc_out)
{
    CTX_CHECK(ctx);
    return ret(ctx, ctx->e->autotune(reps, by_out, zc_out));
}

static int bret(ecsg_batch *b, int rc)
{
    if (b && rc != ECSG_OK) b->err = b->b->st.msg;
    return rc;
}

int ecsg_batch_create(ecsg_ctx *const *members, int n, ecsg_batch **out)
{
    if (!members || n < 1 || !out) return ECSG_EINVAL;
    *out = nullptr;
    ecsg_batch *b = new ecsg_batch();
    int prec = -1;
    for (int m = 0; m < n; m++) {
        if (!members[m] || !members[m]->e) { delete b; return ECSG_EINVAL; }
        const int p = dynamic_cast<Engine<double> *>(members[m]->e.get()) ? ECSG_F64 : ECSG_F32;
        if (prec >= 0 && p != prec) {
            b->err = "members must share one precision";
            *out = b;
            return ECSG_EINVAL;
        }
        prec = p;
    }
    int rc;
    if (prec == ECSG_F64) {
        auto *x = new Batch<double>();
        b->b.reset(x);
        std::vector<Engine<double> *> v;
        for (int m = 0; m < n; m++) v.push_back(static_cast<Engine<double> *>(members[m]->e.get()));
        rc = x->init(v);
    } else {
        auto *x = new Batch<float>();
        b->b.reset(x);
        std::vector<Engine<float> *> v;
        for (int m = 0; m < n; m++) v.push_back(static_cast<Engine<float> *>(members[m]->e.get()));
        rc = x->init(v);
    }
    if (rc != ECSG_OK) b->err = b->b->st.msg;
    *out = b;
    return rc;
}
void ecsg_batch_destroy(ecsg_batch *b) { delete b; }
const char *ecsg_batch_last_error(const ecsg_batch *b) { return b ? b->err.c_str() : "null batch"; }
#define BATCH_CHECK(b) \
    if (!(b) || !(b)->b) return ECSG_EINVAL
int ecsg_batch_step_n(ecsg_batch *b, int nsteps, int use_graph) { BATCH_CHECK(b); return bret(b, b->b->step_n(nsteps, use_graph)); }
int ecsg_batch_synchronize(ecsg_batch *b) { BATCH_CHECK(b); return bret(b, b->b->sync()); }
int ecsg_batch_set_tiling(ecsg_batch *b, int by, int zc) { BATCH_CHECK(b); return bret(b, b->b->set_tiling(by, zc)); }
int ecsg_batch_autotune(ecsg_batch *b, int reps, int *by, int *zc) { BATCH_CHECK(b); return bret(b, b->b->autotune(reps, by, zc)); }
int ecsg_batch_gradient(ecsg_batch *b, int nsteps, const double *const *observed, double *misfit, const int *group)
{
    BATCH_CHECK(b);
    return bret(b, b->b->gradient(nsteps, observed, misfit, group));
}
int ecsg_batch_checkpoint_interval(const ecsg_batch *b) { return b && b->b ? b->b->last_interval() : 0; }

}  // extern "C"
