/*
 * ecsg_hip.h -- C ABI of the MI355X (gfx950) ECSG time-stepping path.
 *
 * The reference (llodds/ecsg) is a C++/OpenMP code whose time loop in
 * main.cpp:165-261 calls free functions over caller-owned Field/Array objects
 * (sim_utils.hpp:128-161).  This library owns device copies of those objects
 * behind an opaque handle and exposes one entry point per line of that loop,
 * plus a fused, graph-captured multi-step entry point.  Every function returns
 * an int status (ECSG_OK = 0, negative on error); ecsg_last_error() gives the
 * message.  No exceptions and no exit() cross this boundary.
 *
 * Conventions
 *   - Arrays are x-fastest, like the reference's Field (field.hpp:84-85).
 *   - Host pointers are borrowed for the duration of a call only.
 *   - One HIP stream per handle; a handle is used from one host thread at a
 *     time.  Distinct handles may run concurrently.
 *   - Precision: ECSG_F64 stores wavefields and parameters in double (the
 *     reference's default build, init.hpp:16-18); ECSG_F32 stores them in
 *     float.  Interface ghost solves, energy and residual reductions always
 *     compute in double.
 */
#ifndef ECSG_HIP_H
#define ECSG_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ECSG_MAX_GRIDS 8

enum {
    ECSG_OK = 0,
    ECSG_EINVAL = -1,   /* bad argument / unsupported configuration */
    ECSG_EHIP = -2,     /* HIP runtime error */
    ECSG_ENOMEM = -3,   /* device allocation failed */
    ECSG_ENOCONV = -4,  /* Jacobi iteration cap reached */
    ECSG_ESTATE = -5,   /* call out of order (e.g. params not set) */
    ECSG_ETIMEOUT = -6  /* a grid barrier of the cooperative Jacobi solve waited past its limit */
};

enum { ECSG_F32 = 0, ECSG_F64 = 1 };

/* Field / parameter selectors for ecsg_get_field / ecsg_set_field.
 * 0..8 follow sim[][] of main.cpp:86-94 (init_del_sim.cpp:161-173). */
enum {
    ECSG_VX = 0, ECSG_VY, ECSG_VZ, ECSG_SXX, ECSG_SYY, ECSG_SZZ, ECSG_SYZ, ECSG_SXZ, ECSG_SXY,
    ECSG_BUOY = 9,          /* buoyancy at primary nodes */
    ECSG_C11 = 10           /* 10..18: staged c11 c22 c33 c23 c13 c12 c44 c55 c66 */
};

typedef struct ecsg_ctx ecsg_ctx;

typedef struct {
    int num_grids;                      /* main.cpp:66-67 */
    int n[ECSG_MAX_GRIDS][3];           /* cells per grid (Grid::size, grid.hpp:8) */
    double h[ECSG_MAX_GRIDS][3];        /* grid steps (Grid::step) */
    double dt;                          /* time step (main.cpp:51) */
    int radius;                         /* stencil radius 1..6 (main.cpp:48, init_del_fdcoeff.cpp:38-85);
                                           above 1 on a single grid only */
    int top_bc;                         /* 0 rigid, 1 free surface (main.cpp:60) */
    int imode;                          /* 1 = energy-conserving interface (ECI), 0 = the
                                           "intuitive" interpolation (II, ratio 2 only), main.cpp:54 */
    int check_energy;                   /* main.cpp:45 */
    int precision;                      /* ECSG_F32 or ECSG_F64 */
    int max_steps;                      /* capacity of the trace / energy buffers */
    int vy_mirror_all;                  /* 0: the reference's ghost-velocity loop ranges, in which
                                           vy is not mirrored across the y walls on plane iz = nz
                                           (update_vel.cpp:211-218, 239-246); 1: mirror it there too,
                                           which makes the composite energy exactly conserved */
    int sxz_solver;                     /* vx-sxz interface system: 0 = Jacobi sweeps as the
                                           reference (update_ghost_sxz.cpp:314-350), one workgroup
                                           per interface for iterates of up to 2048 entries, a
                                           cooperative tiled solve above; 1 = the same sweeps with
                                           Chebyshev acceleration (same fixed point and exit test,
                                           fewer sweeps); 2 = the Jacobi sweeps of 0, one workgroup
                                           per interface up to 9728 entries (graph-capturable; for
                                           many models stepped at once on their own streams);
                                           3 = a direct solve (no counterpart in the reference):
                                           block elimination over the coarse rows with the inverse
                                           Schur complements factored once per parameter set, one
                                           workgroup per interface, graph-capturable; coarse rows of
                                           at most 1024 unknowns and at most 512 MiB of factors per
                                           interface.  The transposed solves of the gradient stay
                                           Jacobi sweeps (to the adjoint tolerance). */
    int ii_vz_xend;                     /* imode 0 only: the x-end edge of the fine vz ghosts
                                           (update_ghost_vz.cpp:98-107).  0: left unwritten, as the
                                           committed version of that line (whose stores fall outside
                                           the field) -- the form the reference's golden II traces
                                           were made with; 1: the working copy's form, including its
                                           read of vz2 at the fine index nx1-1 (flat Field storage:
                                           the element one coarse row length further on) */
} ecsg_config;

/* Creates the device state: wavefields (InitSim, init_del_sim.cpp:3-29),
 * interface containers (InitGhostContainer, init_del_ghost_container.cpp:4-44),
 * FD coefficients (InitFDcoeff, init_del_fdcoeff.cpp:7-22).  Fine grid k and
 * coarse grid k+1 must have sizes in an integer ratio r >= 2 in x and y. */
int ecsg_create(const ecsg_config *cfg, ecsg_ctx **out);
void ecsg_destroy(ecsg_ctx *ctx);
const char *ecsg_last_error(const ecsg_ctx *ctx);
const char *ecsg_version(void);

/* Parameter staging of grid g (InitParField, init_del_par_field.cpp:77-150 and
 * ReshapeField, field.cpp:428-476).  buoy and stiff[0..8] are primary-node
 * arrays of (nx+1)(ny+1)(nz+1) values in the order c11 c22 c33 c23 c13 c12
 * c44 c55 c66, code convention c44 = 2*C44(Voigt).  iso != 0 selects the
 * reference's input mode: only stiff[0] (s11) and stiff[4] (s13) are read and
 * c44 = c55 = c66 = s11 - s13 (init_del_par_field.cpp:119).  Also builds the
 * time-invariant interface matrices of the interfaces touching g. */
int ecsg_set_params(ecsg_ctx *ctx, int g, const double *buoy, const double *const stiff[9], int iso);

/* Time step from the energy-method CFL bound the reference's flows use
 * (tests/test1-cmp-eci-ii/eci/SConstruct:107-114, where main.cpp takes dt as
 * given, main.cpp:51): dt = frac * min over grids g of
 * min(h_g) / (vp_g sqrt 3), vp_g^2 = the largest max(c11, c22, c33) * b over
 * the primary nodes of grid g, reduced on the device in FP64.  Inputs as for
 * ecsg_set_params (primary nodes, per grid buoy[g] and stiff[g][0..8]; iso:
 * only stiff[g][0] = s11 is read); cfg supplies num_grids, n and h (its dt is
 * not read), so the call comes before ecsg_create.  Needs no handle. */
int ecsg_cfl_dt(const ecsg_config *cfg, const double *const buoy[], const double *const stiff[][9], int iso,
                double frac, double *dt_out);

/* ecsg_cfl_dt for callers that cannot allocate (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc); no counterpart in
 * the reference.  Inputs as for ecsg_cfl_dt but resident in device memory
 * (buoy[g], stiff[g][0..8]: device pointers to primary-node arrays); dt is
 * written to the device double dt_dev by work enqueued on `stream` (a
 * hipStream_t; null = the null stream).  No device or host allocation: the
 * block maxima live in a fixed scratch array of the library, 65.6 KB per GPU
 * (at most ECSG_MAX_GRIDS grids, else ECSG_EINVAL), shared by all calls on
 * that GPU -- calls that overlap on one GPU race.  The result equals
 * ecsg_cfl_dt's to within one ulp (its square roots are the device's). */
int ecsg_cfl_dt_noalloc(const ecsg_config *cfg, const double *const buoy[], const double *const stiff[][9], int iso,
                        double frac, double *dt_dev, void *stream);

/* Grid g's parameters as main.cpp holds them after InitParField, i.e. what
 * UpdateVel / UpdateStress / UpdateGhost* receive (sim_utils.hpp:128-143):
 * buoy = buoys[g] (primary nodes) and stiff[m] = stiffs_ptr[g][m]->ptr1 for
 * m = 0..8 (order c11 c22 c33 c23 c13 c12 c44 c55 c66), each Field at its
 * staggered position -- c11..c12 at (1,0,0), c44 at (1,1,1), c55 at (0,0,1),
 * c66 at (0,1,0) (init_del_par_field.cpp:82-88, 129-149) -- with halo 0, x
 * fastest, and n or n+1 entries per dimension for a staggered or primary
 * direction (field.cpp:85-98).  Values are taken as given (already averaged).
 * Pointers may alias, as InitParField's do (stiff[0..2] one Field, stiff[3..5]
 * another): then the energy uses the reference's isotropic compliance
 * (init_del_par_field.cpp:259-278).  Gradients of such a grid
 * (ecsg_get_gradient) are w.r.t. the staged Fields, in their layout; a caller
 * whose pointers alias sums the gradients of the aliased entries. */
int ecsg_set_params_staged(ecsg_ctx *ctx, int g, const double *buoy, const double *const stiff[9]);

/* Source (InitSrc, init_src.cpp:7-45, resolved to grid indices by the caller).
 * type 0: vx += 1000*trace[n] in float (main.cpp:194-195);
 * type 1: sxx, syy, szz += trace[n] (main.cpp:197-202). */
int ecsg_set_source(ecsg_ctx *ctx, int g, int ix, int iy, int iz, int type, const float *trace, int len);

/* Receivers (InitRecTrace, init_del_rec_trace.cpp:16-55): loc[4*i..] =
 * grid, ix, iy, iz; each records vx once per step. */
int ecsg_set_receivers(ecsg_ctx *ctx, int n, const int *loc);

/* --- one entry point per phase of the time loop, main.cpp:165-261 ------ */
int ecsg_update_vel(ecsg_ctx *ctx, int g);            /* UpdateVel, update_vel.cpp:4-249; main.cpp:179-180 */
int ecsg_record_receivers(ecsg_ctx *ctx);             /* main.cpp:183-184 */
int ecsg_add_source(ecsg_ctx *ctx);                   /* main.cpp:194-195 */
int ecsg_update_ghost_vz(ecsg_ctx *ctx, int k);       /* UpdateGhostVZ, update_ghost_vz.cpp:124-250 (imode 0: 36-123); main.cpp:206-209 */
int ecsg_compute_energy(ecsg_ctx *ctx, double *e);    /* ComputeEnergy, compute_energy.cpp:34-54; main.cpp:217-225 */
int ecsg_update_stress(ecsg_ctx *ctx, int g);         /* UpdateStress, update_stress.cpp:4-311; main.cpp:228-229 */
int ecsg_update_ghost_sxz(ecsg_ctx *ctx, int k);      /* UpdateGhostSXZ + Jacobi2D, update_ghost_sxz.cpp:86-350 (imode 0: 36-84); main.cpp:242-246 */
int ecsg_update_ghost_syz(ecsg_ctx *ctx, int k);      /* UpdateGhostSYZ, update_ghost_syz.cpp:151-240 (imode 0: 23-150); main.cpp:247-250 */
int ecsg_end_step(ecsg_ctx *ctx);                     /* niter++ of main.cpp:165 */

/* Interface residuals (compute_res_*, compute_res.cpp:4-95) with the factor 4
 * generalised to r^2: which 0 = vx-sxz, 1 = vy-syz, 2 = vz-szz; out has one
 * value per interface. */
int ecsg_compute_res(ecsg_ctx *ctx, int which, double *out);

/* Field::AvgAbs (field.cpp:215-228) of wavefield sel (0..8) of grid g, over
 * the field's own range without halo: the companion values of the residual
 * trace (compute_res.cpp:21-23, 52-54, 84-86).  Summed on the device in a
 * fixed order. */
int ecsg_field_avgabs(ecsg_ctx *ctx, int g, int sel, double *out);

/* TestSBP (test_sbp.cpp:5-199; called after UpdateStress in the committed
 * main.cpp): the seven summation-by-parts residuals of grid g,
 * (u, D s)_u + (D u, s)_s for (vx, sxz, z), (vx, sxx, x), (vx, sxy, y),
 * (vy, syz, z), (vz, szz, z), (vz, sxz, x), (vz, syz, y), with the
 * reference's boundary weights and index ranges.  Summed in FP64 on the
 * device in a fixed order. */
int ecsg_sbp_residuals(ecsg_ctx *ctx, int g, double out[7]);

/* Runs nsteps full steps of main.cpp:165-261 on the device (receivers and
 * energy recorded on the device), captured once into a hipGraph and replayed
 * when use_graph != 0. */
int ecsg_step_n(ecsg_ctx *ctx, int nsteps, int use_graph);

int ecsg_steps_done(const ecsg_ctx *ctx);
int ecsg_synchronize(ecsg_ctx *ctx);

/* Field transfer in the reference's Field layout with halo 1 for wavefields
 * (len = n+1+2 non-staggered, n+2 staggered, field.cpp:85-98) and no halo for
 * parameters (sel >= ECSG_BUOY). */
int ecsg_field_dims(const ecsg_ctx *ctx, int g, int sel, int len[3]);
int ecsg_get_field(ecsg_ctx *ctx, int g, int sel, double *out);
int ecsg_set_field(ecsg_ctx *ctx, int g, int sel, const double *in);

/* Recorded data: traces[nrec][steps_done], energy[steps_done]. */
int ecsg_get_traces(ecsg_ctx *ctx, double *out);
int ecsg_get_energy(ecsg_ctx *ctx, double *out);

/* Jacobi statistics of interface k: total iterations, iterations of the last
 * solve, and whether the iteration cap was hit. */
int ecsg_jacobi_stats(ecsg_ctx *ctx, int k, long long *total, int *last, int *capped);

/* Measurement hook for the benchmark: runs nsteps steps without a graph and
 * times each kernel launch with HIP events on the stream it runs on.
 * ms_out[kind] receives the average duration of kernel kind over the run and
 * count_out[kind] the number of launches: kind 0 = velocity stencil of grid 0,
 * 1 = stress stencil of grid 0, 2 = all velocity stencils, 3 = all stress
 * stencils, 4 = interface kernels, 5 = receivers/source/energy kernels.
 * With the interface overlap active (ecsg_overlap_active) kinds 0 and 1 are
 * the launches over the interior planes of grid 0: velocities z = 0 .. nz-1,
 * stresses z = 0 .. nz-1 (mode 1) or z = 0 .. nz-2 (mode 2); its other
 * planes are launches of their own, counted in kinds 2 and 3 only. */
int ecsg_profile_kernels(ecsg_ctx *ctx, int nsteps, double ms_out[6], int count_out[6]);

/* Measurement hook for the benchmark's timed region: runs nsteps steps as
 * ecsg_step_n without a graph, with HIP events on the launch stream around
 * the grid-0 velocity and stress stencil launches only (the kinds 0 and 1 of
 * ecsg_profile_kernels); ms_out receives their average durations.  Blocks
 * until the steps are done. */
int ecsg_step_n_timed(ecsg_ctx *ctx, int nsteps, double ms_out[2]);

/* --- Adjoint-state gradient.  No counterpart in the reference, which has no
 * inversion path; built on the same time step as the entry points above. ---
 *
 * Misfit J = 1/2 sum_n sum_r (vx_r(n) - observed[r*nsteps + n])^2 over the
 * first nsteps steps from the zero state, vx_r(n) as recorded by the
 * receivers.  The gradient is taken w.r.t. the arrays given to
 * ecsg_set_params (buoyancy and stiff[0..8] at primary nodes; for iso grids
 * w.r.t. stiff[0] = s11 and stiff[4] = s13).  The forward pass stores a
 * checkpoint of the whole state every `every` steps; the backward pass
 * recomputes each segment from its checkpoint and runs the transposed steps
 * (time-reversed transposed stencils, transposed interface solves, misfit
 * injection at the receivers) over it. */
enum { ECSG_CKPT_NATIVE = 0, ECSG_CKPT_FP16 = 1, ECSG_CKPT_BF16 = 2 };

/* Zero every wavefield, interface warm start, trace and the step counter. */
int ecsg_reset(ecsg_ctx *ctx);

/* Checkpoint interval (default 50; 0: from the free device memory, below)
 * and storage: native precision, or IEEE half / bfloat16 of each array
 * divided by its maximum magnitude.  With interval K over N steps the first
 * segment takes the remainder and the last one K steps.  One level: a ring
 * of K + 1 saved states; the last segment is never recomputed (the forward
 * pass leaves its states in the ring), so N - K steps are.  Two levels, when
 * the ring is shorter than the interval (ecsg_set_ring_steps, or the memory
 * does not hold K + 1 states): the backward pass recomputes each segment
 * once from its checkpoint, leaving checkpoints every R steps (R = ring
 * length) and its last R steps in the ring, then recomputes each other
 * sub-segment from its checkpoint into the ring.  Interval 0: the largest
 * one-level interval that fits in the free device memory (less a reserve),
 * else the two-level interval and ring that fit with the fewest recomputed
 * steps.  Results do not depend on the interval or the ring length (native
 * checkpoints: bit for bit). */
int ecsg_set_checkpointing(ecsg_ctx *ctx, int every, int format);

/* The checkpoint interval the last ecsg_gradient call used (0 before any). */
int ecsg_checkpoint_interval(const ecsg_ctx *ctx);

/* Longest ring of saved states a gradient may hold (0, the default: as long
 * as the interval, or as fits in device memory).  No counterpart in the
 * reference (its time loop has no adjoint). */
int ecsg_set_ring_steps(ecsg_ctx *ctx, int steps);

/* Checkpoint interval, ring length and recomputed forward steps of the last
 * ecsg_gradient call (any pointer may be null). */
int ecsg_checkpoint_levels(const ecsg_ctx *ctx, int *interval, int *ring, long long *recomputed);

/* Runs the forward pass and the adjoint pass; misfit receives J.  After the
 * call the forward state is that of an intermediate step: ecsg_reset before
 * stepping again. */
int ecsg_gradient(ecsg_ctx *ctx, int nsteps, const double *observed, double *misfit);

/* dJ/d(input) of the last ecsg_gradient: sel = ECSG_BUOY for buoyancy or
 * ECSG_C11 + m for stiff[m]; out has (nx+1)(ny+1)(nz+1) values, or for a
 * grid set by ecsg_set_params_staged the length of staged Field m. */
int ecsg_get_gradient(ecsg_ctx *ctx, int g, int sel, double *out);

/* Test hooks: one transposed time step applied to the adjoint wavefields
 * (no misfit injection; gradient terms use the current forward state), and
 * transfer of the adjoint wavefields in the layout of ecsg_get_field. */
int ecsg_adjoint_step(ecsg_ctx *ctx);
int ecsg_get_adjoint_field(ecsg_ctx *ctx, int g, int sel, double *out);
int ecsg_set_adjoint_field(ecsg_ctx *ctx, int g, int sel, const double *in);

/* Total Jacobi sweeps of the transposed vx-sxz solves since the last gradient. */
int ecsg_adjoint_stats(ecsg_ctx *ctx, long long *sweeps);

/* Interface overlap (no counterpart in the reference, whose step is
 * serial): in ecsg_step_n, ecsg_profile_kernels and the gradient's forward
 * steps of a composite model (radius 1, imode 1, energy check off) the
 * interface ghost updates run on a second stream beside the stencil planes
 * that do not touch the interfaces, and the planes next to each interface
 * are updated after the ghosts they read.  Results are identical to the
 * serial step.  on: 1 = on, 0 = off, -1 (default) = on when the steps use
 * the multi-workgroup Jacobi solve (vx-sxz iterates above 2048 entries;
 * smaller ones are solved by one workgroup in graph-captured steps).
 * ecsg_overlap_active returns the schedule the
 * next steps use: 0 = serial; 1 = overlap; 2 = overlap with the stress
 * interior of grid 0 also beside the sxz/syz ghost solves (the source and
 * every receiver at least two planes away from each interface).  The
 * overlap is used only when every cooperative (grid-barrier) Jacobi launch of
 * the step has at most one workgroup per compute unit; steps with larger
 * launches are serial whatever `on` says. */
int ecsg_set_overlap(ecsg_ctx *ctx, int on);
int ecsg_overlap_active(const ecsg_ctx *ctx);

/* Longest wait, in milliseconds, of one workgroup at one grid barrier of the
 * cooperative Jacobi solves (default 10000; no counterpart in the reference,
 * whose Jacobi2D is a host loop, update_ghost_sxz.cpp:314-350).  A wait past
 * it ends the solve early and the next ecsg_sync (or gradient) returns
 * ECSG_ETIMEOUT instead of the step never finishing; the wavefield is then
 * invalid.  0 makes the first barrier of every solve time out (a test of
 * the error path). */
int ecsg_set_barrier_timeout(ecsg_ctx *ctx, int ms);

/* Exit tolerance of the forward vx-sxz Jacobi sweeps (sxz_solver 0, 1, 2):
 * the sweeps stop when the update norm is at most
 * min(tol_cap, tol_fac * AvgAbs(B)), as update_ghost_sxz.cpp:229-236 forms it.
 * The defaults are that line's constants of the storage precision: 1e-3 and
 * 1e-4 for ECSG_F32, 1e-12 and 1e-5 for ECSG_F64.  The solve runs in double
 * in both precisions, so an ECSG_F32 handle may take the double-precision
 * pair (1e-12, 1e-5).  Members of one batch must share the tolerance. */
int ecsg_set_jacobi_tol(ecsg_ctx *ctx, double tol_cap, double tol_fac);

/* ecsg_step_n's velocity and stress updates in one fused pass (default off;
 * used when the energy check is off): the pass reads the state before the
 * step and writes the state after it into a second set of field arrays
 * (allocated by the first call with on = 1), then the sets swap.  Results are
 * identical to the separate in-place passes. */
int ecsg_set_fused(ecsg_ctx *ctx, int on);

/* Stencil tiling of each grid (workgroup height by_out[g] in {1,2,4,8} rows
 * of 64, z chunk zc_out[g] in {1,2,4,8,16} planes per thread) chosen by timing
 * the velocity and stress launches of every candidate reps times on a saved
 * copy of the state, which is restored.  New (the reference has no tiling). */
int ecsg_autotune(ecsg_ctx *ctx, int reps, int *by_out, int *zc_out);

/* --- Batches.  No counterpart in the reference, which runs one model per
 * process.  A batch advances many handles (shots, models of any sizes; one
 * precision; no energy check; Jacobi sweeps with interfaces of at most 9728
 * iterate entries) with one launch per phase of the time step: the stencil
 * tiles of every grid of every member in one launch, the interface kernels
 * over every interface, one workgroup per member for receivers/source and
 * per interface for the Jacobi solve.  With use_graph the batched step is
 * captured once in a hipGraph and replayed.  When a member has an
 * interface the step runs the interface solves on a second stream beside
 * the stencil planes away from the interfaces (eight steps per graph; the
 * same results bit for bit; the environment variable ECSG_BATCH_OVERLAP=0
 * at ecsg_batch_create selects the serial step; ECSG_BATCH_PERSIST=k > 0
 * makes the stencil launches persistent, k workgroups per CU looping over
 * the tiles, slower on the workloads measured).  Members keep their state,
 * traces and accessors; calls on a member after ecsg_batch_step_n are
 * ordered after the batch's work.  Members must outlive the batch. --- */
typedef struct ecsg_batch ecsg_batch;
int ecsg_batch_create(ecsg_ctx *const *members, int n, ecsg_batch **out);
void ecsg_batch_destroy(ecsg_batch *b);
const char *ecsg_batch_last_error(const ecsg_batch *b);
int ecsg_batch_step_n(ecsg_batch *b, int nsteps, int use_graph);
int ecsg_batch_synchronize(ecsg_batch *b);
/* one tiling for the whole batch, or the fastest of 4 heights x 4 z chunks
 * by the time of the step's stencil launches alone (on a saved copy of the
 * state; with the overlap this does not see the solves beside them: time
 * real steps per candidate to choose among tilings, as ecsg_amd.Batch's
 * tune_steps does) */
int ecsg_batch_set_tiling(ecsg_batch *b, int by, int zc);
int ecsg_batch_autotune(ecsg_batch *b, int reps, int *by, int *zc);

/* Gradients of every member at once (no counterpart in the reference): each
 * member's ecsg_gradient -- misfit against observed[m] (member m's nrec x
 * nsteps traces, receiver-major), forward pass from rest with checkpoints,
 * recomputation, transposed steps -- with one launch per phase of the
 * forward and of the transposed step across all members; every member's
 * result equals its own ecsg_gradient bit for bit.  misfit[m] receives
 * member m's misfit.  Members share one checkpoint interval and format
 * (ecsg_set_checkpointing; interval 0: the largest that fits every member's
 * ring and checkpoints in free device memory together).  group (may be
 * NULL): group[m] names member m's model; the gradient accumulators of the
 * other members of a group are then added, in member order and on the
 * device, into its first member, so ecsg_get_gradient on that member returns
 * the model's gradient (the sum over its shots); the others keep their own. */
int ecsg_batch_gradient(ecsg_batch *b, int nsteps, const double *const *observed, double *misfit, const int *group);
/* the checkpoint interval the last ecsg_batch_gradient used */
int ecsg_batch_checkpoint_interval(const ecsg_batch *b);

/* --- Measurement.  No counterpart in the reference. --- */
/* Copies n FP32 values from src to dst (device pointers, disjoint) on stream
 * (a hipStream_t, NULL = the null stream), one value per lane as the stencils
 * access them: hints 0 = plain loads and stores, 3 = non-temporal loads and
 * stores (the stencils' streaming hints).  src == dst reads and writes back
 * every value in place; otherwise the ranges must not overlap.  A copy of
 * known size scales the HBM byte counters (FETCH_SIZE / WRITE_SIZE) of
 * kernels that use the same hints.  Asynchronous; returns ECSG_EINVAL for
 * other hints or null pointers. */
int ecsg_copy_probe(float *dst, const float *src, uint64_t n, int hints, void *stream);
/* A stencil sweep's memory traffic without its arithmetic, for the
 * streaming bound of its traversal.  pattern 0 (the stress sweep): 18 FP32
 * arrays are read and arrays 0..5 written back in place; pattern 1 (the
 * velocity sweep): 10 arrays read, 0..2 written back.  Array k (of a radius-1
 * grid of nx x ny x nz cells in the device layout: x fastest, ix = 0 on a
 * 128-byte line, a halo of 1, rows padded to 128 bytes; array k at base + k *
 * ecsg_stream_probe_floats(...) / 18) receives a[k] + a[k+W] + a[k+2W] + ...
 * over the arrays read (W = the number written back), summed left to right
 * in FP32, with the sweeps' non-temporal hints.  zc >= 1: over the (nx+1)(ny+1)(nz+1) nodes in the stencil tiling
 * (64 x by rows x zc planes per workgroup, by = 1, 2, 4 or 8, workgroups in
 * the sweeps' banded XCD order); zc = 0: one flat pass over every entry of planes 0..nz (halo
 * rows and padding included).  Asynchronous on stream (NULL = the null
 * stream); ECSG_EINVAL for other by or pattern, negative zc, empty grids or
 * a null base.
 * ecsg_stream_probe_floats returns the buffer size in floats (0 for empty
 * grids). */
uint64_t ecsg_stream_probe_floats(int nx, int ny, int nz);
int ecsg_stream_probe(float *base, int nx, int ny, int nz, int by, int zc, int pattern, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* ECSG_HIP_H */
