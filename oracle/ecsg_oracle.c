/*
 * ecsg_oracle.c -- CPU restatement of the ECSG composite staggered-grid
 * elastic time step, used as the checker for the MI355X path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load this library, and only as the checker
 * or as the timed CPU baseline.  The product library (libecsg_hip.so)
 * neither links nor calls it.
 *
 * What it restates (file:line of the reference, llodds/ecsg):
 *   field layout / staggering ........ field.cpp:77-149, init_del_sim.cpp:12-29
 *   parameter staging ................ init_del_par_field.cpp:77-150, field.cpp:428-476
 *   compliance / density for energy .. init_del_par_field.cpp:224-290, field.cpp:395-426
 *   FD coefficients .................. init_del_fdcoeff.cpp:7-22, 38-85
 *   UpdateVel ........................ update_vel.cpp:4-249
 *   UpdateStress ..................... update_stress.cpp:4-311
 *   UpdateGhostVZ (mode 1) ........... update_ghost_vz.cpp:124-250
 *   UpdateGhostSXZ (mode 1), Jacobi2D, Norm2D
 *                                      update_ghost_sxz.cpp:86-297, 314-350, 357-388
 *   UpdateGhostSYZ (mode 1) .......... update_ghost_syz.cpp:151-240
 *   UpdateGhostVZ/SXZ/SYZ (mode 0, II)  update_ghost_vz.cpp:36-123,
 *                                      update_ghost_sxz.cpp:36-84,
 *                                      update_ghost_syz.cpp:23-150
 *   ComputeEnergy / TriMulFields ..... compute_energy.cpp:34-54, field.cpp:482-518
 *   compute_res_* .................... compute_res.cpp:4-95
 *   time loop / phase order .......... main.cpp:165-261
 *
 * Two interface ("ghost") formulations are provided:
 *   eci_mode 0: the reference's ratio-2 formulas, term for term, in the
 *               reference's order of floating-point operations;
 *   eci_mode 1: the ratio-generic energy-conserving interface systems derived
 *               for any integer refinement ratio r (bilinear prolongation of
 *               node-centred sums, injection of cell-centred sums, flux balance
 *               fine_sum = r^2 coarse_sum).  At r = 2 it equals mode 0 up to
 *               rounding.
 *
 * Precision: compiled twice.  Default real = double (the reference's default
 * build, init.hpp:16-18); -DORACLE_FP32 gives real = float (the reference's
 * single-precision build, init.hpp:11-13).
 *
 * Deviations from the reference that do not change results:
 *   - reductions (energy, Jacobi norm, AvgAbs) are summed in a fixed order, so
 *     they are deterministic for any thread count (the reference's OpenMP
 *     reductions are not);
 *   - the Jacobi loop has an iteration cap (the reference's do/while has none);
 *     hitting it is reported through orc_jacobi_stats.
 */
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <stdint.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifdef ORACLE_FP32
typedef float real;
#define JACOBI_TOL(avgabs) ((1e-3 < (avgabs) * 1e-4) ? 1e-3 : (avgabs) * 1e-4)
#else
typedef double real;
#define JACOBI_TOL(avgabs) ((1e-12 < (avgabs) * 1e-5) ? 1e-12 : (avgabs) * 1e-5)
#endif

#define MAXG 8
#define JACOBI_CAP 200000

/* ------------------------------------------------------------------ */
/* Field: x fastest, halo r on every side (field.cpp:77-149).          */
typedef struct {
    int pos[3];      /* 1 = staggered in that dimension */
    int n[3];        /* cells of the grid */
    int r;           /* halo */
    int len[3];
    long s1, s2;     /* strides of dims 1, 2 */
    real *base;      /* allocation */
    real *o;         /* address of index (0,0,0) */
} fld;

#define AT(f, i0, i1, i2) ((f)->o[(long)(i2) * (f)->s2 + (long)(i1) * (f)->s1 + (long)(i0)])

static int fld_end(const fld *f, int d) { return f->pos[d] ? f->n[d] - 1 : f->n[d]; }

static int fld_init(fld *f, const int pos[3], const int n[3], int r)
{
    for (int d = 0; d < 3; d++) {
        f->pos[d] = pos[d];
        f->n[d] = n[d];
        f->len[d] = pos[d] ? n[d] + 2 * r : n[d] + 1 + 2 * r;
    }
    f->r = r;
    f->s1 = f->len[0];
    f->s2 = (long)f->len[0] * f->len[1];
    long sz = f->s2 * f->len[2];
    f->base = (real *)calloc((size_t)sz, sizeof(real));
    if (!f->base) return -1;
    f->o = f->base + (long)r * f->s2 + (long)r * f->s1 + r;
    return 0;
}

static void fld_free(fld *f) { free(f->base); f->base = f->o = NULL; }
static long fld_size(const fld *f) { return f->s2 * f->len[2]; }

/* component positions, init_del_sim.cpp:163-173 */
static const int SIM_POS[9][3] = {
    {0, 0, 0}, {1, 1, 0}, {1, 0, 1}, {1, 0, 0}, {1, 0, 0}, {1, 0, 0}, {1, 1, 1}, {0, 0, 1}, {0, 1, 0}};
enum { VX = 0, VY, VZ, SXX, SYY, SZZ, SYZ, SXZ, SXY };
/* stiffness order of stiffs_ptr (init_del_par_field.cpp:129-149):
 * c11 c22 c33 c23 c13 c12 c44 c55 c66, staged at these positions */
static const int ST_POS[9][3] = {
    {1, 0, 0}, {1, 0, 0}, {1, 0, 0}, {1, 0, 0}, {1, 0, 0}, {1, 0, 0}, {1, 1, 1}, {0, 0, 1}, {0, 1, 0}};
enum { C11 = 0, C22, C33, C23, C13, C12, C44, C55, C66 };

typedef struct {
    int n[3];
    real h[3];
    real coeff[3][6];
    fld sim[9];
    fld buoy;           /* (0,0,0), no halo */
    fld st[9];          /* staged stiffness */
    fld dens[3];        /* energy weights for vx, vy, vz */
    fld comp[9];        /* compliance for energy (c11..c66 order) */
    fld vcopy[3];
    int have_params;
    int iso;
} grid_t;

typedef struct {
    int r;               /* refinement ratio */
    real *gx, *gtmp;     /* Jacobi iterate, (nx2+1)(ny2+1), zero border, warm start */
    real *A;             /* 9 bands, (nx2-1)(ny2-1) each */
    real *B;
    long iters_total;
    int iters_last;
    int capped;
} iface_t;

typedef struct {
    int ng, radius, top_bc, imode, eci_mode, check_energy;
    int vy_mirror_all;   /* 1: mirror vy across the y walls on every z plane */
    int ii_xend;         /* imode 0: form of the x-end vz ghost edge (ghost_vz_ii) */
    real dt;
    grid_t g[MAXG];
    iface_t itf[MAXG];
    /* source */
    int src_g, src_loc[3], src_type, src_len;
    float *src_trace;
    /* receivers */
    int nrec;
    int *rec_loc;        /* nrec x 4: grid, ix, iy, iz */
    int max_steps;
    real *traces;        /* nrec x max_steps */
    double *energy;      /* max_steps */
    int steps_done;
} osim;

/* ------------------------------------------------------------------ */
/* FD coefficients, init_del_fdcoeff.cpp:38-85 */
static int fd_coeff(int radius, real *c)
{
    switch (radius) {
    case 1: c[0] = (real)1; break;
    case 2: c[0] = (real)9 / 8; c[1] = -(real)1 / 24; break;
    case 3: c[0] = (real)75 / 64; c[1] = -(real)25 / 384; c[2] = (real)3 / 640; break;
    case 4: c[0] = (real)1225 / 1024; c[1] = -(real)245 / 3072; c[2] = (real)49 / 5120; c[3] = -(real)5 / 7168; break;
    case 5: c[0] = (real)711 / 587; c[1] = -(real)158 / 1761; c[2] = (real)296 / 21383; c[3] = -(real)25 / 14159; c[4] = (real)17 / 143243; break;
    case 6: c[0] = (real)1225 / 1003; c[1] = -(real)338 / 3487; c[2] = (real)35 / 2006; c[3] = -(real)127 / 42800; c[4] = (real)19 / 52924; c[5] = -(real)6 / 274627; break;
    default: return -1;
    }
    return 0;
}

int orc_fd_coeff(int radius, double *out)
{
    real c[6];
    if (fd_coeff(radius, c)) return -1;
    for (int i = 0; i < radius; i++) out[i] = (double)c[i];
    return 0;
}

/* ------------------------------------------------------------------ */
/* Creation                                                           */
void *orc_create(int ng, const int *n, const double *h, int radius, int top_bc, int imode,
                 int eci_mode, double dt, int check_energy, int max_steps)
{
    if (ng < 1 || ng > MAXG || radius < 1 || radius > 6) return NULL;
    if (ng > 1 && radius != 1) return NULL;
    osim *s = (osim *)calloc(1, sizeof(osim));
    if (!s) return NULL;
    s->ng = ng; s->radius = radius; s->top_bc = top_bc; s->imode = imode;
    s->eci_mode = eci_mode; s->dt = (real)dt; s->check_energy = check_energy;
    s->max_steps = max_steps;
    real fd[6];
    fd_coeff(radius, fd);
    for (int g = 0; g < ng; g++) {
        grid_t *G = &s->g[g];
        for (int d = 0; d < 3; d++) { G->n[d] = n[3 * g + d]; G->h[d] = (real)h[3 * g + d]; }
        /* coeff[g][j][k] = fd_coeff[k]*dt/step[j], init_del_fdcoeff.cpp:19 */
        for (int d = 0; d < 3; d++)
            for (int k = 0; k < radius; k++) G->coeff[d][k] = fd[k] * s->dt / G->h[d];
        for (int c = 0; c < 9; c++)
            if (fld_init(&G->sim[c], SIM_POS[c], G->n, radius)) return NULL;
        static const int p0[3] = {0, 0, 0};
        if (fld_init(&G->buoy, p0, G->n, 0)) return NULL;
        for (int c = 0; c < 9; c++)
            if (fld_init(&G->st[c], ST_POS[c], G->n, 0)) return NULL;
        if (check_energy) {
            for (int c = 0; c < 3; c++) {
                if (fld_init(&G->dens[c], SIM_POS[c], G->n, 0)) return NULL;
                if (fld_init(&G->vcopy[c], SIM_POS[c], G->n, radius)) return NULL;
            }
            for (int c = 0; c < 9; c++)
                if (fld_init(&G->comp[c], ST_POS[c], G->n, 0)) return NULL;
        }
    }
    for (int k = 0; k + 1 < ng; k++) {
        grid_t *F = &s->g[k], *C = &s->g[k + 1];
        iface_t *I = &s->itf[k];
        int r = F->n[0] / C->n[0];
        if (r < 2 || F->n[0] != r * C->n[0] || F->n[1] != r * C->n[1]) return NULL;
        if ((eci_mode == 0 || imode == 0) && r != 2) return NULL;
        I->r = r;
        long nx2 = C->n[0], ny2 = C->n[1];
        I->gx = (real *)calloc((size_t)((nx2 + 1) * (ny2 + 1)), sizeof(real));
        I->gtmp = (real *)calloc((size_t)((nx2 + 1) * (ny2 + 1)), sizeof(real));
        I->A = (real *)calloc((size_t)(9 * (nx2 - 1) * (ny2 - 1)), sizeof(real));
        I->B = (real *)calloc((size_t)((nx2 - 1) * (ny2 - 1)), sizeof(real));
        if (!I->gx || !I->gtmp || !I->A || !I->B) return NULL;
    }
    if (max_steps > 0) s->energy = (double *)calloc((size_t)max_steps, sizeof(double));
    return s;
}

void orc_destroy(void *h)
{
    osim *s = (osim *)h;
    if (!s) return;
    for (int g = 0; g < s->ng; g++) {
        grid_t *G = &s->g[g];
        for (int c = 0; c < 9; c++) { fld_free(&G->sim[c]); fld_free(&G->st[c]); fld_free(&G->comp[c]); }
        for (int c = 0; c < 3; c++) { fld_free(&G->dens[c]); fld_free(&G->vcopy[c]); }
        fld_free(&G->buoy);
    }
    for (int k = 0; k + 1 < s->ng; k++) {
        free(s->itf[k].gx); free(s->itf[k].gtmp); free(s->itf[k].A); free(s->itf[k].B);
    }
    free(s->src_trace); free(s->rec_loc); free(s->traces); free(s->energy);
    free(s);
}

/* ------------------------------------------------------------------ */
/* Parameter staging.                                                  */
/* ReshapeField (field.cpp:428-476): 8-point sum in the reference's    */
/* term order, divided by 8.                                           */
static void reshape(const real *prim, const int n[3], fld *out)
{
    const int *p = out->pos;
    long px = n[0] + 1, py = n[1] + 1;
#define PR(i, j, k) prim[((long)(k) * py + (j)) * px + (i)]
    int e0 = fld_end(out, 0), e1 = fld_end(out, 1), e2 = fld_end(out, 2);
#pragma omp parallel for schedule(static)
    for (int k = 0; k <= e2; k++)
        for (int j = 0; j <= e1; j++)
            for (int i = 0; i <= e0; i++) {
                if (!p[0] && !p[1] && !p[2]) { AT(out, i, j, k) = PR(i, j, k); continue; }
                AT(out, i, j, k) = (PR(i, j, k) + PR(i, j, k + p[2]) + PR(i, j + p[1], k) + PR(i, j + p[1], k + p[2])
                                    + PR(i + p[0], j, k) + PR(i + p[0], j, k + p[2]) + PR(i + p[0], j + p[1], k)
                                    + PR(i + p[0], j + p[1], k + p[2])) / 8;
            }
#undef PR
}

static void field_energy_weights(grid_t *G)
{
    /* densities: InvertParameters (init_del_par_field.cpp:231-247) */
    long np = (long)(G->n[0] + 1) * (G->n[1] + 1) * (G->n[2] + 1);
    real *prim = (real *)malloc(sizeof(real) * np);
    long px = G->n[0] + 1, py = G->n[1] + 1;
    for (int k = 0; k <= G->n[2]; k++)
        for (int j = 0; j <= G->n[1]; j++)
            for (int i = 0; i <= G->n[0]; i++) prim[((long)k * py + j) * px + i] = AT(&G->buoy, i, j, k);
    for (int c = 0; c < 3; c++) {
        fld *d = &G->dens[c];
        reshape(prim, G->n, d);
        for (int k = 0; k <= fld_end(d, 2); k++)
            for (int j = 0; j <= fld_end(d, 1); j++)
                for (int i = 0; i <= fld_end(d, 0); i++) AT(d, i, j, k) = 1 / AT(d, i, j, k);
    }
    free(prim);
    /* compliance.  Isotropic input: the reference's closed form
     * (init_del_par_field.cpp:259-278).  Otherwise: inverse of the 3x3
     * normal-stress block (no counterpart in the reference, which exits on
     * non-isotropic media at init_del_par_field.cpp:285-289). */
    fld *s11 = &G->st[C11], *s13 = &G->st[C13];
    int e0 = fld_end(s11, 0), e1 = fld_end(s11, 1), e2 = fld_end(s11, 2);
    for (int k = 0; k <= e2; k++)
        for (int j = 0; j <= e1; j++)
            for (int i = 0; i <= e0; i++) {
                if (G->iso) {
                    real a = AT(s11, i, j, k), b = AT(s13, i, j, k);
                    real base = a * b * b * 3 - b * b * b * 2 - a * a * a;
                    base = 1 / base;
                    real cc11 = (b * b - a * a) * base;
                    real cc13 = (b * a - b * b) * base;
                    for (int c = 0; c < 3; c++) AT(&G->comp[c], i, j, k) = cc11;
                    for (int c = 3; c < 6; c++) AT(&G->comp[c], i, j, k) = cc13;
                } else {
                    double m11 = AT(&G->st[C11], i, j, k), m22 = AT(&G->st[C22], i, j, k), m33 = AT(&G->st[C33], i, j, k);
                    double m23 = AT(&G->st[C23], i, j, k), m13 = AT(&G->st[C13], i, j, k), m12 = AT(&G->st[C12], i, j, k);
                    double a11 = m22 * m33 - m23 * m23, a22 = m11 * m33 - m13 * m13, a33 = m11 * m22 - m12 * m12;
                    double a23 = m12 * m13 - m11 * m23, a13 = m12 * m23 - m13 * m22, a12 = m13 * m23 - m12 * m33;
                    double det = m11 * a11 + m12 * a12 + m13 * a13;
                    AT(&G->comp[C11], i, j, k) = (real)(a11 / det);
                    AT(&G->comp[C22], i, j, k) = (real)(a22 / det);
                    AT(&G->comp[C33], i, j, k) = (real)(a33 / det);
                    AT(&G->comp[C23], i, j, k) = (real)(a23 / det);
                    AT(&G->comp[C13], i, j, k) = (real)(a13 / det);
                    AT(&G->comp[C12], i, j, k) = (real)(a12 / det);
                }
            }
    for (int c = 6; c < 9; c++) {
        fld *f = &G->comp[c];
        for (int k = 0; k <= fld_end(f, 2); k++)
            for (int j = 0; j <= fld_end(f, 1); j++)
                for (int i = 0; i <= fld_end(f, 0); i++) AT(f, i, j, k) = 1 / AT(&G->st[c], i, j, k);
    }
}

/* stiff9: primary-node arrays (nx+1)(ny+1)(nz+1), x fastest, in the order
 * c11 c22 c33 c23 c13 c12 c44 c55 c66 with code convention c44 = 2*C44(Voigt).
 * iso != 0: reference input mode; only stiff9[0] (s11) and stiff9[4] (s13) are
 * read and c44 = c55 = c66 = s11 - s13 (init_del_par_field.cpp:119). */
int orc_set_params(void *h, int g, const double *buoy, const double *const *stiff9, int iso)
{
    osim *s = (osim *)h;
    if (g < 0 || g >= s->ng) return -1;
    grid_t *G = &s->g[g];
    long np = (long)(G->n[0] + 1) * (G->n[1] + 1) * (G->n[2] + 1);
    long px = G->n[0] + 1, py = G->n[1] + 1;
    for (int k = 0; k <= G->n[2]; k++)
        for (int j = 0; j <= G->n[1]; j++)
            for (int i = 0; i <= G->n[0]; i++) AT(&G->buoy, i, j, k) = (real)buoy[((long)k * py + j) * px + i];
    real *prim = (real *)malloc(sizeof(real) * np);
    G->iso = iso;
    for (int c = 0; c < 9; c++) {
        if (iso) {
            if (c == C44 || c == C55 || c == C66) {
                for (long q = 0; q < np; q++) prim[q] = (real)stiff9[0][q] - (real)stiff9[4][q];
            } else if (c == C11 || c == C22 || c == C33) {
                for (long q = 0; q < np; q++) prim[q] = (real)stiff9[0][q];
            } else {
                for (long q = 0; q < np; q++) prim[q] = (real)stiff9[4][q];
            }
        } else {
            for (long q = 0; q < np; q++) prim[q] = (real)stiff9[c][q];
        }
        reshape(prim, G->n, &G->st[c]);
    }
    free(prim);
    G->have_params = 1;
    if (s->check_energy) field_energy_weights(G);
    return 0;
}

int orc_set_source(void *h, int g, int ix, int iy, int iz, int type, const float *trace, int len)
{
    osim *s = (osim *)h;
    if (g < 0 || g >= s->ng || len < 0) return -1;
    s->src_g = g; s->src_loc[0] = ix; s->src_loc[1] = iy; s->src_loc[2] = iz;
    s->src_type = type; s->src_len = len;
    free(s->src_trace);
    s->src_trace = (float *)malloc(sizeof(float) * (len > 0 ? len : 1));
    if (len > 0) memcpy(s->src_trace, trace, sizeof(float) * len);
    return 0;
}

int orc_set_receivers(void *h, int n, const int *loc4)
{
    osim *s = (osim *)h;
    free(s->rec_loc); free(s->traces);
    s->nrec = n;
    s->rec_loc = (int *)malloc(sizeof(int) * 4 * (n > 0 ? n : 1));
    if (n > 0) memcpy(s->rec_loc, loc4, sizeof(int) * 4 * n);
    s->traces = (real *)calloc((size_t)(n > 0 ? n : 1) * (s->max_steps > 0 ? s->max_steps : 1), sizeof(real));
    return 0;
}

/* ------------------------------------------------------------------ */
/* UpdateVel, update_vel.cpp:4-249                                     */
static void update_vel(osim *s, int g)
{
    grid_t *G = &s->g[g];
    int top = (g == 0), bottom = (g == s->ng - 1), top_bc = s->top_bc, radius = s->radius;
    fld *vx = &G->sim[VX], *vy = &G->sim[VY], *vz = &G->sim[VZ];
    fld *sxx = &G->sim[SXX], *syy = &G->sim[SYY], *szz = &G->sim[SZZ];
    fld *syz = &G->sim[SYZ], *sxz = &G->sim[SXZ], *sxy = &G->sim[SXY];
    fld *b = &G->buoy;
    real(*co)[6] = G->coeff;
    int nx = G->n[0], ny = G->n[1], nz = G->n[2];
    if (top && top_bc == 1) top = 0;

    /* vx */
#pragma omp parallel for schedule(static)
    for (int iz = 0 + top; iz <= nz - bottom; iz++)
        for (int iy = 1; iy < ny; iy++)
            for (int ir = 0; ir < radius; ir++)
                for (int ix = 1; ix < nx; ix++)
                    AT(vx, ix, iy, iz) += AT(b, ix, iy, iz)
                        * (co[0][ir] * (AT(sxx, ix + ir, iy, iz) - AT(sxx, ix - ir - 1, iy, iz))
                           + co[1][ir] * (AT(sxy, ix, iy + ir, iz) - AT(sxy, ix, iy - ir - 1, iz))
                           + co[2][ir] * (AT(sxz, ix, iy, iz + ir) - AT(sxz, ix, iy, iz - ir - 1)));
    /* vy */
#pragma omp parallel for schedule(static)
    for (int iz = 0 + top; iz <= nz - bottom; iz++)
        for (int iy = 0; iy <= ny - 1; iy++)
            for (int ir = 0; ir < radius; ir++)
                for (int ix = 0; ix <= nx - 1; ix++)
                    AT(vy, ix, iy, iz) += 0.25 * (AT(b, ix, iy, iz) + AT(b, ix + 1, iy, iz) + AT(b, ix, iy + 1, iz) + AT(b, ix + 1, iy + 1, iz))
                        * (co[0][ir] * (AT(sxy, ix + ir + 1, iy, iz) - AT(sxy, ix - ir, iy, iz))
                           + co[2][ir] * (AT(syz, ix, iy, iz + ir) - AT(syz, ix, iy, iz - ir - 1))
                           + co[1][ir] * (AT(syy, ix, iy + ir + 1, iz) - AT(syy, ix, iy - ir, iz)));
    /* vz */
#pragma omp parallel for schedule(static)
    for (int iz = 0; iz <= nz - 1; iz++)
        for (int iy = 1; iy < ny; iy++)
            for (int ir = 0; ir < radius; ir++)
                for (int ix = 0; ix <= nx - 1; ix++)
                    AT(vz, ix, iy, iz) += 0.25 * (AT(b, ix, iy, iz) + AT(b, ix, iy, iz + 1) + AT(b, ix + 1, iy, iz) + AT(b, ix + 1, iy, iz + 1))
                        * (co[0][ir] * (AT(sxz, ix + ir + 1, iy, iz) - AT(sxz, ix - ir, iy, iz))
                           + co[2][ir] * (AT(szz, ix, iy, iz + ir + 1) - AT(szz, ix, iy, iz - ir))
                           + co[1][ir] * (AT(syz, ix, iy + ir, iz) - AT(syz, ix, iy - ir - 1, iz)));

    /* ghost velocities (update_vel.cpp:73-246); loop ranges kept as written */
    if (top) {
        for (int iy = 0; iy <= ny; iy++)
            for (int ir = 1; ir <= radius - top_bc; ir++)
                for (int ix = 0; ix <= nx - 1; ix++) AT(vz, ix, iy, 0 - ir) = -AT(vz, ix, iy, 0 + ir - 1);
        for (int iy = 0; iy <= ny; iy++)
            for (int ir = 1; ir < radius; ir++)
                for (int ix = 0; ix <= nx - 1; ix++) AT(vx, ix, iy, 0 - ir) = -AT(vx, ix, iy, 0 + ir);
        for (int iy = 0; iy <= ny - 1; iy++)
            for (int ir = 1; ir < radius; ir++)
                for (int ix = 0; ix <= nx - 1; ix++) AT(vy, ix, iy, 0 - ir) = -AT(vy, ix, iy, 0 + ir);
    }
    if (bottom) {
        int iz = nz - 1;
        for (int iy = 0; iy <= ny; iy++)
            for (int ir = 1; ir <= radius; ir++)
                for (int ix = 0; ix <= nx - 1; ix++) AT(vz, ix, iy, iz + ir) = -AT(vz, ix, iy, iz - ir + 1);
        iz = nz;
        for (int iy = 0; iy <= ny; iy++)
            for (int ir = 1; ir < radius; ir++)
                for (int ix = 0; ix <= nx - 1; ix++) AT(vx, ix, iy, iz + ir) = -AT(vx, ix, iy, iz - ir);
        for (int iy = 0; iy <= ny - 1; iy++)
            for (int ir = 1; ir < radius; ir++)
                for (int ix = 0; ix <= nx - 1; ix++) AT(vy, ix, iy, iz + ir) = -AT(vy, ix, iy, iz - ir);
    }
    /* x start */
    for (int iz = 0; iz <= nz - 1; iz++)
        for (int ir = 1; ir <= radius; ir++)
            for (int iy = 0; iy <= ny; iy++) AT(vz, 0 - ir, iy, iz) = -AT(vz, 0 + ir - 1, iy, iz);
    for (int iz = 0; iz <= nz; iz++)
        for (int ir = 1; ir < radius; ir++)
            for (int iy = 0; iy <= ny; iy++) AT(vx, 0 - ir, iy, iz) = -AT(vx, 0 + ir, iy, iz);
    for (int iz = 0; iz <= nz; iz++)
        for (int ir = 1; ir <= radius; ir++)
            for (int iy = 0; iy <= ny; iy++) AT(vy, 0 - ir, iy, iz) = -AT(vy, 0 + ir - 1, iy, iz);
    /* x end */
    for (int iz = 0; iz <= nz - 1; iz++)
        for (int ir = 1; ir <= radius; ir++)
            for (int iy = 0; iy <= ny; iy++) AT(vz, nx - 1 + ir, iy, iz) = -AT(vz, nx - 1 - ir + 1, iy, iz);
    for (int iz = 0; iz <= nz; iz++)
        for (int ir = 1; ir < radius; ir++)
            for (int iy = 0; iy <= ny; iy++) AT(vx, nx + ir, iy, iz) = -AT(vx, nx - ir, iy, iz);
    for (int iz = 0; iz <= nz; iz++)
        for (int ir = 1; ir <= radius; ir++)
            for (int iy = 0; iy <= ny; iy++) AT(vy, nx - 1 + ir, iy, iz) = -AT(vy, nx - 1 - ir + 1, iy, iz);
    /* y start */
    for (int iz = 0; iz <= nz - 1; iz++)
        for (int ir = 1; ir < radius; ir++)
            for (int ix = 0; ix <= nx - 1; ix++) AT(vz, ix, 0 - ir, iz) = -AT(vz, ix, 0 + ir, iz);
    for (int iz = 0; iz <= nz - 1; iz++)
        for (int ir = 1; ir < radius; ir++)
            for (int ix = 0; ix <= nx; ix++) AT(vx, ix, 0 - ir, iz) = -AT(vx, ix, 0 + ir, iz);
    /* y walls for vy: the reference loops over vz's z range [0, nz-1]
     * (update_vel.cpp:215, 243), which leaves vy's plane iz = nz unmirrored;
     * vy_mirror_all extends the loop to that plane. */
    int zy = s->vy_mirror_all ? nz : nz - 1;
    for (int iz = 0; iz <= zy; iz++)
        for (int ir = 1; ir <= radius; ir++)
            for (int ix = 0; ix <= nx - 1; ix++) AT(vy, ix, 0 - ir, iz) = -AT(vy, ix, 0 + ir - 1, iz);
    /* y end */
    for (int iz = 0; iz <= nz - 1; iz++)
        for (int ir = 1; ir < radius; ir++)
            for (int ix = 0; ix <= nx - 1; ix++) AT(vz, ix, ny + ir, iz) = -AT(vz, ix, ny - ir, iz);
    for (int iz = 0; iz <= nz - 1; iz++)
        for (int ir = 1; ir < radius; ir++)
            for (int ix = 0; ix <= nx; ix++) AT(vx, ix, ny + ir, iz) = -AT(vx, ix, ny - ir, iz);
    for (int iz = 0; iz <= zy; iz++)
        for (int ir = 1; ir <= radius; ir++)
            for (int ix = 0; ix <= nx - 1; ix++) AT(vy, ix, ny - 1 + ir, iz) = -AT(vy, ix, ny - 1 - ir + 1, iz);
}

/* ------------------------------------------------------------------ */
/* UpdateStress, update_stress.cpp:4-311                               */
static void update_stress(osim *s, int g)
{
    grid_t *G = &s->g[g];
    int top = (g == 0), bottom = (g == s->ng - 1), top_bc = s->top_bc, radius = s->radius;
    fld *vx = &G->sim[VX], *vy = &G->sim[VY], *vz = &G->sim[VZ];
    fld *sxx = &G->sim[SXX], *syy = &G->sim[SYY], *szz = &G->sim[SZZ];
    fld *syz = &G->sim[SYZ], *sxz = &G->sim[SXZ], *sxy = &G->sim[SXY];
    fld *c11 = &G->st[C11], *c22 = &G->st[C22], *c33 = &G->st[C33];
    fld *c23 = &G->st[C23], *c13 = &G->st[C13], *c12 = &G->st[C12];
    fld *c44 = &G->st[C44], *c55 = &G->st[C55], *c66 = &G->st[C66];
    real(*co)[6] = G->coeff;
    int nx = G->n[0], ny = G->n[1], nz = G->n[2];
    int flag = (top && top_bc) ? 1 : 0;

    if (flag) {
        int i2 = 0;
#pragma omp parallel for schedule(static)
        for (int i1 = 0; i1 <= ny; i1++)
            for (int ir = 0; ir < radius; ir++)
                for (int i0 = 0; i0 <= nx - 1; i0++) {
                    AT(sxx, i0, i1, i2) += (AT(c11, i0, i1, i2) - AT(c13, i0, i1, i2) * AT(c13, i0, i1, i2) / AT(c33, i0, i1, i2))
                        * co[0][ir] * (AT(vx, i0 + ir + 1, i1, i2) - AT(vx, i0 - ir, i1, i2))
                        + (AT(c12, i0, i1, i2) - AT(c13, i0, i1, i2) * AT(c23, i0, i1, i2) / AT(c33, i0, i1, i2))
                        * co[1][ir] * (AT(vy, i0, i1 + ir, i2) - AT(vy, i0, i1 - ir - 1, i2));
                    AT(syy, i0, i1, i2) += (AT(c12, i0, i1, i2) - AT(c23, i0, i1, i2) * AT(c13, i0, i1, i2) / AT(c33, i0, i1, i2))
                        * co[0][ir] * (AT(vx, i0 + ir + 1, i1, i2) - AT(vx, i0 - ir, i1, i2))
                        + (AT(c22, i0, i1, i2) - AT(c23, i0, i1, i2) * AT(c23, i0, i1, i2) / AT(c33, i0, i1, i2))
                        * co[1][ir] * (AT(vy, i0, i1 + ir, i2) - AT(vy, i0, i1 - ir - 1, i2));
                }
    }
#pragma omp parallel for schedule(static)
    for (int i2 = 0 + flag; i2 <= nz; i2++)
        for (int i1 = 0; i1 <= ny; i1++)
            for (int ir = 0; ir < radius; ir++)
                for (int i0 = 0; i0 <= nx - 1; i0++) {
                    real dvx = AT(vx, i0 + ir + 1, i1, i2) - AT(vx, i0 - ir, i1, i2);
                    real dvy = AT(vy, i0, i1 + ir, i2) - AT(vy, i0, i1 - ir - 1, i2);
                    real dvz = AT(vz, i0, i1, i2 + ir) - AT(vz, i0, i1, i2 - ir - 1);
                    AT(sxx, i0, i1, i2) += AT(c11, i0, i1, i2) * co[0][ir] * dvx + AT(c12, i0, i1, i2) * co[1][ir] * dvy
                        + AT(c13, i0, i1, i2) * co[2][ir] * dvz;
                    AT(syy, i0, i1, i2) += AT(c12, i0, i1, i2) * co[0][ir] * dvx + AT(c22, i0, i1, i2) * co[1][ir] * dvy
                        + AT(c23, i0, i1, i2) * co[2][ir] * dvz;
                    AT(szz, i0, i1, i2) += AT(c13, i0, i1, i2) * co[0][ir] * dvx + AT(c23, i0, i1, i2) * co[1][ir] * dvy
                        + AT(c33, i0, i1, i2) * co[2][ir] * dvz;
                }
#pragma omp parallel for schedule(static)
    for (int i2 = 0; i2 <= nz - 1; i2++)
        for (int i1 = 0; i1 <= ny - 1; i1++)
            for (int ir = 0; ir < radius; ir++)
                for (int i0 = 0; i0 <= nx - 1; i0++)
                    AT(syz, i0, i1, i2) += 0.5 * AT(c44, i0, i1, i2)
                        * (co[2][ir] * (AT(vy, i0, i1, i2 + ir + 1) - AT(vy, i0, i1, i2 - ir))
                           + co[1][ir] * (AT(vz, i0, i1 + ir + 1, i2) - AT(vz, i0, i1 - ir, i2)));
#pragma omp parallel for schedule(static)
    for (int i2 = 0; i2 <= nz - 1; i2++)
        for (int i1 = 0; i1 <= ny; i1++)
            for (int ir = 0; ir < radius; ir++)
                for (int i0 = 0; i0 <= nx; i0++)
                    AT(sxz, i0, i1, i2) += 0.5 * AT(c55, i0, i1, i2)
                        * (co[2][ir] * (AT(vx, i0, i1, i2 + ir + 1) - AT(vx, i0, i1, i2 - ir))
                           + co[0][ir] * (AT(vz, i0 + ir, i1, i2) - AT(vz, i0 - ir - 1, i1, i2)));
#pragma omp parallel for schedule(static)
    for (int i2 = 0; i2 <= nz; i2++)
        for (int i1 = 0; i1 <= ny - 1; i1++)
            for (int ir = 0; ir < radius; ir++)
                for (int i0 = 0; i0 <= nx; i0++)
                    AT(sxy, i0, i1, i2) += 0.5 * AT(c66, i0, i1, i2)
                        * (co[1][ir] * (AT(vx, i0, i1 + ir + 1, i2) - AT(vx, i0, i1 - ir, i2))
                           + co[0][ir] * (AT(vy, i0 + ir, i1, i2) - AT(vy, i0 - ir - 1, i1, i2)));

    /* ghost stresses (update_stress.cpp:135-309) */
    for (int i2 = 0; i2 <= nz; i2++)
        for (int ir = 1; ir < radius; ir++)
            for (int i1 = 0; i1 <= ny; i1++) AT(sxx, 0 - ir, i1, i2) = AT(sxx, 0 + ir - 1, i1, i2);
    for (int i2 = 0; i2 <= nz - 1; i2++)
        for (int ir = 1; ir < radius; ir++)
            for (int i1 = 0; i1 <= ny; i1++) AT(sxz, 0 - ir, i1, i2) = AT(sxz, 0 + ir, i1, i2);
    for (int i2 = 0; i2 <= nz; i2++)
        for (int ir = 1; ir < radius; ir++)
            for (int i1 = 0; i1 <= ny - 1; i1++) AT(sxy, 0 - ir, i1, i2) = AT(sxy, 0 + ir, i1, i2);
    for (int i2 = 0; i2 <= nz; i2++)
        for (int ir = 1; ir < radius; ir++)
            for (int i1 = 0; i1 <= ny; i1++) AT(sxx, nx - 1 + ir, i1, i2) = AT(sxx, nx - 1 - ir + 1, i1, i2);
    for (int i2 = 0; i2 <= nz - 1; i2++)
        for (int ir = 1; ir < radius; ir++)
            for (int i1 = 0; i1 <= ny; i1++) AT(sxz, nx + ir, i1, i2) = AT(sxz, nx - ir, i1, i2);
    for (int i2 = 0; i2 <= nz; i2++)
        for (int ir = 1; ir < radius; ir++)
            for (int i1 = 0; i1 <= ny - 1; i1++) AT(sxy, nx + ir, i1, i2) = AT(sxy, nx - ir, i1, i2);
    for (int i2 = 0; i2 <= nz; i2++)
        for (int ir = 1; ir < radius; ir++)
            for (int i0 = 0; i0 <= nx - 1; i0++) AT(syy, i0, 0 - ir, i2) = AT(syy, i0, 0 + ir, i2);
    for (int i2 = 0; i2 <= nz - 1; i2++)
        for (int ir = 1; ir < radius; ir++)
            for (int i0 = 0; i0 <= nx - 1; i0++) AT(syz, i0, 0 - ir, i2) = AT(syz, i0, 0 + ir - 1, i2);
    for (int i2 = 0; i2 <= nz; i2++)
        for (int ir = 1; ir < radius; ir++)
            for (int i0 = 0; i0 <= nx; i0++) AT(sxy, i0, 0 - ir, i2) = AT(sxy, i0, 0 + ir - 1, i2);
    for (int i2 = 0; i2 <= nz; i2++)
        for (int ir = 1; ir < radius; ir++)
            for (int i0 = 0; i0 <= nx - 1; i0++) AT(syy, i0, ny + ir, i2) = AT(syy, i0, ny - ir, i2);
    for (int i2 = 0; i2 <= nz - 1; i2++)
        for (int ir = 1; ir < radius; ir++)
            for (int i0 = 0; i0 <= nx - 1; i0++) AT(syz, i0, ny - 1 + ir, i2) = AT(syz, i0, ny - 1 - ir + 1, i2);
    for (int i2 = 0; i2 <= nz; i2++)
        for (int ir = 1; ir < radius; ir++)
            for (int i0 = 0; i0 <= nx; i0++) AT(sxy, i0, ny - 1 + ir, i2) = AT(sxy, i0, ny - 1 - ir + 1, i2);
    if (top) {
        for (int i1 = 1; i1 < ny; i1++)
            for (int ir = 1; ir < radius; ir++)
                for (int i0 = 0; i0 <= nx - 1; i0++) AT(szz, i0, i1, 0 - ir) = AT(szz, i0, i1, 0 + ir);
        for (int i1 = 0; i1 <= ny - 1; i1++)
            for (int ir = 1; ir < radius + flag; ir++)
                for (int i0 = 0; i0 <= nx - 1; i0++) AT(syz, i0, i1, 0 - ir) = (1 - 2 * flag) * AT(syz, i0, i1, 0 + ir - 1);
        for (int i1 = 1; i1 < ny; i1++)
            for (int ir = 1; ir < radius + flag; ir++)
                for (int i0 = 1; i0 < nx; i0++) AT(sxz, i0, i1, 0 - ir) = (1 - 2 * flag) * AT(sxz, i0, i1, 0 + ir - 1);
    }
    if (bottom) {
        for (int i1 = 1; i1 < ny; i1++)
            for (int ir = 1; ir < radius; ir++)
                for (int i0 = 0; i0 <= nx - 1; i0++) AT(szz, i0, i1, nz + ir) = AT(szz, i0, i1, nz - ir);
        for (int i1 = 0; i1 <= ny - 1; i1++)
            for (int ir = 1; ir < radius; ir++)
                for (int i0 = 0; i0 <= nx - 1; i0++) AT(syz, i0, i1, nz - 1 + ir) = AT(syz, i0, i1, nz - 1 - ir + 1);
        for (int i1 = 1; i1 < ny; i1++)
            for (int ir = 1; ir < radius; ir++)
                for (int i0 = 1; i0 < nx; i0++) AT(sxz, i0, i1, nz - 1 + ir) = AT(sxz, i0, i1, nz - 1 - ir + 1);
    }
}

/* ------------------------------------------------------------------ */
/* Interface, imode 0: the reference's "intuitive" (II) interpolation, */
/* ratio 2 only, term for term.  Every formula reads the fine planes   */
/* nz1-2 / nz1-1 and the coarse plane 0 only, so the loop order does   */
/* not matter.                                                         */

/* UpdateGhostVZ mode 0, update_ghost_vz.cpp:36-123.
 * The x-end edge (fine x = nx1-1, update_ghost_vz.cpp:98-107) has two forms:
 *   ii_xend 0: not written.  This is the committed version of that line
 *              (`int i1 = vz1.GetEnd(0);`, which leaves i0 = 0 and shadows
 *              i1, so its stores go to y indices beyond the field -- outside
 *              the allocation for the reference flow's grids -- and are not
 *              reproduced); the reference's golden II traces were made with
 *              it and are matched by this form;
 *   ii_xend 1: the working copy's form (`i0 = vz1.GetEnd(0);`), which reads
 *              vz2 at x index nx1-1, beyond the coarse row: Field storage is
 *              one flat array (field.cpp:123-136), so the read lands in a
 *              following row.  AT() computes the same flat offset (the
 *              oracle's fields have the reference's lengths and halo). */
static void ghost_vz_ii(osim *s, int k)
{
    grid_t *F = &s->g[k], *C = &s->g[k + 1];
    fld *vz1 = &F->sim[VZ], *vz2 = &C->sim[VZ];
    const int nx2 = C->n[0], ny2 = C->n[1], nz1 = F->n[2];
#pragma omp parallel for schedule(static)
    for (int i1 = 1; i1 < ny2; i1++)
        for (int i0 = 0; i0 < nx2; i0++)
            AT(vz2, i0, i1, -1) = 0.25 * (AT(vz1, 2 * i0, 2 * i1, nz1 - 2) + AT(vz1, 2 * i0 + 1, 2 * i1, nz1 - 2)
                                          + AT(vz1, 2 * i0, 2 * i1, nz1 - 1) + AT(vz1, 2 * i0 + 1, 2 * i1, nz1 - 1));
#pragma omp parallel for schedule(static)
    for (int i1 = 0; i1 < ny2; i1++)
        for (int i0 = 0; i0 < nx2 - 1; i0++) {
            const real a = AT(vz2, i0, i1, 0), b = AT(vz2, i0 + 1, i1, 0);
            const real c = AT(vz2, i0, i1 + 1, 0), d = AT(vz2, i0 + 1, i1 + 1, 0);
            AT(vz1, 2 * i0 + 1, 2 * i1, nz1) = AT(vz1, 2 * i0 + 1, 2 * i1, nz1 - 1) / 3 + a / 2 + b / 6;
            AT(vz1, 2 * i0 + 2, 2 * i1, nz1) = AT(vz1, 2 * i0 + 2, 2 * i1, nz1 - 1) / 3 + a / 6 + b / 2;
            AT(vz1, 2 * i0 + 1, 2 * i1 + 1, nz1) = AT(vz1, 2 * i0 + 1, 2 * i1 + 1, nz1 - 1) / 3 + a / 4 + b / 12 + c / 4 + d / 12;
            AT(vz1, 2 * i0 + 2, 2 * i1 + 1, nz1) = AT(vz1, 2 * i0 + 2, 2 * i1 + 1, nz1 - 1) / 3 + a / 12 + b / 4 + c / 12 + d / 4;
        }
    /* extrapolated x edges (option 1 of the reference) */
    const int e0 = fld_end(vz1, 0);
#pragma omp parallel for schedule(static)
    for (int i1 = 0; i1 < ny2; i1++) {
        int i0 = 0;
        AT(vz1, i0, 2 * i1, nz1) = AT(vz1, i0, 2 * i1, nz1 - 1) * 2 / 3 - AT(vz1, i0 + 1, 2 * i1, nz1 - 1) / 3
                                   + AT(vz2, i0, i1, 0) * 2 / 3;
        AT(vz1, i0, 2 * i1 + 1, nz1) = AT(vz1, i0, 2 * i1 + 1, nz1 - 1) * 2 / 3 - AT(vz1, i0 + 1, 2 * i1, nz1 - 1) / 3
                                       + AT(vz2, i0, i1, 0) / 3 + AT(vz2, i0, i1 + 1, 0) / 3;
        if (!s->ii_xend) continue;
        i0 = e0;  /* a fine index, used on vz2 as the reference does */
        AT(vz1, i0, 2 * i1, nz1) = AT(vz1, i0, 2 * i1, nz1 - 1) * 2 / 3 - AT(vz1, i0 - 1, 2 * i1, nz1 - 1) / 3
                                   + AT(vz2, i0, i1, 0) * 2 / 3;
        AT(vz1, i0, 2 * i1 + 1, nz1) = AT(vz1, i0, 2 * i1 + 1, nz1 - 1) * 2 / 3 - AT(vz1, i0 - 1, 2 * i1, nz1 - 1) / 3
                                       + AT(vz2, i0, i1, 0) / 3 + AT(vz2, i0, i1 + 1, 0) / 3;
    }
}

/* UpdateGhostSXZ mode 0, update_ghost_sxz.cpp:36-84 */
static void ghost_sxz_ii(osim *s, int k)
{
    grid_t *F = &s->g[k], *C = &s->g[k + 1];
    fld *sxz1 = &F->sim[SXZ], *sxz2 = &C->sim[SXZ];
    const int nx2 = C->n[0], ny2 = C->n[1], nz1 = F->n[2];
#pragma omp parallel for schedule(static)
    for (int i1 = 1; i1 < ny2; i1++)
        for (int i0 = 1; i0 < nx2; i0++)
            AT(sxz2, i0, i1, -1) = 0.5 * (AT(sxz1, 2 * i0, 2 * i1, nz1 - 2) + AT(sxz1, 2 * i0, 2 * i1, nz1 - 1));
#pragma omp parallel for schedule(static)
    for (int i1 = 0; i1 <= ny2; i1++)
        for (int i0 = 0; i0 <= nx2; i0++) {
            if (i0 >= 1 && i0 < nx2 && i1 >= 1 && i1 < ny2)
                AT(sxz1, 2 * i0, 2 * i1, nz1) = AT(sxz1, 2 * i0, 2 * i1, nz1 - 1) / 3 + AT(sxz2, i0, i1, 0) * 2 / 3.0;
            if (i0 < nx2 && i1 >= 1 && i1 < ny2)
                AT(sxz1, 2 * i0 + 1, 2 * i1, nz1) = AT(sxz1, 2 * i0 + 1, 2 * i1, nz1 - 1) / 3 + AT(sxz2, i0, i1, 0) / 3
                                                    + AT(sxz2, i0 + 1, i1, 0) / 3;
            if (i0 >= 1 && i0 < nx2 && i1 < ny2)
                AT(sxz1, 2 * i0, 2 * i1 + 1, nz1) = AT(sxz1, 2 * i0, 2 * i1 + 1, nz1 - 1) / 3 + AT(sxz2, i0, i1, 0) / 3
                                                    + AT(sxz2, i0, i1 + 1, 0) / 3;
            if (i0 < nx2 && i1 < ny2)
                AT(sxz1, 2 * i0 + 1, 2 * i1 + 1, nz1) = AT(sxz1, 2 * i0 + 1, 2 * i1 + 1, nz1 - 1) / 3
                                                        + AT(sxz2, i0, i1, 0) / 6 + AT(sxz2, i0, i1 + 1, 0) / 6
                                                        + AT(sxz2, i0 + 1, i1, 0) / 6 + AT(sxz2, i0 + 1, i1 + 1, 0) / 6;
        }
}

/* UpdateGhostSYZ mode 0, update_ghost_syz.cpp:23-150 */
static void ghost_syz_ii(osim *s, int k)
{
    grid_t *F = &s->g[k], *C = &s->g[k + 1];
    fld *syz1 = &F->sim[SYZ], *syz2 = &C->sim[SYZ];
    const int nx2 = C->n[0], ny2 = C->n[1], nz1 = F->n[2];
#pragma omp parallel for schedule(static)
    for (int i1 = 0; i1 <= fld_end(syz2, 1); i1++)
        for (int i0 = 0; i0 <= fld_end(syz2, 0); i0++)
            AT(syz2, i0, i1, -1) = 0.125 * (AT(syz1, 2 * i0, 2 * i1, nz1 - 2) + AT(syz1, 2 * i0 + 1, 2 * i1, nz1 - 2)
                                            + AT(syz1, 2 * i0, 2 * i1 + 1, nz1 - 2) + AT(syz1, 2 * i0 + 1, 2 * i1 + 1, nz1 - 2)
                                            + AT(syz1, 2 * i0, 2 * i1, nz1 - 1) + AT(syz1, 2 * i0 + 1, 2 * i1, nz1 - 1)
                                            + AT(syz1, 2 * i0, 2 * i1 + 1, nz1 - 1) + AT(syz1, 2 * i0 + 1, 2 * i1 + 1, nz1 - 1));
#pragma omp parallel for schedule(static)
    for (int i1 = 0; i1 < ny2 - 1; i1++)
        for (int i0 = 0; i0 < nx2 - 1; i0++) {
            const real a = AT(syz2, i0, i1, 0), b = AT(syz2, i0 + 1, i1, 0);
            const real c = AT(syz2, i0, i1 + 1, 0), d = AT(syz2, i0 + 1, i1 + 1, 0);
            AT(syz1, 2 * i0 + 1, 2 * i1 + 1, nz1) = AT(syz1, 2 * i0 + 1, 2 * i1 + 1, nz1 - 1) / 3 + 0.375 * a + 0.125 * b + 0.125 * c + d / 24;
            AT(syz1, 2 * i0 + 2, 2 * i1 + 1, nz1) = AT(syz1, 2 * i0 + 2, 2 * i1 + 1, nz1 - 1) / 3 + 0.125 * a + 0.375 * b + c / 24 + 0.125 * d;
            AT(syz1, 2 * i0 + 1, 2 * i1 + 2, nz1) = AT(syz1, 2 * i0 + 1, 2 * i1 + 2, nz1 - 1) / 3 + 0.125 * a + b / 24 + 0.375 * c + 0.125 * d;
            AT(syz1, 2 * i0 + 2, 2 * i1 + 2, nz1) = AT(syz1, 2 * i0 + 2, 2 * i1 + 2, nz1 - 1) / 3 + a / 24 + 0.125 * b + 0.125 * c + 0.375 * d;
        }
    /* extrapolated edges (option 1 of the reference) */
    const int xe = fld_end(syz1, 0), ye = fld_end(syz1, 1);
#pragma omp parallel for schedule(static)
    for (int i0 = 0; i0 <= xe; i0++) {
        for (int side = 0; side < 2; side++) {
            const int j = side ? ye : 0, jn = side ? ye - 1 : 1;
            real v;
            if (i0 == 0)
                v = AT(syz1, i0, j, nz1 - 1) - AT(syz1, i0 + 1, j, nz1 - 1) / 3 - AT(syz1, i0, jn, nz1 - 1) / 3
                    + AT(syz2, i0 / 2, j / 2, 0) * 2 / 3;
            else if (i0 == xe)
                v = AT(syz1, i0, j, nz1 - 1) - AT(syz1, i0 - 1, j, nz1 - 1) / 3 - AT(syz1, i0, jn, nz1 - 1) / 3
                    + AT(syz2, i0 / 2, j / 2, 0) * 2 / 3;
            else
                v = AT(syz1, i0, j, nz1 - 1) / 3 + AT(syz1, i0 - 1 + 2 * (i0 % 2), j, nz1 - 1) / 3
                    - AT(syz1, i0, jn, nz1 - 1) / 3 + AT(syz2, i0 / 2, j / 2, 0) * 2 / 3;
            AT(syz1, i0, j, nz1) = v;
        }
    }
#pragma omp parallel for schedule(static)
    for (int i1 = 1; i1 < ye; i1++) {
        AT(syz1, 0, i1, nz1) = AT(syz1, 0, i1, nz1 - 1) / 3 + AT(syz1, 0, i1 - 1 + 2 * (i1 % 2), nz1 - 1) / 3
                               - AT(syz1, 1, i1, nz1 - 1) / 3 + AT(syz2, 0, i1 / 2, 0) * 2 / 3;
        AT(syz1, xe, i1, nz1) = AT(syz1, xe, i1, nz1 - 1) / 3 + AT(syz1, xe, i1 - 1 + 2 * (i1 % 2), nz1 - 1) / 3
                                - AT(syz1, xe - 1, i1, nz1 - 1) / 3 + AT(syz2, xe / 2, i1 / 2, 0) * 2 / 3;
    }
}

/* ------------------------------------------------------------------ */
/* Interface: reference ratio-2 formulas                              */

/* UpdateGhostVZ mode 1, update_ghost_vz.cpp:124-250 */
static void ghost_vz_ref(osim *s, int k)
{
    grid_t *F = &s->g[k], *C = &s->g[k + 1];
    fld *vx1 = &F->sim[VX], *vy1 = &F->sim[VY], *vz1 = &F->sim[VZ], *szz1 = &F->sim[SZZ];
    fld *vx2 = &C->sim[VX], *vy2 = &C->sim[VY], *vz2 = &C->sim[VZ], *szz2 = &C->sim[SZZ];
    fld *c33_1 = &F->st[C33], *c33_2 = &C->st[C33], *c23_1 = &F->st[C23], *c23_2 = &C->st[C23];
    fld *c13_1 = &F->st[C13], *c13_2 = &C->st[C13];
    real(*k1)[6] = F->coeff, (*k2)[6] = C->coeff;
    int nx2 = C->n[0], ny2 = C->n[1], nz1 = F->n[2];
    int nl = ny2 - 1;
#pragma omp parallel
    {
        real *A0 = (real *)malloc(sizeof(real) * (nl > 0 ? nl : 1));
        real *A1 = (real *)malloc(sizeof(real) * (nl > 0 ? nl : 1));
        real *A2 = (real *)malloc(sizeof(real) * (nl > 0 ? nl : 1));
        real *Bv = (real *)malloc(sizeof(real) * (nl > 0 ? nl : 1));
#pragma omp for schedule(static)
        for (int i0 = 0; i0 < nx2; i0++) {
            for (int i1 = 0; i1 < ny2 - 1; i1++) {
                A0[i1] = 0.5 * k1[2][0] * (AT(c33_1, 2 * i0, 2 * i1 + 1, nz1) + AT(c33_1, 2 * i0 + 1, 2 * i1 + 1, nz1));
                A2[i1] = 0.5 * k1[2][0] * (AT(c33_1, 2 * i0, 2 * i1 + 3, nz1) + AT(c33_1, 2 * i0 + 1, 2 * i1 + 3, nz1));
                A1[i1] = 8 * k2[2][0] * AT(c33_2, i0, i1 + 1, 0)
                    + 2 * k1[2][0] * (AT(c33_1, 2 * i0, 2 * i1 + 2, nz1) + AT(c33_1, 2 * i0 + 1, 2 * i1 + 2, nz1))
                    + A0[i1] + A2[i1];
                Bv[i1] = 8 * (AT(szz2, i0, i1 + 1, 0)
                              + AT(c13_2, i0, i1 + 1, 0) * k2[0][0] * (AT(vx2, i0 + 1, i1 + 1, 0) - AT(vx2, i0, i1 + 1, 0))
                              + AT(c23_2, i0, i1 + 1, 0) * k2[1][0] * (AT(vy2, i0, i1 + 1, 0) - AT(vy2, i0, i1, 0))
                              + AT(c33_2, i0, i1 + 1, 0) * k2[2][0] * AT(vz2, i0, i1 + 1, 0))
                    - 2 * (AT(szz1, 2 * i0, 2 * i1 + 2, nz1)
                           + AT(c13_1, 2 * i0, 2 * i1 + 2, nz1) * k1[0][0] * (AT(vx1, 2 * i0 + 1, 2 * i1 + 2, nz1) - AT(vx1, 2 * i0, 2 * i1 + 2, nz1))
                           + AT(c23_1, 2 * i0, 2 * i1 + 2, nz1) * k1[1][0] * (AT(vy1, 2 * i0, 2 * i1 + 2, nz1) - AT(vy1, 2 * i0, 2 * i1 + 1, nz1))
                           + AT(c33_1, 2 * i0, 2 * i1 + 2, nz1) * k1[2][0] * (AT(vz2, i0, i1 + 1, 0) - 2 * AT(vz1, 2 * i0, 2 * i1 + 2, nz1 - 1)))
                    - 2 * (AT(szz1, 2 * i0 + 1, 2 * i1 + 2, nz1)
                           + AT(c13_1, 2 * i0 + 1, 2 * i1 + 2, nz1) * k1[0][0] * (AT(vx1, 2 * i0 + 2, 2 * i1 + 2, nz1) - AT(vx1, 2 * i0 + 1, 2 * i1 + 2, nz1))
                           + AT(c23_1, 2 * i0 + 1, 2 * i1 + 2, nz1) * k1[1][0] * (AT(vy1, 2 * i0 + 1, 2 * i1 + 2, nz1) - AT(vy1, 2 * i0 + 1, 2 * i1 + 1, nz1))
                           + AT(c33_1, 2 * i0 + 1, 2 * i1 + 2, nz1) * k1[2][0] * (AT(vz2, i0, i1 + 1, 0) - 2 * AT(vz1, 2 * i0 + 1, 2 * i1 + 2, nz1 - 1)))
                    - (AT(szz1, 2 * i0, 2 * i1 + 1, nz1)
                       + AT(c13_1, 2 * i0, 2 * i1 + 1, nz1) * k1[0][0] * (AT(vx1, 2 * i0 + 1, 2 * i1 + 1, nz1) - AT(vx1, 2 * i0, 2 * i1 + 1, nz1))
                       + AT(c23_1, 2 * i0, 2 * i1 + 1, nz1) * k1[1][0] * (AT(vy1, 2 * i0, 2 * i1 + 1, nz1) - AT(vy1, 2 * i0, 2 * i1, nz1))
                       + AT(c33_1, 2 * i0, 2 * i1 + 1, nz1) * k1[2][0] * (0.5 * (AT(vz2, i0, i1 + 1, 0) + AT(vz2, i0, i1, 0)) - 2 * AT(vz1, 2 * i0, 2 * i1 + 1, nz1 - 1)))
                    - (AT(szz1, 2 * i0, 2 * i1 + 3, nz1)
                       + AT(c13_1, 2 * i0, 2 * i1 + 3, nz1) * k1[0][0] * (AT(vx1, 2 * i0 + 1, 2 * i1 + 3, nz1) - AT(vx1, 2 * i0, 2 * i1 + 3, nz1))
                       + AT(c23_1, 2 * i0, 2 * i1 + 3, nz1) * k1[1][0] * (AT(vy1, 2 * i0, 2 * i1 + 3, nz1) - AT(vy1, 2 * i0, 2 * i1 + 2, nz1))
                       + AT(c33_1, 2 * i0, 2 * i1 + 3, nz1) * k1[2][0] * (0.5 * (AT(vz2, i0, i1 + 2, 0) + AT(vz2, i0, i1 + 1, 0)) - 2 * AT(vz1, 2 * i0, 2 * i1 + 3, nz1 - 1)))
                    - (AT(szz1, 2 * i0 + 1, 2 * i1 + 1, nz1)
                       + AT(c13_1, 2 * i0 + 1, 2 * i1 + 1, nz1) * k1[0][0] * (AT(vx1, 2 * i0 + 2, 2 * i1 + 1, nz1) - AT(vx1, 2 * i0 + 1, 2 * i1 + 1, nz1))
                       + AT(c23_1, 2 * i0 + 1, 2 * i1 + 1, nz1) * k1[1][0] * (AT(vy1, 2 * i0 + 1, 2 * i1 + 1, nz1) - AT(vy1, 2 * i0 + 1, 2 * i1, nz1))
                       + AT(c33_1, 2 * i0 + 1, 2 * i1 + 1, nz1) * k1[2][0] * (0.5 * (AT(vz2, i0, i1 + 1, 0) + AT(vz2, i0, i1, 0)) - 2 * AT(vz1, 2 * i0 + 1, 2 * i1 + 1, nz1 - 1)))
                    - (AT(szz1, 2 * i0 + 1, 2 * i1 + 3, nz1)
                       + AT(c13_1, 2 * i0 + 1, 2 * i1 + 3, nz1) * k1[0][0] * (AT(vx1, 2 * i0 + 2, 2 * i1 + 3, nz1) - AT(vx1, 2 * i0 + 1, 2 * i1 + 3, nz1))
                       + AT(c23_1, 2 * i0 + 1, 2 * i1 + 3, nz1) * k1[1][0] * (AT(vy1, 2 * i0 + 1, 2 * i1 + 3, nz1) - AT(vy1, 2 * i0 + 1, 2 * i1 + 2, nz1))
                       + AT(c33_1, 2 * i0 + 1, 2 * i1 + 3, nz1) * k1[2][0] * (0.5 * (AT(vz2, i0, i1 + 2, 0) + AT(vz2, i0, i1 + 1, 0)) - 2 * AT(vz1, 2 * i0 + 1, 2 * i1 + 3, nz1 - 1)));
            }
            /* Thomas, update_ghost_vz.cpp:211-220 */
            for (int i1 = 1; i1 < ny2 - 1; i1++) {
                real m = A0[i1] / A1[i1 - 1];
                A1[i1] -= m * A2[i1 - 1];
                Bv[i1] -= m * Bv[i1 - 1];
            }
            AT(vz2, i0, ny2 - 1, -1) = Bv[ny2 - 2] / A1[ny2 - 2];
            for (int i1 = ny2 - 2; i1 >= 1; i1--)
                AT(vz2, i0, i1, -1) = (Bv[i1 - 1] - A2[i1 - 1] * AT(vz2, i0, i1 + 1, -1)) / A1[i1 - 1];
        }
        free(A0); free(A1); free(A2); free(Bv);
    }
    /* fine ghosts, update_ghost_vz.cpp:231-248 */
#pragma omp parallel for schedule(static)
    for (int i1 = 1; i1 < ny2; i1++)
        for (int i0 = 0; i0 < nx2; i0++) {
            real tmp = AT(vz2, i0, i1, 0) + AT(vz2, i0, i1, -1);
            AT(vz1, 2 * i0, 2 * i1, nz1) = -AT(vz1, 2 * i0, 2 * i1, nz1 - 1) + tmp;
            AT(vz1, 2 * i0 + 1, 2 * i1, nz1) = -AT(vz1, 2 * i0 + 1, 2 * i1, nz1 - 1) + tmp;
        }
#pragma omp parallel for schedule(static)
    for (int i1 = 0; i1 < ny2; i1++)
        for (int i0 = 0; i0 < nx2; i0++) {
            real tmp = 0.5 * (AT(vz2, i0, i1, 0) + AT(vz2, i0, i1, -1) + AT(vz2, i0, i1 + 1, 0) + AT(vz2, i0, i1 + 1, -1));
            AT(vz1, 2 * i0, 2 * i1 + 1, nz1) = -AT(vz1, 2 * i0, 2 * i1 + 1, nz1 - 1) + tmp;
            AT(vz1, 2 * i0 + 1, 2 * i1 + 1, nz1) = -AT(vz1, 2 * i0 + 1, 2 * i1 + 1, nz1 - 1) + tmp;
        }
}

/* AvgAbs over an n0 x n1 array, array.cpp:172-210 (fixed order) */
static real avg_abs(const real *B, long n)
{
    real sum = 0;
    for (long q = 0; q < n; q++) sum += fabs(B[q]);
    return sum / n;
}

/* Jacobi2D + Norm2D, update_ghost_sxz.cpp:314-388.
 * X, Xt: (n0+2) x (n1+2), index (i0, i1) -> i1*(n0+2)+i0, zero border.
 * A: band k at A[k*n0*n1 + (i1-1)*n0 + (i0-1)].  Returns iterations.
 *
 * Exit rule: the reference's  norm(X_new - X_old) <= tol,  plus a
 * floating-point stagnation exit that the reference lacks: once the update
 * norm is at the rounding floor (<= JACOBI_FLOOR * ||X_new||) and no longer
 * decreases, further sweeps only cycle through rounding noise.  With
 * stresses in SI units the reference's absolute 1e-12 is below that floor and
 * its do/while would not end.  A hard cap reports non-convergence. */
#ifdef ORACLE_FP32
#define JACOBI_FLOOR (1024 * 1.1920928955078125e-07)
#else
#define JACOBI_FLOOR (1024 * 2.220446049250313e-16)
#endif
static int jacobi2d(real *X, const real *A, const real *B, int n0, int n1, real tol, real *Xt, int *capped)
{
    long w = n0 + 2, nn = (long)n0 * n1;
    int it = 0;
    real *rowsum = (real *)malloc(sizeof(real) * (n1 + 1));
    real *rowx = (real *)malloc(sizeof(real) * (n1 + 1));
    real norm, prev = 0;
    for (;;) {
#pragma omp parallel for schedule(static)
        for (int i1 = 1; i1 <= n1; i1++)
            for (int i0 = 1; i0 <= n0; i0++) Xt[i1 * w + i0] = X[i1 * w + i0];
#pragma omp parallel for schedule(static)
        for (int i1 = 1; i1 <= n1; i1++)
            for (int i0 = 1; i0 <= n0; i0++) {
                long q = (long)(i1 - 1) * n0 + (i0 - 1);
                X[i1 * w + i0] = (B[q]
                                  - A[0 * nn + q] * Xt[(i1 - 1) * w + (i0 - 1)]
                                  - A[1 * nn + q] * Xt[(i1) * w + (i0 - 1)]
                                  - A[2 * nn + q] * Xt[(i1 + 1) * w + (i0 - 1)]
                                  - A[3 * nn + q] * Xt[(i1 - 1) * w + (i0)]
                                  - A[5 * nn + q] * Xt[(i1 + 1) * w + (i0)]
                                  - A[6 * nn + q] * Xt[(i1 - 1) * w + (i0 + 1)]
                                  - A[7 * nn + q] * Xt[(i1) * w + (i0 + 1)]
                                  - A[8 * nn + q] * Xt[(i1 + 1) * w + (i0 + 1)]) / A[4 * nn + q];
            }
        it++;
#pragma omp parallel for schedule(static)
        for (int i1 = 1; i1 <= n1; i1++) {
            real acc = 0, ax = 0;
            for (int i0 = 1; i0 <= n0; i0++) {
                real d = X[i1 * w + i0] - Xt[i1 * w + i0];
                acc += d * d;
                ax += X[i1 * w + i0] * X[i1 * w + i0];
            }
            rowsum[i1] = acc;
            rowx[i1] = ax;
        }
        real res = 0, rx = 0;
        for (int i1 = 1; i1 <= n1; i1++) { res += rowsum[i1]; rx += rowx[i1]; }
        norm = sqrt(res);
        if (!(norm > tol)) break;
        if (it > 1 && norm >= prev && norm <= JACOBI_FLOOR * sqrt(rx)) break;
        if (it >= JACOBI_CAP) { *capped = 1; break; }
        prev = norm;
    }
    free(rowsum);
    free(rowx);
    return it;
}

/* UpdateGhostSXZ mode 1, update_ghost_sxz.cpp:86-297 */
static void ghost_sxz_ref(osim *s, int k)
{
    grid_t *F = &s->g[k], *C = &s->g[k + 1];
    iface_t *I = &s->itf[k];
    fld *vx1 = &F->sim[VX], *vx2 = &C->sim[VX], *sxx1 = &F->sim[SXX], *sxx2 = &C->sim[SXX];
    fld *sxz1 = &F->sim[SXZ], *sxz2 = &C->sim[SXZ], *sxy1 = &F->sim[SXY], *sxy2 = &C->sim[SXY];
    fld *b1 = &F->buoy, *b2 = &C->buoy;
    real(*k1)[6] = F->coeff, (*k2)[6] = C->coeff;
    int nx2 = C->n[0], ny2 = C->n[1], nz1 = F->n[2];
    int n0 = nx2 - 1, n1 = ny2 - 1;
    long nn = (long)n0 * n1;
    real *A = I->A, *B = I->B;
    real c1z = k1[2][0];
#define AB(kk, i0, i1) A[(kk) * nn + (long)(i1) * n0 + (i0)]
#pragma omp parallel for schedule(static)
    for (int i1 = 0; i1 < ny2 - 1; i1++)
        for (int i0 = 0; i0 < nx2 - 1; i0++) {
            AB(0, i0, i1) = (i1 != 0) * (i0 != 0) * 0.25 * AT(b1, 2 * i0 + 1, 2 * i1 + 1, nz1) * c1z;
            AB(1, i0, i1) = (i0 != 0) * c1z * (AT(b1, 2 * i0 + 1, 2 * i1 + 2, nz1) + 0.25 * AT(b1, 2 * i0 + 1, 2 * i1 + 3, nz1)
                                               + 0.25 * AT(b1, 2 * i0 + 1, 2 * i1 + 1, nz1));
            AB(2, i0, i1) = (i1 != ny2 - 2) * (i0 != 0) * 0.25 * AT(b1, 2 * i0 + 1, 2 * i1 + 3, nz1) * c1z;
            AB(3, i0, i1) = (i1 != 0) * c1z * (AT(b1, 2 * i0 + 2, 2 * i1 + 1, nz1) + 0.25 * AT(b1, 2 * i0 + 3, 2 * i1 + 1, nz1)
                                               + 0.25 * AT(b1, 2 * i0 + 1, 2 * i1 + 1, nz1));
            AB(4, i0, i1) = 16 * AT(b2, i0 + 1, i1 + 1, 0) * k2[2][0]
                + c1z * (4 * AT(b1, 2 * i0 + 2, 2 * i1 + 2, nz1) + AT(b1, 2 * i0 + 3, 2 * i1 + 2, nz1) + AT(b1, 2 * i0 + 1, 2 * i1 + 2, nz1)
                         + AT(b1, 2 * i0 + 2, 2 * i1 + 3, nz1) + AT(b1, 2 * i0 + 2, 2 * i1 + 1, nz1)
                         + 0.25 * AT(b1, 2 * i0 + 3, 2 * i1 + 3, nz1) + 0.25 * AT(b1, 2 * i0 + 1, 2 * i1 + 3, nz1)
                         + 0.25 * AT(b1, 2 * i0 + 3, 2 * i1 + 1, nz1) + 0.25 * AT(b1, 2 * i0 + 1, 2 * i1 + 1, nz1));
            AB(5, i0, i1) = (i1 != ny2 - 2) * c1z * (AT(b1, 2 * i0 + 2, 2 * i1 + 3, nz1) + 0.25 * AT(b1, 2 * i0 + 3, 2 * i1 + 3, nz1)
                                                     + 0.25 * AT(b1, 2 * i0 + 1, 2 * i1 + 3, nz1));
            AB(6, i0, i1) = (i1 != 0) * (i0 != nx2 - 2) * 0.25 * AT(b1, 2 * i0 + 3, 2 * i1 + 1, nz1) * c1z;
            AB(7, i0, i1) = (i0 != nx2 - 2) * c1z * (AT(b1, 2 * i0 + 3, 2 * i1 + 2, nz1) + 0.25 * AT(b1, 2 * i0 + 3, 2 * i1 + 3, nz1)
                                                     + 0.25 * AT(b1, 2 * i0 + 3, 2 * i1 + 1, nz1));
            AB(8, i0, i1) = (i1 != ny2 - 2) * (i0 != nx2 - 2) * c1z * (0.25 * AT(b1, 2 * i0 + 3, 2 * i1 + 3, nz1));
        }
#undef AB
    /* B, update_ghost_sxz.cpp:139-226.  KF: fine-node term with interpolated
     * coarse sxz(.,.,0) value S; KC: coarse-node term. */
#define KF(i, j, S)                                                                                    \
    (AT(vx1, i, j, nz1) + AT(b1, i, j, nz1) * (k1[0][0] * (AT(sxx1, i, j, nz1) - AT(sxx1, (i) - 1, j, nz1)) \
                                              + k1[1][0] * (AT(sxy1, i, j, nz1) - AT(sxy1, i, (j) - 1, nz1)) \
                                              + k1[2][0] * ((S) - 2 * AT(sxz1, i, j, nz1 - 1))))
#define SC(i, j) AT(sxz2, i, j, 0)
#pragma omp parallel for schedule(static)
    for (int i1 = 0; i1 < ny2 - 1; i1++)
        for (int i0 = 0; i0 < nx2 - 1; i0++) {
            int mx0 = (i0 != 0), mx1 = (i0 != nx2 - 2), my0 = (i1 != 0), my1 = (i1 != ny2 - 2);
            B[(long)i1 * n0 + i0] =
                16 * (AT(vx2, i0 + 1, i1 + 1, 0)
                      + AT(b2, i0 + 1, i1 + 1, 0) * (k2[0][0] * (AT(sxx2, i0 + 1, i1 + 1, 0) - AT(sxx2, i0, i1 + 1, 0))
                                                     + k2[1][0] * (AT(sxy2, i0 + 1, i1 + 1, 0) - AT(sxy2, i0 + 1, i1, 0))
                                                     + k2[2][0] * AT(sxz2, i0 + 1, i1 + 1, 0)))
                - 4 * KF(2 * i0 + 2, 2 * i1 + 2, SC(i0 + 1, i1 + 1))
                - 2 * KF(2 * i0 + 3, 2 * i1 + 2, 0.5 * (SC(i0 + 1, i1 + 1) + mx1 * SC(i0 + 2, i1 + 1)))
                - 2 * KF(2 * i0 + 1, 2 * i1 + 2, 0.5 * (mx0 * SC(i0, i1 + 1) + SC(i0 + 1, i1 + 1)))
                - 2 * KF(2 * i0 + 2, 2 * i1 + 3, 0.5 * (SC(i0 + 1, i1 + 1) + my1 * SC(i0 + 1, i1 + 2)))
                - 2 * KF(2 * i0 + 2, 2 * i1 + 1, 0.5 * (SC(i0 + 1, i1 + 1) + my0 * SC(i0 + 1, i1)))
                - KF(2 * i0 + 3, 2 * i1 + 3, 0.25 * (SC(i0 + 1, i1 + 1) + mx1 * SC(i0 + 2, i1 + 1) + my1 * SC(i0 + 1, i1 + 2)
                                                     + (my1 * mx1) * SC(i0 + 2, i1 + 2)))
                - KF(2 * i0 + 1, 2 * i1 + 3, 0.25 * (mx0 * SC(i0, i1 + 1) + SC(i0 + 1, i1 + 1) + (mx0 * my1) * SC(i0, i1 + 2)
                                                     + my1 * SC(i0 + 1, i1 + 2)))
                - KF(2 * i0 + 3, 2 * i1 + 1, 0.25 * (my0 * SC(i0 + 1, i1) + (my0 * mx1) * SC(i0 + 2, i1) + SC(i0 + 1, i1 + 1)
                                                     + mx1 * SC(i0 + 2, i1 + 1)))
                - KF(2 * i0 + 1, 2 * i1 + 1, 0.25 * ((mx0 * my0) * SC(i0, i1) + my0 * SC(i0 + 1, i1) + mx0 * SC(i0, i1 + 1)
                                                     + SC(i0 + 1, i1 + 1)));
        }
#undef KF
#undef SC
    real tol = JACOBI_TOL(avg_abs(B, nn));
    int capped = 0;
    I->iters_last = jacobi2d(I->gx, A, B, n0, n1, tol, I->gtmp, &capped);
    I->iters_total += I->iters_last;
    I->capped |= capped;
    long w = nx2 + 1;
    real *X = I->gx;
#pragma omp parallel for schedule(static)
    for (int i1 = 1; i1 < ny2; i1++)
        for (int i0 = 1; i0 < nx2; i0++) AT(sxz2, i0, i1, -1) = X[i1 * w + i0];
#define T2(i, j) (AT(sxz2, i, j, -1) + AT(sxz2, i, j, 0))
#pragma omp parallel for schedule(static)
    for (int i1 = 1; i1 < ny2; i1++)
        for (int i0 = 1; i0 < nx2; i0++)
            AT(sxz1, 2 * i0, 2 * i1, nz1) = -AT(sxz1, 2 * i0, 2 * i1, nz1 - 1) + AT(sxz2, i0, i1, -1) + AT(sxz2, i0, i1, 0);
#pragma omp parallel for schedule(static)
    for (int i1 = 1; i1 < ny2; i1++)
        for (int i0 = 0; i0 < nx2; i0++)
            AT(sxz1, 2 * i0 + 1, 2 * i1, nz1) = -AT(sxz1, 2 * i0 + 1, 2 * i1, nz1 - 1)
                + 0.5 * ((i0 != 0) * T2(i0, i1) + (i0 != nx2 - 1) * T2(i0 + 1, i1));
#pragma omp parallel for schedule(static)
    for (int i1 = 0; i1 < ny2; i1++)
        for (int i0 = 1; i0 < nx2; i0++)
            AT(sxz1, 2 * i0, 2 * i1 + 1, nz1) = -AT(sxz1, 2 * i0, 2 * i1 + 1, nz1 - 1)
                + 0.5 * ((i1 != 0) * T2(i0, i1) + (i1 != ny2 - 1) * T2(i0, i1 + 1));
#pragma omp parallel for schedule(static)
    for (int i1 = 0; i1 < ny2; i1++)
        for (int i0 = 0; i0 < nx2; i0++)
            AT(sxz1, 2 * i0 + 1, 2 * i1 + 1, nz1) = -AT(sxz1, 2 * i0 + 1, 2 * i1 + 1, nz1 - 1)
                + 0.25 * ((i1 != 0) * ((i0 != 0) * T2(i0, i1) + (i0 != nx2 - 1) * T2(i0 + 1, i1))
                          + (i1 != ny2 - 1) * ((i0 != 0) * T2(i0, i1 + 1) + (i0 != nx2 - 1) * T2(i0 + 1, i1 + 1)));
#undef T2
}

/* UpdateGhostSYZ mode 1, update_ghost_syz.cpp:151-240 */
static void ghost_syz_ref(osim *s, int k)
{
    grid_t *F = &s->g[k], *C = &s->g[k + 1];
    fld *syy1 = &F->sim[SYY], *syy2 = &C->sim[SYY], *syz1 = &F->sim[SYZ], *syz2 = &C->sim[SYZ];
    fld *sxy1 = &F->sim[SXY], *sxy2 = &C->sim[SXY], *vy1 = &F->sim[VY], *vy2 = &C->sim[VY];
    fld *b1 = &F->buoy, *b2 = &C->buoy;
    real(*k1)[6] = F->coeff, (*k2)[6] = C->coeff;
    int nx2 = C->n[0], ny2 = C->n[1], nz1 = F->n[2];
#pragma omp parallel for schedule(static)
    for (int i1 = 0; i1 <= ny2 - 1; i1++)
        for (int i0 = 0; i0 <= nx2 - 1; i0++) {
            real c0 = 0.25 * (AT(b2, i0, i1, 0) + AT(b2, i0 + 1, i1, 0) + AT(b2, i0, i1 + 1, 0) + AT(b2, i0 + 1, i1 + 1, 0));
            real c1 = 0.25 * (AT(b1, 2 * i0, 2 * i1, nz1) + AT(b1, 2 * i0 + 1, 2 * i1, nz1) + AT(b1, 2 * i0, 2 * i1 + 1, nz1) + AT(b1, 2 * i0 + 1, 2 * i1 + 1, nz1));
            real c2 = 0.25 * (AT(b1, 2 * i0 + 1, 2 * i1, nz1) + AT(b1, 2 * i0 + 2, 2 * i1, nz1) + AT(b1, 2 * i0 + 1, 2 * i1 + 1, nz1) + AT(b1, 2 * i0 + 2, 2 * i1 + 1, nz1));
            real c3 = 0.25 * (AT(b1, 2 * i0, 2 * i1 + 1, nz1) + AT(b1, 2 * i0 + 1, 2 * i1 + 1, nz1) + AT(b1, 2 * i0, 2 * i1 + 2, nz1) + AT(b1, 2 * i0 + 1, 2 * i1 + 2, nz1));
            real c4 = 0.25 * (AT(b1, 2 * i0 + 1, 2 * i1 + 1, nz1) + AT(b1, 2 * i0 + 2, 2 * i1 + 1, nz1) + AT(b1, 2 * i0 + 1, 2 * i1 + 2, nz1) + AT(b1, 2 * i0 + 2, 2 * i1 + 2, nz1));
#define FY(c, i, j)                                                                                          \
    (AT(vy1, i, j, nz1) + (c) * (k1[0][0] * (AT(sxy1, (i) + 1, j, nz1) - AT(sxy1, i, j, nz1))                     \
                                 + k1[1][0] * (AT(syy1, i, (j) + 1, nz1) - AT(syy1, i, j, nz1))                     \
                                 + k1[2][0] * (AT(syz2, i0, i1, 0) - 2 * AT(syz1, i, j, nz1 - 1))))
            AT(syz2, i0, i1, -1) = 4 * (AT(vy2, i0, i1, 0)
                                        + c0 * (k2[0][0] * (AT(sxy2, i0 + 1, i1, 0) - AT(sxy2, i0, i1, 0))
                                                + k2[1][0] * (AT(syy2, i0, i1 + 1, 0) - AT(syy2, i0, i1, 0))
                                                + k2[2][0] * AT(syz2, i0, i1, 0)))
                - FY(c1, 2 * i0, 2 * i1) - FY(c2, 2 * i0 + 1, 2 * i1) - FY(c3, 2 * i0, 2 * i1 + 1) - FY(c4, 2 * i0 + 1, 2 * i1 + 1);
#undef FY
            AT(syz2, i0, i1, -1) /= (4 * c0 * k2[2][0] + c1 * k1[2][0] + c2 * k1[2][0] + c3 * k1[2][0] + c4 * k1[2][0]);
        }
#pragma omp parallel for schedule(static)
    for (int i1 = 0; i1 <= ny2 - 1; i1++)
        for (int i0 = 0; i0 <= nx2 - 1; i0++) {
            real tmp = AT(syz2, i0, i1, -1) + AT(syz2, i0, i1, 0);
            AT(syz1, 2 * i0, 2 * i1, nz1) = -AT(syz1, 2 * i0, 2 * i1, nz1 - 1) + tmp;
            AT(syz1, 2 * i0 + 1, 2 * i1, nz1) = -AT(syz1, 2 * i0 + 1, 2 * i1, nz1 - 1) + tmp;
            AT(syz1, 2 * i0, 2 * i1 + 1, nz1) = -AT(syz1, 2 * i0, 2 * i1 + 1, nz1 - 1) + tmp;
            AT(syz1, 2 * i0 + 1, 2 * i1 + 1, nz1) = -AT(syz1, 2 * i0 + 1, 2 * i1 + 1, nz1 - 1) + tmp;
        }
}

/* ------------------------------------------------------------------ */
/* Interface: ratio-generic formulation (derived; see header).          */
/* Node-centred direction: hat weight of fine offset a (|a| < r).       */
static inline real hatw(int a, int r) { return (real)(r - abs(a)) / (real)r; }

/* vz-szz pair: per coarse x-line a tridiagonal system in J. */
static void ghost_vz_gen(osim *s, int k)
{
    grid_t *F = &s->g[k], *C = &s->g[k + 1];
    int r = s->itf[k].r;
    fld *vx1 = &F->sim[VX], *vy1 = &F->sim[VY], *vz1 = &F->sim[VZ], *szz1 = &F->sim[SZZ];
    fld *vx2 = &C->sim[VX], *vy2 = &C->sim[VY], *vz2 = &C->sim[VZ], *szz2 = &C->sim[SZZ];
    fld *c33_1 = &F->st[C33], *c33_2 = &C->st[C33], *c23_1 = &F->st[C23], *c23_2 = &C->st[C23];
    fld *c13_1 = &F->st[C13], *c13_2 = &C->st[C13];
    real(*k1)[6] = F->coeff, (*k2)[6] = C->coeff;
    int nx2 = C->n[0], ny2 = C->n[1], nz1 = F->n[2];
    int nl = ny2 - 1;
    real r3 = (real)(r * r * r);
#pragma omp parallel
    {
        real *L = (real *)malloc(sizeof(real) * (nl + 1));
        real *D = (real *)malloc(sizeof(real) * (nl + 1));
        real *U = (real *)malloc(sizeof(real) * (nl + 1));
        real *Bv = (real *)malloc(sizeof(real) * (nl + 1));
        real *X = (real *)malloc(sizeof(real) * (nl + 2));
#pragma omp for schedule(static)
        for (int I = 0; I < nx2; I++) {
            for (int J = 1; J <= ny2 - 1; J++) {
                int q = J - 1;
                real d = 0, lo = 0, up = 0, fs = 0;
                for (int b = -(r - 1); b <= r - 1; b++) {
                    int j = r * J + b;
                    real wj = (real)(r - abs(b));
                    real pj = hatw(b, r);
                    /* interpolated coarse vz(.,.,0) on fine row j (hat in y) */
                    int Jl = (b >= 0) ? J : J - 1;
                    int bb = j - r * Jl;
                    real pv = hatw(bb, r) * AT(vz2, I, Jl, 0) + hatw(bb - r, r) * AT(vz2, I, Jl + 1, 0);
                    for (int a = 0; a < r; a++) {
                        int i = r * I + a;
                        real cc = AT(c33_1, i, j, nz1);
                        d += wj * pj * cc;
                        if (b < 0) lo += wj * hatw(b + r, r) * cc;
                        if (b > 0) up += wj * hatw(b - r, r) * cc;
                        real kf = AT(szz1, i, j, nz1)
                            + AT(c13_1, i, j, nz1) * k1[0][0] * (AT(vx1, i + 1, j, nz1) - AT(vx1, i, j, nz1))
                            + AT(c23_1, i, j, nz1) * k1[1][0] * (AT(vy1, i, j, nz1) - AT(vy1, i, j - 1, nz1))
                            + cc * k1[2][0] * (pv - 2 * AT(vz1, i, j, nz1 - 1));
                        fs += wj * kf;
                    }
                }
                real kc = AT(szz2, I, J, 0)
                    + AT(c13_2, I, J, 0) * k2[0][0] * (AT(vx2, I + 1, J, 0) - AT(vx2, I, J, 0))
                    + AT(c23_2, I, J, 0) * k2[1][0] * (AT(vy2, I, J, 0) - AT(vy2, I, J - 1, 0))
                    + AT(c33_2, I, J, 0) * k2[2][0] * AT(vz2, I, J, 0);
                D[q] = r3 * k2[2][0] * AT(c33_2, I, J, 0) + k1[2][0] * d;
                L[q] = k1[2][0] * lo;
                U[q] = k1[2][0] * up;
                Bv[q] = r3 * kc - fs;
            }
            for (int q = 1; q < nl; q++) {
                real m = L[q] / D[q - 1];
                D[q] -= m * U[q - 1];
                Bv[q] -= m * Bv[q - 1];
            }
            X[nl - 1] = Bv[nl - 1] / D[nl - 1];
            for (int q = nl - 2; q >= 0; q--) X[q] = (Bv[q] - U[q] * X[q + 1]) / D[q];
            for (int J = 1; J <= ny2 - 1; J++) AT(vz2, I, J, -1) = X[J - 1];
        }
        free(L); free(D); free(U); free(Bv); free(X);
    }
    /* fine ghosts: injection in x, hat in y */
#pragma omp parallel for schedule(static)
    for (int J = 0; J < ny2; J++)
        for (int I = 0; I < nx2; I++) {
            real t0 = AT(vz2, I, J, 0) + AT(vz2, I, J, -1);
            real t1 = AT(vz2, I, J + 1, 0) + AT(vz2, I, J + 1, -1);
            for (int b = (J == 0) ? 1 : 0; b < r; b++) {
                real tmp = (b == 0) ? t0 : (hatw(b, r) * t0 + hatw(b - r, r) * t1);
                for (int a = 0; a < r; a++)
                    AT(vz1, r * I + a, r * J + b, nz1) = -AT(vz1, r * I + a, r * J + b, nz1 - 1) + tmp;
            }
        }
}

/* vx-sxz pair: 9-point system on interior coarse nodes, Jacobi. */
static void ghost_sxz_gen(osim *s, int k)
{
    grid_t *F = &s->g[k], *C = &s->g[k + 1];
    iface_t *It = &s->itf[k];
    int r = It->r;
    fld *vx1 = &F->sim[VX], *vx2 = &C->sim[VX], *sxx1 = &F->sim[SXX], *sxx2 = &C->sim[SXX];
    fld *sxz1 = &F->sim[SXZ], *sxz2 = &C->sim[SXZ], *sxy1 = &F->sim[SXY], *sxy2 = &C->sim[SXY];
    fld *b1 = &F->buoy, *b2 = &C->buoy;
    real(*k1)[6] = F->coeff, (*k2)[6] = C->coeff;
    int nx2 = C->n[0], ny2 = C->n[1], nz1 = F->n[2];
    int n0 = nx2 - 1, n1 = ny2 - 1;
    long nn = (long)n0 * n1;
    real *A = It->A, *B = It->B;
    real r2 = (real)(r * r), r4 = r2 * r2;
#pragma omp parallel for schedule(static)
    for (int J = 1; J <= ny2 - 1; J++)
        for (int I = 1; I <= nx2 - 1; I++) {
            long q = (long)(J - 1) * n0 + (I - 1);
            real band[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
            real fs = 0;
            for (int b = -(r - 1); b <= r - 1; b++)
                for (int a = -(r - 1); a <= r - 1; a++) {
                    int i = r * I + a, j = r * J + b;
                    real w = (real)((r - abs(a)) * (r - abs(b)));
                    real bb = AT(b1, i, j, nz1);
                    /* prolongation weights of this fine node onto the coarse
                     * nodes I+dI, J+dJ, masked to interior coarse nodes */
                    real S = 0;
                    for (int dJ = -1; dJ <= 1; dJ++)
                        for (int dI = -1; dI <= 1; dI++) {
                            int Ic = I + dI, Jc = J + dJ;
                            int ai = i - r * Ic, bj = j - r * Jc;
                            if (abs(ai) >= r || abs(bj) >= r) continue;
                            if (Ic < 1 || Ic > nx2 - 1 || Jc < 1 || Jc > ny2 - 1) continue;
                            real p = hatw(ai, r) * hatw(bj, r);
                            band[3 * (dI + 1) + (dJ + 1)] += w * p * bb;
                            S += p * AT(sxz2, Ic, Jc, 0);
                        }
                    real kf = AT(vx1, i, j, nz1)
                        + bb * (k1[0][0] * (AT(sxx1, i, j, nz1) - AT(sxx1, i - 1, j, nz1))
                                + k1[1][0] * (AT(sxy1, i, j, nz1) - AT(sxy1, i, j - 1, nz1))
                                + k1[2][0] * (S - 2 * AT(sxz1, i, j, nz1 - 1)));
                    fs += w * kf;
                }
            for (int m = 0; m < 9; m++) A[m * nn + q] = k1[2][0] * band[m];
            A[4 * nn + q] += r4 * AT(b2, I, J, 0) * k2[2][0];
            real kc = AT(vx2, I, J, 0)
                + AT(b2, I, J, 0) * (k2[0][0] * (AT(sxx2, I, J, 0) - AT(sxx2, I - 1, J, 0))
                                     + k2[1][0] * (AT(sxy2, I, J, 0) - AT(sxy2, I, J - 1, 0))
                                     + k2[2][0] * AT(sxz2, I, J, 0));
            B[q] = r4 * kc - fs;
        }
    real tol = JACOBI_TOL(avg_abs(B, nn));
    int capped = 0;
    It->iters_last = jacobi2d(It->gx, A, B, n0, n1, tol, It->gtmp, &capped);
    It->iters_total += It->iters_last;
    It->capped |= capped;
    long w = nx2 + 1;
    real *X = It->gx;
#pragma omp parallel for schedule(static)
    for (int J = 1; J < ny2; J++)
        for (int I = 1; I < nx2; I++) AT(sxz2, I, J, -1) = X[J * w + I];
    /* fine ghosts on fine interior nodes */
    int nx1 = F->n[0], ny1 = F->n[1];
#pragma omp parallel for schedule(static)
    for (int j = 1; j < ny1; j++)
        for (int i = 1; i < nx1; i++) {
            int I0 = i / r, J0 = j / r, a = i - r * I0, b = j - r * J0;
            real S = 0;
            for (int dJ = 0; dJ <= 1; dJ++)
                for (int dI = 0; dI <= 1; dI++) {
                    int Ic = I0 + dI, Jc = J0 + dJ;
                    if (Ic < 1 || Ic > nx2 - 1 || Jc < 1 || Jc > ny2 - 1) continue;
                    real p = hatw(a - r * dI, r) * hatw(b - r * dJ, r);
                    if (p == 0) continue;
                    S += p * (AT(sxz2, Ic, Jc, -1) + AT(sxz2, Ic, Jc, 0));
                }
            AT(sxz1, i, j, nz1) = -AT(sxz1, i, j, nz1 - 1) + S;
        }
}

/* vy-syz pair: pointwise (injection in x and y). */
static void ghost_syz_gen(osim *s, int k)
{
    grid_t *F = &s->g[k], *C = &s->g[k + 1];
    int r = s->itf[k].r;
    fld *syy1 = &F->sim[SYY], *syy2 = &C->sim[SYY], *syz1 = &F->sim[SYZ], *syz2 = &C->sim[SYZ];
    fld *sxy1 = &F->sim[SXY], *sxy2 = &C->sim[SXY], *vy1 = &F->sim[VY], *vy2 = &C->sim[VY];
    fld *b1 = &F->buoy, *b2 = &C->buoy;
    real(*k1)[6] = F->coeff, (*k2)[6] = C->coeff;
    int nx2 = C->n[0], ny2 = C->n[1], nz1 = F->n[2];
    real r2 = (real)(r * r);
#pragma omp parallel for schedule(static)
    for (int J = 0; J <= ny2 - 1; J++)
        for (int I = 0; I <= nx2 - 1; I++) {
            real c0 = 0.25 * (AT(b2, I, J, 0) + AT(b2, I + 1, J, 0) + AT(b2, I, J + 1, 0) + AT(b2, I + 1, J + 1, 0));
            real num = r2 * (AT(vy2, I, J, 0)
                             + c0 * (k2[0][0] * (AT(sxy2, I + 1, J, 0) - AT(sxy2, I, J, 0))
                                     + k2[1][0] * (AT(syy2, I, J + 1, 0) - AT(syy2, I, J, 0))
                                     + k2[2][0] * AT(syz2, I, J, 0)));
            real den = r2 * c0 * k2[2][0];
            for (int b = 0; b < r; b++)
                for (int a = 0; a < r; a++) {
                    int i = r * I + a, j = r * J + b;
                    real c = 0.25 * (AT(b1, i, j, nz1) + AT(b1, i + 1, j, nz1) + AT(b1, i, j + 1, nz1) + AT(b1, i + 1, j + 1, nz1));
                    num -= AT(vy1, i, j, nz1)
                        + c * (k1[0][0] * (AT(sxy1, i + 1, j, nz1) - AT(sxy1, i, j, nz1))
                               + k1[1][0] * (AT(syy1, i, j + 1, nz1) - AT(syy1, i, j, nz1))
                               + k1[2][0] * (AT(syz2, I, J, 0) - 2 * AT(syz1, i, j, nz1 - 1)));
                    den += c * k1[2][0];
                }
            AT(syz2, I, J, -1) = num / den;
            real tmp = AT(syz2, I, J, -1) + AT(syz2, I, J, 0);
            for (int b = 0; b < r; b++)
                for (int a = 0; a < r; a++)
                    AT(syz1, r * I + a, r * J + b, nz1) = -AT(syz1, r * I + a, r * J + b, nz1 - 1) + tmp;
        }
}

/* ------------------------------------------------------------------ */
/* Energy, compute_energy.cpp:34-54 with TriMulFields field.cpp:482-518 */
static double trimul(const fld *fa, const fld *fb, const fld *fc)
{
    int e0 = fld_end(fa, 0), e1 = fld_end(fa, 1), e2 = fld_end(fa, 2);
    double *plane = (double *)malloc(sizeof(double) * (e2 + 1));
#pragma omp parallel for schedule(static)
    for (int i2 = 0; i2 <= e2; i2++) {
        real res = 0;
        real mask2 = 1 - (fa->pos[2] ? 0 : 0.5 * ((i2 == 0) || (i2 == e2)));
        for (int i1 = 0; i1 <= e1; i1++) {
            real mask1 = 1 - (fa->pos[1] ? 0 : 0.5 * ((i1 == 0) || (i1 == e1)));
            for (int i0 = 0; i0 <= e0; i0++) {
                real mask0 = 1 - (fa->pos[0] ? 0 : 0.5 * ((i0 == 0) || (i0 == e0)));
                res += mask2 * mask1 * mask0 * AT(fa, i0, i1, i2) * AT(fb, i0, i1, i2) * AT(fc, i0, i1, i2);
            }
        }
        plane[i2] = (double)res;
    }
    double sum = 0;
    for (int i2 = 0; i2 <= e2; i2++) sum += plane[i2];
    free(plane);
    return sum;
}

static double grid_energy(grid_t *G)
{
    double res = 0;
    for (int i = 0; i < 3; i++) res += trimul(&G->dens[i], &G->sim[i], &G->vcopy[i]);
    for (int i = 0; i < 3; i++) res += trimul(&G->comp[i], &G->sim[i + 3], &G->sim[i + 3]);
    res += 2 * trimul(&G->comp[C23], &G->sim[SYY], &G->sim[SZZ]);
    res += 2 * trimul(&G->comp[C13], &G->sim[SXX], &G->sim[SZZ]);
    res += 2 * trimul(&G->comp[C12], &G->sim[SXX], &G->sim[SYY]);
    for (int i = 6; i < 9; i++) res += 2 * trimul(&G->comp[i], &G->sim[i], &G->sim[i]);
    res *= 0.5 * G->h[2] * G->h[1] * G->h[0];
    return res;
}

/* compute_res_* (compute_res.cpp:4-95) with the factor 4 generalised to
 * r^2.  out[3*k + 0..2] = vx-sxz, vy-syz, vz-szz balance of interface k. */
static void residuals(osim *s, int which, double *out)
{
    for (int k = 0; k + 1 < s->ng; k++) {
        grid_t *F = &s->g[k], *C = &s->g[k + 1];
        int r = s->itf[k].r;
        double rr = (double)(r * r);
        int nz = F->n[2];
        double res = 0.0;
        if (which == 0) {
            for (int i1 = 1; i1 < F->n[1]; i1++)
                for (int i0 = 1; i0 < F->n[0]; i0++)
                    res += AT(&F->sim[VX], i0, i1, nz) * (AT(&F->sim[SXZ], i0, i1, nz - 1) + AT(&F->sim[SXZ], i0, i1, nz));
            for (int i1 = 1; i1 < C->n[1]; i1++)
                for (int i0 = 1; i0 < C->n[0]; i0++)
                    res -= rr * AT(&C->sim[VX], i0, i1, 0) * (AT(&C->sim[SXZ], i0, i1, -1) + AT(&C->sim[SXZ], i0, i1, 0));
        } else if (which == 1) {
            for (int i1 = 0; i1 < F->n[1]; i1++)
                for (int i0 = 0; i0 < F->n[0]; i0++)
                    res += AT(&F->sim[VY], i0, i1, nz) * (AT(&F->sim[SYZ], i0, i1, nz - 1) + AT(&F->sim[SYZ], i0, i1, nz));
            for (int i1 = 0; i1 < C->n[1]; i1++)
                for (int i0 = 0; i0 < C->n[0]; i0++)
                    res -= rr * AT(&C->sim[VY], i0, i1, 0) * (AT(&C->sim[SYZ], i0, i1, -1) + AT(&C->sim[SYZ], i0, i1, 0));
        } else {
            for (int i1 = 1; i1 < F->n[1]; i1++)
                for (int i0 = 0; i0 < F->n[0]; i0++)
                    res += AT(&F->sim[SZZ], i0, i1, nz) * (AT(&F->sim[VZ], i0, i1, nz - 1) + AT(&F->sim[VZ], i0, i1, nz));
            for (int i1 = 1; i1 < C->n[1]; i1++)
                for (int i0 = 0; i0 < C->n[0]; i0++)
                    res -= rr * AT(&C->sim[SZZ], i0, i1, 0) * (AT(&C->sim[VZ], i0, i1, -1) + AT(&C->sim[VZ], i0, i1, 0));
        }
        out[k] = res;
    }
}

/* ------------------------------------------------------------------ */
/* Phase entry points and the time loop (main.cpp:165-261)             */
int orc_update_vel(void *h, int g) { update_vel((osim *)h, g); return 0; }
int orc_update_stress(void *h, int g) { update_stress((osim *)h, g); return 0; }
int orc_ghost_vz(void *h, int k)
{
    osim *s = (osim *)h;
    if (s->imode == 0) ghost_vz_ii(s, k);
    else if (s->eci_mode == 0) ghost_vz_ref(s, k);
    else ghost_vz_gen(s, k);
    return 0;
}
int orc_ghost_sxz(void *h, int k)
{
    osim *s = (osim *)h;
    if (s->imode == 0) ghost_sxz_ii(s, k);
    else if (s->eci_mode == 0) ghost_sxz_ref(s, k);
    else ghost_sxz_gen(s, k);
    return 0;
}
int orc_ghost_syz(void *h, int k)
{
    osim *s = (osim *)h;
    if (s->imode == 0) ghost_syz_ii(s, k);
    else if (s->eci_mode == 0) ghost_syz_ref(s, k);
    else ghost_syz_gen(s, k);
    return 0;
}
int orc_residuals(void *h, int which, double *out) { residuals((osim *)h, which, out); return 0; }

static void copy_vel(osim *s)
{
    for (int g = 0; g < s->ng; g++)
        for (int c = 0; c < 3; c++)
            memcpy(s->g[g].vcopy[c].base, s->g[g].sim[c].base, sizeof(real) * fld_size(&s->g[g].sim[c]));
}

/* energy of the current state given the velocity copy taken at step start */
double orc_energy_now(void *h)
{
    osim *s = (osim *)h;
    double e = 0;
    for (int g = 0; g < s->ng; g++) e += grid_energy(&s->g[g]);
    return e;
}

int orc_copy_vel(void *h) { copy_vel((osim *)h); return 0; }

int orc_record(void *h)
{
    osim *s = (osim *)h;
    int n = s->steps_done;
    if (n >= s->max_steps) return -1;
    for (int q = 0; q < s->nrec; q++) {
        int *L = s->rec_loc + 4 * q;
        s->traces[(long)q * s->max_steps + n] = AT(&s->g[L[0]].sim[VX], L[1], L[2], L[3]);
    }
    return 0;
}

int orc_add_source(void *h)
{
    osim *s = (osim *)h;
    int n = s->steps_done;
    if (n < s->src_len) {
        grid_t *G = &s->g[s->src_g];
        int *p = s->src_loc;
        if (s->src_type == 0) {
            /* main.cpp:194-195: float product, then added to real */
            AT(&G->sim[VX], p[0], p[1], p[2]) += 1000 * s->src_trace[n];
        } else {
            /* explosive variant, main.cpp:197-202 */
            AT(&G->sim[SXX], p[0], p[1], p[2]) += s->src_trace[n];
            AT(&G->sim[SYY], p[0], p[1], p[2]) += s->src_trace[n];
            AT(&G->sim[SZZ], p[0], p[1], p[2]) += s->src_trace[n];
        }
    }
    return 0;
}

int orc_run(void *h, int nsteps)
{
    osim *s = (osim *)h;
    for (int it = 0; it < nsteps; it++) {
        if (s->steps_done >= s->max_steps) return -1;
        if (s->check_energy) copy_vel(s);
        for (int g = 0; g < s->ng; g++) update_vel(s, g);
        orc_record(s);
        orc_add_source(s);
        for (int k = 0; k + 1 < s->ng; k++) orc_ghost_vz(s, k);
        if (s->check_energy) s->energy[s->steps_done] = orc_energy_now(s);
        for (int g = 0; g < s->ng; g++) update_stress(s, g);
        for (int k = 0; k + 1 < s->ng; k++) { orc_ghost_sxz(s, k); orc_ghost_syz(s, k); }
        s->steps_done++;
    }
    return 0;
}

int orc_set_vy_mirror_all(void *h, int flag) { ((osim *)h)->vy_mirror_all = flag ? 1 : 0; return 0; }
int orc_set_ii_xend(void *h, int flag) { ((osim *)h)->ii_xend = flag ? 1 : 0; return 0; }

int orc_steps_done(void *h) { return ((osim *)h)->steps_done; }
int orc_set_steps_done(void *h, int n) { ((osim *)h)->steps_done = n; return 0; }

/* field access: c 0..8 wavefields, 9 buoyancy, 10..18 staged stiffness,
 * 19..21 energy densities, 22..30 compliances */
static fld *pick(osim *s, int g, int c)
{
    if (g < 0 || g >= s->ng) return NULL;
    grid_t *G = &s->g[g];
    if (c >= 0 && c < 9) return &G->sim[c];
    if (c == 9) return &G->buoy;
    if (c >= 10 && c < 19) return &G->st[c - 10];
    if (c >= 19 && c < 22) return s->check_energy ? &G->dens[c - 19] : NULL;
    if (c >= 22 && c < 31) return s->check_energy ? &G->comp[c - 22] : NULL;
    return NULL;
}

int orc_field_dims(void *h, int g, int c, int *len3)
{
    fld *f = pick((osim *)h, g, c);
    if (!f) return -1;
    for (int d = 0; d < 3; d++) len3[d] = f->len[d];
    return f->r;
}

int orc_get_field(void *h, int g, int c, double *out)
{
    fld *f = pick((osim *)h, g, c);
    if (!f) return -1;
    long n = fld_size(f);
    for (long q = 0; q < n; q++) out[q] = (double)f->base[q];
    return 0;
}

int orc_set_field(void *h, int g, int c, const double *in)
{
    fld *f = pick((osim *)h, g, c);
    if (!f) return -1;
    long n = fld_size(f);
    for (long q = 0; q < n; q++) f->base[q] = (real)in[q];
    return 0;
}

int orc_get_traces(void *h, double *out)
{
    osim *s = (osim *)h;
    for (int q = 0; q < s->nrec; q++)
        for (int n = 0; n < s->steps_done; n++) out[(long)q * s->steps_done + n] = (double)s->traces[(long)q * s->max_steps + n];
    return 0;
}

int orc_get_energy(void *h, double *out)
{
    osim *s = (osim *)h;
    for (int n = 0; n < s->steps_done; n++) out[n] = s->energy ? s->energy[n] : 0.0;
    return 0;
}

int orc_jacobi_stats(void *h, int k, long *total, int *last, int *capped)
{
    osim *s = (osim *)h;
    if (k < 0 || k + 1 >= s->ng) return -1;
    *total = s->itf[k].iters_total;
    *last = s->itf[k].iters_last;
    *capped = s->itf[k].capped;
    return 0;
}

/* Jacobi iterate (warm-start state) of interface k, (nx2+1)(ny2+1) */
int orc_get_jacobi_state(void *h, int k, double *out)
{
    osim *s = (osim *)h;
    if (k < 0 || k + 1 >= s->ng) return -1;
    long n = (long)(s->g[k + 1].n[0] + 1) * (s->g[k + 1].n[1] + 1);
    for (long q = 0; q < n; q++) out[q] = (double)s->itf[k].gx[q];
    return 0;
}

int orc_get_sxz_system(void *h, int k, double *A, double *B)
{
    osim *s = (osim *)h;
    if (k < 0 || k + 1 >= s->ng) return -1;
    long nn = (long)(s->g[k + 1].n[0] - 1) * (s->g[k + 1].n[1] - 1);
    for (long q = 0; q < 9 * nn; q++) A[q] = (double)s->itf[k].A[q];
    for (long q = 0; q < nn; q++) B[q] = (double)s->itf[k].B[q];
    return 0;
}

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* threads of the following parallel regions (the CPU baseline's one-thread leg) */
void orc_set_num_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

int orc_precision_bytes(void) { return (int)sizeof(real); }
