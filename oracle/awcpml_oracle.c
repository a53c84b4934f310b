/*
 * awcpml_oracle.c -- CPU FP64 restatement of the AcousticWaveCPML.jl xPU
 * time-stepping path (2D and 3D, CPML + free top).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle: the tests, the
 * smoke check and the CPU-baseline leg of bench.py call it as the checker.
 * The product (the HIP library) never links, loads or calls it.
 *
 * The restatement follows the reference's Julia sources statement by
 * statement, in the same evaluation order, in IEEE FP64 with no contraction
 * (build with -ffp-contract=off, no -ffast-math):
 *   calc_Kab_CPML ........ src/utils.jl:19-57
 *   rickersource1D ....... src/utils.jl:64-66
 *   scalar setup ......... src/shared/acoustic_2D_xPU.jl:155-201,
 *                          src/shared/acoustic_3D_xPU.jl:191-248
 *   source wavelets ...... src/shared/acoustic_2D_xPU.jl:219-226, 3D:271-275
 *   position snapping .... src/shared/acoustic_2D_xPU.jl:228-240, 3D:277-285
 *   update_psi_{x,y,z}! .. src/shared/acoustic_2D_xPU.jl:7-31, 3D:8-45
 *   update_p! ............ src/shared/acoustic_2D_xPU.jl:33-75, 3D:47-104
 *   inject_sources! ...... src/shared/acoustic_2D_xPU.jl:77-83, 3D:106-113
 *   record_receivers! .... src/shared/acoustic_2D_xPU.jl:85-91, 3D:115-122
 *   forward! + time loop . src/shared/acoustic_2D_xPU.jl:93-131,327-345,
 *                          3D:124-167,392-407 (buffer rotation)
 *   outputs .............. src/shared/acoustic_2D_xPU.jl:441,444; 3D:541,544
 *   FWI gradient ......... discrete adjoint of the loop above (oracle_fwi_gradient)
 *
 * Parallelism mirrors ParallelStencil's Threads backend: the outermost grid
 * index is split over OpenMP threads, x innermost.  Results do not depend
 * on the thread count (no reductions).
 *
 * Known bit-level gaps (all <= 1 ulp in setup values, i.e. ~1e-16 relative
 * in the fields, far below the 1.49e-8 fixture tolerance and the 1e-12
 * parity bar): glibc exp/log instead of Julia's own libm, and LinRange
 * element evaluation written as (1-t)*a + t*b.
 *
 * Arrays are column-major (x fastest) exactly as in Julia; every index in
 * the public entry points that refers to a grid node is 1-based, as in the
 * reference.
 */
#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORACLE_OK 0
#define ORACLE_ERR_ARG -1
#define ORACLE_ERR_PPW -2      /* "Not enough points per wavelength!"   */
#define ORACLE_ERR_SRC -3      /* "At least one source is not inside"   */
#define ORACLE_ERR_REC -4      /* "At least one receiver is not inside" */
#define ORACLE_ERR_ONWHERE -5  /* "Wrong onwhere parameter!"            */
#define ORACLE_ERR_NOMEM -6

static const double PI_F64 = 3.141592653589793; /* Float64(pi) */

/* ------------------------------------------------------------------ */
/* setup helpers                                                      */
/* ------------------------------------------------------------------ */

/* round(Int, x, RoundNearestTiesUp) for the finite grid coordinates used
 * by the position snapping (src/shared/acoustic_2D_xPU.jl:230,238).
 * x - floor(x) is exact for |x| < 2^52, so the tie test is exact. */
long oracle_round_ties_up(double x)
{
    double t = floor(x);
    return (long)((x - t >= 0.5) ? t + 1.0 : t);
}

/* rickersource1D, src/utils.jl:64-66 */
double oracle_ricker(double t, double t0, double f0)
{
    double s = PI_F64 * f0 * (t - t0);
    double s2 = s * s;
    return (1.0 - 2.0 * s2) * exp(-s2);
}

/* x ^ npower as the reference evaluates it; for an integer-valued
 * exponent Julia's Float64 ^ reduces to the correctly rounded power,
 * which for 2.0 is x*x. */
static double powf64(double x, double p)
{
    if (p == 2.0) return x * x;
    if (p == 1.0) return x;
    return pow(x, p);
}

/* calc_Kab_CPML, src/utils.jl:19-57.  halfgrd = 0 -> "ongrd" (halo values),
 * halfgrd = 1 -> "halfgrd" (halo+1 values), any other value -> the
 * reference's error("Wrong onwhere parameter!").  Outputs a_l, a_r, b_K_l,
 * b_K_r, each of length halo (+1 for halfgrd). */
int oracle_calc_kab_cpml(int halo, double dt, double npower, double d0,
                         double alpha_max, double K_max, int halfgrd,
                         double *a_l, double *a_r, double *bK_l, double *bK_r)
{
    if (halo < 0) return ORACLE_ERR_ARG;
    if (halfgrd != 0 && halfgrd != 1) return ORACLE_ERR_ONWHERE;
    int n = halo;
    double shift = 0.0;
    if (halfgrd) { n += 1; shift = 0.5; }
    if (n == 0) return ORACLE_OK;
    double *dist = (double *)malloc(sizeof(double) * (size_t)n);
    if (!dist) return ORACLE_ERR_NOMEM;
    /* LinRange(0-shift, Kab_size-shift-1, Kab_size), utils.jl:35 */
    double start = 0.0 - shift;
    double stop = ((double)n - shift) - 1.0;
    double lendiv = (n > 1) ? (double)(n - 1) : 1.0;
    for (int i = 0; i < n; ++i) {
        double t = (double)i / lendiv;
        dist[i] = (1.0 - t) * start + t * stop;
    }
    if (halfgrd) dist[0] = 0.0; /* utils.jl:36-38 */
    for (int side = 0; side < 2; ++side) {
        double *a = side ? a_r : a_l;
        double *bK = side ? bK_r : bK_l;
        for (int i = 0; i < n; ++i) {
            /* left uses reverse(dist), right uses dist (utils.jl:39-40) */
            double nd = (side ? dist[i] : dist[n - 1 - i]) / (double)halo;
            double d = d0 * powf64(nd, npower);
            double alpha = alpha_max * (1.0 - nd);
            double K = 1.0 + (K_max - 1.0) * powf64(nd, npower);
            double b = exp((-(d / K + alpha)) * dt);
            a[i] = d * (b - 1.0) / (K * (d + K * alpha));
            bK[i] = b / K;
        }
    }
    free(dist);
    return ORACLE_OK;
}

/* ------------------------------------------------------------------ */
/* problem state                                                      */
/* ------------------------------------------------------------------ */

typedef struct {
    int ndim;
    long n[3];       /* nx, ny, nz (nz = 1 in 2D) */
    int halo;
    int freetop;
    double d[3];     /* dx, dy, dz */
    double dt;
    long nt;
    double inv_d[3];  /* _dx, _dy, _dz   */
    double inv_d2[3]; /* _dx2, _dy2, _dz2 */
    /* CPML coefficients per axis: [0]=a_l [1]=a_r [2]=b_K_l [3]=b_K_r (halo)
     * [4]=a_hl [5]=a_hr [6]=b_K_hl [7]=b_K_hr (halo+1) */
    double *coef[3][8];
    double *fact;
    double *p[3];     /* pold, pcur, pnew */
    double *psi[3][2];/* per axis: left, right */
    double *xi[3][2];
    int nsrcs, nrecs;
    long *possrcs;    /* nsrcs x ndim, column-major, 1-based */
    long *posrecs;
    double *dt2srctf; /* nt x nsrcs column-major */
    double *traces;   /* nt x nrecs column-major */
    long it_done;
} oracle_t;

static double *zalloc(size_t count)
{
    if (count == 0) count = 1;
    return (double *)calloc(count, sizeof(double));
}

void oracle_destroy(oracle_t *o)
{
    if (!o) return;
    for (int a = 0; a < 3; ++a) {
        for (int c = 0; c < 8; ++c) free(o->coef[a][c]);
        for (int s = 0; s < 2; ++s) { free(o->psi[a][s]); free(o->xi[a][s]); }
    }
    for (int b = 0; b < 3; ++b) free(o->p[b]);
    free(o->fact);
    free(o->possrcs);
    free(o->posrecs);
    free(o->dt2srctf);
    free(o->traces);
    free(o);
}

/* Size (in elements) of the psi/xi slab of `axis` (halo+1 or halo layers
 * across the axis, full extent along the others). */
static size_t slab_size(const oracle_t *o, int axis, long layers)
{
    size_t s = (size_t)layers;
    for (int a = 0; a < o->ndim; ++a)
        if (a != axis) s *= (size_t)o->n[a];
    return s;
}

/*
 * Build the full problem exactly as solve2D / solve3D do before their time
 * loop (src/shared/acoustic_2D_xPU.jl:155-240, 3D:191-285).
 *   n[ndim], l[ndim]: grid points and model size per axis [m]
 *   vel: velocity model, column-major, prod(n) values
 *   srcpos: nsrcs x ndim column-major positions [m]; t0s: activation times
 *   recpos: nrecs x ndim column-major positions [m]
 * Returns NULL and sets *status on the reference's assertion failures.
 */
oracle_t *oracle_create(int ndim, const long *n, const double *l, double lt,
                        const double *vel, int nsrcs, const double *srcpos,
                        const double *t0s, double f0, int nrecs,
                        const double *recpos, int halo, double rcoef,
                        double ppw, int freetop, int *status)
{
    int st = ORACLE_OK;
    if (ndim != 2 && ndim != 3) { st = ORACLE_ERR_ARG; goto fail_early; }
    if (halo < 0 || nsrcs < 1 || nrecs < 1) { st = ORACLE_ERR_ARG; goto fail_early; }
    for (int a = 0; a < ndim; ++a)
        if (n[a] < 3) { st = ORACLE_ERR_ARG; goto fail_early; }

    oracle_t *o = (oracle_t *)calloc(1, sizeof(oracle_t));
    if (!o) { st = ORACLE_ERR_NOMEM; goto fail_early; }
    o->ndim = ndim;
    o->halo = halo;
    o->freetop = freetop;
    o->n[0] = o->n[1] = o->n[2] = 1;
    for (int a = 0; a < ndim; ++a) o->n[a] = n[a];
    size_t N = (size_t)o->n[0] * (size_t)o->n[1] * (size_t)o->n[2];

    /* vel_max = maximum(vel) */
    double vmax = vel[0];
    for (size_t i = 1; i < N; ++i)
        if (vel[i] > vmax) vmax = vel[i];

    for (int a = 0; a < ndim; ++a) o->d[a] = l[a] / (double)(n[a] - 1);
    double dx = o->d[0], dy = o->d[1], dz = (ndim == 3) ? o->d[2] : 0.0;
    /* CFL time step, acoustic_2D_xPU.jl:163 / acoustic_3D_xPU.jl:200 */
    if (ndim == 2)
        o->dt = sqrt(2.0) / (vmax * (1.0 / dx + 1.0 / dy)) / 2.0;
    else
        o->dt = sqrt(3.0) / (vmax * (1.0 / dx + 1.0 / dy + 1.0 / dz)) / 2.0;
    o->nt = (long)ceil(lt / o->dt);
    if (o->nt < 0) o->nt = 0;
    double dt = o->dt;

    /* CPML numerics, acoustic_2D_xPU.jl:167-185 */
    double alpha_max = PI_F64 * f0;
    double npower = 2.0, K_max = 1.0;
    for (int a = 0; a < ndim; ++a) {
        double thickness = (double)halo * o->d[a];
        double d0 = -(npower + 1.0) * vmax * log(rcoef) / (2.0 * thickness);
        for (int c = 0; c < 4; ++c) o->coef[a][c] = zalloc((size_t)halo);
        for (int c = 4; c < 8; ++c) o->coef[a][c] = zalloc((size_t)halo + 1);
        oracle_calc_kab_cpml(halo, dt, npower, d0, alpha_max, K_max, 0,
                             o->coef[a][0], o->coef[a][1], o->coef[a][2], o->coef[a][3]);
        oracle_calc_kab_cpml(halo, dt, npower, d0, alpha_max, K_max, 1,
                             o->coef[a][4], o->coef[a][5], o->coef[a][6], o->coef[a][7]);
    }
    if (freetop) { /* top = y-low side, acoustic_2D_xPU.jl:180-185 */
        for (int i = 0; i < halo; ++i) { o->coef[1][0][i] = 0.0; o->coef[1][2][i] = 1.0; }
        for (int i = 0; i < halo + 1; ++i) { o->coef[1][4][i] = 0.0; o->coef[1][6][i] = 1.0; }
    }

    /* precomputations, acoustic_2D_xPU.jl:190-196 */
    for (int a = 0; a < ndim; ++a) {
        o->inv_d[a] = 1.0 / o->d[a];
        o->inv_d2[a] = 1.0 / (o->d[a] * o->d[a]);
    }
    o->fact = zalloc(N);
    double dt2 = dt * dt;
    for (size_t i = 0; i < N; ++i) o->fact[i] = dt2 * (vel[i] * vel[i]);

    /* PPW assertion, acoustic_2D_xPU.jl:201 */
    {
        double h2 = dx * dx + dy * dy;
        if (ndim == 3) h2 = dx * dx + dy * dy + dz * dz;
        if (!(sqrt(h2) <= vmax / (ppw * f0))) { st = ORACLE_ERR_PPW; goto fail; }
    }

    /* state arrays, acoustic_2D_xPU.jl:208-215 */
    for (int b = 0; b < 3; ++b) o->p[b] = zalloc(N);
    for (int a = 0; a < ndim; ++a)
        for (int s = 0; s < 2; ++s) {
            o->psi[a][s] = zalloc(slab_size(o, a, halo + 1));
            o->xi[a][s] = zalloc(slab_size(o, a, halo));
        }

    /* sources, acoustic_2D_xPU.jl:219-232 */
    o->nsrcs = nsrcs;
    o->dt2srctf = zalloc((size_t)o->nt * (size_t)nsrcs);
    {
        double cell = dx * dy;
        if (ndim == 3) cell = dx * dy * dz;
        double scale = dt2 / cell;
        for (int s = 0; s < nsrcs; ++s)
            for (long it = 0; it < o->nt; ++it)
                o->dt2srctf[(size_t)s * o->nt + it] =
                    scale * oracle_ricker((double)it * dt, t0s[s], f0);
    }
    o->possrcs = (long *)calloc((size_t)nsrcs * ndim, sizeof(long));
    for (int s = 0; s < nsrcs; ++s)
        for (int a = 0; a < ndim; ++a) {
            long v = oracle_round_ties_up(srcpos[(size_t)a * nsrcs + s] / o->d[a] + 1.0);
            o->possrcs[(size_t)a * nsrcs + s] = v;
            if (v < 1 || v > o->n[a]) { st = ORACLE_ERR_SRC; goto fail; }
        }
    /* receivers, acoustic_2D_xPU.jl:233-240 */
    o->nrecs = nrecs;
    o->traces = zalloc((size_t)o->nt * (size_t)nrecs);
    o->posrecs = (long *)calloc((size_t)nrecs * ndim, sizeof(long));
    for (int r = 0; r < nrecs; ++r)
        for (int a = 0; a < ndim; ++a) {
            long v = oracle_round_ties_up(recpos[(size_t)a * nrecs + r] / o->d[a] + 1.0);
            o->posrecs[(size_t)a * nrecs + r] = v;
            if (v < 1 || v > o->n[a]) { st = ORACLE_ERR_REC; goto fail; }
        }
    o->it_done = 0;
    if (status) *status = ORACLE_OK;
    return o;

fail:
    oracle_destroy(o);
fail_early:
    if (status) *status = st;
    return NULL;
}

/* ------------------------------------------------------------------ */
/* kernels, 2D                                                        */
/* ------------------------------------------------------------------ */

/* 1-based column-major accessors */
#define A2(arr, ldx, i, j) (arr)[((size_t)(j) - 1) * (size_t)(ldx) + ((size_t)(i) - 1)]
#define A3(arr, ldx, ldy, i, j, k) \
    (arr)[(((size_t)(k) - 1) * (size_t)(ldy) + ((size_t)(j) - 1)) * (size_t)(ldx) + ((size_t)(i) - 1)]
#define C1(arr, i) (arr)[(i) - 1]

/* update_psi_x! and update_psi_y!, acoustic_2D_xPU.jl:7-31, launched over
 * (1:halo+1, 1:ny) and (1:nx, 1:halo+1) (acoustic_2D_xPU.jl:107,111).
 * Exposed separately for the known-answer test of test/test2D.jl:13-43. */
void oracle_update_psi_2d(long nx, long ny, int halo, double _dx, double _dy,
                          const double *pcur,
                          double *psi_x_l, double *psi_x_r,
                          double *psi_y_l, double *psi_y_r,
                          const double *a_x_hl, const double *a_x_hr,
                          const double *b_K_x_hl, const double *b_K_x_hr,
                          const double *a_y_hl, const double *a_y_hr,
                          const double *b_K_y_hl, const double *b_K_y_hr)
{
    long hp1 = halo + 1;
    #pragma omp parallel for schedule(static)
    for (long j = 1; j <= ny; ++j)
        for (long i = 1; i <= hp1; ++i) {
            long ii = i + nx - halo - 2;
            A2(psi_x_l, hp1, i, j) = C1(b_K_x_hl, i) * A2(psi_x_l, hp1, i, j) +
                C1(a_x_hl, i) * (A2(pcur, nx, i + 1, j) - A2(pcur, nx, i, j)) * _dx;
            A2(psi_x_r, hp1, i, j) = C1(b_K_x_hr, i) * A2(psi_x_r, hp1, i, j) +
                C1(a_x_hr, i) * (A2(pcur, nx, ii + 1, j) - A2(pcur, nx, ii, j)) * _dx;
        }
    #pragma omp parallel for schedule(static)
    for (long j = 1; j <= hp1; ++j)
        for (long i = 1; i <= nx; ++i) {
            long jj = j + ny - halo - 2;
            A2(psi_y_l, nx, i, j) = C1(b_K_y_hl, j) * A2(psi_y_l, nx, i, j) +
                C1(a_y_hl, j) * (A2(pcur, nx, i, j + 1) - A2(pcur, nx, i, j)) * _dy;
            A2(psi_y_r, nx, i, j) = C1(b_K_y_hr, j) * A2(psi_y_r, nx, i, j) +
                C1(a_y_hr, j) * (A2(pcur, nx, i, jj + 1) - A2(pcur, nx, i, jj)) * _dy;
        }
}

/* update_p!, acoustic_2D_xPU.jl:33-75, over (2:nx-1, 2:ny-1) */
static void update_p_2d(oracle_t *o)
{
    const long nx = o->n[0], ny = o->n[1];
    const long halo = o->halo, hp1 = halo + 1;
    const double *pold = o->p[0], *pcur = o->p[1];
    double *pnew = o->p[2];
    const double *fact = o->fact;
    const double _dx = o->inv_d[0], _dy = o->inv_d[1];
    const double _dx2 = o->inv_d2[0], _dy2 = o->inv_d2[1];
    double *const *cx = o->coef[0];
    double *const *cy = o->coef[1];
    const double *psi_x_l = o->psi[0][0], *psi_x_r = o->psi[0][1];
    const double *psi_y_l = o->psi[1][0], *psi_y_r = o->psi[1][1];
    double *xi_x_l = o->xi[0][0], *xi_x_r = o->xi[0][1];
    double *xi_y_l = o->xi[1][0], *xi_y_r = o->xi[1][1];
    #pragma omp parallel for schedule(static)
    for (long j = 2; j <= ny - 1; ++j)
        for (long i = 2; i <= nx - 1; ++i) {
            double d2p_dx2 = (A2(pcur, nx, i + 1, j) - 2.0 * A2(pcur, nx, i, j) + A2(pcur, nx, i - 1, j)) * _dx2;
            double d2p_dy2 = (A2(pcur, nx, i, j + 1) - 2.0 * A2(pcur, nx, i, j) + A2(pcur, nx, i, j - 1)) * _dy2;
            double damp = 0.0;
            if (i <= halo + 1) {
                double dpsi = (A2(psi_x_l, hp1, i, j) - A2(psi_x_l, hp1, i - 1, j)) * _dx;
                A2(xi_x_l, halo, i - 1, j) = C1(cx[2], i - 1) * A2(xi_x_l, halo, i - 1, j) +
                                             C1(cx[0], i - 1) * (d2p_dx2 + dpsi);
                damp += A2(fact, nx, i, j) * (dpsi + A2(xi_x_l, halo, i - 1, j));
            } else if (i >= nx - halo) {
                long ii = i - (nx - halo) + 2;
                double dpsi = (A2(psi_x_r, hp1, ii, j) - A2(psi_x_r, hp1, ii - 1, j)) * _dx;
                A2(xi_x_r, halo, ii - 1, j) = C1(cx[3], ii - 1) * A2(xi_x_r, halo, ii - 1, j) +
                                              C1(cx[1], ii - 1) * (d2p_dx2 + dpsi);
                damp += A2(fact, nx, i, j) * (dpsi + A2(xi_x_r, halo, ii - 1, j));
            }
            if (j <= halo + 1) {
                double dpsi = (A2(psi_y_l, nx, i, j) - A2(psi_y_l, nx, i, j - 1)) * _dy;
                A2(xi_y_l, nx, i, j - 1) = C1(cy[2], j - 1) * A2(xi_y_l, nx, i, j - 1) +
                                           C1(cy[0], j - 1) * (d2p_dy2 + dpsi);
                damp += A2(fact, nx, i, j) * (dpsi + A2(xi_y_l, nx, i, j - 1));
            } else if (j >= ny - halo) {
                long jj = j - (ny - halo) + 2;
                double dpsi = (A2(psi_y_r, nx, i, jj) - A2(psi_y_r, nx, i, jj - 1)) * _dy;
                A2(xi_y_r, nx, i, jj - 1) = C1(cy[3], jj - 1) * A2(xi_y_r, nx, i, jj - 1) +
                                            C1(cy[1], jj - 1) * (d2p_dy2 + dpsi);
                damp += A2(fact, nx, i, j) * (dpsi + A2(xi_y_r, nx, i, jj - 1));
            }
            A2(pnew, nx, i, j) = 2.0 * A2(pcur, nx, i, j) - A2(pold, nx, i, j) +
                                 A2(fact, nx, i, j) * (d2p_dx2 + d2p_dy2) + damp;
        }
}

/* ------------------------------------------------------------------ */
/* kernels, 3D                                                        */
/* ------------------------------------------------------------------ */

/* update_psi_{x,y,z}!, acoustic_3D_xPU.jl:8-45, launched over
 * (1:halo+1,1:ny,1:nz), (1:nx,1:halo+1,1:nz), (1:nx,1:ny,1:halo+1)
 * (acoustic_3D_xPU.jl:138,142,146).  Exposed for the known-answer test of
 * test/test3D.jl:13-71. */
void oracle_update_psi_3d(long nx, long ny, long nz, int halo,
                          double _dx, double _dy, double _dz,
                          const double *pcur,
                          double *psi_x_l, double *psi_x_r,
                          double *psi_y_l, double *psi_y_r,
                          double *psi_z_l, double *psi_z_r,
                          const double *a_x_hl, const double *a_x_hr,
                          const double *b_K_x_hl, const double *b_K_x_hr,
                          const double *a_y_hl, const double *a_y_hr,
                          const double *b_K_y_hl, const double *b_K_y_hr,
                          const double *a_z_hl, const double *a_z_hr,
                          const double *b_K_z_hl, const double *b_K_z_hr)
{
    long hp1 = halo + 1;
    #pragma omp parallel for schedule(static)
    for (long k = 1; k <= nz; ++k)
        for (long j = 1; j <= ny; ++j)
            for (long i = 1; i <= hp1; ++i) {
                long ii = i + nx - halo - 2;
                A3(psi_x_l, hp1, ny, i, j, k) = C1(b_K_x_hl, i) * A3(psi_x_l, hp1, ny, i, j, k) +
                    C1(a_x_hl, i) * (A3(pcur, nx, ny, i + 1, j, k) - A3(pcur, nx, ny, i, j, k)) * _dx;
                A3(psi_x_r, hp1, ny, i, j, k) = C1(b_K_x_hr, i) * A3(psi_x_r, hp1, ny, i, j, k) +
                    C1(a_x_hr, i) * (A3(pcur, nx, ny, ii + 1, j, k) - A3(pcur, nx, ny, ii, j, k)) * _dx;
            }
    #pragma omp parallel for schedule(static)
    for (long k = 1; k <= nz; ++k)
        for (long j = 1; j <= hp1; ++j)
            for (long i = 1; i <= nx; ++i) {
                long jj = j + ny - halo - 2;
                A3(psi_y_l, nx, hp1, i, j, k) = C1(b_K_y_hl, j) * A3(psi_y_l, nx, hp1, i, j, k) +
                    C1(a_y_hl, j) * (A3(pcur, nx, ny, i, j + 1, k) - A3(pcur, nx, ny, i, j, k)) * _dy;
                A3(psi_y_r, nx, hp1, i, j, k) = C1(b_K_y_hr, j) * A3(psi_y_r, nx, hp1, i, j, k) +
                    C1(a_y_hr, j) * (A3(pcur, nx, ny, i, jj + 1, k) - A3(pcur, nx, ny, i, jj, k)) * _dy;
            }
    #pragma omp parallel for schedule(static)
    for (long k = 1; k <= hp1; ++k)
        for (long j = 1; j <= ny; ++j)
            for (long i = 1; i <= nx; ++i) {
                long kk = k + nz - halo - 2;
                A3(psi_z_l, nx, ny, i, j, k) = C1(b_K_z_hl, k) * A3(psi_z_l, nx, ny, i, j, k) +
                    C1(a_z_hl, k) * (A3(pcur, nx, ny, i, j, k + 1) - A3(pcur, nx, ny, i, j, k)) * _dz;
                A3(psi_z_r, nx, ny, i, j, k) = C1(b_K_z_hr, k) * A3(psi_z_r, nx, ny, i, j, k) +
                    C1(a_z_hr, k) * (A3(pcur, nx, ny, i, j, kk + 1) - A3(pcur, nx, ny, i, j, kk)) * _dz;
            }
}

/* update_p!, acoustic_3D_xPU.jl:47-104, over (2:nx-1, 2:ny-1, 2:nz-1) */
static void update_p_3d(oracle_t *o)
{
    const long nx = o->n[0], ny = o->n[1], nz = o->n[2];
    const long halo = o->halo, hp1 = halo + 1;
    const double *pold = o->p[0], *pcur = o->p[1];
    double *pnew = o->p[2];
    const double *fact = o->fact;
    const double _dx = o->inv_d[0], _dy = o->inv_d[1], _dz = o->inv_d[2];
    const double _dx2 = o->inv_d2[0], _dy2 = o->inv_d2[1], _dz2 = o->inv_d2[2];
    double *const *cx = o->coef[0];
    double *const *cy = o->coef[1];
    double *const *cz = o->coef[2];
    const double *psi_x_l = o->psi[0][0], *psi_x_r = o->psi[0][1];
    const double *psi_y_l = o->psi[1][0], *psi_y_r = o->psi[1][1];
    const double *psi_z_l = o->psi[2][0], *psi_z_r = o->psi[2][1];
    double *xi_x_l = o->xi[0][0], *xi_x_r = o->xi[0][1];
    double *xi_y_l = o->xi[1][0], *xi_y_r = o->xi[1][1];
    double *xi_z_l = o->xi[2][0], *xi_z_r = o->xi[2][1];
    #pragma omp parallel for schedule(static)
    for (long k = 2; k <= nz - 1; ++k)
        for (long j = 2; j <= ny - 1; ++j)
            for (long i = 2; i <= nx - 1; ++i) {
                double pc = A3(pcur, nx, ny, i, j, k);
                double d2p_dx2 = (A3(pcur, nx, ny, i + 1, j, k) - 2.0 * pc + A3(pcur, nx, ny, i - 1, j, k)) * _dx2;
                double d2p_dy2 = (A3(pcur, nx, ny, i, j + 1, k) - 2.0 * pc + A3(pcur, nx, ny, i, j - 1, k)) * _dy2;
                double d2p_dz2 = (A3(pcur, nx, ny, i, j, k + 1) - 2.0 * pc + A3(pcur, nx, ny, i, j, k - 1)) * _dz2;
                double f = A3(fact, nx, ny, i, j, k);
                double damp = 0.0;
                if (i <= halo + 1) {
                    double dpsi = (A3(psi_x_l, hp1, ny, i, j, k) - A3(psi_x_l, hp1, ny, i - 1, j, k)) * _dx;
                    A3(xi_x_l, halo, ny, i - 1, j, k) = C1(cx[2], i - 1) * A3(xi_x_l, halo, ny, i - 1, j, k) +
                                                        C1(cx[0], i - 1) * (d2p_dx2 + dpsi);
                    damp += f * (dpsi + A3(xi_x_l, halo, ny, i - 1, j, k));
                } else if (i >= nx - halo) {
                    long ii = i - (nx - halo) + 2;
                    double dpsi = (A3(psi_x_r, hp1, ny, ii, j, k) - A3(psi_x_r, hp1, ny, ii - 1, j, k)) * _dx;
                    A3(xi_x_r, halo, ny, ii - 1, j, k) = C1(cx[3], ii - 1) * A3(xi_x_r, halo, ny, ii - 1, j, k) +
                                                         C1(cx[1], ii - 1) * (d2p_dx2 + dpsi);
                    damp += f * (dpsi + A3(xi_x_r, halo, ny, ii - 1, j, k));
                }
                if (j <= halo + 1) {
                    double dpsi = (A3(psi_y_l, nx, hp1, i, j, k) - A3(psi_y_l, nx, hp1, i, j - 1, k)) * _dy;
                    A3(xi_y_l, nx, halo, i, j - 1, k) = C1(cy[2], j - 1) * A3(xi_y_l, nx, halo, i, j - 1, k) +
                                                        C1(cy[0], j - 1) * (d2p_dy2 + dpsi);
                    damp += f * (dpsi + A3(xi_y_l, nx, halo, i, j - 1, k));
                } else if (j >= ny - halo) {
                    long jj = j - (ny - halo) + 2;
                    double dpsi = (A3(psi_y_r, nx, hp1, i, jj, k) - A3(psi_y_r, nx, hp1, i, jj - 1, k)) * _dy;
                    A3(xi_y_r, nx, halo, i, jj - 1, k) = C1(cy[3], jj - 1) * A3(xi_y_r, nx, halo, i, jj - 1, k) +
                                                         C1(cy[1], jj - 1) * (d2p_dy2 + dpsi);
                    damp += f * (dpsi + A3(xi_y_r, nx, halo, i, jj - 1, k));
                }
                if (k <= halo + 1) {
                    double dpsi = (A3(psi_z_l, nx, ny, i, j, k) - A3(psi_z_l, nx, ny, i, j, k - 1)) * _dz;
                    A3(xi_z_l, nx, ny, i, j, k - 1) = C1(cz[2], k - 1) * A3(xi_z_l, nx, ny, i, j, k - 1) +
                                                      C1(cz[0], k - 1) * (d2p_dz2 + dpsi);
                    damp += f * (dpsi + A3(xi_z_l, nx, ny, i, j, k - 1));
                } else if (k >= nz - halo) {
                    long kk = k - (nz - halo) + 2;
                    double dpsi = (A3(psi_z_r, nx, ny, i, j, kk) - A3(psi_z_r, nx, ny, i, j, kk - 1)) * _dz;
                    A3(xi_z_r, nx, ny, i, j, kk - 1) = C1(cz[3], kk - 1) * A3(xi_z_r, nx, ny, i, j, kk - 1) +
                                                       C1(cz[1], kk - 1) * (d2p_dz2 + dpsi);
                    damp += f * (dpsi + A3(xi_z_r, nx, ny, i, j, kk - 1));
                }
                A3(pnew, nx, ny, i, j, k) = 2.0 * pc - A3(pold, nx, ny, i, j, k) +
                                            f * (d2p_dx2 + d2p_dy2 + d2p_dz2) + damp;
            }
}

/* ------------------------------------------------------------------ */
/* forward! and the time loop                                         */
/* ------------------------------------------------------------------ */

/* One forward! call for step `it` (1-based), followed by the caller's
 * rotation (pold, pcur, pnew) <- (pcur, pnew, pold), acoustic_2D_xPU.jl:332. */
void oracle_step(oracle_t *o, long it)
{
    const long nx = o->n[0], ny = o->n[1], nz = o->n[2];
    const int halo = o->halo;
    if (halo > 0) { /* with halo = 0 the psi values are never consumed */
        if (o->ndim == 2)
            oracle_update_psi_2d(nx, ny, halo, o->inv_d[0], o->inv_d[1], o->p[1],
                                 o->psi[0][0], o->psi[0][1], o->psi[1][0], o->psi[1][1],
                                 o->coef[0][4], o->coef[0][5], o->coef[0][6], o->coef[0][7],
                                 o->coef[1][4], o->coef[1][5], o->coef[1][6], o->coef[1][7]);
        else
            oracle_update_psi_3d(nx, ny, nz, halo, o->inv_d[0], o->inv_d[1], o->inv_d[2], o->p[1],
                                 o->psi[0][0], o->psi[0][1], o->psi[1][0], o->psi[1][1],
                                 o->psi[2][0], o->psi[2][1],
                                 o->coef[0][4], o->coef[0][5], o->coef[0][6], o->coef[0][7],
                                 o->coef[1][4], o->coef[1][5], o->coef[1][6], o->coef[1][7],
                                 o->coef[2][4], o->coef[2][5], o->coef[2][6], o->coef[2][7]);
    }
    if (o->ndim == 2) update_p_2d(o);
    else update_p_3d(o);

    double *pnew = o->p[2];
    /* inject_sources!, sequential in source order (the plain-loop twin
     * src/shared/acoustic_2D.jl:81-88 sums duplicates in this order) */
    for (int s = 0; s < o->nsrcs; ++s) {
        long i = o->possrcs[s], j = o->possrcs[(size_t)o->nsrcs + s];
        size_t idx;
        if (o->ndim == 2) idx = (size_t)(j - 1) * nx + (size_t)(i - 1);
        else {
            long k = o->possrcs[(size_t)2 * o->nsrcs + s];
            idx = ((size_t)(k - 1) * ny + (size_t)(j - 1)) * nx + (size_t)(i - 1);
        }
        pnew[idx] += o->dt2srctf[(size_t)s * o->nt + (it - 1)];
    }
    /* record_receivers! */
    for (int r = 0; r < o->nrecs; ++r) {
        long i = o->posrecs[r], j = o->posrecs[(size_t)o->nrecs + r];
        size_t idx;
        if (o->ndim == 2) idx = (size_t)(j - 1) * nx + (size_t)(i - 1);
        else {
            long k = o->posrecs[(size_t)2 * o->nrecs + r];
            idx = ((size_t)(k - 1) * ny + (size_t)(j - 1)) * nx + (size_t)(i - 1);
        }
        o->traces[(size_t)r * o->nt + (it - 1)] = pnew[idx];
    }
    /* rotation: forward! returns (pcur, pnew, pold) -> (pold, pcur, pnew) */
    double *t0 = o->p[0];
    o->p[0] = o->p[1];
    o->p[1] = o->p[2];
    o->p[2] = t0;
    o->it_done = it;
}

/* OpenMP threads of the following calls (n <= 0: the runtime's default,
 * OMP_NUM_THREADS or all cores); returns the count now in effect.  For the
 * CPU baseline's one-core figure; results do not depend on it. */
int oracle_set_threads(int n)
{
#ifdef _OPENMP
    static int dflt = 0;
    if (!dflt) dflt = omp_get_max_threads();
    omp_set_num_threads(n > 0 ? n : dflt);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

/* Run steps it_done+1 .. min(it_done+nsteps, nt). Returns steps run. */
long oracle_run(oracle_t *o, long nsteps)
{
    long done = 0;
    while (done < nsteps && o->it_done < o->nt) {
        oracle_step(o, o->it_done + 1);
        ++done;
    }
    return done;
}

/* ------------------------------------------------------------------ */
/* discrete adjoint (FWI gradient), FP64, 2D and 3D                   */
/* ------------------------------------------------------------------ */
/*
 * The reference has no gradient ("WIP", README.md:297).  This is the exact
 * reverse-mode transpose of the time step above, written operation by
 * operation in scatter form, for the misfit J = 1/2 sum_{it,r}
 * (trace[it,r] - obs[it,r])^2.  It returns dJ/dfact at every node with dt
 * and the CPML coefficients held fixed (fact = dt^2 vel^2, so
 * dJ/dvel = dJ/dfact * 2 dt^2 vel).  Dirichlet nodes never depend on the
 * state or on fact (they only accumulate sources), so the transpose runs
 * over interior nodes only.  The whole forward history is kept in memory:
 * meant for the small grids of the gradient checks.
 */

/* zero state and traces, step counter back to 0 */
void oracle_reset(oracle_t *o)
{
    size_t N = (size_t)o->n[0] * o->n[1] * o->n[2];
    for (int b = 0; b < 3; ++b) memset(o->p[b], 0, sizeof(double) * N);
    for (int a = 0; a < o->ndim; ++a)
        for (int s = 0; s < 2; ++s) {
            memset(o->psi[a][s], 0, sizeof(double) * slab_size(o, a, o->halo + 1));
            memset(o->xi[a][s], 0, sizeof(double) * slab_size(o, a, o->halo));
        }
    memset(o->traces, 0, sizeof(double) * (size_t)o->nt * o->nrecs);
    o->it_done = 0;
}

/* replace fact (e.g. for finite-difference checks with dt fixed) */
void oracle_set_fact(oracle_t *o, const double *fact)
{
    size_t N = (size_t)o->n[0] * o->n[1] * o->n[2];
    memcpy(o->fact, fact, sizeof(double) * N);
}

/* linear index of node c with coordinate `axis` replaced by `layer` in an
 * array whose extent along `axis` is `layers` */
static size_t slab_lin(const oracle_t *o, int axis, const long *c, long layer, long layers)
{
    size_t idx = 0, stride = 1;
    for (int a = 0; a < o->ndim; ++a) {
        long v = (a == axis) ? layer : c[a];
        long ext = (a == axis) ? layers : o->n[a];
        idx += (size_t)v * stride;
        stride *= (size_t)ext;
    }
    return idx;
}

/*
 * FWI gradient by the discrete adjoint of the time loop above (reverse-mode
 * transpose of every statement, scatter form, full forward history kept in
 * memory -- small cases only).
 *
 *   J(fact) = 1/2 sum_{it,r} (traces[it,r] - obs[it,r])^2     (obs: nt x nrecs)
 *   grad[i] = dJ/dfact[i] with dt, the CPML coefficients and the source
 *             wavelets held fixed (they depend on vmax, not on each node).
 *
 * The reference has no gradient yet (README.md:297 lists the adjoint and
 * checkpointing as work in progress); this is the exact gradient of the
 * reference's discrete forward map restated above.  dJ/dvel = grad * 2 dt^2
 * vel.  Pinned by central finite differences of oracle_run's misfit
 * (tests/test_oracle_adjoint.py).
 * Returns 0, or -1 when the history does not fit in memory.
 */
int oracle_fwi_gradient(oracle_t *o, const double *obs, double *grad, double *misfit)
{
    const int nd = o->ndim;
    const long h = o->halo, hp1 = h + 1;
    const long nt = o->nt;
    const size_t N = (size_t)o->n[0] * o->n[1] * o->n[2];
    size_t stride[3] = {1, (size_t)o->n[0], (size_t)o->n[0] * o->n[1]};
    size_t psz[3] = {0, 0, 0}, xsz[3] = {0, 0, 0};
    for (int a = 0; a < nd; ++a) {
        psz[a] = slab_size(o, a, hp1);
        xsz[a] = slab_size(o, a, h);
    }
    /* forward with history: pcur at the start of each step, psi/xi after */
    double *Bh = (double *)malloc(sizeof(double) * N * (size_t)(nt > 0 ? nt : 1));
    double *PSh[3][2] = {{0}}, *XIh[3][2] = {{0}};
    int ok = Bh != NULL;
    for (int a = 0; a < nd && ok; ++a)
        for (int s = 0; s < 2 && ok; ++s) {
            PSh[a][s] = (double *)malloc(sizeof(double) * psz[a] * (size_t)(nt > 0 ? nt : 1));
            XIh[a][s] = (double *)malloc(sizeof(double) * (xsz[a] ? xsz[a] : 1) * (size_t)(nt > 0 ? nt : 1));
            ok = PSh[a][s] && XIh[a][s];
        }
    double *lamA = (double *)calloc(N, sizeof(double));
    double *lamB = (double *)calloc(N, sizeof(double));
    double *lamC = (double *)calloc(N, sizeof(double));
    double *lamAn = (double *)calloc(N, sizeof(double));
    double *lamBn = (double *)calloc(N, sizeof(double));
    double *lP[3][2] = {{0}}, *lPn[3][2] = {{0}}, *lX[3][2] = {{0}}, *lXn[3][2] = {{0}};
    ok = ok && lamA && lamB && lamC && lamAn && lamBn;
    for (int a = 0; a < nd && ok; ++a)
        for (int s = 0; s < 2 && ok; ++s) {
            lP[a][s] = zalloc(psz[a]);
            lPn[a][s] = zalloc(psz[a]);
            lX[a][s] = zalloc(xsz[a]);
            lXn[a][s] = zalloc(xsz[a]);
            ok = lP[a][s] && lPn[a][s] && lX[a][s] && lXn[a][s];
        }
    int status = ok ? ORACLE_OK : ORACLE_ERR_NOMEM;
    if (!ok) goto done;

    oracle_reset(o);
    for (long it = 1; it <= nt; ++it) {
        memcpy(Bh + (size_t)(it - 1) * N, o->p[1], sizeof(double) * N);
        oracle_step(o, it);
        for (int a = 0; a < nd; ++a)
            for (int s = 0; s < 2; ++s) {
                memcpy(PSh[a][s] + (size_t)(it - 1) * psz[a], o->psi[a][s], sizeof(double) * psz[a]);
                if (xsz[a])
                    memcpy(XIh[a][s] + (size_t)(it - 1) * xsz[a], o->xi[a][s], sizeof(double) * xsz[a]);
            }
    }
    double J = 0.0;
    for (size_t q = 0; q < (size_t)nt * o->nrecs; ++q) {
        double r = o->traces[q] - obs[q];
        J += 0.5 * r * r;
    }
    if (misfit) *misfit = J;
    memset(grad, 0, sizeof(double) * N);

    for (long it = nt; it >= 1; --it) {
        const double *B = Bh + (size_t)(it - 1) * N;
        /* adjoint of this step's output C: from the next step's pcur slot
         * plus the receiver residuals (record_receivers! transposed) */
        memcpy(lamC, lamBn, sizeof(double) * N);
        for (int r = 0; r < o->nrecs; ++r) {
            size_t idx = 0;
            for (int a = 0; a < nd; ++a) idx += (size_t)(o->posrecs[(size_t)a * o->nrecs + r] - 1) * stride[a];
            lamC[idx] += o->traces[(size_t)r * nt + (it - 1)] - obs[(size_t)r * nt + (it - 1)];
        }
        /* B is passed on as the next step's pold */
        memcpy(lamB, lamAn, sizeof(double) * N);
        memset(lamA, 0, sizeof(double) * N);
        for (int a = 0; a < nd; ++a)
            for (int s = 0; s < 2; ++s) {
                memset(lPn[a][s], 0, sizeof(double) * psz[a]);
                if (xsz[a]) memset(lXn[a][s], 0, sizeof(double) * xsz[a]);
            }
        /* update_p! transposed */
        long c[3] = {0, 0, 0};
        const long lo[3] = {1, 1, nd == 3 ? 1 : 0};
        const long hi[3] = {o->n[0] - 1, o->n[1] - 1, nd == 3 ? o->n[2] - 1 : 1};
        for (c[2] = lo[2]; c[2] < hi[2]; ++c[2])
            for (c[1] = lo[1]; c[1] < hi[1]; ++c[1])
                for (c[0] = lo[0]; c[0] < hi[0]; ++c[0]) {
                    size_t x = (size_t)c[0] + (size_t)c[1] * stride[1] + (size_t)c[2] * stride[2];
                    const double lc = lamC[x];
                    const double f = o->fact[x];
                    lamA[x] -= lc;
                    lamB[x] += 2.0 * lc;
                    double qsum = 0.0;
                    for (int a = 0; a < nd; ++a) {
                        const double d2 = (B[x + stride[a]] - 2.0 * B[x] + B[x - stride[a]]) * o->inv_d2[a];
                        qsum += d2;
                        double lam_d2 = f * lc;
                        long q = c[a], m = -1;
                        int side = -1;
                        if (q <= h) { side = 0; m = q; }
                        else if (q >= o->n[a] - h - 1) { side = 1; m = q - (o->n[a] - h - 2); }
                        if (side >= 0) {
                            const double *Ps = PSh[a][side] + (size_t)(it - 1) * psz[a];
                            const double *Xs = XIh[a][side] + (size_t)(it - 1) * xsz[a];
                            const size_t pm = slab_lin(o, a, c, m, hp1), pm1 = slab_lin(o, a, c, m - 1, hp1);
                            const size_t xm1 = slab_lin(o, a, c, m - 1, h);
                            const double dpsi = (Ps[pm] - Ps[pm1]) * o->inv_d[a];
                            qsum += dpsi + Xs[xm1];
                            const double bK = o->coef[a][2 + side][m - 1], aa = o->coef[a][side][m - 1];
                            /* xi' is read by damp (f) and by the next step */
                            const double lxi = lX[a][side][xm1] + f * lc;
                            lXn[a][side][xm1] += bK * lxi;
                            lam_d2 += aa * lxi;
                            const double ldpsi = f * lc + aa * lxi;
                            lP[a][side][pm] += ldpsi * o->inv_d[a];
                            lP[a][side][pm1] -= ldpsi * o->inv_d[a];
                        }
                        lamB[x + stride[a]] += lam_d2 * o->inv_d2[a];
                        lamB[x] -= 2.0 * lam_d2 * o->inv_d2[a];
                        lamB[x - stride[a]] += lam_d2 * o->inv_d2[a];
                    }
                    grad[x] += lc * qsum;
                }
        /* update_psi_*! transposed (all layers, all rows, as launched) */
        if (h > 0)
            for (int a = 0; a < nd; ++a)
                for (int s = 0; s < 2; ++s) {
                    const double *ah = o->coef[a][4 + s], *bKh = o->coef[a][6 + s];
                    const long shift = s ? o->n[a] - h - 2 : 0;
                    long e[3] = {0, 0, 0};
                    long ext[3] = {o->n[0], o->n[1], nd == 3 ? o->n[2] : 1};
                    ext[a] = hp1;
                    for (e[2] = 0; e[2] < ext[2]; ++e[2])
                        for (e[1] = 0; e[1] < ext[1]; ++e[1])
                            for (e[0] = 0; e[0] < ext[0]; ++e[0]) {
                                const long m = e[a];
                                const size_t pi = slab_lin(o, a, e, m, hp1);
                                const double lp = lP[a][s][pi];
                                if (lp == 0.0) continue;
                                lPn[a][s][pi] += bKh[m] * lp;
                                long nc[3] = {e[0], e[1], e[2]};
                                nc[a] = m + shift;
                                size_t k = 0;
                                for (int b = 0; b < nd; ++b) k += (size_t)nc[b] * stride[b];
                                lamB[k + stride[a]] += ah[m] * lp * o->inv_d[a];
                                lamB[k] -= ah[m] * lp * o->inv_d[a];
                            }
                }
        memcpy(lamAn, lamA, sizeof(double) * N);
        memcpy(lamBn, lamB, sizeof(double) * N);
        for (int a = 0; a < nd; ++a)
            for (int s = 0; s < 2; ++s) {
                memcpy(lP[a][s], lPn[a][s], sizeof(double) * psz[a]);
                if (xsz[a]) memcpy(lX[a][s], lXn[a][s], sizeof(double) * xsz[a]);
            }
    }

done:
    free(Bh);
    free(lamA);
    free(lamB);
    free(lamC);
    free(lamAn);
    free(lamBn);
    for (int a = 0; a < 3; ++a)
        for (int s = 0; s < 2; ++s) {
            free(PSh[a][s]);
            free(XIh[a][s]);
            free(lP[a][s]);
            free(lPn[a][s]);
            free(lX[a][s]);
            free(lXn[a][s]);
        }
    return status;
}

/* ------------------------------------------------------------------ */
/* accessors                                                          */
/* ------------------------------------------------------------------ */

long oracle_nt(const oracle_t *o) { return o->nt; }
double oracle_dt(const oracle_t *o) { return o->dt; }
long oracle_it_done(const oracle_t *o) { return o->it_done; }
void oracle_get_traces(const oracle_t *o, double *out)
{
    memcpy(out, o->traces, sizeof(double) * (size_t)o->nt * (size_t)o->nrecs);
}
/* which = 0: pold, 1: pcur (the field solve returns), 2: pnew */
void oracle_get_field(const oracle_t *o, int which, double *out)
{
    size_t N = (size_t)o->n[0] * o->n[1] * o->n[2];
    memcpy(out, o->p[which], sizeof(double) * N);
}
void oracle_get_fact(const oracle_t *o, double *out)
{
    size_t N = (size_t)o->n[0] * o->n[1] * o->n[2];
    memcpy(out, o->fact, sizeof(double) * N);
}
/* axis 0..ndim-1; which: 0 a_l, 1 a_r, 2 b_K_l, 3 b_K_r, 4 a_hl, 5 a_hr,
 * 6 b_K_hl, 7 b_K_hr */
void oracle_get_coef(const oracle_t *o, int axis, int which, double *out)
{
    size_t len = (size_t)o->halo + (which >= 4 ? 1 : 0);
    memcpy(out, o->coef[axis][which], sizeof(double) * len);
}
void oracle_get_dt2srctf(const oracle_t *o, double *out)
{
    memcpy(out, o->dt2srctf, sizeof(double) * (size_t)o->nt * (size_t)o->nsrcs);
}
void oracle_get_positions(const oracle_t *o, long *possrcs, long *posrecs)
{
    memcpy(possrcs, o->possrcs, sizeof(long) * (size_t)o->nsrcs * o->ndim);
    memcpy(posrecs, o->posrecs, sizeof(long) * (size_t)o->nrecs * o->ndim);
}
/* psi/xi state, for white-box tests of the CPML slabs (layout as in the
 * reference arrays) */
void oracle_get_psi(const oracle_t *o, int axis, int side, double *out)
{
    memcpy(out, o->psi[axis][side], sizeof(double) * slab_size(o, axis, o->halo + 1));
}
void oracle_get_xi(const oracle_t *o, int axis, int side, double *out)
{
    memcpy(out, o->xi[axis][side], sizeof(double) * slab_size(o, axis, o->halo));
}
