/*
 * organoid_oracle.c -- CPU oracle of the organoid timestep.  TEST
 * INFRASTRUCTURE: loaded only by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py.  Never linked into the product library.
 *
 * A plain-C restatement of the reference's per-timestep functions, kept on
 * full N x N planes with the reference's loop bounds, so that what the
 * reference computes (including its quirks: mode-indexed viscosity, dropped
 * imaginary parts, discarded last reshaping iterate, integer centroids) is
 * what this computes.  Citations are file:line in the reference repository.
 *
 * Deliberate, documented deviations (same in the GPU library):
 *   - reductions (cell areas s, sum_ecm, reshaping rdev) accumulate in
 *     double; the reference accumulates in float;
 *   - the reference's window loops read out of range when a window touches
 *     the grid edge; here that is an error (ORG_EWINDOW) checked on the
 *     window plus one pixel of stencil halo;
 *   - a cell with no phi>0.5 pixel is an error (ORG_EEMPTY_CELL) instead of
 *     a division by zero; fewer than two cells is an error (the reference's
 *     ring of neighbours reads cell 1 when Ntot==1, rho_functions.cpp:298-302).
 */
#include "organoid_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef ORC_REAL
#define ORC_REAL double
#endif
typedef ORC_REAL real;
typedef struct { real x, y; } cplx;

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* The eleven FFT-box buffers of declare_fft_tables.cpp:10-41 plus the
 * transform's double work array.  The reference holds one set and loops over
 * cells serially (rho_functions.cpp:413-467); here each OpenMP thread owns a
 * set, so cells run in parallel with the same per-cell arithmetic. */
typedef struct {
    cplx *temp_fft, *temp_fft_real, *fft_rho, *fft_react, *fft_pi_xx, *fft_pi_xy,
         *fft_pi_yy, *fft_vx, *fft_vy, *fft_rho_vx, *fft_rho_vy;
    double *fftwork;
} orc_spec;

struct orc_state {
    org_params p;
    int N, nc, zs, M;
    size_t NN, P;
    /* per-cell planes (declare_tables.cpp:97-213) */
    real *phi, *phin, *rho, *rhon, *hfunc, *temp_phase, *ngp, *pnd, *Jx, *Jy;
    real *deg, *gen, *atan1, *aort1, *visc, *b, *pixx, *pixy, *piyy;
    /* shared planes */
    real *lum, *lumn, *ECM, *ECMn, *psi, *hlum, *hECM, *nglum, *Jlx, *Jly;
    /* FFT-box buffers (declare_fft_tables.cpp:10-41) */
    orc_spec *spec;          /* one FFT-box workspace per OpenMP thread */
    int nspec;
    real *vx_box, *vy_box;
    real *f0;
    double *s, *s0_cell;
    int32_t *xc, *yc, *count_ri;
    int32_t count_lum, count_ecm;
    int64_t relax_failures;   /* reshaping loops that passed rimax */
    int err;
};

/* ------------------------------------------------------------------------ */
/* DFT: radix-2 Cooley-Tukey for powers of two, Bluestein's chirp-z for any
 * other length (cuFFT's plans for M = 2*zone_size-2 = 398 at the template
 * defaults need a non-power-of-two length).  Written from the DFT
 * definition X_k = sum_j x_j exp(sign*2*pi*i*j*k/n).                        */
static int is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

static void fft_pow2(double *re, double *im, int n, int sign) {
    for (int i = 1, j = 0; i < n; i++) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            double t = re[i]; re[i] = re[j]; re[j] = t;
            t = im[i]; im[i] = im[j]; im[j] = t;
        }
    }
    double *cw = (double *)malloc(sizeof(double) * (n / 2 + 1));
    double *sw = (double *)malloc(sizeof(double) * (n / 2 + 1));
    for (int k = 0; k < n / 2; k++) {
        cw[k] = cos(2.0 * M_PI * k / n);
        sw[k] = sign * sin(2.0 * M_PI * k / n);
    }
    for (int len = 2; len <= n; len <<= 1) {
        int half = len / 2, step = n / len;
        for (int i = 0; i < n; i += len) {
            for (int k = 0; k < half; k++) {
                double wr = cw[k * step], wi = sw[k * step];
                int a = i + k, b = i + k + half;
                double xr = re[b] * wr - im[b] * wi;
                double xi = re[b] * wi + im[b] * wr;
                re[b] = re[a] - xr; im[b] = im[a] - xi;
                re[a] += xr; im[a] += xi;
            }
        }
    }
    free(cw); free(sw);
}

static void fft_any(double *re, double *im, int n, int sign) {
    if (n <= 1) return;
    if (is_pow2(n)) { fft_pow2(re, im, n, sign); return; }
    int m = 1;
    while (m < 2 * n - 1) m <<= 1;
    double *wr = (double *)malloc(sizeof(double) * n), *wi = (double *)malloc(sizeof(double) * n);
    double *ar = (double *)calloc(m, sizeof(double)), *ai = (double *)calloc(m, sizeof(double));
    double *br = (double *)calloc(m, sizeof(double)), *bi = (double *)calloc(m, sizeof(double));
    for (int k = 0; k < n; k++) {
        long long k2 = ((long long)k * k) % (2LL * n);   /* exact argument reduction */
        double ang = sign * M_PI * (double)k2 / n;
        wr[k] = cos(ang); wi[k] = sin(ang);
    }
    for (int k = 0; k < n; k++) {          /* a_k = x_k w_k */
        ar[k] = re[k] * wr[k] - im[k] * wi[k];
        ai[k] = re[k] * wi[k] + im[k] * wr[k];
    }
    br[0] = wr[0]; bi[0] = -wi[0];          /* b_k = conj(w_k), b_{-k} = b_k */
    for (int k = 1; k < n; k++) {
        br[k] = wr[k]; bi[k] = -wi[k];
        br[m - k] = wr[k]; bi[m - k] = -wi[k];
    }
    fft_pow2(ar, ai, m, -1);
    fft_pow2(br, bi, m, -1);
    for (int k = 0; k < m; k++) {
        double r = ar[k] * br[k] - ai[k] * bi[k];
        double i = ar[k] * bi[k] + ai[k] * br[k];
        ar[k] = r; ai[k] = i;
    }
    fft_pow2(ar, ai, m, +1);
    for (int k = 0; k < n; k++) {          /* X_k = w_k (a*b)_k / m */
        double cr = ar[k] / m, ci = ai[k] / m;
        re[k] = cr * wr[k] - ci * wi[k];
        im[k] = cr * wi[k] + ci * wr[k];
    }
    free(wr); free(wi); free(ar); free(ai); free(br); free(bi);
}

/* 2D transform on interleaved doubles: rows, then columns. */
static void fft2d_inplace(int M, double *d, int sign) {
    double *re = (double *)malloc(sizeof(double) * M), *im = (double *)malloc(sizeof(double) * M);
    for (int i = 0; i < M; i++) {
        for (int j = 0; j < M; j++) { re[j] = d[2 * (i * M + j)]; im[j] = d[2 * (i * M + j) + 1]; }
        fft_any(re, im, M, sign);
        for (int j = 0; j < M; j++) { d[2 * (i * M + j)] = re[j]; d[2 * (i * M + j) + 1] = im[j]; }
    }
    for (int j = 0; j < M; j++) {
        for (int i = 0; i < M; i++) { re[i] = d[2 * (i * M + j)]; im[i] = d[2 * (i * M + j) + 1]; }
        fft_any(re, im, M, sign);
        for (int i = 0; i < M; i++) { d[2 * (i * M + j)] = re[i]; d[2 * (i * M + j) + 1] = im[i]; }
    }
    free(re); free(im);
}

int orc_fft1d(int n, double *data, int sign) {
    if (n <= 0 || !data || (sign != 1 && sign != -1)) return ORG_EINVAL;
    double *re = (double *)malloc(sizeof(double) * n), *im = (double *)malloc(sizeof(double) * n);
    for (int k = 0; k < n; k++) { re[k] = data[2 * k]; im[k] = data[2 * k + 1]; }
    fft_any(re, im, n, sign);
    for (int k = 0; k < n; k++) { data[2 * k] = re[k]; data[2 * k + 1] = im[k]; }
    free(re); free(im);
    return ORG_OK;
}

int orc_fft2d(int M, double *data, int sign) {
    if (M <= 0 || !data || (sign != 1 && sign != -1)) return ORG_EINVAL;
    fft2d_inplace(M, data, sign);
    return ORG_OK;
}

/* cufftExecC2C(plan, in, out, dir) as called by compute_fft_forward /
 * compute_fft_backward (useful_functions.cpp:646-660): out-of-place,
 * unnormalised.  Computed in double, stored in the buffer type. */
static void exec_c2c(const orc_state *st, orc_spec *sp, const cplx *in, cplx *out, int sign) {
    int M = st->M;
    double *w = sp->fftwork;
    for (int k = 0; k < M * M; k++) { w[2 * k] = in[k].x; w[2 * k + 1] = in[k].y; }
    fft2d_inplace(M, w, sign);
    for (int k = 0; k < M * M; k++) { out[k].x = (real)w[2 * k]; out[k].y = (real)w[2 * k + 1]; }
}

/* ------------------------------------------------------------------------ */
#define PL(n, i, j) ((size_t)(n) * st->NN + (size_t)(i) * st->N + (size_t)(j))
#define G(i, j) ((size_t)(i) * st->N + (size_t)(j))

static void *xcalloc(size_t n, size_t sz) { return calloc(n ? n : 1, sz); }

orc_state *orc_create(const org_params *p, const double *s0_cell) {
    if (!p || p->N < 8 || p->ncell < 2 || p->zone_size < 3) return NULL;
    orc_state *st = (orc_state *)calloc(1, sizeof(orc_state));
    st->p = *p;
    st->N = p->N; st->nc = p->ncell; st->zs = p->zone_size; st->M = 2 * p->zone_size - 2;
    st->NN = (size_t)st->N * st->N; st->P = st->NN * st->nc;
    size_t P = st->P, NN = st->NN, MM = (size_t)st->M * st->M;
    real **cellp[] = {&st->phi, &st->phin, &st->rho, &st->rhon, &st->hfunc, &st->temp_phase,
                      &st->ngp, &st->pnd, &st->Jx, &st->Jy, &st->deg, &st->gen, &st->atan1,
                      &st->aort1, &st->visc, &st->b, &st->pixx, &st->pixy, &st->piyy};
    for (size_t k = 0; k < sizeof(cellp) / sizeof(cellp[0]); k++) *cellp[k] = (real *)xcalloc(P, sizeof(real));
    real **gp[] = {&st->lum, &st->lumn, &st->ECM, &st->ECMn, &st->psi, &st->hlum, &st->hECM,
                   &st->nglum, &st->Jlx, &st->Jly};
    for (size_t k = 0; k < sizeof(gp) / sizeof(gp[0]); k++) *gp[k] = (real *)xcalloc(NN, sizeof(real));
#ifdef _OPENMP
    st->nspec = omp_get_max_threads();
#else
    st->nspec = 1;
#endif
    st->spec = (orc_spec *)xcalloc((size_t)st->nspec, sizeof(orc_spec));
    for (int t = 0; t < st->nspec; t++) {
        orc_spec *sp = &st->spec[t];
        cplx **fp[] = {&sp->temp_fft, &sp->temp_fft_real, &sp->fft_rho, &sp->fft_react, &sp->fft_pi_xx,
                       &sp->fft_pi_xy, &sp->fft_pi_yy, &sp->fft_vx, &sp->fft_vy, &sp->fft_rho_vx,
                       &sp->fft_rho_vy};
        for (size_t k = 0; k < sizeof(fp) / sizeof(fp[0]); k++) *fp[k] = (cplx *)xcalloc(MM, sizeof(cplx));
        sp->fftwork = (double *)xcalloc(2 * MM, sizeof(double));
    }
    st->vx_box = (real *)xcalloc(MM * st->nc, sizeof(real));
    st->vy_box = (real *)xcalloc(MM * st->nc, sizeof(real));
    st->f0 = (real *)xcalloc(st->nc, sizeof(real));
    st->s = (double *)xcalloc(st->nc, sizeof(double));
    st->s0_cell = (double *)xcalloc(st->nc, sizeof(double));
    for (int n = 0; n < st->nc; n++)
        st->s0_cell[n] = (p->use_s0_cell && s0_cell) ? s0_cell[n] : p->s0;
    st->xc = (int32_t *)xcalloc(st->nc, sizeof(int32_t));
    st->yc = (int32_t *)xcalloc(st->nc, sizeof(int32_t));
    st->count_ri = (int32_t *)xcalloc(st->nc, sizeof(int32_t));
    return st;
}

void orc_destroy(orc_state *st) {
    if (!st) return;
    real *cellp[] = {st->phi, st->phin, st->rho, st->rhon, st->hfunc, st->temp_phase, st->ngp,
                     st->pnd, st->Jx, st->Jy, st->deg, st->gen, st->atan1, st->aort1, st->visc,
                     st->b, st->pixx, st->pixy, st->piyy, st->lum, st->lumn, st->ECM, st->ECMn,
                     st->psi, st->hlum, st->hECM, st->nglum, st->Jlx, st->Jly, st->vx_box,
                     st->vy_box, st->f0};
    for (size_t k = 0; k < sizeof(cellp) / sizeof(cellp[0]); k++) free(cellp[k]);
    for (int t = 0; t < st->nspec; t++) {
        orc_spec *sp = &st->spec[t];
        cplx *fp[] = {sp->temp_fft, sp->temp_fft_real, sp->fft_rho, sp->fft_react, sp->fft_pi_xx,
                      sp->fft_pi_xy, sp->fft_pi_yy, sp->fft_vx, sp->fft_vy, sp->fft_rho_vx, sp->fft_rho_vy};
        for (size_t k = 0; k < sizeof(fp) / sizeof(fp[0]); k++) free(fp[k]);
        free(sp->fftwork);
    }
    free(st->spec); free(st->s); free(st->s0_cell); free(st->xc); free(st->yc); free(st->count_ri);
    free(st);
}

int orc_real_bytes(void) { return (int)sizeof(real); }
int orc_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
int orc_params_bytes(void) { return (int)sizeof(org_params); }

static real *field_ptr(const orc_state *st, int f, size_t *n) {
    size_t P = st->P, NN = st->NN, MB = (size_t)st->M * st->M * st->nc;
    switch (f) {
    case ORG_F_PHI: *n = P; return st->phi;
    case ORG_F_RHO: *n = P; return st->rho;
    case ORG_F_HFUNC: *n = P; return st->hfunc;
    case ORG_F_LUM: *n = NN; return st->lum;
    case ORG_F_ECM: *n = NN; return st->ECM;
    case ORG_F_PSI: *n = NN; return st->psi;
    case ORG_F_HLUM: *n = NN; return st->hlum;
    case ORG_F_HECM: *n = NN; return st->hECM;
    case ORG_F_LUMN: *n = NN; return st->lumn;
    case ORG_F_ECMN: *n = NN; return st->ECMn;
    case ORC_F_PHIN: *n = P; return st->phin;
    case ORC_F_RHON: *n = P; return st->rhon;
    case ORC_F_NGP: *n = P; return st->ngp;
    case ORC_F_PND: *n = P; return st->pnd;
    case ORC_F_JX: *n = P; return st->Jx;
    case ORC_F_JY: *n = P; return st->Jy;
    case ORC_F_TEMP: *n = P; return st->temp_phase;
    case ORC_F_DEG: *n = P; return st->deg;
    case ORC_F_GEN: *n = P; return st->gen;
    case ORC_F_ATAN: *n = P; return st->atan1;
    case ORC_F_AORT: *n = P; return st->aort1;
    case ORC_F_VISC: *n = P; return st->visc;
    case ORC_F_B: *n = P; return st->b;
    case ORC_F_PIXX: *n = P; return st->pixx;
    case ORC_F_PIXY: *n = P; return st->pixy;
    case ORC_F_PIYY: *n = P; return st->piyy;
    case ORC_F_NGLUM: *n = NN; return st->nglum;
    case ORC_F_JLX: *n = NN; return st->Jlx;
    case ORC_F_JLY: *n = NN; return st->Jly;
    case ORC_F_VX: *n = MB; return st->vx_box;
    case ORC_F_VY: *n = MB; return st->vy_box;
    default: *n = 0; return NULL;
    }
}

int64_t orc_field_size(const orc_state *st, int field) {
    size_t n; return field_ptr(st, field, &n) ? (int64_t)n : -1;
}
int orc_set_field(orc_state *st, int field, const double *src) {
    size_t n; real *d = field_ptr(st, field, &n);
    if (!d || !src) return ORG_EINVAL;
    for (size_t k = 0; k < n; k++) d[k] = (real)src[k];
    return ORG_OK;
}
int orc_get_field(const orc_state *st, int field, double *dst) {
    size_t n; real *d = field_ptr(st, field, &n);
    if (!d || !dst) return ORG_EINVAL;
    for (size_t k = 0; k < n; k++) dst[k] = (double)d[k];
    return ORG_OK;
}
int orc_get_centres(const orc_state *st, int32_t *xc, int32_t *yc) {
    memcpy(xc, st->xc, sizeof(int32_t) * st->nc); memcpy(yc, st->yc, sizeof(int32_t) * st->nc);
    return ORG_OK;
}
int orc_set_centres(orc_state *st, const int32_t *xc, const int32_t *yc) {
    memcpy(st->xc, xc, sizeof(int32_t) * st->nc); memcpy(st->yc, yc, sizeof(int32_t) * st->nc);
    return ORG_OK;
}
int orc_get_areas(const orc_state *st, double *s, double *f0) {
    for (int n = 0; n < st->nc; n++) { s[n] = st->s[n]; f0[n] = (double)st->f0[n]; }
    return ORG_OK;
}
int orc_get_iterations(const orc_state *st, int32_t *cells, int32_t *lum, int32_t *ecm) {
    memcpy(cells, st->count_ri, sizeof(int32_t) * st->nc);
    *lum = st->count_lum; *ecm = st->count_ecm;
    return ORG_OK;
}
int orc_last_error(const orc_state *st) { return st->err; }
int64_t orc_relax_failures(const orc_state *st) { return st->relax_failures; }

static int set_err(orc_state *st, int e) { if (!st->err) st->err = e; return e; }

/* run_with_active_gels.cpp:92-98 */
int orc_init_like_driver(orc_state *st) {
    memcpy(st->ECMn, st->ECM, st->NN * sizeof(real));
    memcpy(st->lumn, st->lum, st->NN * sizeof(real));
    memcpy(st->phin, st->phi, st->P * sizeof(real));
    memcpy(st->rhon, st->rho, st->P * sizeof(real));
    return ORG_OK;
}

/* Window [xc-zs, xc+zs) plus one pixel of stencil halo must stay inside the
 * interior rows/columns 1..N-2 (see file header). */
static int check_windows(orc_state *st) {
    for (int n = 0; n < st->nc; n++) {
        int x0 = st->xc[n] - st->zs, y0 = st->yc[n] - st->zs;
        if (x0 - 1 < 1 || x0 + 2 * st->zs > st->N - 2 || y0 - 1 < 1 || y0 + 2 * st->zs > st->N - 2)
            return set_err(st, ORG_EWINDOW);
    }
    return ORG_OK;
}

/* ---- useful_functions.cpp ---------------------------------------------- */

/* find_centres_cells, useful_functions.cpp:193-227: integer sums of the row
 * and column indices of phi>0.5 pixels, truncating division. */
int orc_find_centres_cells(orc_state *st) {
    int N = st->N, empty = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(|: empty)
    for (int n = 0; n < st->nc; n++) {
        long long xs = 0, ys = 0, cnt = 0;
        for (int i = 0; i < N; i++)
            for (int j = 0; j < N; j++)
                if (st->phi[PL(n, i, j)] > 0.5) { xs += i; ys += j; cnt += 1; }
        if (cnt == 0) { empty = 1; continue; }
        st->xc[n] = (int32_t)(xs / cnt);
        st->yc[n] = (int32_t)(ys / cnt);
    }
    if (empty) return set_err(st, ORG_EEMPTY_CELL);
    return check_windows(st);
}

/* compute_hfunc_3d_tables_kernels_loop_n, useful_functions.cpp:233-251 */
int orc_compute_hfunc_cells(orc_state *st) {
    if (check_windows(st)) return st->err;
    int zs = st->zs;
#pragma omp parallel for schedule(static)
    for (int n = 0; n < st->nc; n++)
        for (int i = st->xc[n] - zs; i < st->xc[n] + zs; i++)
            for (int j = st->yc[n] - zs; j < st->yc[n] + zs; j++) {
                real v = st->phi[PL(n, i, j)];
                st->hfunc[PL(n, i, j)] = v * v * (3.0 - 2.0 * v);
            }
    return ORG_OK;
}

/* compute_all_hfunc_2d_tables_kernels, useful_functions.cpp:272-285 */
int orc_compute_hfunc_lum_ecm(orc_state *st) {
#pragma omp parallel for schedule(static)
    for (size_t k = 0; k < st->NN; k++) {
        real a = st->lum[k], e = st->ECM[k];
        st->hlum[k] = a * a * (3.0 - 2.0 * a);
        st->hECM[k] = e * e * (3.0 - 2.0 * e);
    }
    return ORG_OK;
}

/* boundary_conditions_cells, useful_functions.cpp:391-420: periodic ghost
 * ring with period N-2 on every cell plane. */
static void bc_cells(int N, int nplanes, real *t) {
    size_t NN = (size_t)N * N;
#pragma omp parallel for schedule(static) if (nplanes > 1)
    for (int n = 0; n < nplanes; n++) {
        real *q = t + (size_t)n * NN;
        for (int i = 1; i < N - 1; i++) {
            q[(size_t)i * N] = q[(size_t)i * N + N - 2];
            q[(size_t)i * N + N - 1] = q[(size_t)i * N + 1];
            q[i] = q[(size_t)N * (N - 2) + i];
            q[(size_t)N * (N - 1) + i] = q[(size_t)N + i];
        }
        q[0] = q[(size_t)N * (N - 2) + N - 2];
        q[N - 1] = q[(size_t)N * (N - 2) + 1];
        q[(size_t)N * (N - 1)] = q[(size_t)N + N - 2];
        q[(size_t)N * (N - 1) + N - 1] = q[(size_t)N + 1];
    }
}

int orc_bc_cells_planes(int N, int nplanes, double *planes) {
    size_t n = (size_t)N * N * nplanes;
    real *t = (real *)malloc(n * sizeof(real));
    for (size_t k = 0; k < n; k++) t[k] = (real)planes[k];
    bc_cells(N, nplanes, t);
    for (size_t k = 0; k < n; k++) planes[k] = t[k];
    free(t);
    return ORG_OK;
}

/* compute_normal_grad_phi, useful_functions.cpp:326-354, then
 * boundary_conditions_cells(norm_grad_phi) (run_with_active_gels.cpp:136). */
int orc_compute_normal_grad_phi(orc_state *st) {
    int N = st->N;
    real dx = (real)st->p.dx;
#pragma omp parallel for collapse(2) schedule(static)
    for (int n = 0; n < st->nc; n++)
        for (int i = 1; i < N - 1; i++)
            for (int j = 1; j < N - 1; j++) {
                real a = st->phi[PL(n, i + 1, j)] - st->phi[PL(n, i - 1, j)];
                real c = st->phi[PL(n, i, j + 1)] - st->phi[PL(n, i, j - 1)];
                st->ngp[PL(n, i, j)] = sqrt((1.0 / dx / dx / 4.0) * (a * a + c * c));
            }
    bc_cells(N, st->nc, st->ngp);
    return ORG_OK;
}

/* compute_psi_kernels, useful_functions.cpp:305-322 */
int orc_compute_psi(orc_state *st) {
#pragma omp parallel for schedule(static)
    for (size_t k = 0; k < st->NN; k++) {
        real acc = 0.0;
        for (int n = 0; n < st->nc; n++) acc += st->hfunc[(size_t)n * st->NN + k];
        st->psi[k] = acc;
    }
    return ORG_OK;
}

/* sum_s_kernels, useful_functions.cpp:168-190: s_n = sum hfunc*dx*dx over
 * the whole plane (double accumulation), f0 = alpha*(s0 - s). */
int orc_sum_s(orc_state *st) {
    real dx = (real)st->p.dx;
#pragma omp parallel for schedule(static)
    for (int n = 0; n < st->nc; n++) {
        double acc = 0.0;
        const real *h = st->hfunc + (size_t)n * st->NN;
        for (size_t k = 0; k < st->NN; k++) acc += (double)(h[k] * dx * dx);
        st->s[n] = acc;
        st->f0[n] = (real)((real)st->p.alpha * ((real)st->s0_cell[n] - (real)acc));
    }
    return ORG_OK;
}

/* compute_normal_to_phi, useful_functions.cpp:356-388, then
 * boundary_conditions_cells(J_x), (J_y) (run_with_active_gels.cpp:164-166). */
int orc_compute_normal_to_phi(orc_state *st) {
    int N = st->N;
    real dx = (real)st->p.dx;
#pragma omp parallel for collapse(2) schedule(static)
    for (int n = 0; n < st->nc; n++)
        for (int i = 1; i < N - 1; i++)
            for (int j = 1; j < N - 1; j++) {
                real g = st->ngp[PL(n, i, j)];
                if (g > 0.00) {
                    st->Jx[PL(n, i, j)] = (1.0 / (2.0 * dx)) * (st->phi[PL(n, i + 1, j)] - st->phi[PL(n, i - 1, j)]) / g;
                    st->Jy[PL(n, i, j)] = (1.0 / (2.0 * dx)) * (st->phi[PL(n, i, j + 1)] - st->phi[PL(n, i, j - 1)]) / g;
                } else {
                    st->Jx[PL(n, i, j)] = 0.0;
                    st->Jy[PL(n, i, j)] = 0.0;
                }
            }
    bc_cells(N, st->nc, st->Jx);
    bc_cells(N, st->nc, st->Jy);
    return ORG_OK;
}

/* ---- rho_functions.cpp --------------------------------------------------- */

/* create_turnover_rates_tensions, rho_functions.cpp:286-405.  Neighbours
 * n1, n2 are index neighbours on a ring (:298-312); label by strict '>'
 * (:328-353). */
int orc_create_turnover_rates_tensions(orc_state *st) {
    if (check_windows(st)) return st->err;
    const org_params *p = &st->p;
    int zs = st->zs, nc = st->nc;
#pragma omp parallel for schedule(static)
    for (int n = 0; n < nc; n++) {
        int n1, n2;
        if (n == 0) { n1 = nc - 1; n2 = n + 1; }
        else if (n == nc - 1) { n1 = n - 1; n2 = 0; }
        else { n1 = n - 1; n2 = n + 1; }
        for (int i = st->xc[n] - zs; i < st->xc[n] + zs; i++)
            for (int j = st->yc[n] - zs; j < st->yc[n] + zs; j++) {
                real ph = st->phi[PL(n, i, j)];
                real c1l = ph * st->lum[G(i, j)];
                real c1e = ph * st->ECM[G(i, j)];
                int index; real mx;
                if (c1l > c1e) { index = 0; mx = c1l; } else { index = 1; mx = c1e; }
                real c1c2 = ph * st->phi[PL(n1, i, j)];
                if (c1c2 > mx) { index = 2; mx = c1c2; }
                real c1c3 = ph * st->phi[PL(n2, i, j)];
                if (c1c3 > mx) { index = 3; mx = c1c3; }
                size_t q = PL(n, i, j);
                if (index == 2 || index == 3) {         /* lateral */
                    st->deg[q] = (real)p->deg_lat; st->gen[q] = (real)p->gen_lat;
                    st->aort1[q] = (real)p->a_ort_lat; st->atan1[q] = (real)p->a_tan_lat;
                    st->visc[q] = (real)p->visc_lat; st->b[q] = (real)p->b_lat;
                } else if (index == 0) {                /* apical */
                    st->deg[q] = (real)p->deg_ap; st->gen[q] = (real)p->gen_ap;
                    st->aort1[q] = (real)p->a_ort_ap; st->atan1[q] = (real)p->a_tan_ap;
                    st->visc[q] = (real)p->visc_ap; st->b[q] = (real)p->b_ap;
                } else {                                /* basal */
                    st->deg[q] = (real)p->deg_bas; st->gen[q] = (real)p->gen_bas;
                    st->aort1[q] = (real)p->a_ort_bas; st->atan1[q] = (real)p->a_tan_bas;
                    st->visc[q] = (real)p->visc_bas; st->b[q] = (real)p->b_bas;
                }
            }
    }
    return ORG_OK;
}

/* compute_prod_norm_delta, rho_functions.cpp:13-30 */
int orc_compute_prod_norm_delta(orc_state *st) {
    if (check_windows(st)) return st->err;
    int zs = st->zs;
    real sigma = (real)st->p.sigma;
#pragma omp parallel for schedule(static)
    for (int n = 0; n < st->nc; n++)
        for (int i = st->xc[n] - zs; i < st->xc[n] + zs; i++)
            for (int j = st->yc[n] - zs; j < st->yc[n] + zs; j++) {
                real ph = st->phi[PL(n, i, j)];
                st->pnd[PL(n, i, j)] = st->ngp[PL(n, i, j)] * exp(-(ph - 0.5) * (ph - 0.5) / 2.0 / sigma / sigma)
                                       / sqrt(2.0 * M_PI) / sigma;
            }
    return ORG_OK;
}

/* compute_pi_xx / compute_pi_yy / compute_pi_xy, rho_functions.cpp:32-115,
 * over the window of cell n. */
static void compute_pi(orc_state *st, int n) {
    int zs = st->zs;
    real dx = (real)st->p.dx;
    for (int i = st->xc[n] - zs; i < st->xc[n] + zs; i++)
        for (int j = st->yc[n] - zs; j < st->yc[n] + zs; j++) {
            size_t q = PL(n, i, j);
            real r = st->rho[q], at = st->atan1[q], ao = st->aort1[q], bb = st->b[q], g = st->ngp[q];
            real dxp = st->phi[PL(n, i + 1, j)] - st->phi[PL(n, i - 1, j)];
            real dyp = st->phi[PL(n, i, j + 1)] - st->phi[PL(n, i, j - 1)];
            real base = -at * r * r * r + bb * r * r * r * r;
            st->pixx[q] = base;                                           /* :47 */
            if (g > 0.0) st->pixx[q] += (at - ao) * r * r * r * (1.0 / (4.0 * dx * dx)) * dxp * dxp / g / g;  /* :51 */
            st->pixy[q] = 0.0;                                            /* :105 */
            if (g > 0.0) st->pixy[q] += (at - ao) * r * r * r * (1.0 / (4.0 * dx * dx)) * dxp * dyp / g / g;  /* :109 */
            st->piyy[q] = base;                                           /* :75 */
            if (g > 0.0) st->piyy[q] += (at - ao) * r * r * r * (1.0 / (4.0 * dx * dx)) * dyp * dyp / g / g;  /* :79 */
        }
}

/* load_temp_fft, useful_functions.cpp:605-625: FFT box starts at +1. */
static void load_temp_fft(orc_state *st, int n, const real *table, cplx *t) {
    int M = st->M, zs = st->zs;
    for (int i = 0; i < M; i++)
        for (int j = 0; j < M; j++) {
            int ig = i + 1 + st->xc[n] - zs, jg = j + 1 + st->yc[n] - zs;
            t[i * M + j].x = table[PL(n, ig, jg)];
            t[i * M + j].y = 0.0;
        }
}

/* create_deg_gen, rho_functions.cpp:193-209 */
static void create_deg_gen(orc_state *st, int n, cplx *t) {
    int M = st->M, zs = st->zs;
    for (int i = 0; i < M; i++)
        for (int j = 0; j < M; j++) {
            size_t q = PL(n, i + 1 + st->xc[n] - zs, j + 1 + st->yc[n] - zs);
            t[i * M + j].x = -st->deg[q] * st->rho[q] + st->gen[q] * st->pnd[q];
            t[i * M + j].y = 0.0;
        }
}

/* The per-mode 2x2 friction-viscosity solve of update_rho_velocity,
 * rho_functions.cpp:164-186, for one mode. */
static void velocity_mode(real eta, real xi, real kx, real ky, cplx pxx, cplx pxy, cplx pyy,
                          cplx *vx, cplx *vy) {
    real D = -(4.0 * eta * eta * kx * kx * ky * ky + xi * xi + 3.0 * eta * xi * (kx * kx + ky * ky)
               + 2.0 * eta * eta * (kx * kx * kx * kx + ky * ky * ky * ky));
    real A = xi + 2.0 * eta * (kx * kx + ky * ky / 2.0);
    cplx y, x;
    y.x = 0.0; y.y = 0.0; x.x = 0.0; x.y = 0.0;
    y.x += eta * kx * ky * (kx * pxx.y + ky * pxy.y);                 /* :170 */
    y.x -= A * (ky * pyy.y + kx * pxy.y);                              /* :171 */
    y.x /= D;                                                          /* :172 */
    y.y -= eta * kx * ky * (kx * pxx.x + ky * pxy.x);                 /* :175 */
    y.y += A * (ky * pyy.x + kx * pxy.x);                              /* :176 */
    y.y /= D;                                                          /* :177 */
    x.x += kx * pxx.y + ky * pxy.y - eta * kx * ky * y.x;             /* :181 */
    x.x /= A;                                                          /* :182 */
    x.y += -kx * pxx.x - ky * pxy.x - eta * kx * ky * y.y;            /* :185 */
    x.y /= A;                                                          /* :186 */
    *vx = x; *vy = y;
}

static void mode_wavenumbers(int M, real dx, int i, int j, real *kx, real *ky) {
    int ir = (i < M / 2) ? i : i - M;                                  /* :159-162 */
    int jr = (j < M / 2) ? j : j - M;
    *kx = (2.0 * M_PI / (M * dx)) * ir;                                /* :164 */
    *ky = (2.0 * M_PI / (M * dx)) * jr;                                /* :165 */
}

/* update_rho_velocity, rho_functions.cpp:133-190.  Viscosity of mode (i,j)
 * is read at the real-space pixel (i+1+xc-zs, j+1+yc-zs) (:149-150). */
static void update_rho_velocity(orc_state *st, orc_spec *sp, int n) {
    int M = st->M, zs = st->zs;
    real dx = (real)st->p.dx, xi = (real)st->p.friction;
    for (int i = 0; i < M; i++)
        for (int j = 0; j < M; j++) {
            real kx, ky;
            mode_wavenumbers(M, dx, i, j, &kx, &ky);
            real eta = st->visc[PL(n, i + 1 + st->xc[n] - zs, j + 1 + st->yc[n] - zs)];
            velocity_mode(eta, xi, kx, ky, sp->fft_pi_xx[i * M + j], sp->fft_pi_xy[i * M + j],
                          sp->fft_pi_yy[i * M + j], &sp->fft_vx[i * M + j], &sp->fft_vy[i * M + j]);
        }
}

int orc_velocity_modes(const org_params *p, int M, const double *pxx, const double *pxy,
                       const double *pyy, const double *eta, double *vx, double *vy) {
    real dx = (real)p->dx, xi = (real)p->friction;
    for (int i = 0; i < M; i++)
        for (int j = 0; j < M; j++) {
            int k = i * M + j;
            real kx, ky;
            mode_wavenumbers(M, dx, i, j, &kx, &ky);
            cplx a = {(real)pxx[2 * k], (real)pxx[2 * k + 1]};
            cplx b = {(real)pxy[2 * k], (real)pxy[2 * k + 1]};
            cplx c = {(real)pyy[2 * k], (real)pyy[2 * k + 1]};
            cplx ox, oy;
            velocity_mode((real)eta[k], xi, kx, ky, a, b, c, &ox, &oy);
            vx[2 * k] = ox.x; vx[2 * k + 1] = ox.y; vy[2 * k] = oy.x; vy[2 * k + 1] = oy.y;
        }
    return ORG_OK;
}

/* load_fft_rho_v, rho_functions.cpp:267-284: rho * Re(F^-1 v)/M^2, Im=0. */
static void load_fft_rho_v(orc_state *st, int n, cplx *t, const cplx *treal, real *vbox) {
    int M = st->M, zs = st->zs;
    for (int i = 0; i < M; i++)
        for (int j = 0; j < M; j++) {
            size_t q = PL(n, i + 1 + st->xc[n] - zs, j + 1 + st->yc[n] - zs);
            t[i * M + j].x = st->rho[q] * treal[i * M + j].x / (M * M);
            t[i * M + j].y = 0.0;
            vbox[(size_t)n * M * M + i * M + j] = treal[i * M + j].x / (M * M);
        }
}

/* update_fft_rho, rho_functions.cpp:213-251 (forward Euler, no diffusion). */
static void update_fft_rho(orc_state *st, orc_spec *sp) {
    int M = st->M;
    real dt = (real)st->p.dt, dx = (real)st->p.dx;
    for (int k = 0; k < M * M; k++) {
        sp->fft_rho[k].x += dt * sp->fft_react[k].x;                  /* :222 */
        sp->fft_rho[k].y += dt * sp->fft_react[k].y;                  /* :223 */
    }
    for (int i = 0; i < M; i++)
        for (int j = 0; j < M; j++) {
            int ir = (i < M / 2) ? i : i - M;
            int jr = (j < M / 2) ? j : j - M;
            int k = i * M + j;
            sp->fft_rho[k].x += dt * (2.0 * M_PI / (M * dx)) * (ir * sp->fft_rho_vx[k].y + jr * sp->fft_rho_vy[k].y);  /* :243 */
            sp->fft_rho[k].y -= dt * (2.0 * M_PI / (M * dx)) * (ir * sp->fft_rho_vx[k].x + jr * sp->fft_rho_vy[k].x);  /* :247 */
        }
}

/* copy_to_rhon, rho_functions.cpp:253-265 */
static void copy_to_rhon(orc_state *st, int n, const cplx *t) {
    int M = st->M, zs = st->zs;
    for (int i = 0; i < M; i++)
        for (int j = 0; j < M; j++)
            st->rhon[PL(n, i + 1 + st->xc[n] - zs, j + 1 + st->yc[n] - zs)] = t[i * M + j].x / (M * M);
}

/* update_all_cortices, rho_functions.cpp:409-471 */
int orc_update_all_cortices(orc_state *st) {
    if (check_windows(st)) return st->err;
    /* cells are independent: each thread runs whole cells on its own set of
     * FFT-box buffers, in the reference's per-cell order */
#pragma omp parallel for schedule(dynamic, 1)
    for (int nc = 0; nc < st->nc; nc++) {
#ifdef _OPENMP
        orc_spec *sp = &st->spec[omp_get_thread_num()];
#else
        orc_spec *sp = &st->spec[0];
#endif
        load_temp_fft(st, nc, st->rho, sp->temp_fft);                  /* :415 */
        exec_c2c(st, sp, sp->temp_fft, sp->fft_rho, -1);               /* :417 */
        create_deg_gen(st, nc, sp->temp_fft);                          /* :420 */
        exec_c2c(st, sp, sp->temp_fft, sp->fft_react, -1);             /* :422 */
        compute_pi(st, nc);                                            /* :428-430 */
        load_temp_fft(st, nc, st->pixx, sp->temp_fft);                 /* :434 */
        exec_c2c(st, sp, sp->temp_fft, sp->fft_pi_xx, -1);
        load_temp_fft(st, nc, st->pixy, sp->temp_fft);                 /* :437 */
        exec_c2c(st, sp, sp->temp_fft, sp->fft_pi_xy, -1);
        load_temp_fft(st, nc, st->piyy, sp->temp_fft);                 /* :440 */
        exec_c2c(st, sp, sp->temp_fft, sp->fft_pi_yy, -1);
        update_rho_velocity(st, sp, nc);                               /* :445 */
        exec_c2c(st, sp, sp->fft_vx, sp->temp_fft_real, +1);           /* :449 */
        load_fft_rho_v(st, nc, sp->temp_fft, sp->temp_fft_real, st->vx_box);
        exec_c2c(st, sp, sp->temp_fft, sp->fft_rho_vx, -1);            /* :451 */
        exec_c2c(st, sp, sp->fft_vy, sp->temp_fft_real, +1);           /* :455 */
        load_fft_rho_v(st, nc, sp->temp_fft, sp->temp_fft_real, st->vy_box);
        exec_c2c(st, sp, sp->temp_fft, sp->fft_rho_vy, -1);            /* :457 */
        update_fft_rho(st, sp);                                        /* :462 */
        exec_c2c(st, sp, sp->fft_rho, sp->temp_fft, +1);               /* :464 */
        copy_to_rhon(st, nc, sp->temp_fft);                            /* :466 */
    }
    bc_cells(st->N, st->nc, st->rhon);                                 /* :469 */
    return ORG_OK;
}

/* ---- cell_phase_functions.cpp ------------------------------------------- */

/* 9-point Laplacian numerator used throughout cell_phase_functions.cpp
 * (e.g. :198): -12 u + 2(sum of edge neighbours) + sum of diagonals. */
#define L9(t, i, j) (-12.0 * (t)[G(i, j)] + 2.0 * ((t)[G(i + 1, j)] + (t)[G(i - 1, j)] + (t)[G(i, j + 1)] + (t)[G(i, j - 1)]) \
                     + (t)[G(i + 1, j + 1)] + (t)[G(i + 1, j - 1)] + (t)[G(i - 1, j - 1)] + (t)[G(i - 1, j + 1)])

/* update_lumen_and_ECM, cell_phase_functions.cpp:372-444, then
 * boundary_conditions_lumen_and_ECM(lumn, ECMn), useful_functions.cpp:439-462. */
int orc_update_lumen_and_ECM(orc_state *st) {
    const org_params *p = &st->p;
    int N = st->N;
    real dx = (real)p->dx, dt = (real)p->dt;
    double sum_ecm = 0.0;                                              /* :376-386 */
    for (size_t k = 0; k < st->NN; k++) sum_ecm += (double)(st->hECM[k] * dx * dx);
    real sum_e = (real)sum_ecm;
    real tau_lum = (real)p->tau_lum, xi_lum = (real)p->xi_lum, beta_lc = (real)p->beta_lum_cells;
    real eta_lc = (real)p->eta_lum_cells, beta_le = (real)p->beta_lum_ECM;
    real tau_E = (real)p->tau_ECM, alpha_E = (real)p->alpha_ECM, xi_E = (real)p->xi_ECM;
    real beta_ec = (real)p->beta_ECM_cells, eta_ec = (real)p->eta_ECM_cells;
    const real *psi = st->psi;
#pragma omp parallel for schedule(static)
    for (int i = 1; i < N - 1; i++)
        for (int j = 1; j < N - 1; j++) {
            size_t q = G(i, j);
            real l = st->lum[q], e = st->ECM[q];
            real lap = L9(psi, i, j);
            st->lumn[q] = l;                                                         /* :408 */
            st->lumn[q] += (1.0 / tau_lum) * l * (1.0 - l) * dt * (                  /* :409 */
                xi_lum - beta_lc * psi[q] + eta_lc * lap / (4.0 * dx * dx) - beta_le * st->hECM[q]);
            st->ECMn[q] = e;                                                         /* :424 */
            st->ECMn[q] += (1.0 / tau_E) * e * (1.0 - e) * dt * (                    /* :425 */
                alpha_E * (N * N * dx * dx - sum_e) + xi_E - beta_ec * psi[q]
                + eta_ec * lap / (4.0 * dx * dx) - beta_le * st->hlum[q]);
        }
    /* boundary_conditions_lumen_and_ECM, useful_functions.cpp:439-462 */
    for (int i = 1; i < N - 1; i++) {
        st->lumn[G(i, 0)] = 0.0; st->lumn[G(i, N - 1)] = 0.0;
        st->lumn[G(0, i)] = 0.0; st->lumn[G(N - 1, i)] = 0.0;
        st->ECMn[G(i, 0)] = 1.0; st->ECMn[G(i, N - 1)] = 1.0;
        st->ECMn[G(0, i)] = 1.0; st->ECMn[G(N - 1, i)] = 1.0;
    }
    st->lumn[0] = 0.0; st->lumn[N - 1] = 0.0; st->lumn[G(N - 1, 0)] = 0.0; st->lumn[G(N - 1, N - 1)] = 0.0;
    st->ECMn[0] = 1.0; st->ECMn[N - 1] = 1.0; st->ECMn[G(N - 1, 0)] = 1.0; st->ECMn[G(N - 1, N - 1)] = 1.0;
    return ORG_OK;
}

/* boundary_conditions_periodic, useful_functions.cpp:505-526 (one plane). */
static void bc_periodic(int N, real *t) { bc_cells(N, 1, t); }

/* boundary_conditions_dirichlet, useful_functions.cpp:422-437: edges get
 * side_value, the four corners get 0.0 whatever side_value is (:432-435). */
static void bc_dirichlet(int N, real *t, real side) {
    for (int i = 1; i < N - 1; i++) {
        t[(size_t)i * N] = side; t[(size_t)i * N + N - 1] = side;
        t[i] = side; t[(size_t)N * (N - 1) + i] = side;
    }
    t[0] = 0.0; t[N - 1] = 0.0; t[(size_t)(N - 1) * N] = 0.0; t[(size_t)N * (N - 1) + N - 1] = 0.0;
}

/* The conservative interface flux used by both reshaping loops
 * (cell_phase_functions.cpp:511-547 and :802-838): returns the new value
 * of phasen at (i,j) given the current phasen value, the relaxed field u
 * (phase / lum) and the normal field (jx, jy), on planes with row stride N. */
static real reshape_update(real un, const real *u, const real *jx, const real *jy, int N, int i, int j,
                           real wi, real rdt, real dx) {
#define U(a, b) u[(size_t)(a) * N + (b)]
#define JX(a, b) jx[(size_t)(a) * N + (b)]
#define JY(a, b) jy[(size_t)(a) * N + (b)]
    int ip = i + 1, im = i - 1, jp = j + 1, jm = j - 1;
    un += 0.5 * (JX(ip, j) + JX(i, j)) * (
        wi * wi * (rdt / dx / dx) * (U(ip, j) - U(i, j)) * 0.5 * (JX(ip, j) + JX(i, j))
        + wi * wi * (rdt / dx / dx) * 0.25 * (U(ip, jp) + U(i, jp) - U(ip, jm) - U(i, jm)) * 0.5 * (JY(ip, j) + JY(i, j))
        - rdt * wi * sqrt(2.0) / dx * 0.5 * (U(ip, j) * (1.0 - U(ip, j)) + U(i, j) * (1.0 - U(i, j))));
    un += 0.5 * (JY(i, jp) + JY(i, j)) * (
        wi * wi * (rdt / dx / dx) * (U(i, jp) - U(i, j)) * 0.5 * (JY(i, jp) + JY(i, j))
        + wi * wi * (rdt / dx / dx) * 0.25 * (U(ip, jp) + U(ip, j) - U(im, jp) - U(im, j)) * 0.5 * (JX(i, jp) + JX(i, j))
        - rdt * wi * sqrt(2.0) / dx * 0.5 * (U(i, jp) * (1.0 - U(i, jp)) + U(i, j) * (1.0 - U(i, j))));
    un -= 0.5 * (JX(i, j) + JX(im, j)) * (
        wi * wi * (rdt / dx / dx) * (U(i, j) - U(im, j)) * 0.5 * (JX(i, j) + JX(im, j))
        + wi * wi * (rdt / dx / dx) * 0.25 * (U(i, jp) + U(im, jp) - U(i, jm) - U(im, jm)) * 0.5 * (JY(i, j) + JY(im, j))
        - rdt * wi * sqrt(2.0) / dx * 0.5 * (U(i, j) * (1.0 - U(i, j)) + U(im, j) * (1.0 - U(im, j))));
    un -= 0.5 * (JY(i, j) + JY(i, jm)) * (
        wi * wi * (rdt / dx / dx) * (U(i, j) - U(i, jm)) * 0.5 * (JY(i, j) + JY(i, jm))
        + wi * wi * (rdt / dx / dx) * 0.25 * (U(ip, j) + U(ip, jm) - U(im, j) - U(im, jm)) * 0.5 * (JX(i, j) + JX(i, jm))
        - rdt * wi * sqrt(2.0) / dx * 0.5 * (U(i, j) * (1.0 - U(i, j)) + U(i, jm) * (1.0 - U(i, jm))));
    return un;
#undef U
#undef JX
#undef JY
}

/* reshaping_lumen_or_ECM, cell_phase_functions.cpp:709-871.  which_ecm = 0:
 * (lum, lumn, side 0.0); 1: (ECM, ECMn, side 1.0).  |grad| from the old
 * field, normals from the new one (:734-763); the last iterate is kept in
 * the n-field (the driver copies it back, run_with_active_gels.cpp:181-182). */
int orc_reshaping_lumen_or_ECM(orc_state *st, int which_ecm) {
    const org_params *p = &st->p;
    int N = st->N;
    real *lum = which_ecm ? st->ECM : st->lum;
    real *lumn = which_ecm ? st->ECMn : st->lumn;
    real side = which_ecm ? 1.0 : 0.0;
    real dx = (real)p->dx, wi = (real)p->wi, rdt = (real)p->rdt, rtol = (real)p->rtol;
    real *ng = st->nglum, *jx = st->Jlx, *jy = st->Jly;
    double rdev = (double)(p->rdev_ini * rtol * rdt);                  /* :720 */
    double thr = (double)(p->rdev_threshold * rtol * rdt);             /* :774 */
#pragma omp parallel for schedule(static)
    for (int i = 1; i < N - 1; i++)
        for (int j = 1; j < N - 1; j++) {
            real a = lum[G(i + 1, j)] - lum[G(i - 1, j)], c = lum[G(i, j + 1)] - lum[G(i, j - 1)];
            ng[G(i, j)] = sqrt((1.0 / dx / dx / 4.0) * (a * a + c * c));  /* :734-736 */
        }
    bc_periodic(N, ng);                                                /* :741 */
#pragma omp parallel for schedule(static)
    for (int i = 1; i < N - 1; i++)
        for (int j = 1; j < N - 1; j++) {
            real g = ng[G(i, j)];
            if (g != 0.0) {                                            /* :757-763 */
                jx[G(i, j)] = (1.0 / (2.0 * dx)) * (lumn[G(i + 1, j)] - lumn[G(i - 1, j)]) / g;
                jy[G(i, j)] = (1.0 / (2.0 * dx)) * (lumn[G(i, j + 1)] - lumn[G(i, j - 1)]) / g;
            } else { jx[G(i, j)] = 0.0; jy[G(i, j)] = 0.0; }
        }
    bc_periodic(N, jx);                                                /* :770-771 */
    bc_periodic(N, jy);
    int ri = 0, count = 0;
    while (rdev > thr) {                                               /* :774 */
        memcpy(lum, lumn, st->NN * sizeof(real));                      /* :777-785 */
#pragma omp parallel for schedule(static)
        for (int i = 1; i < N - 1; i++)
            for (int j = 1; j < N - 1; j++)
                lumn[G(i, j)] = reshape_update(lumn[G(i, j)], lum, jx, jy, N, i, j, wi, rdt, dx);
        bc_dirichlet(N, lumn, side);                                   /* :844 */
        rdev = 0.0;                                                    /* :848-855 */
        for (size_t k = 0; k < st->NN; k++) rdev += (double)(fabs(lumn[k] - lum[k]) * dx * dx);
        count++;
        if (ri > p->rimax) { st->relax_failures++; break; }             /* :862-865 */
        ri++;
    }
    if (which_ecm) st->count_ecm = count; else st->count_lum = count;
    return st->err;
}

/* update_all_cells, cell_phase_functions.cpp:126-217, then
 * boundary_conditions_cells(phin) (run_with_active_gels.cpp:170). */
int orc_update_all_cells(orc_state *st) {
    if (check_windows(st)) return st->err;
    const org_params *p = &st->p;
    int N = st->N, zs = st->zs;
    real dx = (real)p->dx, dt = (real)p->dt, tau = (real)p->tau, sigma = (real)p->sigma;
    real cpl = (real)p->coupling_PF_AG, beta = (real)p->beta, eta = (real)p->eta;
    real beta_lc = (real)p->beta_lum_cells, eta_lc = (real)p->eta_lum_cells;
    real beta_ec = (real)p->beta_ECM_cells, eta_ec = (real)p->eta_ECM_cells;
#pragma omp parallel for schedule(dynamic, 1)
    for (int n = 0; n < st->nc; n++) {
        real *tp = st->temp_phase + (size_t)n * st->NN;
        const real *ph = st->phi + (size_t)n * st->NN;
        const real *rh = st->rho + (size_t)n * st->NN;
        const real *ng = st->ngp + (size_t)n * st->NN;
        real *pn = st->phin + (size_t)n * st->NN;
        const real *hl = st->hlum, *he = st->hECM;
        for (size_t k = 0; k < st->NN; k++) tp[k] = st->psi[k] - st->hfunc[(size_t)n * st->NN + k];  /* :150 */
        real f0 = st->f0[n];
        for (int i = st->xc[n] - zs; i < st->xc[n] + zs; i++)
            for (int j = st->yc[n] - zs; j < st->yc[n] + zs; j++) {
                size_t q = G(i, j);
                real v = ph[q];
                pn[q] = v;                                                              /* :170 */
                if (ng[q] > 0.0) {                                                      /* :177-186 */
                    pn[q] -= cpl / tau * (dt / dx / dx / 4.0) * exp(-(v - 0.5) * (v - 0.5) / 2.0 / sigma / sigma)
                             / sqrt(2.0 * M_PI) / sigma
                             * ((rh[G(i + 1, j)] - rh[G(i - 1, j)]) * (ph[G(i + 1, j)] - ph[G(i - 1, j)])
                                + (rh[G(i, j + 1)] - rh[G(i, j - 1)]) * (ph[G(i, j + 1)] - ph[G(i, j - 1)]))
                             / ng[q];
                }
                pn[q] += (1.0 / tau) * v * (1.0 - v) * dt * (                           /* :190-211 */
                    f0
                    - beta * tp[q]
                    + eta * L9(tp, i, j) / (4.0 * dx * dx)
                    - beta_lc * hl[q]
                    + eta_lc * L9(hl, i, j) / (4.0 * dx * dx)
                    - beta_ec * he[q]
                    + eta_ec * L9(he, i, j) / (4.0 * dx * dx));
            }
    }
    bc_cells(N, st->nc, st->phin);
    return ORG_OK;
}

/* reshaping_cells, cell_phase_functions.cpp:447-580.  Per cell: copy
 * phase<-phasen on the window, update phasen with the interface flux, stop
 * when sum |phasen-phase| dx^2 / N_current <= rtol*2*pi*R*wi*rdt.  phase
 * ends at the second-to-last iterate. */
int orc_reshaping_cells(orc_state *st) {
    if (check_windows(st)) return st->err;
    const org_params *p = &st->p;
    int N = st->N, zs = st->zs;
    real dx = (real)p->dx, wi = (real)p->wi, rdt = (real)p->rdt, rtol = (real)p->rtol, R = (real)p->R;
    double thr = (double)(rtol * 2.0 * M_PI * R * wi * rdt);           /* :475 */
    int relax_fail = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+: relax_fail)
    for (int n = 0; n < st->nc; n++) {
        real *ph = st->phi + (size_t)n * st->NN;
        real *pn = st->phin + (size_t)n * st->NN;
        const real *jx = st->Jx + (size_t)n * st->NN, *jy = st->Jy + (size_t)n * st->NN;
        int x0 = st->xc[n] - zs, x1 = st->xc[n] + zs, y0 = st->yc[n] - zs, y1 = st->yc[n] + zs;
        double rdev = (double)(2.0 * rtol * 2.0 * M_PI * R * wi * rdt); /* :473 */
        int ri = 0, count = 0;
        while (rdev > thr) {
            for (int i = x0; i < x1; i++)                              /* :480-492 */
                for (int j = y0; j < y1; j++) ph[G(i, j)] = pn[G(i, j)];
            for (int i = x0; i < x1; i++)                              /* :498-549 */
                for (int j = y0; j < y1; j++)
                    pn[G(i, j)] = reshape_update(pn[G(i, j)], ph, jx, jy, N, i, j, wi, rdt, dx);
            rdev = 0.0;                                                /* :554-562 */
            for (int i = x0; i < x1; i++)
                for (int j = y0; j < y1; j++)
                    rdev += (double)(fabs(pn[G(i, j)] - ph[G(i, j)]) * dx * dx / p->n_current);
            count++;
            if (ri > p->rimax) { relax_fail++; break; }               /* :565-568, one report per cell */
            ri++;
        }
        st->count_ri[n] = count;
    }
    st->relax_failures += relax_fail;
    return st->err;
}

/* copy_cell_tables_1_by_1 (rho<-rhon) and copy_tables_ECM_or_lumen
 * (ECM<-ECMn, lum<-lumn), useful_functions.cpp:679-709, as called at
 * run_with_active_gels.cpp:175-182. */
int orc_commit_step(orc_state *st) {
    memcpy(st->rho, st->rhon, st->P * sizeof(real));
    memcpy(st->ECM, st->ECMn, st->NN * sizeof(real));
    memcpy(st->lum, st->lumn, st->NN * sizeof(real));
    return ORG_OK;
}

/* One timestep, run_with_active_gels.cpp:118-184. */
int orc_step(orc_state *st, int it) {
    (void)it;
    int e;
    if ((e = orc_find_centres_cells(st))) return e;                    /* :123 */
    orc_create_turnover_rates_tensions(st);                            /* :130 */
    orc_compute_hfunc_cells(st);                                       /* :133 */
    orc_compute_hfunc_lum_ecm(st);                                     /* :134 */
    orc_compute_normal_grad_phi(st);                                   /* :135-136 */
    orc_compute_prod_norm_delta(st);                                   /* :137 */
    orc_compute_psi(st);                                               /* :139 */
    orc_sum_s(st);                                                     /* :142 */
    orc_update_all_cortices(st);                                       /* :154 */
    orc_update_lumen_and_ECM(st);                                      /* :157-158 */
    if ((e = orc_reshaping_lumen_or_ECM(st, 0))) return e;             /* :160 */
    if ((e = orc_reshaping_lumen_or_ECM(st, 1))) return e;             /* :161 */
    orc_compute_normal_to_phi(st);                                     /* :164-166 */
    orc_update_all_cells(st);                                          /* :169-170 */
    if ((e = orc_reshaping_cells(st))) return e;                       /* :172 */
    orc_commit_step(st);                                               /* :175-182 */
    return st->err;
}
