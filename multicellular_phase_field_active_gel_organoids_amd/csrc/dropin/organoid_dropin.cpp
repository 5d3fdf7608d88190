/*
 * organoid_dropin.cpp -- the reference's timestep functions (declared in
 * include/organoid_dropin.h with the reference's signatures) on top of the
 * device-resident C-ABI of liborganoid_hip.so.
 *
 * Host arrays keep the reference's roles and ownership: malloc_all and
 * ini_fft_all hand out malloc'd buffers that the driver frees
 * (run_with_active_gels.cpp:275-329).  The device copy of the state lives in
 * the library context; it is created and filled from phi, rho, lum, ECM and
 * hfunc_cell at the first timestep call, which is where the driver's
 * `#pragma acc data copyin` (run_with_active_gels.cpp:83) takes effect.
 * Host arrays the driver reads back are written back eagerly: the output
 * slabs (saving_out_functions) and the areas s every ndisplay steps
 * (run_with_active_gels.cpp:144-151).
 *
 * Compiled twice: org_real_t float (liborganoid_dropin.so, FP32 device
 * state) and, with ORGANOID_DROPIN_F64, double (liborganoid_dropin_f64.so,
 * FP64 device state through org_create_f64 and the _f64 field I/O).
 */
#include "organoid_dropin.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

/* ---- declare_tables.cpp:97-157 ------------------------------------------- */
org_real_t *__restrict phi, *__restrict phin, *__restrict rho, *__restrict rho_to_compare, *__restrict rhon;
org_real_t *__restrict lum, *__restrict ECM, *__restrict lumn, *__restrict ECMn;
org_real_t *__restrict psi, *__restrict temp_phase;
org_real_t *__restrict hfunc_cell, *__restrict hfunc_lum, *__restrict hfunc_ECM;
org_real_t *__restrict norm_grad_phi, *__restrict prod_norm_delta, *__restrict norm_grad_lum, *__restrict norm_grad_ecm;
org_real_t *__restrict a_tan_1, *__restrict a_ort_1, *__restrict b, *__restrict viscosity;
org_real_t *__restrict pi_xx, *__restrict pi_yy, *__restrict pi_xy;
org_real_t *__restrict phiout, *__restrict rhoout, *__restrict lumout, *__restrict ECMout;
org_real_t *__restrict Jxout, *__restrict Jyout, *__restrict vol_out;
org_real_t *__restrict J_x, *__restrict J_y;
int *__restrict xcenters, *__restrict ycenters;
org_real_t *__restrict matgenrates, *__restrict degrates, *__restrict matout, *__restrict degout;

/* ---- declare_fft_tables.cpp:10-22 ---------------------------------------- */
cufftComplex *__restrict temp_fft, *__restrict temp_fft_real, *__restrict fft_pi_xx, *__restrict fft_pi_yy;
cufftComplex *__restrict fft_pi_xy, *__restrict fft_rho_vx, *__restrict fft_rho_vy, *__restrict fft_vx;
cufftComplex *__restrict fft_vy, *__restrict fft_rho, *__restrict fft_react;
cufftHandle planf, planr;

namespace {

org_ctx *g_ctx = nullptr;
bool g_phin_is_phi = false;    /* phin <- phi copied since the last update_all_cells */

[[noreturn]] void die(const char *what, int rc) {
    fprintf(stderr, "organoid drop-in: %s failed: %s (code %d)\n", what, org_error_string(rc), rc);
    std::exit(1);
}
void chk(int rc, const char *what) {
    if (rc) die(what, rc);
}

org_real_t *fmalloc(size_t n) { return (org_real_t *)malloc(n * sizeof(org_real_t)); }

#ifdef ORGANOID_DROPIN_F64
int create_ctx(const org_params *p, org_ctx **c) { return org_create_f64(p, nullptr, c); }
int put(org_ctx *c, int f, const double *h) { return org_set_field_f64(c, f, h); }
int get(org_ctx *c, int f, double *h) { return org_get_field_f64(c, f, h); }
#else
int create_ctx(const org_params *p, org_ctx **c) { return org_create(p, nullptr, c); }
int put(org_ctx *c, int f, const float *h) { return org_set_field(c, f, h); }
int get(org_ctx *c, int f, float *h) { return org_get_field(c, f, h); }
#endif

size_t planes() { return (size_t)Ntot * N_s * N_s; }
size_t plane() { return (size_t)N_s * N_s; }

org_params params_from_globals() {
    org_params p{};
    p.N = N_s; p.ncell = Ntot; p.zone_size = zone_size; p.rimax = rimax;
    p.rdev_threshold = rdev_threshold; p.rdev_ini = rdev_ini; p.n_current = N_current; p.use_s0_cell = 0;
    p.dt = dt; p.dx = dx; p.R = R; p.s0 = s0; p.sigma = sigma; p.coupling_PF_AG = coupling_PF_AG;
    p.tau = tau; p.alpha = alpha; p.beta = beta; p.eta = eta;
    p.tau_lum = tau_lum; p.beta_lum_cells = beta_lum_cells; p.eta_lum_cells = eta_lum_cells; p.xi_lum = xi_lum;
    p.alpha_ECM = alpha_ECM; p.tau_ECM = tau_ECM; p.beta_ECM_cells = beta_ECM_cells;
    p.eta_ECM_cells = eta_ECM_cells; p.xi_ECM = xi_ECM; p.beta_lum_ECM = beta_lum_ECM;
    p.wi = wi; p.rtol = rtol; p.rdt = rdt; p.friction = friction;
    p.visc_ap = viscosity_ap; p.visc_bas = viscosity_bas; p.visc_lat = viscosity_lat;
    p.a_ort_ap = a_ort_apical; p.a_ort_bas = a_ort_basal; p.a_ort_lat = a_ort_lateral;
    p.a_tan_ap = a_tan_apical; p.a_tan_bas = a_tan_basal; p.a_tan_lat = a_tan_lateral;
    p.b_ap = b_ap; p.b_bas = b_bas; p.b_lat = b_lat;
    p.gen_ap = matgenrate_apical; p.gen_bas = matgenrate_basal; p.gen_lat = matgenrate_lateral;
    p.deg_ap = degrate_apical; p.deg_bas = degrate_basal; p.deg_lat = degrate_lateral;
    return p;
}

void release() {
    if (g_ctx) { org_destroy(g_ctx); g_ctx = nullptr; }
}

/* Device state on first use: the data region's copyin. */
org_ctx *ctx() {
    if (g_ctx) return g_ctx;
    org_params p = params_from_globals();
    chk(create_ctx(&p, &g_ctx), "org_create");
    chk(put(g_ctx, ORG_F_PHI, phi), "copyin phi");
    chk(put(g_ctx, ORG_F_RHO, rho), "copyin rho");
    chk(put(g_ctx, ORG_F_LUM, lum), "copyin lum");
    chk(put(g_ctx, ORG_F_ECM, ECM), "copyin ECM");
    chk(put(g_ctx, ORG_F_HFUNC, hfunc_cell), "copyin hfunc_cell");
    chk(org_prepare(g_ctx), "org_prepare");
    std::atexit(release);
    return g_ctx;
}

}  // namespace

org_ctx *organoid_dropin_context(void) { return ctx(); }

int cufftPlan2d(cufftHandle *plan, int nx, int ny, int type) {
#ifdef ORGANOID_DROPIN_F64
    const bool type_ok = type == CUFFT_Z2Z || type == CUFFT_C2C;
#else
    const bool type_ok = type == CUFFT_C2C;
#endif
    if (!plan || nx != 2 * zone_size - 2 || ny != nx || !type_ok) {
        fprintf(stderr, "organoid drop-in: plan %dx%d type %d unsupported (box is %d)\n", nx, ny, type,
                2 * zone_size - 2);
        std::exit(1);
    }
    static int next = 1;
    *plan = next++;
    return CUFFT_SUCCESS;
}

/* declare_tables.cpp:160-214 */
void malloc_all(void) {
    size_t P = planes(), G = plane(), S = (size_t)Ntotsave;
    phiout = fmalloc(P * S); rhoout = fmalloc(P * S); lumout = fmalloc(G * S); ECMout = fmalloc(G * S);
    Jxout = fmalloc(P * S); Jyout = fmalloc(P * S); matout = fmalloc(P * S); degout = fmalloc(P * S);
    phi = fmalloc(P); phin = fmalloc(P); rho = fmalloc(P); rho_to_compare = fmalloc(P); rhon = fmalloc(P);
    hfunc_cell = fmalloc(P); norm_grad_phi = fmalloc(P); prod_norm_delta = fmalloc(P);
    norm_grad_lum = fmalloc(G); norm_grad_ecm = fmalloc(G); J_x = fmalloc(P); J_y = fmalloc(P);
    psi = fmalloc(G); temp_phase = fmalloc(P); lum = fmalloc(G); lumn = fmalloc(G); ECM = fmalloc(G);
    ECMn = fmalloc(G); hfunc_lum = fmalloc(G); hfunc_ECM = fmalloc(G); a_tan_1 = fmalloc(P); a_ort_1 = fmalloc(P);
    viscosity = fmalloc(P); b = fmalloc(P); pi_xx = fmalloc(P); pi_yy = fmalloc(P); pi_xy = fmalloc(P);
    xcenters = (int *)malloc(Ntot * sizeof(int)); ycenters = (int *)malloc(Ntot * sizeof(int));
    matgenrates = fmalloc(P); degrates = fmalloc(P);
}

/* declare_fft_tables.cpp:26-41 */
void ini_fft_all(void) {
    size_t MM = (size_t)(2 * zone_size - 2) * (2 * zone_size - 2);
    auto cm = [MM]() { return (cufftComplex *)malloc(MM * sizeof(cufftComplex)); };
    temp_fft = cm(); fft_pi_xx = cm(); fft_pi_yy = cm(); fft_pi_xy = cm(); fft_rho_vx = cm(); fft_rho_vy = cm();
    fft_vx = cm(); fft_vy = cm(); temp_fft_real = cm(); fft_rho = cm(); fft_react = cm();
}

/* useful_functions.cpp:66-99: snapshot into slab it/nsaving of the host
 * output arrays, which the driver reads after the loop. */
void saving_out_functions(int it_, org_real_t *phi_, org_real_t *rho_, org_real_t *ECM_, org_real_t *lum_, org_real_t *phiout_,
                          org_real_t *rhoout_, org_real_t *lumout_, org_real_t *ECMout_) {
    org_ctx *c = ctx();
    size_t slab = (size_t)(it_ / nsaving);
    chk(get(c, ORG_F_PHI, phiout_ + slab * planes()), "saving phi");
    chk(get(c, ORG_F_RHO, rhoout_ + slab * planes()), "saving rho");
    chk(get(c, ORG_F_LUM, lumout_ + slab * plane()), "saving lum");
    chk(get(c, ORG_F_ECM, ECMout_ + slab * plane()), "saving ECM");
    (void)phi_; (void)rho_; (void)ECM_; (void)lum_;
}

void find_centres_cells(org_real_t *, int *, int *) { chk(org_find_centres_cells(ctx()), "find_centres_cells"); }

/* Labels are evaluated where they are consumed (update_all_cortices). */
void create_turnover_rates_tensions(org_real_t *, org_real_t *, org_real_t *, org_real_t *, org_real_t *, org_real_t *, org_real_t *, org_real_t *, org_real_t *,
                                    int *, int *) {}

void compute_hfunc_3d_tables_kernels_loop_n(org_real_t *, org_real_t *, int *, int *) {
    chk(org_compute_hfunc_cells(ctx()), "compute_hfunc_3d_tables_kernels_loop_n");
}
void compute_all_hfunc_2d_tables_kernels(org_real_t *, org_real_t *, org_real_t *, org_real_t *) {
    chk(org_compute_hfunc_lum_ecm(ctx()), "compute_all_hfunc_2d_tables_kernels");
}
/* |grad phi|, prod_norm_delta, the normals and their ghost rings are only
 * read on the windows; the consuming kernels recompute them there. */
void compute_normal_grad_phi(org_real_t *, org_real_t *) {}
void compute_prod_norm_delta(org_real_t *, org_real_t *, org_real_t *, int *, int *) {}
void compute_normal_to_phi(org_real_t *, org_real_t *, org_real_t *, org_real_t *) {}
void boundary_conditions_cells(org_real_t *) {}
void boundary_conditions_lumen_and_ECM(org_real_t *, org_real_t *) {}

void compute_psi_kernels(org_real_t *, org_real_t *) { chk(org_compute_psi(ctx()), "compute_psi_kernels"); }

void sum_s_kernels(org_real_t *, org_real_t *s, org_real_t *) {
    org_ctx *c = ctx();
    chk(org_sum_s(c), "sum_s_kernels");
    if (ndisplay > 0 && it % ndisplay == 0) {      /* acc update host(s), run_with_active_gels.cpp:146 */
        std::vector<double> sv(Ntot), fv(Ntot);
        chk(org_get_areas(c, sv.data(), fv.data()), "sum_s_kernels readback");
        for (int q = 0; q < Ntot; q++) s[q] = (org_real_t)sv[q];
    }
}

void update_all_cortices(org_real_t *, org_real_t *, org_real_t *, int *, int *, org_real_t *, org_real_t *, org_real_t *, org_real_t *, org_real_t *, org_real_t *,
                         org_real_t *, org_real_t *, cufftHandle, cufftHandle, cufftComplex *, cufftComplex *, cufftComplex *,
                         cufftComplex *, cufftComplex *, cufftComplex *, cufftComplex *, cufftComplex *,
                         cufftComplex *, cufftComplex *, cufftComplex *) {
    chk(org_update_all_cortices(ctx()), "update_all_cortices");
}

void update_lumen_and_ECM(org_real_t *, org_real_t *, org_real_t *, org_real_t *, org_real_t *, org_real_t *, org_real_t *) {
    chk(org_update_lumen_and_ECM(ctx()), "update_lumen_and_ECM");
}

void reshaping_lumen_or_ECM(int, org_real_t *field, org_real_t *, org_real_t) {
    chk(org_reshaping_lumen_or_ECM(ctx(), field == ECM ? 1 : 0), "reshaping_lumen_or_ECM");
}

void update_all_cells(int, org_real_t *, org_real_t *, org_real_t *, org_real_t *, org_real_t *, org_real_t *, org_real_t *, int *, int *, org_real_t *,
                      org_real_t *) {
    g_phin_is_phi = false;
    chk(org_update_all_cells(ctx()), "update_all_cells");
}

void reshaping_cells(int it_, org_real_t *, org_real_t *, org_real_t *, org_real_t *, org_real_t *, int *, int *) {
    org_ctx *c = ctx();
    if (g_phin_is_phi) chk(org_reshaping_cells_from_phi(c), "reshaping_cells");
    else chk(org_reshaping_cells(c), "reshaping_cells");
    /* the reference's diagnostics every ndisplay steps: each cell's window
     * bounds and its reshaping count (cell_phase_functions.cpp:463-471,
     * 575-578) */
    if (ndisplay > 0 && it_ % ndisplay == 0) {
        std::vector<int32_t> xc(Ntot), yc(Ntot), cnt(Ntot);
        int32_t l = 0, e = 0;
        chk(org_get_centres(c, xc.data(), yc.data()), "reshaping_cells diagnostics");
        chk(org_get_iterations(c, cnt.data(), &l, &e), "reshaping_cells diagnostics");
        for (int q = 0; q < Ntot; q++) {
            fprintf(stderr, "n = %d \n \n", q);
            fprintf(stderr, "xcenters[%d]-%d = %d  \n", q, zone_size, xc[q] - zone_size);
            fprintf(stderr, "xcenters[%d]+%d = %d  \n", q, zone_size, xc[q] + zone_size);
            fprintf(stderr, "ycenters[%d]-%d = %d  \n", q, zone_size, yc[q] - zone_size);
            fprintf(stderr, "ycenters[%d]+%d = %d  \n", q, zone_size, yc[q] + zone_size);
            fprintf(stderr, "count_ri = %d \n \n", cnt[q]);
        }
    }
}

/* cell_phase_functions.cpp:42-62 */
void update_rho_to_shape(org_real_t *, org_real_t *, org_real_t rho0_, org_real_t tol_) {
#ifdef ORGANOID_DROPIN_F64
    chk(org_update_rho_to_shape_f64(ctx(), rho0_, tol_), "update_rho_to_shape");
#else
    chk(org_update_rho_to_shape(ctx(), rho0_, tol_), "update_rho_to_shape");
#endif
}

/* useful_functions.cpp:11-25: NaNs in the first two cell planes (host) */
int check_nan(org_real_t *table) {
    int c = 0;
    const size_t pl = plane();
    for (size_t k = 0; k < pl; k++) c += std::isnan(table[k]) + std::isnan(table[pl + k]);
    return c;
}

/* useful_functions.cpp:679-693.  Before the loop the driver copies
 * phin<-phi, rhon<-rho (run_with_active_gels.cpp:94-98): the device state
 * needs no copy; phin<-phi only marks that a reshaping_cells before the next
 * update_all_cells starts from phi (the IC program's order,
 * prepare_initial_files.cpp:87-100).  In the loop the driver copies
 * rho<-rhon cell by cell (:175-178); the device commits all cells with the
 * last one. */
void copy_cell_tables_1_by_1(int n_, org_real_t *phase, org_real_t *phasen) {
    if (phase == phin && phasen == phi && n_ == Ntot - 1) g_phin_is_phi = true;
    if (!g_ctx) return;
    if (phase == rho && n_ == Ntot - 1) chk(org_commit_rho(g_ctx), "copy rho<-rhon");
}

/* useful_functions.cpp:695-709: ECM<-ECMn then lum<-lumn at the end of a
 * step (:181-182) commit the lumen/ECM iterate; the initial copies in the
 * other direction (:92-93) need nothing. */
void copy_tables_ECM_or_lumen(org_real_t *phase, org_real_t *) {
    if (!g_ctx) return;
    if (phase == ECM) chk(org_commit_lumecm(g_ctx), "copy ECM/lum <- ECMn/lumn");
}
