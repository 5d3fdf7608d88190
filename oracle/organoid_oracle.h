/*
 * organoid_oracle.h -- CPU oracle of the organoid timestep (TEST
 * INFRASTRUCTURE ONLY).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library, and only as the checker or the timed CPU baseline.  The
 * product library never links, loads or calls it.
 *
 * It restates, function by function and on full N x N planes exactly as the
 * reference does, the per-timestep path of
 * tristanguyomar/multicellular_phase_field_active_gel_organoids (files under
 * functions/ and main_code/; each function in organoid_oracle.c cites the
 * file:line it follows).  The reference itself cannot be built here: it
 * includes <cufft.h>, which this image does not have, and its constants.cpp
 * is generated from a jinja template; so this restatement is the oracle.
 *
 * Precision: built twice from one source.  ORC_REAL=double (default) is the
 * FP64 oracle; ORC_REAL=float reproduces the reference's FP32 build, where
 * double literals promote intermediate expressions to double before the
 * store back to float.  Reductions (areas, sum_ecm, rdev) accumulate in
 * double in both builds.  The 2D DFT is computed in double in both builds.
 */
#ifndef ORGANOID_ORACLE_H
#define ORGANOID_ORACLE_H

#include "../include/organoid.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Oracle-only field ids extend ORG_F_* (organoid.h). */
enum {
    ORC_F_PHIN = 10, ORC_F_RHON = 11, ORC_F_NGP = 12, ORC_F_PND = 13,
    ORC_F_JX = 14, ORC_F_JY = 15, ORC_F_TEMP = 16,
    ORC_F_DEG = 17, ORC_F_GEN = 18, ORC_F_ATAN = 19, ORC_F_AORT = 20,
    ORC_F_VISC = 21, ORC_F_B = 22,
    ORC_F_PIXX = 23, ORC_F_PIXY = 24, ORC_F_PIYY = 25,
    ORC_F_NGLUM = 26, ORC_F_JLX = 27, ORC_F_JLY = 28,
    ORC_F_VX = 29, ORC_F_VY = 30,           /* [ncell][M][M] FFT-box velocity */
    ORC_F_COUNT = 31
};

typedef struct orc_state orc_state;

orc_state *orc_create(const org_params *p, const double *s0_cell);
void orc_destroy(orc_state *st);
int orc_real_bytes(void);                 /* 8 (FP64 build) or 4 (FP32 build) */
int orc_params_bytes(void);               /* sizeof(org_params) as compiled */
int64_t orc_field_size(const orc_state *st, int field);
int orc_set_field(orc_state *st, int field, const double *src);
int orc_get_field(const orc_state *st, int field, double *dst);
int orc_get_centres(const orc_state *st, int32_t *xc, int32_t *yc);
int orc_set_centres(orc_state *st, const int32_t *xc, const int32_t *yc);
int orc_get_areas(const orc_state *st, double *s, double *f0);
int orc_get_iterations(const orc_state *st, int32_t *cells, int32_t *lum, int32_t *ecm);
int orc_last_error(const orc_state *st);
/* OpenMP threads the oracle's loops use (OMP_NUM_THREADS); 1 without OpenMP. */
int orc_threads(void);
/* Reshaping loops that passed rimax and stopped there, as the reference's
 * "Relaxation failure!" break does (the step goes on). */
int64_t orc_relax_failures(const orc_state *st);

/* run_with_active_gels.cpp:92-98: lumn<-lum, ECMn<-ECM, phin<-phi, rhon<-rho.
 * hfunc is left as uploaded (the driver zeroes it, run_with_active_gels.cpp:74). */
int orc_init_like_driver(orc_state *st);

/* One full timestep in the driver's order (run_with_active_gels.cpp:118-184). */
int orc_step(orc_state *st, int it);

/* The individual functions, in driver order. */
int orc_find_centres_cells(orc_state *st);
int orc_create_turnover_rates_tensions(orc_state *st);
int orc_compute_hfunc_cells(orc_state *st);
int orc_compute_hfunc_lum_ecm(orc_state *st);
int orc_compute_normal_grad_phi(orc_state *st);   /* + boundary_conditions_cells */
int orc_compute_prod_norm_delta(orc_state *st);
int orc_compute_psi(orc_state *st);
int orc_sum_s(orc_state *st);
int orc_update_all_cortices(orc_state *st);
int orc_update_lumen_and_ECM(orc_state *st);       /* + boundary_conditions_lumen_and_ECM */
int orc_reshaping_lumen_or_ECM(orc_state *st, int which_ecm);
int orc_compute_normal_to_phi(orc_state *st);      /* + boundary_conditions_cells x2 */
int orc_update_all_cells(orc_state *st);           /* + boundary_conditions_cells(phin) */
int orc_reshaping_cells(orc_state *st);
int orc_commit_step(orc_state *st);

/* Stand-alone helpers for known-answer tests. */
/* Unnormalised 2D DFT of an M x M complex array (interleaved re,im, double),
 * sign -1 forward / +1 inverse, as cufftExecC2C (useful_functions.cpp:650,658). */
int orc_fft2d(int M, double *data, int sign);
int orc_fft1d(int n, double *data, int sign);
/* boundary_conditions_cells on an arbitrary stack of nplanes N x N planes. */
int orc_bc_cells_planes(int N, int nplanes, double *planes);
/* update_rho_velocity for one cell on explicit spectra (interleaved complex
 * M*M arrays) and a per-mode viscosity array eta[M*M] (already gathered). */
int orc_velocity_modes(const org_params *p, int M, const double *pxx, const double *pxy,
                       const double *pyy, const double *eta, double *vx, double *vy);

#ifdef __cplusplus
}
#endif
#endif
