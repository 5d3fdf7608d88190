/*
 * organoid_dropin.h -- the reference's own C++ interface for the timestep
 * path, implemented by liborganoid_dropin.so on top of the C-ABI in
 * organoid.h.
 *
 * A caller compiled against the reference's headers (declare_tables.h,
 * declare_fft_tables.h, useful_functions.h, rho_functions.h,
 * cell_phase_functions.h, constants.h) links against this library unchanged:
 * every declaration below keeps the reference's name, parameter types and
 * C++ linkage.  Line citations are into the reference's functions/ and
 * main_code/ directories.
 *
 * FFT types.  The reference's signatures carry cufftComplex and cufftHandle
 * (rho_functions.h:10-26, declare_fft_tables.h:12-24), so the mangled names
 * of update_all_cortices and of the FFT buffer globals depend on them.  They
 * are defined here as the library's own types (8-byte {float x, y} and an
 * int plan id): signature types, not a wrapper over any vendor FFT.  A caller
 * that includes <cufft.h> gets them through a one-line cufft.h that includes
 * this header (see the repository's integration notes).
 *
 * Error behaviour: the reference's functions return void and have no error
 * channel.  These abort the process with a message when the library reports
 * an error (empty cell, window leaving the grid, relaxation failure, HIP
 * error); org_last_error() keeps the code.
 */
#ifndef ORGANOID_DROPIN_H
#define ORGANOID_DROPIN_H

#include "organoid.h"

#ifndef ORGANOID_FFT_TYPES
#define ORGANOID_FFT_TYPES
typedef struct org_complex32 { float x, y; } cufftComplex;   /* rho_functions.h:10 */
typedef int cufftHandle;                                      /* declare_fft_tables.h:24 */
enum { CUFFT_SUCCESS = 0, CUFFT_C2C = 0x29, CUFFT_FORWARD = -1, CUFFT_INVERSE = 1 };
#endif

/* Plan creation called directly by the driver (run_with_active_gels.cpp:87-88).
 * Returns 0; the plan records the box size (must be 2*zone_size-2). */
int cufftPlan2d(cufftHandle *plan, int nx, int ny, int type);

/* ---- constants.h:4-124 (defined by the caller's rendered constants.cpp) ---- */
extern int Nsteps, zone_size, threshold_reshaping_steps, rdev_ini, rdev_threshold;
extern int ncontinue, ndisplay, nsaving, riprint, Ntotsave;
extern float dt, dx;
extern const int N_s;
extern const int L;
extern float R0, R, s0, Rint, Rout, sigma;
extern int N_ini;
extern const int Ntot;
extern int N_current;
extern float coupling_PF_AG;
extern int random_ontime, gen_ontime, active_gel_ontime, contract_ontime;
extern float seuil_gel;
extern float b_ap, b_bas, b_lat, a_tan_lateral, a_tan_apical, a_tan_basal;
extern float a_ort_lateral, a_ort_apical, a_ort_basal;
extern float matgenrate_lateral, degrate_lateral, matgenrate_apical, degrate_apical;
extern float matgenrate_basal, degrate_basal;
extern float friction, viscosity_ap, viscosity_bas, viscosity_lat;
extern float D0, tau, alpha, beta, eta, gamma_cell;
extern float D_lum, tau_lum, beta_lum_cells, eta_lum_cells, gamma_lum_cells, xi_lum;
extern float alpha_ECM, D_ECM, tau_ECM, beta_ECM_cells, eta_ECM_cells, gamma_ECM_cells, xi_ECM;
extern float beta_lum_ECM, wi, rho0, tol;
extern int ri, rimax;
extern float rdev, rtol, rdt, ans, sum_s;
extern int i, j, it, n, k, l, ir, jr, c;
extern int nelse, ip, im, jp, jm;

/* ---- declare_tables.h:5-70 ------------------------------------------------ */
extern void malloc_all(void);
extern float *__restrict phi, *__restrict phin, *__restrict rho, *__restrict rho_to_compare, *__restrict rhon;
extern float *__restrict lum, *__restrict ECM, *__restrict lumn, *__restrict ECMn;
extern float *__restrict psi, *__restrict temp_phase;
extern float *__restrict hfunc_cell, *__restrict hfunc_lum, *__restrict hfunc_ECM;
extern float *__restrict norm_grad_phi, *__restrict prod_norm_delta, *__restrict norm_grad_lum, *__restrict norm_grad_ecm;
extern float *__restrict a_tan_1, *__restrict a_ort_1, *__restrict b, *__restrict viscosity;
extern float *__restrict pi_xx, *__restrict pi_yy, *__restrict pi_xy;
extern float *__restrict phiout, *__restrict rhoout, *__restrict lumout, *__restrict ECMout;
extern float *__restrict Jxout, *__restrict Jyout, *__restrict vol_out;
extern float *__restrict J_x, *__restrict J_y;
extern int *__restrict xcenters, *__restrict ycenters;
extern float *__restrict matgenrates, *__restrict degrates, *__restrict matout, *__restrict degout;

/* ---- declare_fft_tables.h:8-24 ------------------------------------------- */
extern void ini_fft_all(void);
extern cufftComplex *__restrict temp_fft, *__restrict temp_fft_real, *__restrict fft_pi_xx, *__restrict fft_pi_yy;
extern cufftComplex *__restrict fft_pi_xy, *__restrict fft_rho_vx, *__restrict fft_rho_vy, *__restrict fft_vx;
extern cufftComplex *__restrict fft_vy, *__restrict fft_rho, *__restrict fft_react;
extern cufftHandle planf, planr;

/* ---- useful_functions.h:15-48 (the helpers the driver calls) -------------- */
extern void saving_out_functions(int it, float *phi, float *rho, float *ECM, float *lum, float *phiout,
                                 float *rhoout, float *lumout, float *ECMout);
extern void find_centres_cells(float *phi, int *xcenters, int *ycenters);
extern void compute_hfunc_3d_tables_kernels_loop_n(float *, float *, int *, int *);
extern void compute_all_hfunc_2d_tables_kernels(float *, float *, float *, float *);
extern void compute_psi_kernels(float *, float *);
extern void compute_normal_grad_phi(float *, float *);
extern void compute_normal_to_phi(float *phi, float *norm_grad_phi, float *J_x, float *J_y);
extern void sum_s_kernels(float *, float *, float *);
extern void boundary_conditions_cells(float *);
extern void boundary_conditions_lumen_and_ECM(float *, float *);
extern void copy_cell_tables_1_by_1(int n, float *phase, float *phasen);
extern void copy_tables_ECM_or_lumen(float *phase, float *phasen);

/* ---- rho_functions.h:10-26 ---------------------------------------------- */
extern void compute_prod_norm_delta(float *, float *, float *, int *, int *);
extern void create_turnover_rates_tensions(float *degrates, float *matgenrates, float *a_tan_1, float *a_ort_1,
                                           float *viscosity, float *b, float *phi, float *lum, float *ECM,
                                           int *xcenters, int *ycenters);
extern void update_all_cortices(float *phi, float *rho, float *rhon, int *xcenters, int *ycenters,
                                float *degrates, float *matgenrates, float *a_tan_1, float *a_ort_1,
                                float *viscosity, float *b, float *norm_grad_phi, float *prod_norm_delta,
                                cufftHandle planf, cufftHandle planr, cufftComplex *temp_fft,
                                cufftComplex *temp_fft_real, cufftComplex *fft_rho, cufftComplex *fft_pi_xx,
                                cufftComplex *fft_pi_xy, cufftComplex *fft_pi_yy, cufftComplex *fft_vx,
                                cufftComplex *fft_vy, cufftComplex *fft_react, cufftComplex *fft_rho_vx,
                                cufftComplex *fft_rho_vy);

/* ---- cell_phase_functions.h:18-39 ---------------------------------------- */
extern void update_all_cells(int it, float *phi, float *phin, float *norm_grad_phi, float *rho, float *hfunc_cell,
                             float *vol_diff, float *psi, int *xcenters, int *ycenters, float *hfunc_lum,
                             float *hfunc_ECM);
extern void update_lumen_and_ECM(float *lum, float *lumn, float *ECM, float *ECMn, float *hfunc_lum,
                                 float *hfunc_ECM, float *psi);
extern void reshaping_cells(int, float *, float *, float *, float *, float *, int *, int *);
extern void reshaping_lumen_or_ECM(int, float *, float *, float);

/* ---- the IC program's other calls (prepare_initial_files.cpp:67, 100) ----
 * update_rho_to_shape runs on the device (cell_phase_functions.cpp:42-62);
 * check_nan counts NaNs in the first two cell planes on the host
 * (useful_functions.cpp:11-25).  reshaping_cells after
 * copy_cell_tables_1_by_1(n, phin, phi) for the last cell (and before any
 * update_all_cells) starts from phi, as the IC program's does. */
extern void update_rho_to_shape(float *rho, float *phi, float rho0, float tol);
extern int check_nan(float *table);

/* ---- initialisation_functions.h:1-12 (host code, organoid_initial.cpp) ----
 * Analytic initial state of the IC program (prepare_initial_files.cpp:56-66)
 * and the readers of the .bin state files (run_with_active_gels.cpp:63-66).
 * The .bin layout keeps the reference's 8-byte records: element k at byte 8k.
 * A missing or short file aborts with a message. */
extern void ini_phin(float *, int n);                        /* initialisation_functions.cpp:36 */
extern void ini_phi(float *);                                 /* :117 */
extern void ini_rho(float *, float *, float rho0, float tol, int n);  /* :64 */
extern void ini_rhon(float *, int n);                         /* :75 */
extern void ini_ECM_lumen(float *, float *, float *);         /* :82 */
extern void ini_phi_files(float *);                           /* :10 */
extern void ini_lumen_files(float *);                         /* :163 */
extern void ini_ECM_files(float *);                           /* :174 */
extern void ini_rho_files(float *);                           /* :43 */

/* Library-side handle of the organoid behind the globals (no counterpart in
 * the reference): created on the first timestep call, destroyed at exit. */
org_ctx *organoid_dropin_context(void);

#endif /* ORGANOID_DROPIN_H */
