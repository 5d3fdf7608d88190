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
 * Precision.  Built as liborganoid_dropin.so, the library serves the
 * reference as written (float tables and constants; org_real_t is float).
 * Built with ORGANOID_DROPIN_F64 as liborganoid_dropin_f64.so, it serves the
 * reference retyped to double (every `float` of its sources read as
 * `double`, the FFT buffers {double x, y}, plans CUFFT_Z2Z or CUFFT_C2C):
 * the same names with double parameters and tables, on FP64 device state
 * (org_create_f64).  A caller defines ORGANOID_DROPIN_F64 before including
 * this header for the FP64 form.
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
#ifdef ORGANOID_DROPIN_F64
/* The FP64 form: the reference with `float` retyped to `double`; the FFT
 * buffers hold {double x, y} under either cuFFT name. */
typedef double org_real_t;
typedef struct org_complex64 { double x, y; } cufftDoubleComplex;
typedef cufftDoubleComplex cufftComplex;
#else
typedef float org_real_t;
typedef struct org_complex32 { float x, y; } cufftComplex;   /* rho_functions.h:10 */
#endif
typedef int cufftHandle;                                      /* declare_fft_tables.h:24 */
enum { CUFFT_SUCCESS = 0, CUFFT_C2C = 0x29, CUFFT_Z2Z = 0x69, CUFFT_FORWARD = -1, CUFFT_INVERSE = 1 };
#endif

/* Plan creation called directly by the driver (run_with_active_gels.cpp:87-88).
 * Returns 0; the plan records the box size (must be 2*zone_size-2). */
int cufftPlan2d(cufftHandle *plan, int nx, int ny, int type);

/* ---- constants.h:4-124 (defined by the caller's rendered constants.cpp) ---- */
extern int Nsteps, zone_size, threshold_reshaping_steps, rdev_ini, rdev_threshold;
extern int ncontinue, ndisplay, nsaving, riprint, Ntotsave;
extern org_real_t dt, dx;
extern const int N_s;
extern const int L;
extern org_real_t R0, R, s0, Rint, Rout, sigma;
extern int N_ini;
extern const int Ntot;
extern int N_current;
extern org_real_t coupling_PF_AG;
extern int random_ontime, gen_ontime, active_gel_ontime, contract_ontime;
extern org_real_t seuil_gel;
extern org_real_t b_ap, b_bas, b_lat, a_tan_lateral, a_tan_apical, a_tan_basal;
extern org_real_t a_ort_lateral, a_ort_apical, a_ort_basal;
extern org_real_t matgenrate_lateral, degrate_lateral, matgenrate_apical, degrate_apical;
extern org_real_t matgenrate_basal, degrate_basal;
extern org_real_t friction, viscosity_ap, viscosity_bas, viscosity_lat;
extern org_real_t D0, tau, alpha, beta, eta, gamma_cell;
extern org_real_t D_lum, tau_lum, beta_lum_cells, eta_lum_cells, gamma_lum_cells, xi_lum;
extern org_real_t alpha_ECM, D_ECM, tau_ECM, beta_ECM_cells, eta_ECM_cells, gamma_ECM_cells, xi_ECM;
extern org_real_t beta_lum_ECM, wi, rho0, tol;
extern int ri, rimax;
extern org_real_t rdev, rtol, rdt, ans, sum_s;
extern int i, j, it, n, k, l, ir, jr, c;
extern int nelse, ip, im, jp, jm;

/* ---- declare_tables.h:5-70 ------------------------------------------------ */
extern void malloc_all(void);
extern org_real_t *__restrict phi, *__restrict phin, *__restrict rho, *__restrict rho_to_compare, *__restrict rhon;
extern org_real_t *__restrict lum, *__restrict ECM, *__restrict lumn, *__restrict ECMn;
extern org_real_t *__restrict psi, *__restrict temp_phase;
extern org_real_t *__restrict hfunc_cell, *__restrict hfunc_lum, *__restrict hfunc_ECM;
extern org_real_t *__restrict norm_grad_phi, *__restrict prod_norm_delta, *__restrict norm_grad_lum, *__restrict norm_grad_ecm;
extern org_real_t *__restrict a_tan_1, *__restrict a_ort_1, *__restrict b, *__restrict viscosity;
extern org_real_t *__restrict pi_xx, *__restrict pi_yy, *__restrict pi_xy;
extern org_real_t *__restrict phiout, *__restrict rhoout, *__restrict lumout, *__restrict ECMout;
extern org_real_t *__restrict Jxout, *__restrict Jyout, *__restrict vol_out;
extern org_real_t *__restrict J_x, *__restrict J_y;
extern int *__restrict xcenters, *__restrict ycenters;
extern org_real_t *__restrict matgenrates, *__restrict degrates, *__restrict matout, *__restrict degout;

/* ---- declare_fft_tables.h:8-24 ------------------------------------------- */
extern void ini_fft_all(void);
extern cufftComplex *__restrict temp_fft, *__restrict temp_fft_real, *__restrict fft_pi_xx, *__restrict fft_pi_yy;
extern cufftComplex *__restrict fft_pi_xy, *__restrict fft_rho_vx, *__restrict fft_rho_vy, *__restrict fft_vx;
extern cufftComplex *__restrict fft_vy, *__restrict fft_rho, *__restrict fft_react;
extern cufftHandle planf, planr;

/* ---- useful_functions.h:15-48 (the helpers the driver calls) -------------- */
extern void saving_out_functions(int it, org_real_t *phi, org_real_t *rho, org_real_t *ECM, org_real_t *lum, org_real_t *phiout,
                                 org_real_t *rhoout, org_real_t *lumout, org_real_t *ECMout);
extern void find_centres_cells(org_real_t *phi, int *xcenters, int *ycenters);
extern void compute_hfunc_3d_tables_kernels_loop_n(org_real_t *, org_real_t *, int *, int *);
extern void compute_all_hfunc_2d_tables_kernels(org_real_t *, org_real_t *, org_real_t *, org_real_t *);
extern void compute_psi_kernels(org_real_t *, org_real_t *);
extern void compute_normal_grad_phi(org_real_t *, org_real_t *);
extern void compute_normal_to_phi(org_real_t *phi, org_real_t *norm_grad_phi, org_real_t *J_x, org_real_t *J_y);
extern void sum_s_kernels(org_real_t *, org_real_t *, org_real_t *);
extern void boundary_conditions_cells(org_real_t *);
extern void boundary_conditions_lumen_and_ECM(org_real_t *, org_real_t *);
extern void copy_cell_tables_1_by_1(int n, org_real_t *phase, org_real_t *phasen);
extern void copy_tables_ECM_or_lumen(org_real_t *phase, org_real_t *phasen);

/* ---- rho_functions.h:10-26 ---------------------------------------------- */
extern void compute_prod_norm_delta(org_real_t *, org_real_t *, org_real_t *, int *, int *);
extern void create_turnover_rates_tensions(org_real_t *degrates, org_real_t *matgenrates, org_real_t *a_tan_1, org_real_t *a_ort_1,
                                           org_real_t *viscosity, org_real_t *b, org_real_t *phi, org_real_t *lum, org_real_t *ECM,
                                           int *xcenters, int *ycenters);
extern void update_all_cortices(org_real_t *phi, org_real_t *rho, org_real_t *rhon, int *xcenters, int *ycenters,
                                org_real_t *degrates, org_real_t *matgenrates, org_real_t *a_tan_1, org_real_t *a_ort_1,
                                org_real_t *viscosity, org_real_t *b, org_real_t *norm_grad_phi, org_real_t *prod_norm_delta,
                                cufftHandle planf, cufftHandle planr, cufftComplex *temp_fft,
                                cufftComplex *temp_fft_real, cufftComplex *fft_rho, cufftComplex *fft_pi_xx,
                                cufftComplex *fft_pi_xy, cufftComplex *fft_pi_yy, cufftComplex *fft_vx,
                                cufftComplex *fft_vy, cufftComplex *fft_react, cufftComplex *fft_rho_vx,
                                cufftComplex *fft_rho_vy);

/* ---- cell_phase_functions.h:18-39 ---------------------------------------- */
extern void update_all_cells(int it, org_real_t *phi, org_real_t *phin, org_real_t *norm_grad_phi, org_real_t *rho, org_real_t *hfunc_cell,
                             org_real_t *vol_diff, org_real_t *psi, int *xcenters, int *ycenters, org_real_t *hfunc_lum,
                             org_real_t *hfunc_ECM);
extern void update_lumen_and_ECM(org_real_t *lum, org_real_t *lumn, org_real_t *ECM, org_real_t *ECMn, org_real_t *hfunc_lum,
                                 org_real_t *hfunc_ECM, org_real_t *psi);
extern void reshaping_cells(int, org_real_t *, org_real_t *, org_real_t *, org_real_t *, org_real_t *, int *, int *);
extern void reshaping_lumen_or_ECM(int, org_real_t *, org_real_t *, org_real_t);

/* ---- the IC program's other calls (prepare_initial_files.cpp:67, 100) ----
 * update_rho_to_shape runs on the device (cell_phase_functions.cpp:42-62);
 * check_nan counts NaNs in the first two cell planes on the host
 * (useful_functions.cpp:11-25).  reshaping_cells after
 * copy_cell_tables_1_by_1(n, phin, phi) for the last cell (and before any
 * update_all_cells) starts from phi, as the IC program's does. */
extern void update_rho_to_shape(org_real_t *rho, org_real_t *phi, org_real_t rho0, org_real_t tol);
extern int check_nan(org_real_t *table);

/* ---- initialisation_functions.h:1-12 (host code, organoid_initial.cpp) ----
 * Analytic initial state of the IC program (prepare_initial_files.cpp:56-66)
 * and the readers of the .bin state files (run_with_active_gels.cpp:63-66).
 * The .bin layout keeps the reference's 8-byte records: element k at byte 8k.
 * A missing or short file aborts with a message. */
extern void ini_phin(org_real_t *, int n);                        /* initialisation_functions.cpp:36 */
extern void ini_phi(org_real_t *);                                 /* :117 */
extern void ini_rho(org_real_t *, org_real_t *, org_real_t rho0, org_real_t tol, int n);  /* :64 */
extern void ini_rhon(org_real_t *, int n);                         /* :75 */
extern void ini_ECM_lumen(org_real_t *, org_real_t *, org_real_t *);         /* :82 */
extern void ini_phi_files(org_real_t *);                           /* :10 */
extern void ini_lumen_files(org_real_t *);                         /* :163 */
extern void ini_ECM_files(org_real_t *);                           /* :174 */
extern void ini_rho_files(org_real_t *);                           /* :43 */

/* Library-side handle of the organoid behind the globals (no counterpart in
 * the reference): created on the first timestep call, destroyed at exit. */
org_ctx *organoid_dropin_context(void);

#endif /* ORGANOID_DROPIN_H */
