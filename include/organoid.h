/*
 * organoid.h -- C-ABI of the MI355X (gfx950) organoid timestep library.
 *
 * Plain C types only: pointers, sizes, int status codes.  Every entry point
 * below is the device-resident counterpart of one function (or one fused
 * group of functions) on the per-timestep path of the reference simulator;
 * the comment on each names the reference function it replaces as
 * file:line (files under functions/ and main_code/ of
 * tristanguyomar/multicellular_phase_field_active_gel_organoids).
 *
 * The same header is included by the CPU oracle (oracle/organoid_oracle.c)
 * for org_params, so both sides read identical constants.
 *
 * Conventions
 *   - Planes are row-major: plane[n][i][j] = base[n*N*N + i*N + j], i is the
 *     row ("x" in the reference, useful_functions.cpp:214), j the column.
 *   - Per-cell fields (phi, rho, hfunc) hold ncell planes; shared fields
 *     (lum, ECM, psi, ...) hold one plane.
 *   - Host buffers passed to org_set_* / org_get_* are FP32 unless named _f64.
 *   - Every entry point returns 0 on success or an ORG_E* code; the first
 *     error is also kept in a sticky slot (org_last_error).  The reference has
 *     no error channel (all its functions return void); see the per-code
 *     comments for which reference behaviour each one turns into an error.
 */
#ifndef ORGANOID_H
#define ORGANOID_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------- */
enum {
    ORG_OK = 0,
    ORG_EINVAL = 1,       /* bad argument / size                                */
    ORG_EHIP = 2,         /* a HIP runtime call failed                          */
    ORG_EEMPTY_CELL = 3,  /* no phi>0.5 pixel: the reference divides by zero
                             (useful_functions.cpp:221)                         */
    ORG_EWINDOW = 4,      /* a cell window (+stencil halo) leaves the grid
                             interior; the reference reads out of range
                             ("careful here", rho_functions.cpp:43)             */
    ORG_ERELAX = 5,       /* a reshaping loop passed rimax ("Relaxation
                             failure!", cell_phase_functions.cpp:565, 862).
                             A warning, as in the reference: the loop stops,
                             stepping goes on, the call returns ORG_OK and
                             org_get_relax_failures counts it.  Only the
                             oracle and org_error_string use the code.       */
    ORG_ENOTBUILT = 6,    /* kernels for this build are missing                 */
    ORG_ESTATE = 7,       /* call order / state error                           */
    ORG_ETIMEOUT = 8      /* a wait between workgroups of one launch passed its
                             time limit (2 s); the step's results are invalid  */
};

/* ---- configuration ---------------------------------------------------------
 * The reference keeps these as compile-time globals rendered from
 * constants_template.cpp:1-130 (declared in constants.h:4-124).  Values are
 * held in double; the FP32 GPU path rounds each to float once, as the
 * reference's float globals do.                                              */
typedef struct org_params {
    int32_t N;                 /* N_s, grid side                 constants_template.cpp:145 */
    int32_t ncell;             /* Ntot = N_ini = N_current       constants_template.cpp:158-160 */
    int32_t zone_size;         /* half window side               constants_template.cpp:162 */
    int32_t rimax;             /* reshaping iteration cap        constants_template.cpp:239 */
    int32_t rdev_threshold;    /* lumen/ECM stop factor (int)    constants_template.cpp:130 */
    int32_t rdev_ini;          /* lumen/ECM start factor (int)   constants_template.cpp:131 */
    int32_t n_current;         /* N_current (divides cell rdev)  cell_phase_functions.cpp:560 */
    int32_t use_s0_cell;       /* 1: per-cell target areas (extension), 0: global s0 */
    double dt, dx;             /* constants_template.cpp:141-142 */
    double R;                  /* R (reshaping tolerance radius) constants_template.cpp:150 */
    double s0;                 /* target area                    constants_template.cpp:151 */
    double sigma;              /* delta width                    constants_template.cpp:155 */
    double coupling_PF_AG;     /*                                 constants_template.cpp:166 */
    double tau, alpha, beta, eta;                       /* constants_template.cpp:208-211 */
    double tau_lum, beta_lum_cells, eta_lum_cells, xi_lum;            /* :215-219 */
    double alpha_ECM, tau_ECM, beta_ECM_cells, eta_ECM_cells, xi_ECM; /* :222-227 */
    double beta_lum_ECM;                                              /* :230 */
    double wi, rtol, rdt;                               /* :233, 241, 242 */
    double friction;                                    /* :201 */
    double visc_ap, visc_bas, visc_lat;                 /* :202-204 */
    double a_ort_ap, a_ort_bas, a_ort_lat;              /* :181-183 */
    double a_tan_ap, a_tan_bas, a_tan_lat;              /* :185-187 */
    double b_ap, b_bas, b_lat;                          /* :189-191 */
    double gen_ap, gen_bas, gen_lat;                    /* :193,196,198 */
    double deg_ap, deg_bas, deg_lat;                    /* :194,197,199 */
} org_params;

/* ---- field identifiers (org_set_field / org_get_field) ------------------- */
enum {
    ORG_F_PHI = 0,     /* ncell planes  declare_tables.cpp:97  */
    ORG_F_RHO = 1,     /* ncell planes  declare_tables.cpp:99  */
    ORG_F_HFUNC = 2,   /* ncell planes  declare_tables.cpp:111 */
    ORG_F_LUM = 3,     /* 1 plane       declare_tables.cpp:102 */
    ORG_F_ECM = 4,     /* 1 plane       declare_tables.cpp:103 */
    ORG_F_PSI = 5,     /* 1 plane       declare_tables.cpp:107 */
    ORG_F_HLUM = 6,    /* 1 plane       declare_tables.cpp:112 */
    ORG_F_HECM = 7,    /* 1 plane       declare_tables.cpp:113 */
    ORG_F_LUMN = 8,    /* 1 plane       declare_tables.cpp:104 */
    ORG_F_ECMN = 9,    /* 1 plane       declare_tables.cpp:105 */
    ORG_F_COUNT = 10
};

/* window-local outputs (org_get_window): [ncell][2*zone_size][2*zone_size]
 * anchored at (xc-zone_size, yc-zone_size) of each cell's current centre, or
 * [ncell][M][M] anchored at (xc-zone_size+1, yc-zone_size+1) for the FFT box
 * (M = 2*zone_size-2, useful_functions.cpp:618-619).                         */
enum {
    ORG_W_PHIN = 0,    /* phin after update_all_cells (window)            */
    ORG_W_RHON = 1,    /* rhon after update_all_cortices (FFT box, M x M) */
    ORG_W_VX = 2,      /* Re(F^-1 v_x)/M^2 (FFT box), rho_functions.cpp:449-450 */
    ORG_W_VY = 3,      /* Re(F^-1 v_y)/M^2 (FFT box), rho_functions.cpp:455-456 */
    ORG_W_COUNT = 4
};

typedef struct org_ctx org_ctx;

/* ---- lifetime ------------------------------------------------------------ */
/* Allocate device state for one organoid.  s0_cell (ncell doubles) is used
 * when p->use_s0_cell is 1, else ignored.  Replaces malloc_all
 * (declare_tables.cpp:89-143) and ini_fft_all (declare_fft_tables.cpp:26-41);
 * device memory is laid out once for the whole run.
 *
 * org_create holds the state in FP32, as the reference's float build
 * (constants.h, `float` tables).  With zone_size 65 (FFT box 128) every
 * entry point below runs the LDS-resident kernels; any other zone_size
 * (3..513: boxes M = 2*zone_size-2 up to 1024, powers of two or not, e.g.
 * the reference template's zone_size 200, M = 398) runs a per-function
 * path on full planes in global memory (one launch per reference function,
 * 2D transforms in double by radix-2 or Bluestein).  org_create_f64 holds
 * the state in FP64 (the reference with `float` retyped to `double`) on
 * that same full-plane path, for any zone_size.  Contexts of the full-plane
 * path step host-driven (org_step runs the per-function sequence; the
 * reshaping loops' stop tests read back each sweep), ignore org_set_tuning,
 * and cannot join a batch or run the surrogate (ORG_EINVAL).                */
int org_create(const org_params *p, const double *s0_cell, org_ctx **out);
int org_create_f64(const org_params *p, const double *s0_cell, org_ctx **out);
int org_destroy(org_ctx *ctx);
/* 4 (FP32 state) or 8 (FP64 state). */
int org_get_precision(org_ctx *ctx, int32_t *real_bytes);

/* Upload / download a whole field (all planes).  Upload plays the role of
 * the driver's `#pragma acc data copyin` (run_with_active_gels.cpp:83).  After
 * uploading phi and hfunc, org_prepare must run before the first step. */
int org_set_field(org_ctx *ctx, int field, const float *host);
int org_get_field(org_ctx *ctx, int field, float *host);
/* Window-local debug outputs of the last step (see ORG_W_*). */
int org_get_window(org_ctx *ctx, int which, float *host);
/* The same with double host buffers: unchanged for FP64 contexts, rounded
 * to float (set) or widened (get) for FP32 ones. */
int org_set_field_f64(org_ctx *ctx, int field, const double *host);
int org_get_field_f64(org_ctx *ctx, int field, double *host);
int org_get_window_f64(org_ctx *ctx, int which, double *host);
/* Centres (ncell ints each), areas s and f0 (ncell doubles each). */
int org_get_centres(org_ctx *ctx, int32_t *xc, int32_t *yc);
int org_get_areas(org_ctx *ctx, double *s, double *f0);
/* Reshaping iteration counts of the last step: cells[ncell], lum, ecm. */
int org_get_iterations(org_ctx *ctx, int32_t *cells, int32_t *lum, int32_t *ecm);

/* Scan the uploaded phi/hfunc once to seed each cell's support box (the
 * rectangle outside which phi<=0.5 and hfunc==0; both fields only change
 * inside the cell's windows afterwards).  Also zeroes psi and copies
 * lum->lumn, ECM->ECMn (run_with_active_gels.cpp:92-95).  rhon and phin
 * need no copy (run_with_active_gels.cpp:96-98): rhon lives only on each
 * cell's FFT box and is committed over rho there, phin only on each
 * cell's window, so outside them both equal rho and phi by construction. */
int org_prepare(org_ctx *ctx);

/* ---- per-function entry points (the plain drop-in form) -------------------
 * Each runs over all cells in one launch where the reference launched one
 * kernel per cell.  Fields the driver never reads back (norm_grad_phi,
 * prod_norm_delta, the six turnover maps, pi_xx/xy/yy, J_x/J_y, temp_phase)
 * are not materialised: they are recomputed inside the kernels that consume
 * them.                                                                       */
int org_find_centres_cells(org_ctx *ctx);          /* useful_functions.cpp:193-227 */
int org_compute_hfunc_cells(org_ctx *ctx);         /* useful_functions.cpp:233-251 */
int org_compute_hfunc_lum_ecm(org_ctx *ctx);       /* useful_functions.cpp:272-285 */
int org_compute_psi(org_ctx *ctx);                 /* useful_functions.cpp:305-322 */
int org_sum_s(org_ctx *ctx);                       /* useful_functions.cpp:168-190 */
/* create_turnover_rates_tensions (rho_functions.cpp:286-405),
 * compute_prod_norm_delta (rho_functions.cpp:13-30), the compute_pi_* family
 * (rho_functions.cpp:32-115) and the whole per-cell spectral loop
 * update_all_cortices (rho_functions.cpp:409-471), fused. */
int org_update_all_cortices(org_ctx *ctx);
int org_update_lumen_and_ECM(org_ctx *ctx);        /* cell_phase_functions.cpp:372-444
                                                      + useful_functions.cpp:439-462 */
int org_reshaping_lumen_or_ECM(org_ctx *ctx, int which_ecm);
                                                   /* cell_phase_functions.cpp:709-871; the
                                                      field ends at the second-to-last iterate,
                                                      lumn / ECMn at the last (:777-785) */
int org_update_all_cells(org_ctx *ctx);            /* cell_phase_functions.cpp:126-217
                                                      (+ compute_normal_grad_phi
                                                      useful_functions.cpp:326-354) */
int org_reshaping_cells(org_ctx *ctx);             /* cell_phase_functions.cpp:447-580
                                                      (+ compute_normal_to_phi
                                                      useful_functions.cpp:356-388) */
/* Initial-condition pipeline (initialisation_code/prepare_initial_files.cpp).
 * reshaping_cells with phin = phi (the program's copy, :87-90);
 * update_rho_to_shape (cell_phase_functions.cpp:42-62) on every plane;
 * org_relax_initial_state: the program's whole device part (:82-110), with
 * psi, h(lum), h(ECM) zero (the program never computes them), ending with
 * phi, lum and ECM at their second-to-last reshaping iterates as the
 * program saves them. */
int org_reshaping_cells_from_phi(org_ctx *ctx);
int org_update_rho_to_shape(org_ctx *ctx, float rho0, float tol);
/* The same with double constants (the FP64 form's double rho0 and tol;
 * an FP32 context rounds them to float). */
int org_update_rho_to_shape_f64(org_ctx *ctx, double rho0, double tol);
int org_relax_initial_state(org_ctx *ctx, float rho0, float tol);
int org_commit_step(org_ctx *ctx);                 /* useful_functions.cpp:679-709 as
                                                      called at run_with_active_gels.cpp:175-182 */
int org_commit_rho(org_ctx *ctx);                  /* rho <- rhon for every cell,
                                                      useful_functions.cpp:679-693 (:175-178) */
int org_commit_lumecm(org_ctx *ctx);               /* ECM <- ECMn, lum <- lumn,
                                                      useful_functions.cpp:695-709 (:181-182) */

/* ---- batched form -----------------------------------------------------------
 * One full timestep = the loop body run_with_active_gels.cpp:118-184 with no
 * host round trip: one launch per stage for all cells, and one launch for
 * the spectral cortex update, the phase-field update and reshaping of every
 * cell and the lumen/ECM reshaping loops together (their stop tests run on
 * the device).  use_graph=1 captures the step into a hipGraph on first use
 * and replays it.                                                            */
int org_step(org_ctx *ctx, int it0, int nsteps, int use_graph);
/* Captures and instantiates the hipGraphs that org_step(ctx, ., nsteps, 1)
 * replays, without running them (the state does not change), so that a
 * later replay of that length starts at once.  No counterpart in the
 * reference (like cufftPlan2d, set-up of launches to come).  A no-op on the
 * host-driven full-plane path. */
int org_graph_prepare(org_ctx *ctx, int nsteps);

/* ---- launch tuning of the batched step ---------------------------------------
 * No counterpart in the reference (its loop bounds are fixed by N_s and the
 * OpenACC defaults).  lum_blocks: lumen/ECM workgroups of the step's main
 * launch (1..min(128, compute units); they wait on each other, so all must
 * fit on the device at once); roles: workgroups of each non-cell role of the
 * launches before and between steps (psi tiles, h(lum)/h(ECM) parts,
 * lumen/ECM copy-back; 1..512); graph_steps: steps per replayed hipGraph
 * (1..64); shared_bracket: 1 = the copy-back launch writes the lumen/ECM
 * part of the cell update's bracket once per step into a shared FP64 plane
 * that every cell's update reads (28 loads per pixel), 0 = each cell's
 * update forms it from h(lum), h(ECM) itself (45 loads per pixel, a shorter
 * copy-back launch).  The defaults are sized from the device's compute-unit
 * count; autotune.py picks them by timing.  None of them changes results:
 * the partitions feed only pointwise updates, per-tile sums in a fixed order
 * and box unions, and both bracket forms evaluate the same expression on the
 * same operands.  Setting a tuning drops the captured graphs.                */
typedef struct org_tuning {
    int32_t lum_blocks;
    int32_t roles;
    int32_t graph_steps;
    int32_t shared_bracket;
} org_tuning;
int org_get_tuning(org_ctx *ctx, org_tuning *out);
int org_set_tuning(org_ctx *ctx, const org_tuning *tuning);

/* ---- the velocity surrogate in the batched step ------------------------------
 * No counterpart in the reference (its velocity always comes from the
 * spectral solve inside update_all_cortices, rho_functions.cpp:134-190);
 * the network is the FNO of organoid_fno.h.  Both run the batched step
 * eagerly (no graph), one step at a time.
 *
 * org_surrogate_samples: nsteps exact steps from it0, recording for step s
 * and cell n, at index s*ncell + n, the network input (phi_n, rho_n) on the
 * cell's FFT box at the start of the step (x_dev, [.][2][M][M] FP32), the
 * solve's velocity (v_dev, [.][2][M][M], vx then vy) and, if res_dev is not
 * NULL, the judge's residual of that exact velocity (res_dev, [.] FP32: the
 * judge's floor on an exact solution).
 *
 * org_step_surrogate: nsteps steps in which the cortex velocity of every
 * cell comes from the network (net), or, with net NULL, from v_given
 * ([ncell][2][M][M] on the device, the same for every step: a caller's own
 * surrogate, and the test of this path), except that (a) every step with
 * (it0 + s) % k_exact == 0 solves every cell exactly, and (b) a cell whose
 * judge residual (ORG_OPT_SURROGATE_JUDGE: spectral by default, or the
 * finite-difference residual of the force balance) exceeds `threshold` is
 * solved exactly in that step.  exact_cells (nsteps int32, may be NULL):
 * cells solved exactly per step; residuals (nsteps*ncell FP32, may be
 * NULL): the judge's residual per step and cell (-1 where the step was
 * exact by (a)).  The network must have M = 128 and batch_max >= ncell; it
 * runs on the organoid's stream for the call. */
typedef struct fno_net fno_net;
int org_surrogate_samples(org_ctx *ctx, int it0, int nsteps, float *x_dev, float *v_dev, float *res_dev);
int org_step_surrogate(org_ctx *ctx, fno_net *net, const float *v_given, int it0, int nsteps, int k_exact,
                       float threshold, int32_t *exact_cells, float *residuals);

/* ---- batches of organoids ----------------------------------------------------
 * The counterpart of a setup_simulation.py parameter sweep
 * (main_code/setup_simulation.py:29-78 renders one constants file and one
 * run directory per combination, each run stepping alone): the organoids of
 * a batch, each an org_ctx with its own grid size, cell count and constants,
 * advance together with one launch per stage for all of them
 * (org_batch_step = org_step over every member, one replayable hipGraph).
 * Every member evolves bit for bit as it would under org_step.
 *
 * org_batch_create takes 1..255 distinct, uploaded contexts that belong to no
 * other batch; it sets their workgroup counts for the shared launches
 * (org_step keeps working on a member).  Destroy the batch before any of its
 * contexts.  Errors: ORG_EINVAL (bad arguments, or more lumen/ECM workgroups
 * than the device can hold at once), ORG_ESTATE (a member not prepared), and
 * the members' device errors after org_batch_step (the first one is
 * returned). */
typedef struct org_batch org_batch;
int org_batch_create(org_ctx *const *ctxs, int n, org_batch **out);
int org_batch_step(org_batch *batch, int nsteps, int use_graph);
int org_batch_sync(org_batch *batch);
int org_batch_destroy(org_batch *batch);

/* Device synchronisation and timing.  org_timer_* brackets work on the
 * library's stream with HIP events; org_kernel_time returns the summed
 * duration (ms) and launch count of one kernel family since the last reset,
 * measured with events recorded around its launches on the same stream. */
int org_sync(org_ctx *ctx);
int org_profile_enable(org_ctx *ctx, int on);
int org_kernel_time(org_ctx *ctx, int kernel_id, double *ms, int64_t *launches);
enum {
    ORG_K_CENTRES = 0, ORG_K_PREP = 1, ORG_K_PSI = 2, ORG_K_LUMECM = 3,
    ORG_K_LUMRESHAPE = 4, ORG_K_SPECTRAL = 5, ORG_K_CELLS = 6, ORG_K_COMMIT = 7,
    ORG_K_STEP = 8,       /* the batched step's main launch (org_step) */
    ORG_K_MID = 9,        /* batched: copy-back merged with the next step's first part */
    ORG_K_COUNT = 10
};

/* Options.  ORG_OPT_WRITE_VELOCITY (default 1): keep the FFT-box velocity of
 * the last step for org_get_window.  ORG_OPT_SURROGATE_JUDGE: the judge of
 * org_step_surrogate and of org_surrogate_samples' residuals,
 * ORG_JUDGE_SPECTRAL (default): the candidate's relative L2 distance from
 * the per-mode solution of update_rho_velocity's system (rho_functions.cpp:
 * 133-190), computed from F(pi) and F(v) (three of the exact solve's five
 * transforms; the exact solve scores ~1e-6), or ORG_JUDGE_FINITE_DIFF: the
 * finite-difference residual of the force balance over |div pi| on the
 * box (cheap, but the exact solve itself scores ~0.5 on sharp cortices). */
enum { ORG_OPT_WRITE_VELOCITY = 0, ORG_OPT_SURROGATE_JUDGE = 1 };
enum { ORG_JUDGE_SPECTRAL = 0, ORG_JUDGE_FINITE_DIFF = 1 };
int org_set_option(org_ctx *ctx, int option, int value);

/* ---- allocation-free form -----------------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the work of org_compute_hfunc_lum_ecm
 * (compute_all_hfunc_2d_tables_kernels, useful_functions.cpp:272-285, and
 * the FP64 sum of h(ECM) dx^2 that update_lumen_and_ECM takes,
 * cell_phase_functions.cpp:376-386) on caller-owned device planes, with no
 * org_ctx.  lum, ecm: N x N device planes (read); hlum, hecm: N x N device
 * planes (written, h(u) = u^2 (3 - 2u)); sum_ecm: one device double
 * (written).  stream: a hipStream_t, or NULL for the default stream.  The
 * results equal org_compute_hfunc_lum_ecm's bit for bit.  It calls no
 * allocating routine: its partial sums and counter live in a static device
 * array (about 33 KB per GPU, held until the process ends), so N must be
 * 8 .. 4096 (ORG_EINVAL with a message otherwise), and calls that overlap
 * on one GPU (other streams, other host threads) race on that array. */
int org_compute_hfunc_lum_ecm_noalloc(int N, double dx, const float *lum, const float *ecm, float *hlum,
                                      float *hecm, double *sum_ecm, void *stream);

/* Test hook (no counterpart in the reference): one unnormalised 128 x 128
 * 2D DFT through the LDS transform used by org_update_all_cortices.  data:
 * 128*128 interleaved (re, im) floats on the host, sign -1 or +1. */
int org_debug_fft2d(float *data, int sign);
/* Test hook (no counterpart in the reference): one unnormalised M x M 2D DFT
 * (M = 2..1024) through the full-plane path's transform: rows then
 * columns, each line in double (radix-2 for powers of two, Bluestein's
 * chirp-z otherwise), the input and output stored as float (f64 = 0) or
 * double (f64 = 1) as that path's contexts store them, with the schedule
 * (fuse, lines_per_wg) of org_set_fft_schedule.  data: M*M interleaved
 * (re, im) doubles on the host, sign -1 or +1. */
int org_debug_fft2d_generic(int M, double *data, int sign, int f64, int fuse, int lines_per_wg);

/* The full-plane path's transform schedule (no counterpart in the reference,
 * whose transforms are cuFFT plans): radix-2 stages fused per LDS round
 * trip (fuse 1, 2, 3: radix 2, 4, 8 passes) and lines per workgroup (1, 2,
 * 4; lines_per_wg x the power-of-two line length <= 2048).  Every schedule
 * gives the same values (each butterfly is the same operation on the same
 * operands); org_create / org_create_f64 time all of them on the step's
 * largest batch of transforms and keep the fastest.  Contexts on the LDS
 * path (one 16 x 8 schedule) report 0, 0 and refuse the setter. */
int org_get_fft_schedule(org_ctx *ctx, int32_t *fuse, int32_t *lines_per_wg);
int org_set_fft_schedule(org_ctx *ctx, int fuse, int lines_per_wg);

/* Calls on this context in which a reshaping loop passed rimax (the
 * reference prints "Relaxation failure!" and goes on,
 * cell_phase_functions.cpp:565-568, 862-865). */
int org_get_relax_failures(org_ctx *ctx, int64_t *count);

/* Sticky first error and a readable message for any code. */
int org_last_error(void);
const char *org_error_string(int code);
/* Library and kernel build identification (gfx target). */
const char *org_build_info(void);

#ifdef __cplusplus
}
#endif
#endif /* ORGANOID_H */
