/*
 * organoid_generic.h -- internal interface between the C-ABI host code
 * (organoid.hip) and the per-function full-plane path (organoid_generic.hip).
 *
 * The generic path serves the contexts the LDS kernels do not cover: FP64
 * state (org_create_f64) and FFT boxes other than 128 (zone_size != 65).
 * Every method is one entry point of include/organoid.h and returns ORG_OK
 * or an ORG_E* code; HIP failures print a message and return ORG_EHIP.
 */
#ifndef ORGANOID_GENERIC_H
#define ORGANOID_GENERIC_H

#include <hip/hip_runtime.h>

#include "organoid.h"

namespace gen {

struct Generic {
    virtual ~Generic() = default;
    virtual int real_bytes() const = 0;
    /* host buffers are float (f64 = 0) or double (f64 = 1) */
    virtual int set_field(int field, const void *host, int f64) = 0;
    virtual int get_field(int field, void *host, int f64) = 0;
    virtual int get_window(int which, void *host, int f64) = 0;
    virtual int get_centres(int32_t *xc, int32_t *yc) = 0;
    virtual int get_areas(double *s, double *f0) = 0;
    virtual int get_iterations(int32_t *cells, int32_t *lum, int32_t *ecm) = 0;
    virtual int prepare() = 0;
    virtual int find_centres() = 0;
    virtual int hfunc_cells() = 0;
    virtual int hfunc_lum_ecm() = 0;
    virtual int psi() = 0;
    virtual int sum_s() = 0;
    virtual int cortices() = 0;
    virtual int lumecm_update() = 0;
    /* *relax_failed is set to 1 when the loop passed rimax */
    virtual int reshape_lumecm(int which_ecm, int *relax_failed) = 0;
    /* the lumen's then the ECM's loop (neither reads the other's field),
     * their host round trips shared; relax_failed[2] */
    virtual int reshape_lumecm_pair(int *relax_failed) = 0;
    virtual int update_cells() = 0;
    virtual int reshape_cells(int from_phi, int *relax_failed) = 0;
    virtual int commit_rho() = 0;
    virtual int commit_lumecm() = 0;
    virtual int rho_to_shape(double rho0, double tol) = 0;
    /* psi, h(lum), h(ECM) set to zero (the initial-condition program never computes them) */
    virtual int zero_gel_inputs() = 0;
    /* the 2D transforms' schedule: stages fused per LDS round trip (1..3) and lines per workgroup (1, 2, 4) */
    virtual int get_fft_schedule(int32_t *fuse, int32_t *lines_per_wg) = 0;
    virtual int set_fft_schedule(int fuse, int lines_per_wg) = 0;
    /* Inside one whole step (from its first call to the end of the cell
     * reshaping) phi does not change, so |grad phi| is formed once there
     * instead of by each of cortices, update_cells and reshape_cells: on
     * from the step's start, off before its copy-back.  The per-function
     * entry points leave it off (each call forms its own, as the reference). */
    virtual void step_scope(bool on) = 0;
    /* Around the steps of one org_step call: from one step's centroid to the
     * next, phi changes only where the cell reshaping writes it (each cell's
     * window), so the centroid's integer sums are carried over (the windows'
     * pixels taken out before that write and put back after it) instead of
     * summed again over the whole planes.  Off, every find_centres sums the
     * whole planes. */
    virtual void step_chain(bool on) = 0;
};

/* A context on `stream` (borrowed: the caller owns and destroys it). */
int create_generic(const org_params *p, const double *s0_cell, int f64, hipStream_t stream, Generic **out);

/* One unnormalised 2D DFT of an M x M plane (interleaved re, im doubles on
 * the host) through the generic path's transform, stored as float (f64 = 0)
 * or double between calls, as its contexts do, with the given schedule. */
int debug_fft2d(int M, double *data, int sign, int f64, int fuse, int lines_per_wg);

}  // namespace gen

#endif
