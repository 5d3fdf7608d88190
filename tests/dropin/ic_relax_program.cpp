/*
 * ic_relax_program.cpp -- a program in the call order of the reference's
 * initial-condition program (initialisation_code/prepare_initial_files.cpp:
 * 56-110), linked against the drop-in library.  Usage: ic_relax_program DIR.
 *
 * It builds the analytic initial state (ini_phi, ini_ECM_lumen, ini_rho;
 * hfunc_cell zeroed as the driver does), dumps it to DIR/ic_*.f32, runs the
 * program's device part (centres, phin <- phi, reshaping_cells,
 * update_rho_to_shape, update_lumen_and_ECM, both reshaping loops) and dumps
 * the frame the program saves (slab 1) to DIR/out_*.f32.
 */
#include <unistd.h>

#include <cstdio>
#include <cstdlib>

#include "organoid_dropin.h"

static void dump(const char *name, const float *a, size_t n) {
    FILE *f = fopen(name, "wb");
    if (!f || fwrite(a, sizeof(float), n, f) != n) { fprintf(stderr, "cannot write %s\n", name); std::exit(1); }
    fclose(f);
}

int main(int argc, char **argv) {
    if (argc < 2 || chdir(argv[1]) != 0) { fprintf(stderr, "usage: ic_relax_program DIR\n"); return 2; }
    malloc_all();
    const size_t pl = (size_t)N_s * N_s, cells = (size_t)Ntot * pl;
    for (size_t k = 0; k < pl; k++) { lum[k] = 0.0f; ECM[k] = 0.0f; }
    ini_phi(phi);
    ini_ECM_lumen(lum, ECM, phi);
    for (int n = 0; n < N_ini; n++) {
        ini_rho(rho, phi, rho0, tol, n);
        ini_phin(phin, n);
        ini_phin(hfunc_cell, n);
    }
    dump("ic_phi.f32", phi, cells);
    dump("ic_rho.f32", rho, cells);
    dump("ic_lum.f32", lum, pl);
    dump("ic_ecm.f32", ECM, pl);
    if (check_nan(phi) != 0) { fprintf(stderr, "NaN in the initial phi\n"); return 3; }

    saving_out_functions(0, phi, rho, ECM, lum, phiout, rhoout, lumout, ECMout);
    find_centres_cells(phi, xcenters, ycenters);
    for (int n = 0; n < Ntot; n++) copy_cell_tables_1_by_1(n, phin, phi);
    compute_normal_grad_phi(norm_grad_phi, phi);
    boundary_conditions_cells(norm_grad_phi);
    compute_normal_to_phi(phi, norm_grad_phi, J_x, J_y);
    boundary_conditions_cells(J_x);
    boundary_conditions_cells(J_y);
    reshaping_cells(0, phi, phin, norm_grad_phi, J_x, J_y, xcenters, ycenters);
    update_rho_to_shape(rho, phi, rho0, tol);
    update_lumen_and_ECM(lum, lumn, ECM, ECMn, hfunc_lum, hfunc_ECM, psi);
    boundary_conditions_lumen_and_ECM(lumn, ECMn);
    reshaping_lumen_or_ECM(0, lum, lumn, 0.0);
    reshaping_lumen_or_ECM(0, ECM, ECMn, 1.0);
    saving_out_functions(nsaving, phi, rho, ECM, lum, phiout, rhoout, lumout, ECMout);

    dump("out_phi.f32", phiout + cells, cells);
    dump("out_rho.f32", rhoout + cells, cells);
    dump("out_lum.f32", lumout + pl, pl);
    dump("out_ecm.f32", ECMout + pl, pl);
    return 0;
}
