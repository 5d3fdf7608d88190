/* Test program: runs timesteps through the drop-in library in the order of
 * the reference driver's loop (main_code/run_with_active_gels.cpp:92-184),
 * reading the initial fields from raw files and writing the final ones.
 * Usage: driver_two_cells <dir> <steps>;  reads <dir>/{phi,rho,lum,ecm}.f32,
 * writes <dir>/out_{phi,rho,lum,ecm}.f32 and <dir>/out_s.f32.  Built with
 * ORGANOID_DROPIN_F64 (driver_two_cells_f64, the FP64 form) the files are
 * .f64. */
#include <cstdio>
#include <cstdlib>
#include <string>

#include "organoid_dropin.h"

static const char *EXT = sizeof(org_real_t) == 8 ? ".f64" : ".f32";

static void read_raw(const std::string &path, org_real_t *dst, size_t n) {
    FILE *f = fopen(path.c_str(), "rb");
    if (!f || fread(dst, sizeof(org_real_t), n, f) != n) { fprintf(stderr, "cannot read %s\n", path.c_str()); exit(2); }
    fclose(f);
}
static void write_raw(const std::string &path, const org_real_t *src, size_t n) {
    FILE *f = fopen(path.c_str(), "wb");
    if (!f || fwrite(src, sizeof(org_real_t), n, f) != n) { fprintf(stderr, "cannot write %s\n", path.c_str()); exit(2); }
    fclose(f);
}

int main(int argc, char **argv) {
    if (argc < 3) { fprintf(stderr, "usage: %s dir steps\n", argv[0]); return 2; }
    std::string dir = argv[1];
    int steps = atoi(argv[2]);
    size_t P = (size_t)Ntot * N_s * N_s, G = (size_t)N_s * N_s;
    malloc_all();
    ini_fft_all();
    org_real_t *s = (org_real_t *)malloc(Ntot * sizeof(org_real_t)), *f0 = (org_real_t *)malloc(Ntot * sizeof(org_real_t));
    read_raw(dir + "/phi" + EXT, phi, P);
    read_raw(dir + "/rho" + EXT, rho, P);
    read_raw(dir + "/lum" + EXT, lum, G);
    read_raw(dir + "/ecm" + EXT, ECM, G);
    for (size_t q = 0; q < P; q++) { phin[q] = 0.0; rhon[q] = 0.0; hfunc_cell[q] = 0.0; }
    cufftPlan2d(&planf, 2 * zone_size - 2, 2 * zone_size - 2, CUFFT_C2C);
    cufftPlan2d(&planr, 2 * zone_size - 2, 2 * zone_size - 2, CUFFT_C2C);
    copy_tables_ECM_or_lumen(ECMn, ECM);
    copy_tables_ECM_or_lumen(lumn, lum);
    for (int q = 0; q < Ntot; q++) { copy_cell_tables_1_by_1(q, phin, phi); copy_cell_tables_1_by_1(q, rhon, rho); }
    for (it = 0; it < steps; it++) {
        if (it % nsaving == 0) saving_out_functions(it, phi, rho, ECM, lum, phiout, rhoout, lumout, ECMout);
        find_centres_cells(phi, xcenters, ycenters);
        create_turnover_rates_tensions(degrates, matgenrates, a_tan_1, a_ort_1, viscosity, b, phi, lum, ECM,
                                       xcenters, ycenters);
        compute_hfunc_3d_tables_kernels_loop_n(hfunc_cell, phi, xcenters, ycenters);
        compute_all_hfunc_2d_tables_kernels(hfunc_lum, lum, hfunc_ECM, ECM);
        compute_normal_grad_phi(norm_grad_phi, phi);
        boundary_conditions_cells(norm_grad_phi);
        compute_prod_norm_delta(prod_norm_delta, norm_grad_phi, phi, xcenters, ycenters);
        compute_psi_kernels(psi, hfunc_cell);
        sum_s_kernels(hfunc_cell, s, f0);
        update_all_cortices(phi, rho, rhon, xcenters, ycenters, degrates, matgenrates, a_tan_1, a_ort_1, viscosity,
                            b, norm_grad_phi, prod_norm_delta, planf, planr, temp_fft, temp_fft_real, fft_rho,
                            fft_pi_xx, fft_pi_xy, fft_pi_yy, fft_vx, fft_vy, fft_react, fft_rho_vx, fft_rho_vy);
        update_lumen_and_ECM(lum, lumn, ECM, ECMn, hfunc_lum, hfunc_ECM, psi);
        boundary_conditions_lumen_and_ECM(lumn, ECMn);
        reshaping_lumen_or_ECM(it, lum, lumn, 0.0);
        reshaping_lumen_or_ECM(it, ECM, ECMn, 1.0);
        compute_normal_to_phi(phi, norm_grad_phi, J_x, J_y);
        boundary_conditions_cells(J_x);
        boundary_conditions_cells(J_y);
        update_all_cells(it, phi, phin, norm_grad_phi, rho, hfunc_cell, f0, psi, xcenters, ycenters, hfunc_lum,
                         hfunc_ECM);
        boundary_conditions_cells(phin);
        reshaping_cells(it, phi, phin, norm_grad_phi, J_x, J_y, xcenters, ycenters);
        for (int q = 0; q < Ntot; q++) copy_cell_tables_1_by_1(q, rho, rhon);
        copy_tables_ECM_or_lumen(ECM, ECMn);
        copy_tables_ECM_or_lumen(lum, lumn);
    }
    /* final state: the snapshot path reads it back into slab 0 */
    saving_out_functions(0, phi, rho, ECM, lum, phiout, rhoout, lumout, ECMout);
    write_raw(dir + "/out_phi" + EXT, phiout, P);
    write_raw(dir + "/out_rho" + EXT, rhoout, P);
    write_raw(dir + "/out_lum" + EXT, lumout, G);
    write_raw(dir + "/out_ecm" + EXT, ECMout, G);
    write_raw(dir + "/out_s" + EXT, s, Ntot);
    printf("driver_two_cells: %d steps done\n", steps);
    return 0;
}
