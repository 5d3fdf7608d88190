/*
 * ic_program.cpp -- host-only test program for the drop-in's initialisation
 * functions and the reference's state-file formats.  Usage: ic_program DIR.
 *
 * In DIR it
 *  1. builds the analytic initial state the way the reference's IC program
 *     does (prepare_initial_files.cpp:56-66: ini_phi, ini_ECM_lumen, ini_rho)
 *     and dumps it as raw float32 (ic_phi.f32, ic_rho.f32, ic_lum.f32,
 *     ic_ecm.f32);
 *  2. writes a two-frame output slab (frame 0 = that state, frame 1 = half
 *     of it) with the driver's post-loop writers
 *     (run_with_active_gels.cpp:191-269): 8 bytes from each element's
 *     address into the .bin files, "%.15f" lines with a blank line after
 *     each row into data_int/.  The slab has one extra zero float at its end,
 *     so the last record's second half is defined;
 *  3. reads the .bin files back with ini_phi_files, ini_rho_files,
 *     ini_lumen_files and ini_ECM_files (what a restarted run does) and dumps
 *     the result (rb_phi.f32, rb_rho.f32, rb_lum.f32, rb_ecm.f32).
 * Nothing here touches the GPU.
 */
#include <sys/stat.h>
#include <unistd.h>

#include <cstdio>
#include <cstdlib>
#include <fstream>

#include "organoid_dropin.h"

static void dump(const char *name, const float *a, size_t n) {
    FILE *f = fopen(name, "wb");
    if (!f || fwrite(a, sizeof(float), n, f) != n) { fprintf(stderr, "cannot write %s\n", name); std::exit(1); }
    fclose(f);
}

static void write_bin(const char *name, const float *slab_at, size_t n) {
    std::ofstream out(name, std::ios::binary);
    for (size_t k = 0; k < n; k++) out.write(reinterpret_cast<const char *>(&slab_at[k]), sizeof(&slab_at[k]));
}

static void write_dat(const char *name, const float *plane) {
    FILE *f = fopen(name, "w");
    if (!f) { fprintf(stderr, "cannot write %s\n", name); std::exit(1); }
    for (int i = 0; i < N_s; i++) {
        for (int j = 0; j < N_s; j++) fprintf(f, "%.15f\n", plane[i * N_s + j]);
        fprintf(f, "\n");
    }
    fclose(f);
}

int main(int argc, char **argv) {
    if (argc < 2 || chdir(argv[1]) != 0) { fprintf(stderr, "usage: ic_program DIR\n"); return 2; }
    mkdir("data_int", 0755);
    const size_t pl = (size_t)N_s * N_s, cells = (size_t)Ntot * pl;
    float *ph = (float *)calloc(cells, sizeof(float)), *rh = (float *)calloc(cells, sizeof(float));
    float *lu = (float *)calloc(pl, sizeof(float)), *ec = (float *)calloc(pl, sizeof(float));
    ini_phi(ph);
    ini_ECM_lumen(lu, ec, ph);
    for (int n = 0; n < N_ini; n++) ini_rho(rh, ph, rho0, tol, n);
    dump("ic_phi.f32", ph, cells);
    dump("ic_rho.f32", rh, cells);
    dump("ic_lum.f32", lu, pl);
    dump("ic_ecm.f32", ec, pl);

    const int frames = 2;
    float *phiout_ = (float *)calloc(frames * cells + 1, sizeof(float));
    float *rhoout_ = (float *)calloc(frames * cells + 1, sizeof(float));
    float *lumout_ = (float *)calloc(frames * pl + 1, sizeof(float));
    float *ecmout_ = (float *)calloc(frames * pl + 1, sizeof(float));
    for (size_t k = 0; k < cells; k++) {
        phiout_[k] = ph[k]; phiout_[cells + k] = ph[k] * 0.5f;
        rhoout_[k] = rh[k]; rhoout_[cells + k] = rh[k] * 0.5f;
    }
    for (size_t k = 0; k < pl; k++) {
        lumout_[k] = lu[k]; lumout_[pl + k] = lu[k] * 0.5f;
        ecmout_[k] = ec[k]; ecmout_[pl + k] = ec[k] * 0.5f;
    }
    char name[96];
    for (int it = 0; it < frames; it++) {
        for (int n = 0; n < N_current; n++) {
            const size_t at = it * cells + n * pl;
            snprintf(name, sizeof(name), "ini_cell_%d.bin", n + 1);
            write_bin(name, phiout_ + at, pl);
            snprintf(name, sizeof(name), "ini_rho_%d.bin", n + 1);
            write_bin(name, rhoout_ + at, pl);
            snprintf(name, sizeof(name), "data_int/phi_cell_%d_step_%d.dat", n + 1, it * nsaving);
            write_dat(name, phiout_ + at);
            snprintf(name, sizeof(name), "data_int/rho_cell_%d_step_%d.dat", n + 1, it * nsaving);
            write_dat(name, rhoout_ + at);
        }
        write_bin("ini_lumen.bin", lumout_ + it * pl, pl);
        write_bin("ini_ECM.bin", ecmout_ + it * pl, pl);
        snprintf(name, sizeof(name), "data_int/lumen_step_%d.dat", it * nsaving);
        write_dat(name, lumout_ + it * pl);
        snprintf(name, sizeof(name), "data_int/ECM_step_%d.dat", it * nsaving);
        write_dat(name, ecmout_ + it * pl);
    }

    float *ph2 = (float *)calloc(cells, sizeof(float)), *rh2 = (float *)calloc(cells, sizeof(float));
    float *lu2 = (float *)calloc(pl, sizeof(float)), *ec2 = (float *)calloc(pl, sizeof(float));
    ini_phi_files(ph2);
    ini_rho_files(rh2);
    ini_lumen_files(lu2);
    ini_ECM_files(ec2);
    dump("rb_phi.f32", ph2, cells);
    dump("rb_rho.f32", rh2, cells);
    dump("rb_lum.f32", lu2, pl);
    dump("rb_ecm.f32", ec2, pl);
    return 0;
}
