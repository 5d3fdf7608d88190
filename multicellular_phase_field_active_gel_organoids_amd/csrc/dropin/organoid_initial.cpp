/*
 * organoid_initial.cpp -- the reference's initial-condition functions
 * (functions/initialisation_functions.h:1-12) for callers linked against the
 * drop-in library: the analytic initialisers used by the IC program
 * (initialisation_code/prepare_initial_files.cpp:56-66) and the readers of
 * the .bin state files used by the driver (run_with_active_gels.cpp:63-66).
 * Host code, like the reference's; nothing here touches the device.
 *
 * Expression types follow the reference: org_real_t operands stay
 * org_real_t (float or double sqrt and exp overloads), a double literal
 * promotes the expression to double, results are stored as org_real_t
 * (float, or double in the FP64 form).
 *
 * .bin layout.  The reference writes and reads each element with
 * sizeof(float*) = 8 bytes at the element's address
 * (run_with_active_gels.cpp:197-201, initialisation_functions.cpp:16-18), so
 * element k sits at byte 8k.  For float, bytes 8k+4..8k+7 repeat element k+1
 * of the writer's slab and the readers below take the first 4 bytes of each
 * 8-byte record; unlike the reference they never store past the element
 * being read (the reference's last read writes 4 bytes beyond the cell's
 * plane).  For double (the FP64 form) the records are the elements.  A
 * missing or short file aborts with a message (the reference leaves the
 * array as it was).
 */
#include "organoid_dropin.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

namespace {

[[noreturn]] void die_file(const char *name, const char *what) {
    fprintf(stderr, "organoid drop-in: %s: %s\n", name, what);
    std::exit(1);
}

/* count elements from an 8-byte-record .bin file into dst */
void read_bin(const char *name, org_real_t *dst, size_t count) {
    FILE *f = fopen(name, "rb");
    if (!f) die_file(name, "cannot open");
    constexpr size_t per = 8 / sizeof(org_real_t);     /* values per record: 2 (float) or 1 (double) */
    std::vector<org_real_t> rec(per * count);
    const size_t got = fread(rec.data(), sizeof(org_real_t), per * count, f);
    fclose(f);
    if (got != per * count) die_file(name, "shorter than 8 bytes per element");
    for (size_t k = 0; k < count; k++) dst[k] = rec[per * k];
}

org_real_t cyst_radius() { return (org_real_t)(N_ini * 0.9 * R / dx / M_PI); }   /* initialisation_functions.cpp:76-77 */

}  // namespace

/* initialisation_functions.cpp:10-19 */
void ini_phi_files(org_real_t *phi_) {
    char name[64];
    const size_t pl = (size_t)N_s * N_s;
    for (int n = 0; n < N_ini; n++) {
        snprintf(name, sizeof(name), "ini_cell_%d.bin", n + 1);
        read_bin(name, phi_ + n * pl, pl);
    }
}

/* initialisation_functions.cpp:43-52 */
void ini_rho_files(org_real_t *rho_) {
    char name[64];
    const size_t pl = (size_t)N_s * N_s;
    for (int n = 0; n < N_ini; n++) {
        snprintf(name, sizeof(name), "ini_rho_%d.bin", n + 1);
        read_bin(name, rho_ + n * pl, pl);
    }
}

/* initialisation_functions.cpp:163-170 */
void ini_lumen_files(org_real_t *lum_) { read_bin("ini_lumen.bin", lum_, (size_t)N_s * N_s); }

/* initialisation_functions.cpp:174-181 */
void ini_ECM_files(org_real_t *ECM_) { read_bin("ini_ECM.bin", ECM_, (size_t)N_s * N_s); }

/* initialisation_functions.cpp:36-40 */
void ini_phin(org_real_t *phin_, int n) {
    const size_t pl = (size_t)N_s * N_s;
    for (size_t k = 0; k < pl; k++) phin_[n * pl + k] = 0.0;
}

/* initialisation_functions.cpp:75-80 */
void ini_rhon(org_real_t *rhon_, int n) { ini_phin(rhon_, n); }

/* initialisation_functions.cpp:64-73: rho = rho0 exp(-1/(4h(1-h)+tol) + 1/(1+tol)),
 * h = phi^2 (3 - 2 phi); zero below rho0 exp(-1/(2 tol) + 1/(1+tol)). */
void ini_rho(org_real_t *rho_, org_real_t *phi_, org_real_t rho0_, org_real_t tol_, int n) {
    const size_t pl = (size_t)N_s * N_s;
    const double floor_v = rho0_ * std::exp(-1.0 / (2.0 * tol_) + 1.0 / (1.0 + tol_));
    for (size_t k = 0; k < pl; k++) {
        const double p = phi_[n * pl + k];
        const double h = p * p * (3.0 - 2.0 * p);
        org_real_t v = (org_real_t)(rho0_ * std::exp(-1.0 / (4.0 * h * (1.0 - h) + tol_) + 1.0 / (1.0 + tol_)));
        if ((double)v < floor_v) v = 0.0;
        rho_[n * pl + k] = v;
    }
}

/* initialisation_functions.cpp:82-114 */
void ini_ECM_lumen(org_real_t *lum_, org_real_t *ECM_, org_real_t *phi_) {
    const org_real_t rc = cyst_radius();
    const size_t pl = (size_t)N_s * N_s;
    for (int i = 0; i < N_s; i++)
        for (int j = 0; j < N_s; j++) {
            const size_t k = (size_t)i * N_s + j;
            org_real_t all = 0.0;
            for (int n = 0; n < N_ini; n++) all += phi_[n * pl + k];
            if (1.0 - all > 0.0) {
                const int di = i - N_s / 2, dj = j - N_s / 2;
                const org_real_t d = (org_real_t)std::sqrt((double)(di * di + dj * dj));
                if (d < rc) lum_[k] = (org_real_t)(1.0 - all);
                else ECM_[k] = (org_real_t)(1.0 - all);
            } else {
                lum_[k] = 0.0;
                ECM_[k] = 0.0;
            }
        }
}

/* initialisation_functions.cpp:117-159 */
void ini_phi(org_real_t *phi_) {
    const org_real_t rc = cyst_radius();
    const size_t pl = (size_t)N_s * N_s;
    const int planes = N_ini > 1 ? N_ini : 1;
    for (int n = 0; n < planes; n++) {
        org_real_t xc, yc;
        if (N_ini > 1) {
            xc = (org_real_t)(N_s / 2 + rc * std::cos(n * 2.0 * M_PI / N_ini));
            yc = (org_real_t)(N_s / 2 + rc * std::sin(n * 2.0 * M_PI / N_ini));
        } else {
            xc = (org_real_t)(N_s / 2);
            yc = (org_real_t)(N_s / 2);
        }
        for (int i = 0; i < N_s; i++)
            for (int j = 0; j < N_s; j++) {
                const org_real_t di = (org_real_t)i - xc, dj = (org_real_t)j - yc;
                const org_real_t r = std::sqrt(di * di + dj * dj);
                const org_real_t e = std::exp((dx * r - R0) / wi);
                phi_[n * pl + (size_t)i * N_s + j] = (org_real_t)(1.0 / (1.0 + e));
            }
    }
}
