/* Constants of the initial-condition test program (ic_program.cpp) in the
 * reference's global form (the role of a rendered constants_template.cpp):
 * 4 cells of radius R0 = 6 on a 96^2 grid, ring radius from R = 10,
 * interface width wi = 1, two saved frames (Nsteps = 8, nsaving = 7).
 * Only the initialisation functions read them; the rest keep the
 * two-cell values so that every global the library references exists. */
#include "organoid_dropin.h"

int threshold_reshaping_steps = 100000;
int riprint = 100000;
int rdev_threshold = 50000;
int rdev_ini = 100000;
int Nsteps = 8;
int ncontinue = 0;
int ndisplay = 5;
int nsaving = 7;
int Ntotsave = 2;
float dx = 1;
float dt = 0.001;
const int N_s = 96;
const int L = 120;
float R0 = 6;
float R = 10;
float s0 = 2827.43339;
float Rint = 1.0;
float Rout = 1.0;
float sigma = 0.2;
int N_ini = 4;
const int Ntot = 4;
int N_current = 4;
int zone_size = 65;
float coupling_PF_AG = 1;
int random_ontime = 1000000;
float seuil_gel = 0.1;
int gen_ontime = 0;
int active_gel_ontime = 0;
int contract_ontime = 0;
float a_ort_lateral = 100;
float a_ort_apical = 100;
float a_ort_basal = 100;
float a_tan_lateral = 100;
float a_tan_apical = 100;
float a_tan_basal = 100;
float b_ap = 1;
float b_bas = 1;
float b_lat = 1;
float matgenrate_apical = 1;
float degrate_apical = 100;
float matgenrate_lateral = 1;
float degrate_lateral = 100;
float matgenrate_basal = 1;
float degrate_basal = 100;
float friction = 1;
float viscosity_ap = 1;
float viscosity_bas = 1;
float viscosity_lat = 1;
float D0 = 0.0;
float tau = 1;
float alpha = 1;
float beta = 1;
float eta = 0.03;
float gamma_cell = 0.045;
float D_lum = 0.0;
float tau_lum = 1;
float beta_lum_cells = 1;
float eta_lum_cells = 0;
float gamma_lum_cells = 0;
float xi_lum = 0.15;
float alpha_ECM = 0.002;
float D_ECM = 0.0;
float tau_ECM = 1;
float beta_ECM_cells = 1;
float eta_ECM_cells = 0.001;
float gamma_ECM_cells = 0.0011;
float xi_ECM = 0;
float beta_lum_ECM = 1;
float wi = 1.0;
float rho0 = 0.001;
float tol = 0.01;
int ri;
int rimax = 1000000;
float rdev;
float rtol = 0.001;
float rdt = 0.0245;
float ans = 0.0;
float sum_s = 0.0;
int it, k, l, ir, jr, c;
int nelse;
int ip, im, jp, jm;
