/*
 * organoid.hip -- gfx950 kernels and C-ABI of the organoid timestep library.
 *
 * One library-owned device state per organoid (org_ctx).  Full N x N planes
 * are kept for the fields that persist between steps or that the caller can
 * read back (phi, rho, hfunc per cell; lum, ECM, psi, h(lum), h(ECM) shared);
 * everything the reference materialises as a full plane only to consume it
 * inside a cell's window (|grad phi|, prod_norm_delta, the six turnover maps,
 * pi_xx/xy/yy, J_x/J_y, temp_phase, the eleven FFT buffers) is recomputed
 * inside the kernel that consumes it, on the window only.
 *
 * Why the window restriction is exact: phi and hfunc of cell n change only
 * inside cell n's window (cell_phase_functions.cpp:160-163, 484-487;
 * useful_functions.cpp:242-245); rho changes only inside the FFT box
 * (rho_functions.cpp:253-265) and on the ghost ring (boundary_conditions_cells,
 * useful_functions.cpp:391-420).  Each cell keeps a "support box": the
 * bounding rectangle of its initial {phi > 0.5} U {hfunc != 0} grown by every
 * window it has used.  The centroid scan (useful_functions.cpp:193-227) and
 * the area sum (useful_functions.cpp:168-190) over the support box equal the
 * full-plane ones; psi = sum_n hfunc_n (useful_functions.cpp:305-322) changes
 * only where some hfunc_n changed, i.e. inside this step's windows, and is
 * recomputed there from every cell whose support box covers the pixel, in
 * cell order, so the FP32 sum has the same operands in the same order.
 *
 * Kernels per step (all cells per launch):
 *   k_centroid      integer centroid + window check + support box update
 *   k_prep          hfunc on the window, FP64 area sum over the support box
 *   k_hlum          h(lum), h(ECM) full plane + FP64 sum of h(ECM)
 *   k_psi           psi on tiles covering the union of this step's windows
 *   k_lumecm_update lumen / ECM explicit step + Dirichlet edges
 *   k_lum_normals   |grad| of the old field, normals of the new one, periodic ring
 *   k_lum_iter      one reshaping sweep of lumen and ECM, FP64 deviation,
 *                   device-side stop test (last-block reduction)
 *   k_spectral      per cell: active stress, 2 forward FFTs, the per-mode
 *                   friction-viscosity solve, Hermitian split, inverse FFT of
 *                   the velocity, rho*v, forward FFT, spectral divergence,
 *                   inverse FFT -> new rho on the FFT box  (4 2D FFTs of
 *                   128 x 128 in LDS instead of the reference's 10)
 *   k_cells         per cell: phase-field update, normals, reshaping loop in
 *                   LDS until the deviation test passes, phi <- second-to-last
 *   k_commit_cells  rho box write-back + ghost ring;  k_commit_lumecm.
 */
#include <hip/hip_runtime.h>

#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "organoid.h"

#define ORG_MB 128                 /* FFT box side supported by k_spectral */
#define ORG_LS (ORG_MB + 1)        /* LDS row stride (float2) */
#define ORG_ZS_MAX 65              /* zone_size supported by k_cells */
#define ORG_WH_MAX (2 * ORG_ZS_MAX + 2)
#define ORG_W (2 * ORG_ZS_MAX)     /* window side */
#define ORG_WH (ORG_W + 2)         /* window plus one-pixel ring */

__constant__ float kCos128[128] = {
    1.000000000e+00f, 9.987954562e-01f, 9.951847267e-01f, 9.891765100e-01f,
    9.807852804e-01f, 9.700312532e-01f, 9.569403357e-01f, 9.415440652e-01f,
    9.238795325e-01f, 9.039892931e-01f, 8.819212643e-01f, 8.577286100e-01f,
    8.314696123e-01f, 8.032075315e-01f, 7.730104534e-01f, 7.409511254e-01f,
    7.071067812e-01f, 6.715589548e-01f, 6.343932842e-01f, 5.956993045e-01f,
    5.555702330e-01f, 5.141027442e-01f, 4.713967368e-01f, 4.275550934e-01f,
    3.826834324e-01f, 3.368898534e-01f, 2.902846773e-01f, 2.429801799e-01f,
    1.950903220e-01f, 1.467304745e-01f, 9.801714033e-02f, 4.906767433e-02f,
    6.123233996e-17f, -4.906767433e-02f, -9.801714033e-02f, -1.467304745e-01f,
    -1.950903220e-01f, -2.429801799e-01f, -2.902846773e-01f, -3.368898534e-01f,
    -3.826834324e-01f, -4.275550934e-01f, -4.713967368e-01f, -5.141027442e-01f,
    -5.555702330e-01f, -5.956993045e-01f, -6.343932842e-01f, -6.715589548e-01f,
    -7.071067812e-01f, -7.409511254e-01f, -7.730104534e-01f, -8.032075315e-01f,
    -8.314696123e-01f, -8.577286100e-01f, -8.819212643e-01f, -9.039892931e-01f,
    -9.238795325e-01f, -9.415440652e-01f, -9.569403357e-01f, -9.700312532e-01f,
    -9.807852804e-01f, -9.891765100e-01f, -9.951847267e-01f, -9.987954562e-01f,
    -1.000000000e+00f, -9.987954562e-01f, -9.951847267e-01f, -9.891765100e-01f,
    -9.807852804e-01f, -9.700312532e-01f, -9.569403357e-01f, -9.415440652e-01f,
    -9.238795325e-01f, -9.039892931e-01f, -8.819212643e-01f, -8.577286100e-01f,
    -8.314696123e-01f, -8.032075315e-01f, -7.730104534e-01f, -7.409511254e-01f,
    -7.071067812e-01f, -6.715589548e-01f, -6.343932842e-01f, -5.956993045e-01f,
    -5.555702330e-01f, -5.141027442e-01f, -4.713967368e-01f, -4.275550934e-01f,
    -3.826834324e-01f, -3.368898534e-01f, -2.902846773e-01f, -2.429801799e-01f,
    -1.950903220e-01f, -1.467304745e-01f, -9.801714033e-02f, -4.906767433e-02f,
    -1.836970199e-16f, 4.906767433e-02f, 9.801714033e-02f, 1.467304745e-01f,
    1.950903220e-01f, 2.429801799e-01f, 2.902846773e-01f, 3.368898534e-01f,
    3.826834324e-01f, 4.275550934e-01f, 4.713967368e-01f, 5.141027442e-01f,
    5.555702330e-01f, 5.956993045e-01f, 6.343932842e-01f, 6.715589548e-01f,
    7.071067812e-01f, 7.409511254e-01f, 7.730104534e-01f, 8.032075315e-01f,
    8.314696123e-01f, 8.577286100e-01f, 8.819212643e-01f, 9.039892931e-01f,
    9.238795325e-01f, 9.415440652e-01f, 9.569403357e-01f, 9.700312532e-01f,
    9.807852804e-01f, 9.891765100e-01f, 9.951847267e-01f, 9.987954562e-01f,
};
__constant__ float kSin128[128] = {
    0.000000000e+00f, 4.906767433e-02f, 9.801714033e-02f, 1.467304745e-01f,
    1.950903220e-01f, 2.429801799e-01f, 2.902846773e-01f, 3.368898534e-01f,
    3.826834324e-01f, 4.275550934e-01f, 4.713967368e-01f, 5.141027442e-01f,
    5.555702330e-01f, 5.956993045e-01f, 6.343932842e-01f, 6.715589548e-01f,
    7.071067812e-01f, 7.409511254e-01f, 7.730104534e-01f, 8.032075315e-01f,
    8.314696123e-01f, 8.577286100e-01f, 8.819212643e-01f, 9.039892931e-01f,
    9.238795325e-01f, 9.415440652e-01f, 9.569403357e-01f, 9.700312532e-01f,
    9.807852804e-01f, 9.891765100e-01f, 9.951847267e-01f, 9.987954562e-01f,
    1.000000000e+00f, 9.987954562e-01f, 9.951847267e-01f, 9.891765100e-01f,
    9.807852804e-01f, 9.700312532e-01f, 9.569403357e-01f, 9.415440652e-01f,
    9.238795325e-01f, 9.039892931e-01f, 8.819212643e-01f, 8.577286100e-01f,
    8.314696123e-01f, 8.032075315e-01f, 7.730104534e-01f, 7.409511254e-01f,
    7.071067812e-01f, 6.715589548e-01f, 6.343932842e-01f, 5.956993045e-01f,
    5.555702330e-01f, 5.141027442e-01f, 4.713967368e-01f, 4.275550934e-01f,
    3.826834324e-01f, 3.368898534e-01f, 2.902846773e-01f, 2.429801799e-01f,
    1.950903220e-01f, 1.467304745e-01f, 9.801714033e-02f, 4.906767433e-02f,
    1.224646799e-16f, -4.906767433e-02f, -9.801714033e-02f, -1.467304745e-01f,
    -1.950903220e-01f, -2.429801799e-01f, -2.902846773e-01f, -3.368898534e-01f,
    -3.826834324e-01f, -4.275550934e-01f, -4.713967368e-01f, -5.141027442e-01f,
    -5.555702330e-01f, -5.956993045e-01f, -6.343932842e-01f, -6.715589548e-01f,
    -7.071067812e-01f, -7.409511254e-01f, -7.730104534e-01f, -8.032075315e-01f,
    -8.314696123e-01f, -8.577286100e-01f, -8.819212643e-01f, -9.039892931e-01f,
    -9.238795325e-01f, -9.415440652e-01f, -9.569403357e-01f, -9.700312532e-01f,
    -9.807852804e-01f, -9.891765100e-01f, -9.951847267e-01f, -9.987954562e-01f,
    -1.000000000e+00f, -9.987954562e-01f, -9.951847267e-01f, -9.891765100e-01f,
    -9.807852804e-01f, -9.700312532e-01f, -9.569403357e-01f, -9.415440652e-01f,
    -9.238795325e-01f, -9.039892931e-01f, -8.819212643e-01f, -8.577286100e-01f,
    -8.314696123e-01f, -8.032075315e-01f, -7.730104534e-01f, -7.409511254e-01f,
    -7.071067812e-01f, -6.715589548e-01f, -6.343932842e-01f, -5.956993045e-01f,
    -5.555702330e-01f, -5.141027442e-01f, -4.713967368e-01f, -4.275550934e-01f,
    -3.826834324e-01f, -3.368898534e-01f, -2.902846773e-01f, -2.429801799e-01f,
    -1.950903220e-01f, -1.467304745e-01f, -9.801714033e-02f, -4.906767433e-02f,
};


/* ------------------------------------------------------------------------ */
/* device-side constants, rounded to float once as the reference's float
 * globals are (constants.h:4-124).  Label order: 0 apical, 1 basal,
 * 2 and 3 lateral (rho_functions.cpp:355-397). */
struct DevP {
    int N, nc, zs, W, M, rimax;
    float dt, dx, sigma, cpl, tau, alpha, beta, eta;
    float tau_lum, beta_lc, eta_lc, xi_lum, alpha_E, tau_E, beta_ec, eta_ec, xi_E, beta_le;
    float wi, rdt, friction, ncur;
    float visc[4], aort[4], atan_[4], bb[4], gen[4], deg[4];
    double thr_cell, rdev0_cell, thr_lum, rdev0_lum;
};

static __device__ __forceinline__ void set_err(int *err, int code) { atomicOr(err, 1 << code); }

static __device__ __forceinline__ float2 cmulf(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
static __device__ __forceinline__ float2 caddf(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
static __device__ __forceinline__ float2 csubf(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }

/* h(u) = u^2 (3 - 2u), useful_functions.cpp:247 */
static __device__ __forceinline__ float hfun(float u) { return u * u * (3.0f - 2.0f * u); }

/* ---- block reductions (blockDim multiple of 64) --------------------------- */
static __device__ __forceinline__ double wave_sum_d(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}
static __device__ __forceinline__ long long wave_sum_ll(long long v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}
/* Sum over the block, result valid in every thread.  red: >= 16 doubles. */
static __device__ double block_sum_d(double v, double *red) {
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    v = wave_sum_d(v);
    __syncthreads();
    if (lane == 0) red[w] = v;
    __syncthreads();
    double s = 0.0;
    for (int k = 0; k < nw; k++) s += red[k];
    return s;
}

/* ---- small in-register DFTs (radix-2, natural order in and out) ----------- */
/* cos / sin of 2*pi*k/16 for k = 0..7 */
#define C16_0 1.0f
#define C16_1 0.923879533f
#define C16_2 0.707106781f
#define C16_3 0.382683432f
template <int S>
static __device__ __forceinline__ float2 w16(int k) {
    /* exp(S * 2 pi i k / 16), k compile-time after unrolling */
    const float c[8] = {C16_0, C16_1, C16_2, C16_3, 0.0f, -C16_3, -C16_2, -C16_1};
    const float s[8] = {0.0f, C16_3, C16_2, C16_1, 1.0f, C16_1, C16_2, C16_3};
    return make_float2(c[k], S * s[k]);
}
template <int N>
static __device__ __forceinline__ int brev(int i) {
    int r = 0;
    for (int b = 1, m = N >> 1; b < N; b <<= 1, m >>= 1)
        if (i & b) r |= m;
    return r;
}
template <int N, int S>
static __device__ __forceinline__ void dft_reg(float2 *v) {
#pragma unroll
    for (int i = 0; i < N; i++) {
        int j = brev<N>(i);
        if (i < j) { float2 t = v[i]; v[i] = v[j]; v[j] = t; }
    }
#pragma unroll
    for (int len = 2; len <= N; len <<= 1) {
#pragma unroll
        for (int i = 0; i < N; i += len) {
#pragma unroll
            for (int k = 0; k < len / 2; k++) {
                float2 w = w16<S>(k * (16 / len));
                float2 b = cmulf(v[i + k + len / 2], w);
                v[i + k + len / 2] = csubf(v[i + k], b);
                v[i + k] = caddf(v[i + k], b);
            }
        }
    }
}

/* One batch of 128-point DFTs over the lines of a 128 x 128 complex array in
 * LDS, 8 threads per line, 1024 threads (128 = 16 x 8):
 *   X[k2 + 16 k1] = sum_j1 W8^(j1 k1) W128^(j1 k2) sum_j2 x[j1 + 8 j2] W16^(j2 k2)
 * es = element stride, ls = line stride (float2 units).  Consecutive lanes
 * take consecutive lines with the same sub-index: along columns that is
 * contiguous, along rows it strides by the odd row pitch, so neither pass
 * has LDS bank conflicts, and the twiddle index is uniform in a wave. */
template <int S>
static __device__ __forceinline__ void fft128_lines(float2 *sm, const float2 *tw, int es, int ls) {
    const int t = threadIdx.x, line = t & 127, sub = t >> 7;
    float2 *base = sm + line * ls;
    float2 v[16];
#pragma unroll
    for (int m = 0; m < 16; m++) v[m] = base[(sub + 8 * m) * es];
    dft_reg<16, S>(v);
#pragma unroll
    for (int k2 = 1; k2 < 16; k2++) {
        float2 w = tw[(sub * k2) & 127];
        w.y *= (float)S;
        v[k2] = cmulf(v[k2], w);
    }
#pragma unroll
    for (int k2 = 0; k2 < 16; k2++) base[(sub + 8 * k2) * es] = v[k2];
    __syncthreads();
    float2 u[16];
#pragma unroll
    for (int h = 0; h < 2; h++)
#pragma unroll
        for (int j1 = 0; j1 < 8; j1++) u[8 * h + j1] = base[(j1 + 8 * (2 * sub + h)) * es];
    __syncthreads();
    dft_reg<8, S>(u);          /* W8 = W16^2: dft_reg<8> uses w16(k * 2) */
    dft_reg<8, S>(u + 8);
#pragma unroll
    for (int h = 0; h < 2; h++)
#pragma unroll
        for (int k1 = 0; k1 < 8; k1++) base[((2 * sub + h) + 16 * k1) * es] = u[8 * h + k1];
    __syncthreads();
}

/* Unnormalised 2D DFT of the 128 x 128 array in LDS (row stride ORG_LS),
 * sign S as cufftExecC2C (useful_functions.cpp:650, 658). */
template <int S>
static __device__ __forceinline__ void fft2d_128(float2 *sm, const float2 *tw) {
    fft128_lines<S>(sm, tw, 1, ORG_LS);    /* rows */
    fft128_lines<S>(sm, tw, ORG_LS, 1);    /* columns */
}

/* ---- per-mode friction-viscosity solve, rho_functions.cpp:164-186 -------- */
static __device__ __forceinline__ void velocity_mode(float eta, float xi, float kx, float ky,
                                                     float2 pxx, float2 pxy, float2 pyy,
                                                     float2 &vx, float2 &vy) {
    float D = -(4.0f * eta * eta * kx * kx * ky * ky + xi * xi + 3.0f * eta * xi * (kx * kx + ky * ky)
                + 2.0f * eta * eta * (kx * kx * kx * kx + ky * ky * ky * ky));
    float A = xi + 2.0f * eta * (kx * kx + ky * ky / 2.0f);
    float ex = eta * kx * ky;
    float2 y, x;
    y.x = (ex * (kx * pxx.y + ky * pxy.y) - A * (ky * pyy.y + kx * pxy.y)) / D;
    y.y = (-ex * (kx * pxx.x + ky * pxy.x) + A * (ky * pyy.x + kx * pxy.x)) / D;
    x.x = (kx * pxx.y + ky * pxy.y - ex * y.x) / A;
    x.y = (-kx * pxx.x - ky * pxy.x - ex * y.y) / A;
    vx = x; vy = y;
}

/* Turnover label of one pixel, rho_functions.cpp:324-353 (strict '>'). */
static __device__ __forceinline__ int turnover_label(float ph, float l, float e, float p1, float p2) {
    float c1l = ph * l, c1e = ph * e;
    int index; float mx;
    if (c1l > c1e) { index = 0; mx = c1l; } else { index = 1; mx = c1e; }
    float c12 = ph * p1;
    if (c12 > mx) { index = 2; mx = c12; }
    float c13 = ph * p2;
    if (c13 > mx) { index = 3; }
    return index;
}

/* ======================================================================== */
/* k_init_box: support box of each cell from the uploaded phi / hfunc.      */
__global__ void k_init_box(DevP P, const float *__restrict__ phi, const float *__restrict__ hfunc,
                           int *__restrict__ box) {
    __shared__ int sred[4];
    const int n = blockIdx.x, N = P.N;
    const size_t NN = (size_t)N * N;
    int i0 = N, i1 = -1, j0 = N, j1 = -1;
    if (threadIdx.x == 0) { sred[0] = N; sred[1] = -1; sred[2] = N; sred[3] = -1; }
    __syncthreads();
    for (int i = threadIdx.x >> 6; i < N; i += blockDim.x >> 6) {
        const float *pr = phi + n * NN + (size_t)i * N, *hr = hfunc + n * NN + (size_t)i * N;
        for (int j = threadIdx.x & 63; j < N; j += 64)
            if (pr[j] > 0.5f || hr[j] != 0.0f) {
                i0 = min(i0, i); i1 = max(i1, i); j0 = min(j0, j); j1 = max(j1, j);
            }
    }
    atomicMin(&sred[0], i0); atomicMax(&sred[1], i1); atomicMin(&sred[2], j0); atomicMax(&sred[3], j1);
    __syncthreads();
    if (threadIdx.x == 0) {
        box[4 * n + 0] = sred[0]; box[4 * n + 1] = sred[1] + 1;   /* half-open */
        box[4 * n + 2] = sred[2]; box[4 * n + 3] = sred[3] + 1;
    }
}

/* k_centroid: find_centres_cells (useful_functions.cpp:193-227) over the
 * support box, window check, support box += window.  Block 0 also resets the
 * union-of-windows box that k_prep fills and k_psi reads. */
__global__ __launch_bounds__(1024) void k_centroid(DevP P, const float *__restrict__ phi, int *__restrict__ xc,
                                                   int *__restrict__ yc, int *__restrict__ box,
                                                   int *__restrict__ uni, int *__restrict__ err) {
    __shared__ long long red[3][16];
    const int n = blockIdx.x, N = P.N;
    const size_t NN = (size_t)N * N;
    const int i0 = box[4 * n], i1 = box[4 * n + 1], j0 = box[4 * n + 2], j1 = box[4 * n + 3];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, nw = blockDim.x >> 6;
    if (n == 0 && threadIdx.x == 0) { uni[0] = N; uni[1] = 0; uni[2] = N; uni[3] = 0; }
    long long sx = 0, sy = 0, c = 0;
    const float *ph = phi + n * NN;
    for (int i = i0 + w; i < i1; i += nw) {
        const float *row = ph + (size_t)i * N;
        for (int j = j0 + lane; j < j1; j += 64)
            if (row[j] > 0.5f) { sx += i; sy += j; c += 1; }
    }
    sx = wave_sum_ll(sx); sy = wave_sum_ll(sy); c = wave_sum_ll(c);
    if (lane == 0) { red[0][w] = sx; red[1][w] = sy; red[2][w] = c; }
    __syncthreads();
    if (threadIdx.x == 0) {
        long long X = 0, Y = 0, C = 0;
        for (int k = 0; k < nw; k++) { X += red[0][k]; Y += red[1][k]; C += red[2][k]; }
        int x = xc[n], y = yc[n];
        if (C == 0) {
            set_err(err, ORG_EEMPTY_CELL);
        } else {
            x = (int)(X / C);
            y = (int)(Y / C);
        }
        /* window [x-zs, x+zs) plus one halo pixel inside rows/cols 1..N-2 */
        int lo = P.zs + 2, hi = N - 2 - P.zs;
        if (x < lo || x > hi || y < lo || y > hi) {
            set_err(err, ORG_EWINDOW);
            x = min(max(x, lo), hi);
            y = min(max(y, lo), hi);
        }
        xc[n] = x; yc[n] = y;
        if (i1 <= i0 || j1 <= j0) {
            box[4 * n + 0] = x - P.zs; box[4 * n + 1] = x + P.zs;
            box[4 * n + 2] = y - P.zs; box[4 * n + 3] = y + P.zs;
        } else {
            box[4 * n + 0] = min(i0, x - P.zs); box[4 * n + 1] = max(i1, x + P.zs);
            box[4 * n + 2] = min(j0, y - P.zs); box[4 * n + 3] = max(j1, y + P.zs);
        }
    }
}

/* k_prep: mode&1 -> hfunc on the window (useful_functions.cpp:233-251);
 * mode&2 -> s_n = sum over the support box of hfunc*dx*dx in FP64 and
 * f0_n = alpha*(s0_n - s_n) (useful_functions.cpp:168-190).  Both modes add
 * the cell's window to the union box read by k_psi. */
__global__ __launch_bounds__(1024) void k_prep(DevP P, const float *__restrict__ phi, float *__restrict__ hfunc,
                                               const int *__restrict__ xc, const int *__restrict__ yc,
                                               const int *__restrict__ box, const double *__restrict__ s0,
                                               double *__restrict__ s, float *__restrict__ f0,
                                               int *__restrict__ uni, int mode) {
    __shared__ double red[16];
    const int n = blockIdx.x, N = P.N;
    constexpr int W = ORG_W;
    const size_t NN = (size_t)N * N;
    const int x0 = xc[n] - P.zs, y0 = yc[n] - P.zs;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, nw = blockDim.x >> 6;
    const float *ph = phi + n * NN;
    float *hf = hfunc + n * NN;
    if (threadIdx.x == 0) {
        atomicMin(&uni[0], x0); atomicMax(&uni[1], x0 + W);
        atomicMin(&uni[2], y0); atomicMax(&uni[3], y0 + W);
    }
    if (mode & 1) {
        for (int i = w; i < W; i += nw) {
            const size_t g = (size_t)(x0 + i) * N + y0;
            for (int j = lane; j < W; j += 64) hf[g + j] = hfun(ph[g + j]);
        }
    }
    if (mode & 2) {
        const int i0 = box[4 * n], i1 = box[4 * n + 1], j0 = box[4 * n + 2], j1 = box[4 * n + 3];
        double acc = 0.0;
        for (int i = i0 + w; i < i1; i += nw) {
            const bool rowin = (mode & 1) && i >= x0 && i < x0 + W;
            const float *pr = ph + (size_t)i * N, *hr = hf + (size_t)i * N;
            for (int j = j0 + lane; j < j1; j += 64) {
                /* in the window use the value this launch stored there */
                float h = (rowin && j >= y0 && j < y0 + W) ? hfun(pr[j]) : hr[j];
                acc += (double)(h * P.dx * P.dx);
            }
        }
        double tot = block_sum_d(acc, red);
        if (threadIdx.x == 0) {
            s[n] = tot;
            f0[n] = P.alpha * ((float)s0[n] - (float)tot);
        }
    }
}

/* k_cell_prep: k_centroid and k_prep (mode 3) of one cell in one workgroup
 * (batched form).  The centre and support box pass through LDS; the union
 * box is reset by k_commit_lumecm at the end of the previous step. */
static __device__ __forceinline__ void cell_prep_body(const DevP &P, const int n, const float *__restrict__ phi,
                                                      float *__restrict__ hfunc, int *__restrict__ xc,
                                                      int *__restrict__ yc, int *__restrict__ box,
                                                      const double *__restrict__ s0, double *__restrict__ s,
                                                      float *__restrict__ f0, int *__restrict__ uni,
                                                      int *__restrict__ err) {
    __shared__ long long red[3][16];
    __shared__ double dred[16];
    __shared__ int sh[6];
    const int N = P.N;
    constexpr int W = ORG_W;
    const size_t NN = (size_t)N * N;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, nw = blockDim.x >> 6;
    const float *ph = phi + n * NN;
    float *hf = hfunc + n * NN;
    /* centroid over the support box (useful_functions.cpp:193-227) */
    {
        const int i0 = box[4 * n], i1 = box[4 * n + 1], j0 = box[4 * n + 2], j1 = box[4 * n + 3];
        long long sx = 0, sy = 0, c = 0;
        for (int i = i0 + w; i < i1; i += nw) {
            const float *row = ph + (size_t)i * N;
            for (int j = j0 + lane; j < j1; j += 64)
                if (row[j] > 0.5f) { sx += i; sy += j; c += 1; }
        }
        sx = wave_sum_ll(sx); sy = wave_sum_ll(sy); c = wave_sum_ll(c);
        if (lane == 0) { red[0][w] = sx; red[1][w] = sy; red[2][w] = c; }
        __syncthreads();
        if (threadIdx.x == 0) {
            long long X = 0, Y = 0, C = 0;
            for (int k = 0; k < nw; k++) { X += red[0][k]; Y += red[1][k]; C += red[2][k]; }
            int x = xc[n], y = yc[n];
            if (C == 0) set_err(err, ORG_EEMPTY_CELL);
            else { x = (int)(X / C); y = (int)(Y / C); }
            int lo = P.zs + 2, hi = N - 2 - P.zs;
            if (x < lo || x > hi || y < lo || y > hi) {
                set_err(err, ORG_EWINDOW);
                x = min(max(x, lo), hi);
                y = min(max(y, lo), hi);
            }
            xc[n] = x; yc[n] = y;
            int b0, b1, b2, b3;
            if (i1 <= i0 || j1 <= j0) { b0 = x - P.zs; b1 = x + P.zs; b2 = y - P.zs; b3 = y + P.zs; }
            else { b0 = min(i0, x - P.zs); b1 = max(i1, x + P.zs); b2 = min(j0, y - P.zs); b3 = max(j1, y + P.zs); }
            box[4 * n + 0] = b0; box[4 * n + 1] = b1; box[4 * n + 2] = b2; box[4 * n + 3] = b3;
            sh[0] = x - P.zs; sh[1] = y - P.zs; sh[2] = b0; sh[3] = b1; sh[4] = b2; sh[5] = b3;
            atomicMin(&uni[0], x - P.zs); atomicMax(&uni[1], x - P.zs + W);
            atomicMin(&uni[2], y - P.zs); atomicMax(&uni[3], y - P.zs + W);
        }
        __syncthreads();
    }
    const int x0 = sh[0], y0 = sh[1], i0 = sh[2], i1 = sh[3], j0 = sh[4], j1 = sh[5];
    /* hfunc on the window (useful_functions.cpp:233-251) */
    for (int i = w; i < W; i += nw) {
        const size_t g = (size_t)(x0 + i) * N + y0;
        for (int j = lane; j < W; j += 64) hf[g + j] = hfun(ph[g + j]);
    }
    /* area over the support box, FP64 (useful_functions.cpp:168-190) */
    double acc = 0.0;
    for (int i = i0 + w; i < i1; i += nw) {
        const bool rowin = i >= x0 && i < x0 + W;
        const float *pr = ph + (size_t)i * N, *hr = hf + (size_t)i * N;
        for (int j = j0 + lane; j < j1; j += 64) {
            float h = (rowin && j >= y0 && j < y0 + W) ? hfun(pr[j]) : hr[j];
            acc += (double)(h * P.dx * P.dx);
        }
    }
    double tot = block_sum_d(acc, dred);
    if (threadIdx.x == 0) {
        s[n] = tot;
        f0[n] = P.alpha * ((float)s0[n] - (float)tot);
    }
}

/* k_hlum: compute_all_hfunc_2d_tables_kernels (useful_functions.cpp:272-285)
 * and the FP64 sum_ecm of update_lumen_and_ECM (cell_phase_functions.cpp:376-386),
 * reduced by the last block to finish (fixed order: deterministic). */
static __device__ __forceinline__ void hlum_body(const DevP &P, const int blk, const int nblk,
                                                 const float *__restrict__ lum, const float *__restrict__ ecm,
                                                 float *__restrict__ hlum, float *__restrict__ hecm,
                                                 double *__restrict__ part, unsigned *__restrict__ counter,
                                                 double *__restrict__ sum_out) {
    __shared__ double red[16];
    __shared__ bool last;
    const int N = P.N;
    double acc = 0.0;
#pragma unroll 4
    for (int i = blk; i < N; i += nblk) {
        const size_t g = (size_t)i * N;
        for (int j = threadIdx.x; j < N; j += blockDim.x) {
            float a = lum[g + j], b = ecm[g + j];
            float hb = hfun(b);
            hlum[g + j] = hfun(a);
            hecm[g + j] = hb;
            acc += (double)(hb * P.dx * P.dx);
        }
    }
    double tot = block_sum_d(acc, red);
    if (threadIdx.x == 0) {
        part[blk] = tot;
        __threadfence();
        unsigned prev = atomicAdd(counter, 1u);
        last = (prev == (unsigned)nblk - 1);
    }
    __syncthreads();
    if (last) {
        __threadfence();
        double v = 0.0;
        for (int k = threadIdx.x; k < nblk; k += blockDim.x)
            v += __hip_atomic_load(&part[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        double t2 = block_sum_d(v, red);
        if (threadIdx.x == 0) { *sum_out = t2; *counter = 0u; }
    }
}

__global__ __launch_bounds__(1024) void k_hlum(DevP P, const float *__restrict__ lum, const float *__restrict__ ecm,
                                               float *__restrict__ hlum, float *__restrict__ hecm,
                                               double *__restrict__ part, unsigned *__restrict__ counter,
                                               double *__restrict__ sum_out) {
    hlum_body(P, blockIdx.x, gridDim.x, lum, ecm, hlum, hecm, part, counter, sum_out);
}

/* k_pre: centroid + prep of every cell (workgroups 0..ncell-1) and
 * h(lum), h(ECM), sum h(ECM) (the other nblk workgroups) in one launch; the
 * two roles are independent. */
__global__ __launch_bounds__(1024) void k_pre(DevP P, const float *__restrict__ phi, float *__restrict__ hfunc,
                                              int *__restrict__ xc, int *__restrict__ yc, int *__restrict__ box,
                                              const double *__restrict__ s0, double *__restrict__ s,
                                              float *__restrict__ f0, int *__restrict__ uni, int *__restrict__ err,
                                              const float *__restrict__ lum, const float *__restrict__ ecm,
                                              float *__restrict__ hlum, float *__restrict__ hecm,
                                              double *__restrict__ part, unsigned *__restrict__ counter,
                                              double *__restrict__ sum_out) {
    if ((int)blockIdx.x < P.nc)
        cell_prep_body(P, blockIdx.x, phi, hfunc, xc, yc, box, s0, s, f0, uni, err);
    else
        hlum_body(P, blockIdx.x - P.nc, gridDim.x - P.nc, lum, ecm, hlum, hecm, part, counter, sum_out);
}

/* k_psi: compute_psi_kernels (useful_functions.cpp:305-322) on 32 x 32 tiles
 * of the union of this step's windows. */
__global__ __launch_bounds__(256) void k_psi(DevP P, const float *__restrict__ hfunc, const int *__restrict__ box,
                                             const int *__restrict__ uni, float *__restrict__ psi) {
    __shared__ int list[1024];
    __shared__ int4 lbox[1024];
    __shared__ int cnt;
    const int N = P.N, ntx = (N + 15) / 16;
    const int ti0 = (blockIdx.x / ntx) * 16, tj0 = (blockIdx.x % ntx) * 16;
    const int u0 = uni[0], u1 = uni[1], v0 = uni[2], v1 = uni[3];
    if (ti0 >= u1 || ti0 + 16 <= u0 || tj0 >= v1 || tj0 + 16 <= v0) return;
    /* ordered list of cells whose support box meets the tile (wave 0) */
    if (threadIdx.x < 64) {
        int base = 0;
        for (int c0 = 0; c0 < P.nc; c0 += 64) {
            int n = c0 + threadIdx.x;
            bool hit = false;
            if (n < P.nc) {
                const int *bx = box + 4 * n;
                hit = !(ti0 >= bx[1] || ti0 + 16 <= bx[0] || tj0 >= bx[3] || tj0 + 16 <= bx[2]);
            }
            unsigned long long m = __ballot(hit);
            if (hit) {
                unsigned long long below = m & ((1ull << threadIdx.x) - 1ull);
                int pos = base + __popcll(below);
                if (pos < 1024) {
                    list[pos] = n;
                    lbox[pos] = make_int4(box[4 * n], box[4 * n + 1], box[4 * n + 2], box[4 * n + 3]);
                }
            }
            base += __popcll(m);
        }
        if (threadIdx.x == 0) cnt = min(base, 1024);
    }
    __syncthreads();
    const size_t NN = (size_t)N * N;
    {
        const int e = threadIdx.x;
        const int i = ti0 + (e >> 4), j = tj0 + (e & 15);
        if (i >= N || j >= N || i < u0 || i >= u1 || j < v0 || j >= v1) return;
        float acc = 0.0f;
        for (int k = 0; k < cnt; k++) {
            const int4 bx = lbox[k];
            if (i >= bx.x && i < bx.y && j >= bx.z && j < bx.w) acc += hfunc[list[k] * NN + (size_t)i * N + j];
        }
        psi[(size_t)i * N + j] = acc;
    }
}

/* 9-point Laplacian numerator, cell_phase_functions.cpp:198 and :416 */
#define L9AT(t, g, N) (-12.0f * (t)[g] + 2.0f * ((t)[(g) + (N)] + (t)[(g) - (N)] + (t)[(g) + 1] + (t)[(g) - 1]) \
                       + (t)[(g) + (N) + 1] + (t)[(g) + (N) - 1] + (t)[(g) - (N) - 1] + (t)[(g) - (N) + 1])

/* lumen / ECM reshaping loop state: stop flag, sweeps done, buffer holding
 * the last iterate */
struct LumState { int done, ri, cur, pad; };

/* k_lumecm_update: update_lumen_and_ECM (cell_phase_functions.cpp:395-443),
 * then boundary_conditions_lumen_and_ECM (useful_functions.cpp:439-462).
 * Writes the first iterate of both reshaping loops into both ping-pong
 * buffers (sweeps rewrite only their active box), resets the loop states and
 * the normals' boxes. */
__global__ __launch_bounds__(1024) void k_lumecm_update(DevP P, const float *__restrict__ lum,
                                                       const float *__restrict__ ecm, const float *__restrict__ psi,
                                                       const float *__restrict__ hlum, const float *__restrict__ hecm,
                                                       const double *__restrict__ sum_ecm, float *__restrict__ lumA,
                                                       float *__restrict__ lumB, float *__restrict__ ecmA,
                                                       float *__restrict__ ecmB, LumState *__restrict__ lst,
                                                       int *__restrict__ abort_flag, unsigned *__restrict__ bar) {
    const int N = P.N;
    if (blockIdx.x == 0 && threadIdx.x == 0) { *abort_flag = 0; bar[0] = 0u; }
    if (blockIdx.x == 0 && threadIdx.x < 2) {
        bool run = P.rdev0_lum > P.thr_lum;
        lst[threadIdx.x].done = run ? 0 : 1;
        lst[threadIdx.x].ri = 0;
        lst[threadIdx.x].cur = 0;
    }
    const float se = (float)(*sum_ecm);
    const float inv4 = 1.0f / (4.0f * P.dx * P.dx);
    const float ecm_area = (float)(N * N) * P.dx * P.dx - se;
#pragma unroll 2
    for (int i = blockIdx.x; i < N; i += gridDim.x) {
        const size_t r = (size_t)i * N;
        for (int j = threadIdx.x; j < N; j += blockDim.x) {
            const size_t e = r + j;
            float ln, en;
            if (i == 0 || j == 0 || i == N - 1 || j == N - 1) {
                ln = 0.0f;       /* lumen edges and corners 0 */
                en = 1.0f;       /* ECM edges and corners 1 (:449-460) */
            } else {
                float l = lum[e], E = ecm[e], ps = psi[e];
                float lap = L9AT(psi, e, (size_t)N) * inv4;
                ln = l + (1.0f / P.tau_lum) * l * (1.0f - l) * P.dt *
                             (P.xi_lum - P.beta_lc * ps + P.eta_lc * lap - P.beta_le * hecm[e]);
                en = E + (1.0f / P.tau_E) * E * (1.0f - E) * P.dt *
                             (P.alpha_E * ecm_area + P.xi_E - P.beta_ec * ps + P.eta_ec * lap - P.beta_le * hlum[e]);
            }
            lumA[e] = ln; lumB[e] = ln;
            ecmA[e] = en; ecmB[e] = en;
        }
    }
}

/* periodic ghost-ring source index of boundary_conditions_periodic /
 * boundary_conditions_cells (useful_functions.cpp:391-420, 505-526) */
static __device__ __forceinline__ int ring_src(int i, int N) { return i == 0 ? N - 2 : (i == N - 1 ? 1 : i); }

/* k_lum_normals: |grad| of the old field (cell_phase_functions.cpp:722-741),
 * normals from the new field's gradient (:744-771) with the periodic ring,
 * and the bounding box of the nonzero normals (jbox, inclusive).  A sweep
 * changes a pixel only if the normal at it or at one of its four neighbours
 * is nonzero (every flux term carries a face average of the normals).
 * blockIdx.y + fbase: 0 lumen, 1 ECM. */
__global__ __launch_bounds__(1024) void k_lum_normals(DevP P, const float *__restrict__ lum_old,
                                                     const float *__restrict__ ecm_old, const float *__restrict__ lumn,
                                                     const float *__restrict__ ecmn, float *__restrict__ jlx,
                                                     float *__restrict__ jly, float *__restrict__ jex,
                                                     float *__restrict__ jey, int *__restrict__ jbox_part,
                                                     int fbase) {
    __shared__ int sb[4];
    const int N = P.N, f = fbase + blockIdx.y;
    const float *old = f ? ecm_old : lum_old;
    const float *nw = f ? ecmn : lumn;
    float *jx = f ? jex : jlx, *jy = f ? jey : jly;
    const float c = 1.0f / P.dx / P.dx / 4.0f, h = 1.0f / (2.0f * P.dx);
    if (threadIdx.x == 0) { sb[0] = N; sb[1] = -1; sb[2] = N; sb[3] = -1; }
    __syncthreads();
    int b0 = N, b1 = -1, b2 = N, b3 = -1;
#pragma unroll 2
    for (int i = blockIdx.x; i < N; i += gridDim.x) {
        const int si = ring_src(i, N);
        for (int j = threadIdx.x; j < N; j += blockDim.x) {
            const size_t g = (size_t)si * N + ring_src(j, N);
            float a = old[g + N] - old[g - N], b = old[g + 1] - old[g - 1];
            float ng = sqrtf(c * (a * a + b * b));
            float vx = 0.0f, vy = 0.0f;
            if (ng != 0.0f) {
                vx = h * (nw[g + N] - nw[g - N]) / ng;
                vy = h * (nw[g + 1] - nw[g - 1]) / ng;
            }
            const size_t e = (size_t)i * N + j;
            jx[e] = vx; jy[e] = vy;
            if (vx != 0.0f || vy != 0.0f) { b0 = min(b0, i); b1 = max(b1, i); b2 = min(b2, j); b3 = max(b3, j); }
        }
    }
    if (b1 >= 0) { atomicMin(&sb[0], b0); atomicMax(&sb[1], b1); atomicMin(&sb[2], b2); atomicMax(&sb[3], b3); }
    __syncthreads();
    if (threadIdx.x < 4) jbox_part[((size_t)f * gridDim.x + blockIdx.x) * 4 + threadIdx.x] = sb[threadIdx.x];
}

/* Bounding box (inclusive) of field f's nonzero normals from the
 * per-workgroup boxes written by k_lum_normals; valid in every thread. */
static __device__ __forceinline__ int4 reduce_jbox(const int *__restrict__ part, int nparts, int f, int N) {
    __shared__ int sb[4];
    if (threadIdx.x == 0) { sb[0] = N; sb[1] = -1; sb[2] = N; sb[3] = -1; }
    __syncthreads();
    int b0 = N, b1 = -1, b2 = N, b3 = -1;
    for (int k = threadIdx.x; k < nparts; k += blockDim.x) {
        const int *q = part + ((size_t)f * nparts + k) * 4;
        if (q[1] >= 0) { b0 = min(b0, q[0]); b1 = max(b1, q[1]); b2 = min(b2, q[2]); b3 = max(b3, q[3]); }
    }
    if (b1 >= 0) { atomicMin(&sb[0], b0); atomicMax(&sb[1], b1); atomicMin(&sb[2], b2); atomicMax(&sb[3], b3); }
    __syncthreads();
    int4 r = make_int4(sb[0], sb[1], sb[2], sb[3]);
    __syncthreads();
    return r;
}

/* Conservative interface flux (cell_phase_functions.cpp:511-547 and
 * :802-838) applied to un at (i,j); u, jx, jy with row strides su, sj. */
static __device__ __forceinline__ float reshape_flux(float un, const float *u, int su, const float *jx,
                                                     const float *jy, int sj, float ka, float kc) {
#define U_(a, b) u[(a) * su + (b)]
#define JX_(a, b) jx[(a) * sj + (b)]
#define JY_(a, b) jy[(a) * sj + (b)]
    float u00 = U_(0, 0);
    float up0 = U_(1, 0), um0 = U_(-1, 0), u0p = U_(0, 1), u0m = U_(0, -1);
    float upp = U_(1, 1), upm = U_(1, -1), ump = U_(-1, 1), umm = U_(-1, -1);
    float jx0 = JX_(0, 0), jy0 = JY_(0, 0);
    float jxp = JX_(1, 0), jyp = JY_(1, 0), jxm = JX_(-1, 0), jym = JY_(-1, 0);
    float jx0p = JX_(0, 1), jy0p = JY_(0, 1), jx0m = JX_(0, -1), jy0m = JY_(0, -1);
    float g00 = u00 * (1.0f - u00);
    float a = 0.5f * (jxp + jx0);
    un += a * (ka * (up0 - u00) * a + ka * 0.25f * (upp + u0p - upm - u0m) * 0.5f * (jyp + jy0)
               - kc * 0.5f * (up0 * (1.0f - up0) + g00));
    a = 0.5f * (jy0p + jy0);
    un += a * (ka * (u0p - u00) * a + ka * 0.25f * (upp + up0 - ump - um0) * 0.5f * (jx0p + jx0)
               - kc * 0.5f * (u0p * (1.0f - u0p) + g00));
    a = 0.5f * (jx0 + jxm);
    un -= a * (ka * (u00 - um0) * a + ka * 0.25f * (u0p + ump - u0m - umm) * 0.5f * (jy0 + jym)
               - kc * 0.5f * (g00 + um0 * (1.0f - um0)));
    a = 0.5f * (jy0 + jy0m);
    un -= a * (ka * (u00 - u0m) * a + ka * 0.25f * (up0 + upm - um0 - umm) * 0.5f * (jx0 + jx0m)
               - kc * 0.5f * (g00 + u0m * (1.0f - u0m)));
    return un;
#undef U_
#undef JX_
#undef JY_
}

/* One reshaping sweep of reshaping_lumen_or_ECM's loop
 * (cell_phase_functions.cpp:774-869) for one field: X_{k+1} = X_k + flux(X_k)
 * inside, Dirichlet edges (value side) and zero corners
 * (useful_functions.cpp:422-437).  Interior pixels are swept only inside
 * the normals' box grown by one pixel (outside it X_{k+1} = X_k exactly);
 * the boundary ring every sweep.  The work is split over `nblk` workgroups
 * of 1024 threads as one flat pixel range; returns this thread's share of
 * the deviation sum |X_{k+1}-X_k| dx^2.  The same decomposition is used by
 * the stand-alone and the fused loop so their sums agree bit for bit. */
#define ORG_LUM_BLOCKS 64
static __device__ __forceinline__ double lum_sweep(const DevP &P, const float *__restrict__ src,
                                                   float *__restrict__ dst, const float *__restrict__ jx,
                                                   const float *__restrict__ jy, const int4 jb,
                                                   float side, int blk, int nblk) {
    const int N = P.N;
    const float ka = P.wi * P.wi * (P.rdt / P.dx / P.dx);
    const float kc = P.rdt * P.wi * 1.41421356237f / P.dx;
    const float dx2 = P.dx * P.dx;
    const int r0 = max(jb.x - 1, 1), r1 = min(jb.y + 1, N - 2);
    const int c0 = max(jb.z - 1, 1), c1 = min(jb.w + 1, N - 2);
    const int bw = c1 - c0 + 1, bh = r1 - r0 + 1;
    const int npx = (bw > 0 && bh > 0) ? bw * bh : 0;
    const int stride = nblk * 1024;
    double acc = 0.0;
    for (int k = blk * 1024 + (int)threadIdx.x; k < npx; k += stride) {
        const int i = r0 + k / bw, j = c0 + k % bw;
        const size_t e = (size_t)i * N + j;
        float u = src[e];
        float v = reshape_flux(u, src + e, N, jx + e, jy + e, N, ka, kc);
        dst[e] = v;
        acc += (double)(fabsf(v - u) * dx2);
    }
    for (int k = blk * 1024 + (int)threadIdx.x; k < 4 * N - 4; k += stride) {
        int i, j;
        if (k < N) { i = 0; j = k; }
        else if (k < 2 * N) { i = N - 1; j = k - N; }
        else if (k < 3 * N - 2) { i = k - 2 * N + 1; j = 0; }
        else { i = k - (3 * N - 2) + 1; j = N - 1; }
        const size_t e = (size_t)i * N + j;
        const bool corner = (i == 0 || i == N - 1) && (j == 0 || j == N - 1);
        float v = corner ? 0.0f : side;
        dst[e] = v;
        acc += (double)(fabsf(v - src[e]) * dx2);
    }
    return acc;
}

/* Sum of one field's per-workgroup deviations in a fixed order, valid in
 * every thread (1024 threads, nblk <= 1024). */
static __device__ __forceinline__ double sum_parts(const double *part, int nblk, double *red) {
    double v = 0.0;
    for (int b = threadIdx.x; b < nblk; b += blockDim.x)
        v += __hip_atomic_load(&part[b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return block_sum_d(v, red);
}

/* Stop test after a sweep (cell_phase_functions.cpp:774, 862-868). */
static __device__ __forceinline__ void lum_stop(const DevP &P, LumState &s, double rdev, int *err) {
    s.cur = (s.ri + 1) & 1;
    if (s.ri > P.rimax) { set_err(err, ORG_ERELAX); s.done = 1; }
    s.ri += 1;
    if (!(rdev > P.thr_lum)) s.done = 1;
}

/* k_lum_iter: one sweep of the lumen (field 0) and / or ECM (1) loop;
 * ORG_LUM_BLOCKS workgroups per field; the last workgroup to finish applies
 * the stop test.  A field whose loop has stopped skips the sweep.  Used by
 * the per-function entry point, whose host loop reads the stop flag after
 * every sweep as the reference does. */
__global__ __launch_bounds__(1024) void k_lum_iter(DevP P, float *__restrict__ bufL0, float *__restrict__ bufL1,
                                                   float *__restrict__ bufE0, float *__restrict__ bufE1,
                                                   const float *__restrict__ jlx, const float *__restrict__ jly,
                                                   const float *__restrict__ jex, const float *__restrict__ jey,
                                                   const int *__restrict__ jbox_part, int nparts,
                                                   LumState *__restrict__ lst, double *__restrict__ part,
                                                   unsigned *__restrict__ counter, int *__restrict__ err, int fbase) {
    __shared__ double red[16];
    __shared__ bool last;
    const int f = fbase + blockIdx.y;
    const LumState s = lst[f];
    if (s.done) return;
    const int4 jb = reduce_jbox(jbox_part, nparts, f, P.N);
    float *b0 = f ? bufE0 : bufL0, *b1 = f ? bufE1 : bufL1;
    const float *src = (s.ri & 1) ? b1 : b0;
    float *dst = (s.ri & 1) ? b0 : b1;
    double acc = lum_sweep(P, src, dst, f ? jex : jlx, f ? jey : jly, jb, f ? 1.0f : 0.0f, blockIdx.x, gridDim.x);
    double tot = block_sum_d(acc, red);
    if (threadIdx.x == 0) {
        part[f * gridDim.x + blockIdx.x] = tot;
        __threadfence();
        unsigned prev = atomicAdd(&counter[f], 1u);
        last = (prev == gridDim.x - 1);
    }
    __syncthreads();
    if (last) {
        __threadfence();
        double rdev = sum_parts(part + f * gridDim.x, gridDim.x, red);
        if (threadIdx.x == 0) {
            LumState o = s;
            lum_stop(P, o, rdev, err);
            lst[f] = o;
            counter[f] = 0u;
        }
    }
}

/* Barrier among the `nblk` loop workgroups of one launch (generation
 * counter, sleeping poll).  Only loop workgroups ever wait, and only on each
 * other, and every other workgroup of the launch finishes on its own, so all
 * loop workgroups become resident eventually.  A wait longer than
 * `limit` wall-clock ticks sets ORG_ERELAX and the abort flag instead of
 * waiting on. */
static __device__ bool lum_barrier(unsigned *bar, int nblk, int *abort_flag, int *err, long long limit) {
    __shared__ int ok;
    __syncthreads();
    if (threadIdx.x == 0) {
        ok = 1;
        unsigned g = __hip_atomic_load(&bar[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __threadfence();
        unsigned arrived = atomicAdd(&bar[0], 1u);
        if (arrived == (unsigned)nblk - 1) {
            __hip_atomic_store(&bar[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __threadfence();
            atomicAdd(&bar[1], 1u);
        } else {
            long long t0 = wall_clock64();
            while (__hip_atomic_load(&bar[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == g) {
                __builtin_amdgcn_s_sleep(2);
                if (__hip_atomic_load(abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ||
                    wall_clock64() - t0 > limit) {
                    atomicExch(abort_flag, 1);
                    set_err(err, ORG_ERELAX);
                    ok = 0;
                    break;
                }
            }
        }
        __threadfence();
    }
    __syncthreads();
    return ok != 0;
}

/* Both reshaping loops (lumen and ECM) run to their stop tests by the loop
 * workgroups of k_cortex_cells, one barrier per sweep.  Every loop workgroup
 * sums the per-workgroup deviations in the same order, so all take the same
 * stop decision after the same sweep.  part: [2 parities][2 fields][nblk]. */
static __device__ void lum_loop_body(const DevP &P, int blk, int nblk, double *red, float *bufL0, float *bufL1,
                                     float *bufE0, float *bufE1, const float *jlx, const float *jly,
                                     const float *jex, const float *jey, const int *jbox_part, int nparts,
                                     LumState *lst, double *part, unsigned *bar, int *abort_flag, int *err,
                                     long long limit) {
    LumState s[2] = {lst[0], lst[1]};
    const int4 jb[2] = {reduce_jbox(jbox_part, nparts, 0, P.N), reduce_jbox(jbox_part, nparts, 1, P.N)};
    for (int k = 0; !(s[0].done && s[1].done); k++) {
        double *pk = part + (size_t)(k & 1) * 2 * nblk;
        for (int f = 0; f < 2; f++) {
            double acc = 0.0;
            if (!s[f].done) {
                float *b0 = f ? bufE0 : bufL0, *b1 = f ? bufE1 : bufL1;
                const float *src = (s[f].ri & 1) ? b1 : b0;
                float *dst = (s[f].ri & 1) ? b0 : b1;
                acc = lum_sweep(P, src, dst, f ? jex : jlx, f ? jey : jly, jb[f], f ? 1.0f : 0.0f, blk, nblk);
            }
            double tot = block_sum_d(acc, red);
            if (threadIdx.x == 0) pk[f * nblk + blk] = tot;
        }
        if (!lum_barrier(bar, nblk, abort_flag, err, limit)) return;
        for (int f = 0; f < 2; f++) {
            if (s[f].done) continue;
            double rdev = sum_parts(pk + f * nblk, nblk, red);
            lum_stop(P, s[f], rdev, err);
        }
    }
    if (blk == 0 && threadIdx.x < 2) lst[threadIdx.x] = s[threadIdx.x];
}

/* k_commit_lumecm: lum <- lumn, ECM <- ECMn (useful_functions.cpp:695-709,
 * run_with_active_gels.cpp:181-182) from the buffer holding the last iterate. */
static __device__ __forceinline__ void commit_lumecm_body(const DevP &P, const int blk, const int nblk,
                                                          float *__restrict__ lum, float *__restrict__ ecm,
                                                          const float *__restrict__ bufL0,
                                                          const float *__restrict__ bufL1,
                                                          const float *__restrict__ bufE0,
                                                          const float *__restrict__ bufE1,
                                                          const LumState *__restrict__ lst, int *__restrict__ err,
                                                          int *__restrict__ uni) {
    const int N = P.N;
    const LumState sl = lst[0], se = lst[1];
    if (blk == 0 && threadIdx.x == 0) { uni[0] = N; uni[1] = 0; uni[2] = N; uni[3] = 0; }
    if (blk == 0 && threadIdx.x == 0 && (!sl.done || !se.done)) set_err(err, ORG_ERELAX);
    const float *L = sl.cur ? bufL1 : bufL0;
    const float *E = se.cur ? bufE1 : bufE0;
#pragma unroll 4
    for (int i = blk; i < N; i += nblk) {
        const size_t r = (size_t)i * N;
        for (int j = threadIdx.x; j < N; j += blockDim.x) {
            lum[r + j] = L[r + j];
            ecm[r + j] = E[r + j];
        }
    }
}

__global__ __launch_bounds__(1024) void k_commit_lumecm(DevP P, float *__restrict__ lum, float *__restrict__ ecm,
                                                        const float *__restrict__ bufL0, const float *__restrict__ bufL1,
                                                        const float *__restrict__ bufE0, const float *__restrict__ bufE1,
                                                        const LumState *__restrict__ lst, int *__restrict__ err,
                                                        int *__restrict__ uni) {
    commit_lumecm_body(P, blockIdx.x, gridDim.x, lum, ecm, bufL0, bufL1, bufE0, bufE1, lst, err, uni);
}

/* ======================================================================== */
/* k_spectral: update_all_cortices (rho_functions.cpp:409-471) for every cell.
 *
 * One workgroup (1024 threads) per cell, one 128 x 128 complex field in LDS,
 * two 128 x 128 complex scratch planes per cell in global memory (L2
 * resident) for the spectra that are needed at k and -k.
 * By linearity of the reference's spectral update, Re(F^-1[F(rho) +
 * dt F(react) + D]) / M^2 = rho + dt*react + Re(F^-1 D) / M^2, so the
 * transforms of rho and of the reaction term (rho_functions.cpp:415-422,
 * 462-464) cancel and are not taken.  Real pairs share one complex
 * transform: F(pi_xx + i pi_xy), F(pi_yy), F^-1(H v_x + i H v_y) where
 * H X(k) = (X(k) + conj X(-k)) / 2 is the part whose inverse is real (the
 * reference keeps only Re, rho_functions.cpp:279), F(rho v_x + i rho v_y),
 * and F^-1(H D).  The per-mode solve uses the viscosity of the real-space
 * pixel with the mode's indices (rho_functions.cpp:149-150, 170). */
static __device__ __forceinline__ void spectral_body(const DevP &P, const int n, float2 *__restrict__ sm,
                                                     float2 *__restrict__ tw, unsigned char *__restrict__ lab,
                                                     const float *__restrict__ phi,
                                                     const float *__restrict__ rho, const float *__restrict__ lum,
                                                     const float *__restrict__ ecm, const int *__restrict__ xc,
                                                     const int *__restrict__ yc, float2 *__restrict__ scratch,
                                                     float *__restrict__ rho_box, float *__restrict__ vxbox,
                                                     float *__restrict__ vybox, int write_v) {
    const int t = threadIdx.x, N = P.N;
    constexpr int MM = ORG_MB * ORG_MB;
    constexpr int NPAIR = 63 * ORG_MB + 2 * 65;      /* mode pairs {k, -k} incl. 4 self-paired modes */
    const size_t NN = (size_t)N * N;
    const int x0 = xc[n] - P.zs + 1, y0 = yc[n] - P.zs + 1;   /* FFT box origin, useful_functions.cpp:618-619 */
    const int n1 = (n == 0) ? P.nc - 1 : n - 1, n2 = (n == P.nc - 1) ? 0 : n + 1;
    if (t < 128) tw[t] = make_float2(kCos128[t], kSin128[t]);
    const float *ph = phi + n * NN, *ph1 = phi + n1 * NN, *ph2 = phi + n2 * NN, *rh = rho + n * NN;
    float2 *SA = scratch + (size_t)n * 2 * MM;
    float *SB = (float *)(SA + MM);
    float *rbox = rho_box + (size_t)n * MM;
    const float inv4 = 1.0f / (4.0f * P.dx * P.dx);
    const float isig2 = 1.0f / (2.0f * P.sigma * P.sigma);
    const float dnorm = 1.0f / (2.50662827463f * P.sigma);   /* 1/(sqrt(2 pi) sigma) */
    const float kscale = 6.28318530718f / (ORG_MB * P.dx);
    const float MMf = (float)MM;

    /* P1: active stress (rho_functions.cpp:32-115), reaction (:193-209) */
#pragma unroll 2
    for (int q = 0; q < 16; q++) {
        const int e = t + 1024 * q, i = e >> 7, j = e & 127;
        const size_t g = (size_t)(x0 + i) * N + (y0 + j);
        float v = ph[g];
        float dxp = ph[g + N] - ph[g - N], dyp = ph[g + 1] - ph[g - 1];
        float ng = sqrtf(inv4 * (dxp * dxp + dyp * dyp));
        int idx = turnover_label(v, lum[g], ecm[g], ph1[g], ph2[g]);
        lab[e] = (unsigned char)idx;
        float r = rh[g];
        float at = P.atan_[idx], ao = P.aort[idx], bb = P.bb[idx];
        float r3 = r * r * r;
        float base = -at * r3 + bb * r3 * r;
        float pxx = base, pxy = 0.0f, pyv = base;
        if (ng > 0.0f) {
            float c = (at - ao) * r3 * inv4 / ng / ng;
            pxx += c * dxp * dxp;
            pxy += c * dxp * dyp;
            pyv += c * dyp * dyp;
        }
        float d = v - 0.5f;
        float pnd = ng * expf(-d * d * isig2) * dnorm;
        float react = -P.deg[idx] * r + P.gen[idx] * pnd;
        rbox[e] = r + P.dt * react;
        SB[e] = pyv;
        sm[i * ORG_LS + j] = make_float2(pxx, pxy);
    }
    __syncthreads();
    fft2d_128<-1>(sm, tw);
#pragma unroll 2
    for (int q = 0; q < 16; q++) {
        const int e = t + 1024 * q, l = (e >> 7) * ORG_LS + (e & 127);
        SA[e] = sm[l];
        sm[l] = make_float2(SB[e], 0.0f);      /* this thread wrote SB[e] in P1 */
    }
    __syncthreads();
    fft2d_128<-1>(sm, tw);
    /* P4: per pair {k, -k}: split F(pi_xx + i pi_xy), solve both modes
     * (rho_functions.cpp:133-190), keep the part of v whose inverse is real,
     * Z = H v_x + i H v_y written in place at k and -k. */
    for (int u = t; u < NPAIR; u += 1024) {
        int i, j;
        if (u < 63 * ORG_MB) { i = 1 + (u >> 7); j = u & 127; }
        else { const int w = u - 63 * ORG_MB; i = (w < 65) ? 0 : 64; j = (w < 65) ? w : w - 65; }
        const int im = (ORG_MB - i) & 127, jm = (ORG_MB - j) & 127;
        const int ek = i * ORG_MB + j, em = im * ORG_MB + jm;
        const int lk = i * ORG_LS + j, lm = im * ORG_LS + jm;
        const float2 A = SA[ek], Am = SA[em], B = sm[lk], Bm = sm[lm];
        const float2 pxx = make_float2(0.5f * (A.x + Am.x), 0.5f * (A.y - Am.y));
        const float2 pxy = make_float2(0.5f * (A.y + Am.y), 0.5f * (Am.x - A.x));
        const float2 pxxm = make_float2(pxx.x, -pxx.y), pxym = make_float2(pxy.x, -pxy.y);
        const int ir = (i < ORG_MB / 2) ? i : i - ORG_MB, jr = (j < ORG_MB / 2) ? j : j - ORG_MB;
        const int irm = (im < ORG_MB / 2) ? im : im - ORG_MB, jrm = (jm < ORG_MB / 2) ? jm : jm - ORG_MB;
        float2 vxk, vyk, vxm, vym;
        velocity_mode(P.visc[lab[ek]], P.friction, kscale * ir, kscale * jr, pxx, pxy, B, vxk, vyk);
        velocity_mode(P.visc[lab[em]], P.friction, kscale * irm, kscale * jrm, pxxm, pxym, Bm, vxm, vym);
        const float2 hx = make_float2(0.5f * (vxk.x + vxm.x), 0.5f * (vxk.y - vxm.y));
        const float2 hy = make_float2(0.5f * (vyk.x + vym.x), 0.5f * (vyk.y - vym.y));
        sm[lk] = make_float2(hx.x - hy.y, hx.y + hy.x);
        if (em != ek) sm[lm] = make_float2(hx.x + hy.y, hy.x - hx.y);
    }
    __syncthreads();
    fft2d_128<1>(sm, tw);
    /* P5: v = Re(.)/M^2 ; rho * v (rho_functions.cpp:267-284) */
#pragma unroll 2
    for (int q = 0; q < 16; q++) {
        const int e = t + 1024 * q, i = e >> 7, j = e & 127;
        float2 z = sm[i * ORG_LS + j];
        float r = rh[(size_t)(x0 + i) * N + (y0 + j)];
        if (write_v) {
            vxbox[(size_t)n * MM + e] = z.x / MMf;
            vybox[(size_t)n * MM + e] = z.y / MMf;
        }
        sm[i * ORG_LS + j] = make_float2(r * z.x / MMf, r * z.y / MMf);
    }
    __syncthreads();
    fft2d_128<-1>(sm, tw);
    /* P6: spectral divergence of update_fft_rho (rho_functions.cpp:230-249)
     * per pair, the part whose inverse is real, in place */
    const float cdt = P.dt * kscale;
    for (int u = t; u < NPAIR; u += 1024) {
        int i, j;
        if (u < 63 * ORG_MB) { i = 1 + (u >> 7); j = u & 127; }
        else { const int w = u - 63 * ORG_MB; i = (w < 65) ? 0 : 64; j = (w < 65) ? w : w - 65; }
        const int im = (ORG_MB - i) & 127, jm = (ORG_MB - j) & 127;
        const int lk = i * ORG_LS + j, lm = im * ORG_LS + jm;
        const float2 Y = sm[lk], Ym = sm[lm];
        const float2 Pk = make_float2(0.5f * (Y.x + Ym.x), 0.5f * (Y.y - Ym.y));   /* F(rho vx)(k) */
        const float2 Qk = make_float2(0.5f * (Y.y + Ym.y), 0.5f * (Ym.x - Y.x));   /* F(rho vy)(k) */
        const float ir = (float)((i < ORG_MB / 2) ? i : i - ORG_MB), jr = (float)((j < ORG_MB / 2) ? j : j - ORG_MB);
        const float irm = (float)((im < ORG_MB / 2) ? im : im - ORG_MB);
        const float jrm = (float)((jm < ORG_MB / 2) ? jm : jm - ORG_MB);
        const float2 D = make_float2(cdt * (ir * Pk.y + jr * Qk.y), -cdt * (ir * Pk.x + jr * Qk.x));
        /* at -k the transforms are conj(Pk), conj(Qk) */
        const float2 Dm = make_float2(cdt * (irm * -Pk.y + jrm * -Qk.y), -cdt * (irm * Pk.x + jrm * Qk.x));
        const float2 H = make_float2(0.5f * (D.x + Dm.x), 0.5f * (D.y - Dm.y));
        sm[lk] = H;
        if (lm != lk) sm[lm] = make_float2(H.x, -H.y);
    }
    __syncthreads();
    fft2d_128<1>(sm, tw);
    /* P7: rho^{n+1} on the box (copy_to_rhon, rho_functions.cpp:253-265) */
#pragma unroll 2
    for (int q = 0; q < 16; q++) {
        const int e = t + 1024 * q;
        rbox[e] = rbox[e] + sm[(e >> 7) * ORG_LS + (e & 127)].x / MMf;
    }
}

__global__ __launch_bounds__(1024) void k_spectral(DevP P, const float *__restrict__ phi,
                                                   const float *__restrict__ rho, const float *__restrict__ lum,
                                                   const float *__restrict__ ecm, const int *__restrict__ xc,
                                                   const int *__restrict__ yc, float2 *__restrict__ scratch,
                                                   float *__restrict__ rho_box, float *__restrict__ vxbox,
                                                   float *__restrict__ vybox, int write_v) {
    __shared__ float2 sm[ORG_MB * ORG_LS];
    __shared__ float2 tw[128];
    __shared__ unsigned char lab[ORG_MB * ORG_MB];
    spectral_body(P, blockIdx.x, sm, tw, lab, phi, rho, lum, ecm, xc, yc, scratch, rho_box, vxbox, vybox, write_v);
}

/* ======================================================================== */
/* k_cells: update_all_cells (cell_phase_functions.cpp:126-217, mode&1) and
 * reshaping_cells (cell_phase_functions.cpp:447-580, mode&2) per cell, one
 * 1024-thread workgroup per cell.  The relaxed field (window plus a
 * one-pixel ring) and the iterate live in LDS, the normals (from the
 * start-of-step phi, useful_functions.cpp:356-388) in a per-cell scratch;
 * the stop test runs in the workgroup. */
static __device__ __forceinline__ void cells_body(const DevP &P, const int n, float *__restrict__ phase,
                                                  float *__restrict__ pnl, double *__restrict__ red,
                                                  float *__restrict__ phi, const float *__restrict__ rho,
                                                  const float *__restrict__ hfunc, const float *__restrict__ psi,
                                                  const float *__restrict__ hlum, const float *__restrict__ hecm,
                                                  const float *__restrict__ f0, const int *__restrict__ xc,
                                                  const int *__restrict__ yc, float *__restrict__ phin_win,
                                                  float *__restrict__ phi_win, float *__restrict__ jwin,
                                                  int *__restrict__ counts, int *__restrict__ err, int mode) {
    constexpr int W = ORG_W, WH = ORG_WH, NPX = ORG_W * ORG_W;
    const int t = threadIdx.x, N = P.N;
    const size_t NN = (size_t)N * N;
    const int x0 = xc[n] - P.zs, y0 = yc[n] - P.zs;
    float *ph = phi + n * NN;
    const float *rh = rho + n * NN, *hf = hfunc + n * NN;
    float *pw = phin_win + (size_t)n * NPX;

    if (mode & 1) {
        const float inv4 = 1.0f / (4.0f * P.dx * P.dx);
        const float isig2 = 1.0f / (2.0f * P.sigma * P.sigma);
        const float cplc = P.cpl / P.tau * (P.dt / P.dx / P.dx / 4.0f) / (2.50662827463f * P.sigma);
        const float fn = f0[n];
        for (int e = t; e < NPX; e += 1024) {
            const int i = e / W, j = e - i * W;
            const size_t g = (size_t)(x0 + i) * N + (y0 + j);
            float v = ph[g];
            float dxp = ph[g + N] - ph[g - N], dyp = ph[g + 1] - ph[g - 1];
            float ng = sqrtf(inv4 * (dxp * dxp + dyp * dyp));
            float out = v;
            if (ng > 0.0f) {
                float d = v - 0.5f;
                out -= cplc * expf(-d * d * isig2) *
                       ((rh[g + N] - rh[g - N]) * dxp + (rh[g + 1] - rh[g - 1]) * dyp) / ng;
            }
            /* temp_phase = psi - hfunc_n (cell_phase_functions.cpp:150) on the 3x3 stencil */
            float tc = psi[g] - hf[g];
            float tp0 = psi[g + N] - hf[g + N], tm0 = psi[g - N] - hf[g - N];
            float t0p = psi[g + 1] - hf[g + 1], t0m = psi[g - 1] - hf[g - 1];
            float tpp = psi[g + N + 1] - hf[g + N + 1], tpm = psi[g + N - 1] - hf[g + N - 1];
            float tmm = psi[g - N - 1] - hf[g - N - 1], tmp = psi[g - N + 1] - hf[g - N + 1];
            float lt = -12.0f * tc + 2.0f * (tp0 + tm0 + t0p + t0m) + tpp + tpm + tmm + tmp;
            float ll = L9AT(hlum, g, (size_t)N), le = L9AT(hecm, g, (size_t)N);
            out += (1.0f / P.tau) * v * (1.0f - v) * P.dt *
                   (fn - P.beta * tc + P.eta * lt * inv4 - P.beta_lc * hlum[g] + P.eta_lc * ll * inv4
                    - P.beta_ec * hecm[g] + P.eta_ec * le * inv4);
            if (mode & 2) pnl[(i + 1) * WH + (j + 1)] = out;
            else pw[e] = out;
        }
        if (!(mode & 2)) return;
    } else {
        for (int e = t; e < NPX; e += 1024) {
            const int i = e / W, j = e - i * W;
            pnl[(i + 1) * WH + (j + 1)] = pw[e];
        }
    }

    /* normals on the window plus ring, from the start-of-step phi */
    float *jx = jwin + (size_t)n * 2 * WH * WH, *jy = jx + WH * WH;
    {
        const float inv4 = 1.0f / (4.0f * P.dx * P.dx), h = 1.0f / (2.0f * P.dx);
        for (int e = t; e < WH * WH; e += 1024) {
            const int i = e / WH - 1, j = e - (i + 1) * WH - 1;
            const size_t g = (size_t)(x0 + i) * N + (y0 + j);
            float dxp = ph[g + N] - ph[g - N], dyp = ph[g + 1] - ph[g - 1];
            float ng = sqrtf(inv4 * (dxp * dxp + dyp * dyp));
            float a = 0.0f, b = 0.0f;
            if (ng > 0.0f) { a = h * dxp / ng; b = h * dyp / ng; }
            jx[e] = a; jy[e] = b;
            if (i < 0 || j < 0 || i >= W || j >= W) phase[e] = ph[g];
        }
    }
    __syncthreads();
    const float ka = P.wi * P.wi * (P.rdt / P.dx / P.dx);
    const float kc = P.rdt * P.wi * 1.41421356237f / P.dx;
    const float dscale = P.dx * P.dx / P.ncur;
    double rdev = P.rdev0_cell;
    int ri = 0, cnt = 0;
    while (rdev > P.thr_cell) {
        for (int e = t; e < NPX; e += 1024) {          /* phase <- phasen (:480-492) */
            const int i = e / W, l = (i + 1) * WH + (e - i * W + 1);
            phase[l] = pnl[l];
        }
        __syncthreads();
        double acc = 0.0;
        for (int e = t; e < NPX; e += 1024) {          /* flux update (:498-549) */
            const int i = e / W, l = (i + 1) * WH + (e - i * W + 1);
            float u = pnl[l];
            float nv = reshape_flux(u, phase + l, WH, jx + l, jy + l, WH, ka, kc);
            acc += (double)(fabsf(nv - u) * dscale);
            pnl[l] = nv;
        }
        rdev = block_sum_d(acc, red);                  /* (:554-562) */
        cnt++;
        if (ri > P.rimax) { if (t == 0) set_err(err, ORG_ERELAX); break; }
        ri++;
    }
    /* phi <- second-to-last iterate (still in phase); phin keeps the last.
     * With phi_win the window goes to a per-cell buffer that k_commit_cells
     * writes back, so other cells' workgroups of the same launch still read
     * the start-of-step phi. */
    float *pwin = phi_win ? phi_win + (size_t)n * NPX : nullptr;
    for (int e = t; e < NPX; e += 1024) {
        const int i = e / W, j = e - i * W, l = (i + 1) * WH + (j + 1);
        if (pwin) pwin[e] = phase[l];
        else ph[(size_t)(x0 + i) * N + (y0 + j)] = phase[l];
        pw[e] = pnl[l];
    }
    if (t == 0) counts[n] = cnt;
}

__global__ __launch_bounds__(1024) void k_cells(DevP P, float *__restrict__ phi, const float *__restrict__ rho,
                                                const float *__restrict__ hfunc, const float *__restrict__ psi,
                                                const float *__restrict__ hlum, const float *__restrict__ hecm,
                                                const float *__restrict__ f0, const int *__restrict__ xc,
                                                const int *__restrict__ yc, float *__restrict__ phin_win,
                                                float *__restrict__ jwin, int *__restrict__ counts,
                                                int *__restrict__ err, int mode) {
    __shared__ float phase[ORG_WH * ORG_WH];
    __shared__ float pnl[ORG_WH * ORG_WH];
    __shared__ double red[16];
    cells_body(P, blockIdx.x, phase, pnl, red, phi, rho, hfunc, psi, hlum, hecm, f0, xc, yc, phin_win, nullptr, jwin,
               counts, err, mode);
}

/* Buffers of the lumen/ECM loop role of k_cortex_cells. */
struct LumBufs {
    float *L0, *L1, *E0, *E1;
    const float *jlx, *jly, *jex, *jey;
    const int *jbox_part;
    int nparts;
    LumState *lst;
    double *part;
    unsigned *bar;
    int *abort_flag;
    long long limit;
};

/* k_cortex_cells: one launch for the rest of the step.
 *   workgroups 0..ncell-1         spectral cortex update of cell n
 *   workgroups ncell..2*ncell-1   phase-field update + reshaping of cell n
 *   workgroups 2*ncell..          the lumen and ECM reshaping loops
 * The three roles read only the start-of-step state (the cell roles read
 * lum/ECM; the loop role writes only its ping-pong buffers), so they run
 * concurrently; phi's new window goes to phi_win.  LDS is shared between
 * the roles. */
#define ORG_LDS_SPECTRAL (ORG_MB * ORG_LS * 8 + 128 * 8 + ORG_MB * ORG_MB)
#define ORG_LDS_CELLS (2 * ORG_WH * ORG_WH * 4 + 16 * 8)
#define ORG_LDS_FUSED (ORG_LDS_SPECTRAL > ORG_LDS_CELLS ? ORG_LDS_SPECTRAL : ORG_LDS_CELLS)
__global__ __launch_bounds__(1024) void k_cortex_cells(DevP P, float *__restrict__ phi, const float *__restrict__ rho,
                                                       const float *__restrict__ lum, const float *__restrict__ ecm,
                                                       const float *__restrict__ hfunc, const float *__restrict__ psi,
                                                       const float *__restrict__ hlum, const float *__restrict__ hecm,
                                                       const float *__restrict__ f0, const int *__restrict__ xc,
                                                       const int *__restrict__ yc, float2 *__restrict__ scratch,
                                                       float *__restrict__ rho_box, float *__restrict__ vxbox,
                                                       float *__restrict__ vybox, int write_v,
                                                       float *__restrict__ phin_win, float *__restrict__ phi_win,
                                                       float *__restrict__ jwin, int *__restrict__ counts,
                                                       int *__restrict__ err, LumBufs lb) {
    __shared__ __attribute__((aligned(16))) unsigned char lds[ORG_LDS_FUSED];
    if ((int)blockIdx.x >= 2 * P.nc) {
        lum_loop_body(P, blockIdx.x - 2 * P.nc, gridDim.x - 2 * P.nc, (double *)lds, lb.L0, lb.L1, lb.E0, lb.E1,
                      lb.jlx, lb.jly, lb.jex, lb.jey, lb.jbox_part, lb.nparts, lb.lst, lb.part, lb.bar, lb.abort_flag,
                      err, lb.limit);
    } else if ((int)blockIdx.x < P.nc) {
        float2 *sm = (float2 *)lds;
        spectral_body(P, blockIdx.x, sm, sm + ORG_MB * ORG_LS, (unsigned char *)(sm + ORG_MB * ORG_LS + 128), phi,
                      rho, lum, ecm, xc, yc, scratch, rho_box, vxbox, vybox, write_v);
    } else {
        float *phase = (float *)lds;
        double *red = (double *)(lds + 2 * ORG_WH * ORG_WH * 4);
        cells_body(P, blockIdx.x - P.nc, phase, phase + ORG_WH * ORG_WH, red, phi, rho, hfunc, psi, hlum, hecm, f0,
                   xc, yc, phin_win, phi_win, jwin, counts, err, 3);
    }
}

/* rho <- rhon on the FFT box and phi's new window (copy_cell_tables_1_by_1,
 * useful_functions.cpp:679-693, run_with_active_gels.cpp:172-178). */
static __device__ __forceinline__ void commit_cell_body(const DevP &P, const int n, float *__restrict__ rho,
                                                        const float *__restrict__ rho_box, float *__restrict__ phi,
                                                        const float *__restrict__ phi_win, const int *__restrict__ xc,
                                                        const int *__restrict__ yc) {
    const int N = P.N;
    constexpr int M = ORG_MB, W = ORG_W;
    const size_t NN = (size_t)N * N;
    float *rh = rho + n * NN;
    const int x0 = xc[n] - P.zs + 1, y0 = yc[n] - P.zs + 1;
    const float *rb = rho_box + (size_t)n * M * M;
#pragma unroll 4
    for (int e = threadIdx.x; e < M * M; e += blockDim.x)
        rh[(size_t)(x0 + (e >> 7)) * N + (y0 + (e & 127))] = rb[e];
    if (phi_win) {        /* phi window from k_cortex_cells (reshaping_cells' phase) */
        const int wx = x0 - 1, wy = y0 - 1;
        const float *pw = phi_win + (size_t)n * W * W;
        float *ph = phi + n * NN;
#pragma unroll 4
        for (int e = threadIdx.x; e < W * W; e += blockDim.x) {
            const int i = e / W;
            ph[(size_t)(wx + i) * N + (wy + e - i * W)] = pw[e];
        }
    }
}

__global__ __launch_bounds__(1024) void k_commit_cells(DevP P, float *__restrict__ rho,
                                                       const float *__restrict__ rho_box, float *__restrict__ phi,
                                                       const float *__restrict__ phi_win, const int *__restrict__ xc,
                                                       const int *__restrict__ yc) {
    commit_cell_body(P, blockIdx.x, rho, rho_box, phi, phi_win, xc, yc);
}

/* k_post: rho/phi copy-back of every cell (workgroups 0..ncell-1) and
 * lum <- lumn, ECM <- ECMn (the other workgroups) in one launch. */
__global__ __launch_bounds__(1024) void k_post(DevP P, float *__restrict__ rho, const float *__restrict__ rho_box,
                                               float *__restrict__ phi, const float *__restrict__ phi_win,
                                               const int *__restrict__ xc, const int *__restrict__ yc,
                                               float *__restrict__ lum, float *__restrict__ ecm,
                                               const float *__restrict__ bufL0, const float *__restrict__ bufL1,
                                               const float *__restrict__ bufE0, const float *__restrict__ bufE1,
                                               const LumState *__restrict__ lst, int *__restrict__ err,
                                               int *__restrict__ uni) {
    if ((int)blockIdx.x < P.nc)
        commit_cell_body(P, blockIdx.x, rho, rho_box, phi, phi_win, xc, yc);
    else
        commit_lumecm_body(P, blockIdx.x - P.nc, gridDim.x - P.nc, lum, ecm, bufL0, bufL1, bufE0, bufE1, lst, err,
                           uni);
}

/* k_ring_rho: periodic ghost ring of every rho plane (the
 * boundary_conditions_cells(rhon) of rho_functions.cpp:469, carried into rho
 * by the copy at run_with_active_gels.cpp:175-178).  Its sources, rows/cols
 * 1 and N-2, lie outside every FFT box (windows keep one pixel inside
 * rows/cols 1..N-2) and are never written, so the ring is the same after
 * every step: the library writes it once, after the first step following
 * org_prepare. */
__global__ __launch_bounds__(1024) void k_ring_rho(DevP P, float *__restrict__ rho) {
    const int n = blockIdx.x, N = P.N;
    float *rh = rho + (size_t)n * N * N;
    for (int i = 1 + threadIdx.x; i < N - 1; i += blockDim.x) {
        rh[(size_t)i * N] = rh[(size_t)i * N + N - 2];
        rh[(size_t)i * N + N - 1] = rh[(size_t)i * N + 1];
        rh[i] = rh[(size_t)N * (N - 2) + i];
        rh[(size_t)N * (N - 1) + i] = rh[(size_t)N + i];
    }
    if (threadIdx.x == 0) {
        rh[0] = rh[(size_t)N * (N - 2) + N - 2];
        rh[N - 1] = rh[(size_t)N * (N - 2) + 1];
        rh[(size_t)N * (N - 1)] = rh[(size_t)N + N - 2];
        rh[(size_t)N * (N - 1) + N - 1] = rh[(size_t)N + 1];
    }
}

/* Test hook: one 128 x 128 2D DFT through the same LDS code path. */
__global__ __launch_bounds__(1024) void k_fft_test(float2 *data, int sign) {
    __shared__ float2 sm[ORG_MB * ORG_LS];
    __shared__ float2 tw[128];
    const int t = threadIdx.x;
    if (t < 128) tw[t] = make_float2(kCos128[t], kSin128[t]);
    for (int e = t; e < ORG_MB * ORG_MB; e += 1024) sm[(e >> 7) * ORG_LS + (e & 127)] = data[e];
    __syncthreads();
    if (sign < 0) fft2d_128<-1>(sm, tw); else fft2d_128<1>(sm, tw);
    for (int e = t; e < ORG_MB * ORG_MB; e += 1024) data[e] = sm[(e >> 7) * ORG_LS + (e & 127)];
}
