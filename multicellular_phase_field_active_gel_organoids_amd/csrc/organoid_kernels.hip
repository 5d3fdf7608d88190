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
 * Per-function kernels (the reference's call sequence, all cells per launch):
 *   k_centroid      integer centroid + window check + support box update
 *   k_prep          hfunc on the window, FP64 area sum over the support box
 *   k_hlum          h(lum), h(ECM) full plane + FP64 sum of h(ECM)
 *   k_psi           psi on tiles covering the union of this step's windows
 *   k_spectral      per cell: active stress, 2 forward FFTs, the per-mode
 *                   friction-viscosity solve, Hermitian split, inverse FFT of
 *                   the velocity, rho*v, forward FFT, spectral divergence,
 *                   inverse FFT -> new rho on the FFT box  (5 2D FFTs of
 *                   128 x 128 in LDS instead of the reference's 10)
 *   k_lumecm_update lumen / ECM explicit step + Dirichlet edges
 *   k_lum_normals   |grad| of the old field, normals of the new one, periodic ring
 *   k_lum_iter      one reshaping sweep of lumen and ECM, FP64 deviation,
 *                   device-side stop test (last-block reduction)
 *   k_cells         per cell: phase-field update, normals, reshaping loop in
 *                   LDS until the deviation test passes, phi <- second-to-last
 *   k_commit_cells  rho box write-back + phi window;  k_commit_lumecm.
 * Batched step (two launches, replayable from one hipGraph):
 *   k_step          everything above up to the copy-back, workgroups in three
 *                   roles that signal each other through device counters
 *   k_post          the copy-back (k_commit_cells + k_commit_lumecm)
 *   k_ring_rho      rho's ghost ring, once after the first step
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <climits>
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

/* Load from a uniform base at a 32-bit byte offset: the global_load form
 * with the base in scalar registers and one vector register per offset
 * (plane[index] with a 64-bit index needs two per load, which limits how many
 * loads a loop can keep in flight). */
static __device__ __forceinline__ float ldb(const float *__restrict__ base, unsigned byte_off) {
    return *(const float *)((const char *)base + byte_off);
}

static __device__ __forceinline__ float2 ldb2(const float2 *__restrict__ base, unsigned byte_off) {
    return *(const float2 *)((const char *)base + byte_off);
}
/* four consecutive floats at a 4-byte-aligned offset: one global_load_dwordx4
 * (multi-dword global loads need dword alignment only) */
typedef float f4a __attribute__((ext_vector_type(4), aligned(4)));
static __device__ __forceinline__ f4a ldb4(const float *__restrict__ base, unsigned byte_off) {
    return *(const f4a *)((const char *)base + byte_off);
}

/* Hardware reciprocal, square root and 2^x (v_rcp_f32, v_sqrt_f32,
 * v_exp_f32, about 1 ulp each) for the per-pixel and per-mode arithmetic,
 * in place of the multi-instruction IEEE sequences of '/', sqrtf and expf:
 * those loops are bound by vector ALU issue, and the results stay within
 * FP32 rounding noise of the reference's float arithmetic. */
static __device__ __forceinline__ float hw_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
static __device__ __forceinline__ float hw_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
static __device__ __forceinline__ float hw_exp(float x) { return __builtin_amdgcn_exp2f(x * 1.44269504089f); }

/* ---- block reductions (blockDim multiple of 64) --------------------------- */
/* v of another lane of the same row of 16 (DPP control CTRL), by halves */
template <int CTRL>
static __device__ __forceinline__ double dpp_d(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xf, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
static __device__ __forceinline__ double readlane_d(double v, int lane) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, lane), hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
static __device__ __forceinline__ double readfirstlane_d(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readfirstlane((int)b), hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
/* Sum over the 64 lanes of a wave (every lane active), the same value in
 * every lane and a fixed order: pairs, quads, halves and whole rows of 16
 * through DPP lane moves (each step adds a lane's value to its partner's,
 * and a + b == b + a, so all lanes of a row hold the same row sum), then
 * (row0 + row1) + (row2 + row3).  No LDS traffic. */
static __device__ __forceinline__ double wave_sum_d(double v) {
    v += dpp_d<0xB1>(v);      /* quad_perm [1, 0, 3, 2] */
    v += dpp_d<0x4E>(v);      /* quad_perm [2, 3, 0, 1] */
    v += dpp_d<0x141>(v);     /* row_half_mirror */
    v += dpp_d<0x140>(v);     /* row_mirror */
    return (readlane_d(v, 0) + readlane_d(v, 16)) + (readlane_d(v, 32) + readlane_d(v, 48));
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

/* ---- synchronisation between workgroups of one launch ----------------------
 * Each XCD has its own L2, so making stores visible to a workgroup on another
 * XCD takes an agent-scope release (the writer's L2 is written back) and an
 * agent-scope acquire on the reader's side (its L2 is invalidated).  Both are
 * costly and disturb every workgroup resident on the XCD, so only the
 * lumen/ECM workgroups of k_step synchronise, through lum_barrier. */
static __device__ __forceinline__ void drain_stores() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_s_waitcnt(0);
}

/* Merge this thread's inclusive box (b1 < 0: empty) into the LDS box sb[0..3]:
 * min/max across the wave first, then one LDS atomic per wave (same-address
 * LDS atomics from every thread serialise). */
static __device__ __forceinline__ void box_to_lds(int b0, int b1, int b2, int b3, int *sb) {
    if (b1 < 0) { b0 = INT_MAX; b2 = INT_MAX; }
    for (int o = 32; o > 0; o >>= 1) {
        b0 = min(b0, __shfl_xor(b0, o, 64)); b1 = max(b1, __shfl_xor(b1, o, 64));
        b2 = min(b2, __shfl_xor(b2, o, 64)); b3 = max(b3, __shfl_xor(b3, o, 64));
    }
    if ((threadIdx.x & 63) == 0 && b1 >= 0) {
        atomicMin(&sb[0], b0); atomicMax(&sb[1], b1); atomicMin(&sb[2], b2); atomicMax(&sb[3], b3);
    }
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
/* a * exp(S * 2 pi i m / 16) for m = 0..7 known at compile time after
 * unrolling: the multiplications by 1, S i and (+-1 + S i)/sqrt(2) are
 * written out (a general complex product would keep the x*0 terms) */
template <int S>
static __device__ __forceinline__ float2 twmul16(float2 a, int m) {
    if (m == 0) return a;
    if (m == 4) return make_float2(-S * a.y, S * a.x);
    if (m == 2) return make_float2(C16_2 * (a.x - S * a.y), C16_2 * (S * a.x + a.y));
    if (m == 6) return make_float2(-C16_2 * (a.x + S * a.y), C16_2 * (S * a.x - a.y));
    return cmulf(a, w16<S>(m));
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
                float2 b = twmul16<S>(v[i + k + len / 2], k * (16 / len));
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
/* Position along its line of output q (0..15) of thread sub in the final
 * stage of fft128_lines: X[k2 + 16 k1] with k2 = 2 sub + q / 8, k1 = q % 8. */
static __device__ __forceinline__ int fft_out_pos(const int sub, const int q) {
    return 2 * sub + (q >> 3) + 16 * (q & 7);
}
/* The default last stage: the transformed lines back into the array. */
struct FftStore {
    float2 *sm;
    int es, ls;
    __device__ void operator()(const float2 *u, const int sub, const int line) const {
        float2 *base = sm + line * ls;
#pragma unroll
        for (int q = 0; q < 16; q++) base[fft_out_pos(sub, q) * es] = u[q];
    }
};
template <int S, class Epi>
static __device__ __forceinline__ void fft128_lines(float2 *sm, const float2 *tw, int es, int ls, const Epi &epi) {
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
    /* every read of this pass is done: the last stage may write anywhere */
    epi(u, sub, line);
    __syncthreads();
}

/* Unnormalised 2D DFT of the 128 x 128 array in LDS (row stride ORG_LS),
 * sign S as cufftExecC2C (useful_functions.cpp:650, 658).  The column pass
 * hands its outputs to epi(u, sub, col): u[q] is X[row][col] with row =
 * fft_out_pos(sub, q); the default stores them in place, a fused consumer
 * may instead use them directly (its own stores to LDS go to its own
 * outputs' places). */
template <int S, class Epi>
static __device__ __forceinline__ void fft2d_128(float2 *sm, const float2 *tw, const Epi &epi) {
    fft128_lines<S>(sm, tw, 1, ORG_LS, FftStore{sm, 1, ORG_LS});    /* rows */
    fft128_lines<S>(sm, tw, ORG_LS, 1, epi);                         /* columns */
}
template <int S>
static __device__ __forceinline__ void fft2d_128(float2 *sm, const float2 *tw) {
    fft2d_128<S>(sm, tw, FftStore{sm, ORG_LS, 1});
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
    const float rD = hw_rcp(D), rA = hw_rcp(A);
    y.x = (ex * (kx * pxx.y + ky * pxy.y) - A * (ky * pyy.y + kx * pxy.y)) * rD;
    y.y = (-ex * (kx * pxx.x + ky * pxy.x) + A * (ky * pyy.x + kx * pxy.x)) * rD;
    x.x = (kx * pxx.y + ky * pxy.y - ex * y.x) * rA;
    x.y = (-kx * pxx.x - ky * pxy.x - ex * y.y) * rA;
    vx = x; vy = y;
}

/* Parameter of a turnover label (0 apical, 1 basal, 2 and 3 lateral, which
 * share their values) from the uniform table a[]: a select chain instead of
 * a per-lane indexed load. */
static __device__ __forceinline__ float by_label(const float *a, int idx) {
    return idx == 0 ? a[0] : (idx == 1 ? a[1] : a[2]);
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

/* ---- centroid and area over a support box ---------------------------------
 * A box [i0,i1) x [j0,j1) is walked as one flat pixel range: pixel e sits at
 * (i0 + e / bw, j0 + e % bw) and thread t takes e = t, t + blockDim, ...;
 * the (row, column) pair advances incrementally, so only one division per
 * thread is needed and consecutive lanes read consecutive columns. */

/* Integer centroid sums of {phi > 0.5} over the box (useful_functions.cpp:
 * 193-227); totals valid in thread 0.  red: >= 48 long longs of LDS. */
static __device__ void centroid_sums(const float *__restrict__ ph, int N, int i0, int i1, int j0, int j1,
                                     long long *red, long long &X, long long &Y, long long &C) {
    const int t = threadIdx.x, lane = t & 63, w = t >> 6, nw = blockDim.x >> 6, nt = blockDim.x;
    const int bw = j1 - j0, bh = i1 - i0;
    const int npx = (bw > 0 && bh > 0) ? bw * bh : 0;
    long long sx = 0, sy = 0, c = 0;
    if (npx > 0) {
        const int di = nt / bw, dj = nt - di * bw;
        int i = t / bw, j = t - i * bw;
#pragma unroll 8
        for (int e = t; e < npx; e += nt) {
            if (ph[(size_t)(i0 + i) * N + (j0 + j)] > 0.5f) { sx += i0 + i; sy += j0 + j; c += 1; }
            i += di; j += dj;
            if (j >= bw) { j -= bw; i++; }
        }
    }
    sx = wave_sum_ll(sx); sy = wave_sum_ll(sy); c = wave_sum_ll(c);
    __syncthreads();
    if (lane == 0) { red[w] = sx; red[16 + w] = sy; red[32 + w] = c; }
    __syncthreads();
    X = 0; Y = 0; C = 0;
    if (t == 0)
        for (int k = 0; k < nw; k++) { X += red[k]; Y += red[16 + k]; C += red[32 + k]; }
}

/* centroid_sums with the pixels of the window [x0, x0+W) x [y0, y0+W) taken
 * from win (LDS, row stride W + 2, one-pixel ring) instead of ph. */
static __device__ void centroid_sums_win(const float *__restrict__ ph, int N, int i0, int i1, int j0, int j1,
                                         int x0, int y0, const float *win, long long *red, long long &X,
                                         long long &Y, long long &C) {
    constexpr int W = ORG_W, WH = ORG_WH;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6, nw = blockDim.x >> 6, nt = blockDim.x;
    const int bw = j1 - j0, bh = i1 - i0;
    const int npx = (bw > 0 && bh > 0) ? bw * bh : 0;
    long long sx = 0, sy = 0, c = 0;
    if (npx > 0) {
        /* 8 pixels per thread at a time: the plane loads only outside the
         * window (the support box normally lies inside it, so most waves
         * issue none) and issued before any use; window pixels take the
         * LDS value */
        const int di = nt / bw, dj = nt - di * bw;
        int i = t / bw, j = t - i * bw;
#pragma unroll 1
        for (int e0 = t; e0 < npx; e0 += 8 * nt) {
            int pi[8], pj[8];
            bool in[8];
            float vg[8];
#pragma unroll
            for (int k = 0; k < 8; k++) {
                pi[k] = i0 + i; pj[k] = j0 + j;
                const int wi = pi[k] - x0, wj = pj[k] - y0;
                in[k] = wi >= 0 && wi < W && wj >= 0 && wj < W;
                vg[k] = 0.0f;
                if (!in[k] && e0 + k * nt < npx)
                    vg[k] = ldb(ph, 4u * ((unsigned)pi[k] * (unsigned)N + (unsigned)pj[k]));
                i += di; j += dj;
                if (j >= bw) { j -= bw; i++; }
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int k = 0; k < 8; k++) {
                if (e0 + k * nt >= npx) break;
                const int wi = pi[k] - x0, wj = pj[k] - y0;
                const float v = in[k] ? win[(wi + 1) * WH + (wj + 1)] : vg[k];
                if (v > 0.5f) { sx += pi[k]; sy += pj[k]; c += 1; }
            }
        }
    }
    sx = wave_sum_ll(sx); sy = wave_sum_ll(sy); c = wave_sum_ll(c);
    __syncthreads();
    if (lane == 0) { red[w] = sx; red[16 + w] = sy; red[32 + w] = c; }
    __syncthreads();
    X = 0; Y = 0; C = 0;
    if (t == 0)
        for (int k = 0; k < nw; k++) { X += red[k]; Y += red[16 + k]; C += red[32 + k]; }
}

/* Centre of cell n from the sums (integer division as the reference's
 * int(sum/count), useful_functions.cpp:221-222; the previous centre when the
 * cell is empty), the window check, and the support box (i0,i1,j0,j1) grown
 * by the window, stored to xo, yo, bo.  Thread 0 only. */
static __device__ void place_window(const DevP &P, int n, long long X, long long Y, long long C, int xprev,
                                    int yprev, int i0, int i1, int j0, int j1, int *__restrict__ xo,
                                    int *__restrict__ yo, int *__restrict__ bo, int *__restrict__ err) {
    const int N = P.N;
    int x = xprev, y = yprev;
    if (C == 0) set_err(err, ORG_EEMPTY_CELL);
    else { x = (int)(X / C); y = (int)(Y / C); }
    /* window [x-zs, x+zs) plus one halo pixel inside rows/cols 1..N-2 */
    const int lo = P.zs + 2, hi = N - 2 - P.zs;
    if (x < lo || x > hi || y < lo || y > hi) {
        set_err(err, ORG_EWINDOW);
        x = min(max(x, lo), hi);
        y = min(max(y, lo), hi);
    }
    xo[n] = x; yo[n] = y;
    int b0, b1, b2, b3;
    if (i1 <= i0 || j1 <= j0) { b0 = x - P.zs; b1 = x + P.zs; b2 = y - P.zs; b3 = y + P.zs; }
    else { b0 = min(i0, x - P.zs); b1 = max(i1, x + P.zs); b2 = min(j0, y - P.zs); b3 = max(j1, y + P.zs); }
    bo[4 * n + 0] = b0; bo[4 * n + 1] = b1; bo[4 * n + 2] = b2; bo[4 * n + 3] = b3;
}

/* Centroid of cell n over its support box box_in (find_centres_cells,
 * useful_functions.cpp:193-227), placed by place_window into xo, yo, bo.
 * red: >= 48 long longs of LDS. */
static __device__ void cell_centroid(const DevP &P, int n, const float *__restrict__ phi,
                                     const int *__restrict__ xin, const int *__restrict__ yin,
                                     const int *__restrict__ box_in, int *__restrict__ xo, int *__restrict__ yo,
                                     int *__restrict__ bo, int *__restrict__ err, long long *red) {
    const int i0 = box_in[4 * n], i1 = box_in[4 * n + 1], j0 = box_in[4 * n + 2], j1 = box_in[4 * n + 3];
    const int xp = xin[n], yp = yin[n];
    long long X, Y, C;
    centroid_sums(phi + n * (size_t)P.N * P.N, P.N, i0, i1, j0, j1, red, X, Y, C);
    if (threadIdx.x == 0) place_window(P, n, X, Y, C, xp, yp, i0, i1, j0, j1, xo, yo, bo, err);
}

/* This thread's share of s_n = sum over the box of h*dx*dx in FP64
 * (sum_s_kernels, useful_functions.cpp:168-190).  With win, h = hfun(phi)
 * inside the window [x0, x0+W) x [y0, y0+W) and is stored to hfunc there
 * (compute_hfunc_3d_tables_kernels_loop_n, useful_functions.cpp:233-251);
 * elsewhere hfunc is read.  Loads are issued 8 pixels at a time. */
static __device__ double area_sum(const DevP &P, const float *__restrict__ ph, float *__restrict__ hf, int i0,
                                  int i1, int j0, int j1, bool win, int x0, int y0) {
    constexpr int B = 8, W = ORG_W;
    const int t = threadIdx.x, N = P.N, nt = blockDim.x;
    const int bw = j1 - j0, bh = i1 - i0;
    const int npx = (bw > 0 && bh > 0) ? bw * bh : 0;
    double acc = 0.0;
    if (npx == 0) return acc;
    const int di = nt / bw, dj = nt - di * bw;
    int i = t / bw, j = t - i * bw;
    for (int e0 = t; e0 < npx; e0 += B * nt) {
        int off[B];
        bool in[B];
        float v[B];
#pragma unroll
        for (int k = 0; k < B; k++) {
            const int gi = i0 + i, gj = j0 + j;
            off[k] = gi * N + gj;
            in[k] = win && gi >= x0 && gi < x0 + W && gj >= y0 && gj < y0 + W;
            v[k] = (e0 + k * nt < npx) ? (in[k] ? ph : hf)[off[k]] : 0.0f;
            i += di; j += dj;
            if (j >= bw) { j -= bw; i++; }
        }
#pragma unroll
        for (int k = 0; k < B; k++) {
            if (e0 + k * nt < npx) {
                const float h = in[k] ? hfun(v[k]) : v[k];
                if (in[k]) hf[off[k]] = h;
                acc += (double)(h * P.dx * P.dx);
            }
        }
    }
    return acc;
}

/* k_centroid: find_centres_cells (useful_functions.cpp:193-227) over the
 * support box, window check, support box += window; the result goes to
 * (xo, yo, bo), which may be the inputs.  Block 0 also resets the
 * union-of-windows box that k_prep fills and k_psi reads. */
__global__ __launch_bounds__(1024) void k_centroid(DevP P, const float *__restrict__ phi, const int *xin,
                                                   const int *yin, const int *box_in, int *xo, int *yo, int *bo,
                                                   int *__restrict__ uni, int *__restrict__ err) {
    __shared__ long long red[48];
    if (blockIdx.x == 0 && threadIdx.x == 0) { uni[0] = P.N; uni[1] = 0; uni[2] = P.N; uni[3] = 0; }
    cell_centroid(P, blockIdx.x, phi, xin, yin, box_in, xo, yo, bo, err, red);
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
        /* the window's hfunc is read back: same values as area_sum's
         * in-register hfun in the batched form */
        __syncthreads();
        const double acc = area_sum(P, ph, hf, box[4 * n], box[4 * n + 1], box[4 * n + 2], box[4 * n + 3], false,
                                    x0, y0);
        const double tot = block_sum_d(acc, red);
        if (threadIdx.x == 0) {
            s[n] = tot;
            f0[n] = P.alpha * ((float)s0[n] - (float)tot);
        }
    }
}

/* hfunc on the window and area of cell n (the batched form of k_prep); the
 * centre and support box come from the previous step's copy-back (xn, yn,
 * bn) and are published to xc, yc, box.  red: >= 16 doubles of LDS. */
static __device__ void cell_prep(const DevP &P, const int n, const float *__restrict__ phi,
                                 float *__restrict__ hfunc, const int *__restrict__ xn, const int *__restrict__ yn,
                                 const int *__restrict__ bn, int *__restrict__ xc, int *__restrict__ yc,
                                 int *__restrict__ box, const double *__restrict__ s0, double *__restrict__ s,
                                 float *__restrict__ f0, double *red) {
    const size_t NN = (size_t)P.N * P.N;
    const int x = xn[n], y = yn[n];
    const int b0 = bn[4 * n], b1 = bn[4 * n + 1], b2 = bn[4 * n + 2], b3 = bn[4 * n + 3];
    if (threadIdx.x == 0) {
        xc[n] = x; yc[n] = y;
        box[4 * n] = b0; box[4 * n + 1] = b1; box[4 * n + 2] = b2; box[4 * n + 3] = b3;
    }
    const double acc = area_sum(P, phi + n * NN, hfunc + n * NN, b0, b1, b2, b3, true, x - P.zs, y - P.zs);
    const double tot = block_sum_d(acc, red);
    if (threadIdx.x == 0) {
        s[n] = tot;
        f0[n] = P.alpha * ((float)s0[n] - (float)tot);
    }
}

/* Sum of per-workgroup partials in a fixed order, valid in every thread
 * (nblk <= blockDim). */
static __device__ __forceinline__ double sum_parts(const double *part, int nblk, double *red) {
    double v = 0.0;
    for (int b = threadIdx.x; b < nblk; b += blockDim.x)
        v += __hip_atomic_load(&part[b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return block_sum_d(v, red);
}

/* ---- h(lum), h(ECM) and the FP64 sum of h(ECM) ---------------------------
 * compute_all_hfunc_2d_tables_kernels (useful_functions.cpp:272-285) and the
 * sum of h(ECM) dx^2 of update_lumen_and_ECM (cell_phase_functions.cpp:
 * 376-386).  The sum is split into fixed parts: one per 64 x 64 tile of the
 * grid (its interior pixels, tile k row-major over ceil(N/64)^2 tiles) and
 * one for the boundary ring (part ntiles); each part is a fixed-order block
 * reduction and the total adds the parts in index order, so a part can be
 * recomputed alone (when its pixels may have changed) and the total stays
 * the same bit for bit. */
#define ORG_HT 64
static __device__ __forceinline__ int hl_ntiles(int N) { const int t = (N + ORG_HT - 1) / ORG_HT; return t * t; }

/* ring pixel k = 0 .. 4N-5: rows 0 and N-1, then columns 0 and N-1 */
static __device__ __forceinline__ void ring_px(int k, int N, int &i, int &j) {
    if (k < N) { i = 0; j = k; }
    else if (k < 2 * N) { i = N - 1; j = k - N; }
    else if (k < 3 * N - 2) { i = k - 2 * N + 1; j = 0; }
    else { i = k - (3 * N - 2) + 1; j = N - 1; }
}

/* part k (tile, or the ring for k == ntiles): stores h(lum), h(ECM) on its
 * pixels, returns this thread's share of its sum. */
static __device__ double hlum_part(const DevP &P, const int k, const float *__restrict__ lum,
                                   const float *__restrict__ ecm, float *__restrict__ hlum,
                                   float *__restrict__ hecm) {
    const int N = P.N;
    double acc = 0.0;
    if (k == hl_ntiles(N)) {
        for (int q = threadIdx.x; q < 4 * N - 4; q += blockDim.x) {
            int i, j;
            ring_px(q, N, i, j);
            const size_t g = (size_t)i * N + j;
            const float a = lum[g], b = ecm[g], hb = hfun(b);
            hlum[g] = hfun(a);
            hecm[g] = hb;
            acc += (double)(hb * P.dx * P.dx);
        }
        return acc;
    }
    const int nt = (N + ORG_HT - 1) / ORG_HT;
    const int ti = (k / nt) * ORG_HT, tj = (k % nt) * ORG_HT;
#pragma unroll 4
    for (int e = threadIdx.x; e < ORG_HT * ORG_HT; e += blockDim.x) {
        const int i = ti + e / ORG_HT, j = tj + e % ORG_HT;
        if (i >= 1 && i < N - 1 && j >= 1 && j < N - 1) {
            const size_t g = (size_t)i * N + j;
            const float a = lum[g], b = ecm[g], hb = hfun(b);
            hlum[g] = hfun(a);
            hecm[g] = hb;
            acc += (double)(hb * P.dx * P.dx);
        }
    }
    return acc;
}

/* The 9-point Laplacian numerator of update_all_cells (cell_phase_functions.
 * cpp:198) as the reference's FP32 build evaluates it: the double literals
 * -12.0 and 2.0 promote everything but the sum of the four edge neighbours to
 * double.  t: the 3 x 3 stencil (index 0 the centre, 1-4 the edges i+1, i-1,
 * j+1, j-1, 5-8 the corners i+1 j+1, i+1 j-1, i-1 j-1, i-1 j+1). */
static __device__ __forceinline__ double l9d(const float *t) {
    return -12.0 * (double)t[0] + 2.0 * (double)(t[1] + t[2] + t[3] + t[4]) + (double)t[5] + (double)t[6] +
           (double)t[7] + (double)t[8];
}

/* The lumen/ECM part of update_all_cells' bracket (cell_phase_functions.cpp:
 * 202-210) at one pixel, the same for every cell, from the 3 x 3 stencils of
 * h(lum) and h(ECM), in double as the reference's FP32 build forms it
 * (inv4 = 1 / (4 dx^2)):
 *   -beta_L h(L) + eta_L L9(h(L)) - beta_E h(E) + eta_E L9(h(E)).
 * The bracket's terms cancel to a small fraction of their size; formed in
 * float, their rounding doubles how fast the phase fields separate from the
 * FP64 solution over 10^4 steps. */
static __device__ __forceinline__ double lumecm_bracket(const DevP &P, const float *l, const float *e,
                                                        const double inv4) {
    return -(double)(P.beta_lc * l[0]) + (double)P.eta_lc * l9d(l) * inv4 - (double)(P.beta_ec * e[0]) +
           (double)P.eta_ec * l9d(e) * inv4;
}

/* The lumen/ECM bracket of part k of the h pass into lbrk, from the fields
 * lsrc, esrc (h = hfun of them, the values the h planes hold): tile k's
 * pixels in rows/cols 1..N-2, with h of the tile grown by one staged in LDS
 * (lds: ORG_LDS_LBRK bytes); the ring parts (k = number of tiles + s, s <
 * ring_parts): slice s of the first interior rows and columns, whose
 * stencils reach the ring (along a column every lane reads its own cache
 * lines, so the ring is split to spread that over workgroups).  The bracket
 * changes only where lum or ECM changed or next to such a pixel. */
#define ORG_LDS_LBRK (2 * (ORG_HT + 2) * (ORG_HT + 2) * 4)
#define ORG_RING_PARTS 16
static __device__ void lbrk_part(const DevP &P, const int k, const float *__restrict__ lsrc,
                                 const float *__restrict__ esrc, double *__restrict__ lbrk, float *lds,
                                 const int ring_parts = 1) {
    const int N = P.N;
    const double inv4 = 1.0 / (4.0 * (double)P.dx * (double)P.dx);
    if (k >= hl_ntiles(N)) {
        /* two pixels per thread and batch, their 36 loads issued before any
         * use (a slot past the slice loads its last pixel and is not stored) */
        const int nr = 2 * (N - 2), nall = nr + 2 * (N - 4), sl = k - hl_ntiles(N);
        const int qa = (int)(((long long)sl * nall) / ring_parts), nq = (int)(((long long)(sl + 1) * nall) / ring_parts);
        if (nq <= qa) return;
        for (int q0 = qa + threadIdx.x; q0 < nq; q0 += 2 * blockDim.x) {
            float l[2][9], e[2][9];
            size_t gs[2];
#pragma unroll
            for (int b = 0; b < 2; b++) {
                const int q = min(q0 + b * (int)blockDim.x, nq - 1);
                int i, j;
                if (q < nr) { i = (q < N - 2) ? 1 : N - 2; j = 1 + (q < N - 2 ? q : q - (N - 2)); }
                else { const int w = q - nr; j = (w < N - 4) ? 1 : N - 2; i = 2 + (w < N - 4 ? w : w - (N - 4)); }
                const size_t g = (size_t)i * N + j;
                gs[b] = g;
                const size_t o[9] = {g, g + N, g - N, g + 1, g - 1, g + N + 1, g + N - 1, g - N - 1, g - N + 1};
#pragma unroll
                for (int m = 0; m < 9; m++) { l[b][m] = lsrc[o[m]]; e[b][m] = esrc[o[m]]; }
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int b = 0; b < 2; b++) {
                if (q0 + b * (int)blockDim.x >= nq) break;
#pragma unroll
                for (int m = 0; m < 9; m++) { l[b][m] = hfun(l[b][m]); e[b][m] = hfun(e[b][m]); }
                lbrk[gs[b]] = lumecm_bracket(P, l[b], e[b], inv4);
            }
        }
        return;
    }
    constexpr int TW = ORG_HT + 2;
    float *hl = lds, *he = lds + TW * TW;
    const int nt = (N + ORG_HT - 1) / ORG_HT;
    const int ti = (k / nt) * ORG_HT, tj = (k % nt) * ORG_HT;
    /* h of the tile grown by one into LDS, up to SQ pixels per thread and
     * batch with their loads issued before any use (a slot off the tile or
     * the grid loads element 0 and is not stored) */
    constexpr int SQ = 5;
    for (int q0 = threadIdx.x; q0 < TW * TW; q0 += SQ * blockDim.x) {
        float lv[SQ], ev[SQ];
        bool ok[SQ];
#pragma unroll
        for (int b = 0; b < SQ; b++) {
            const int q = q0 + b * (int)blockDim.x, i = ti - 1 + q / TW, j = tj - 1 + q % TW;
            ok[b] = q < TW * TW && i >= 0 && i < N && j >= 0 && j < N;
            const size_t g = ok[b] ? (size_t)i * N + j : 0;
            lv[b] = lsrc[g];
            ev[b] = esrc[g];
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int b = 0; b < SQ; b++)
            if (ok[b]) {
                hl[q0 + b * blockDim.x] = hfun(lv[b]);
                he[q0 + b * blockDim.x] = hfun(ev[b]);
            }
    }
    __syncthreads();
#pragma unroll 2
    for (int e0 = threadIdx.x; e0 < ORG_HT * ORG_HT; e0 += blockDim.x) {
        const int r = e0 / ORG_HT, c = e0 % ORG_HT, i = ti + r, j = tj + c;
        if (i >= 1 && i <= N - 2 && j >= 1 && j <= N - 2) {
            const int q = (r + 1) * TW + (c + 1);
            const int o[9] = {q, q + TW, q - TW, q + 1, q - 1, q + TW + 1, q + TW - 1, q - TW - 1, q - TW + 1};
            float l[9], e[9];
#pragma unroll
            for (int m = 0; m < 9; m++) { l[m] = hl[o[m]]; e[m] = he[o[m]]; }
            lbrk[(size_t)i * N + j] = lumecm_bracket(P, l, e, inv4);
        }
    }
    __syncthreads();      /* the next part reuses the LDS */
}

/* k_hlum: every part (grid = ntiles + 1), the total reduced by the last
 * block to finish (fixed order: deterministic). */
__global__ __launch_bounds__(1024) void k_hlum(DevP P, const float *__restrict__ lum, const float *__restrict__ ecm,
                                               float *__restrict__ hlum, float *__restrict__ hecm,
                                               double *__restrict__ part, unsigned *__restrict__ counter,
                                               double *__restrict__ sum_out) {
    __shared__ double red[16];
    __shared__ bool last;
    const int blk = blockIdx.x, nblk = gridDim.x;
    const double tot = block_sum_d(hlum_part(P, blk, lum, ecm, hlum, hecm), red);
    if (threadIdx.x == 0) {
        part[blk] = tot;
        __threadfence();
        unsigned prev = atomicAdd(counter, 1u);
        last = (prev == (unsigned)nblk - 1);
    }
    __syncthreads();
    if (last) {
        __threadfence();
        const double t2 = sum_parts(part, nblk, red);
        if (threadIdx.x == 0) { *sum_out = t2; *counter = 0u; }
    }
}

/* ---- live boxes of the lumen and ECM fields --------------------------------
 * A pixel of a committed field X is live if X is not exactly 0 or 1 there, or
 * differs from one of its four neighbours.  Elsewhere the update leaves X
 * unchanged (the update is X + c X (1 - X) (...)) and both central
 * differences of X vanish, so the normals are zero.  live[4f..4f+3] is a
 * half-open box (rows i0..i1, columns j0..j1) holding every pixel that has
 * been live since the lumen/ECM state was last initialised (it only grows).
 * With the box inside [3, N-3)^2 (box mode) the update, the normals and the
 * copy-back of field f touch only the box grown by one pixel plus the
 * boundary ring, and the ring's normals are zero (their periodic sources,
 * rows/columns 1 and N-2, are not live); otherwise every pixel is processed
 * (full mode).  Invariant kept between steps: outside the live box grown by
 * one and the ring, both ping-pong buffers hold the committed field. */
static __device__ __forceinline__ int4 ld_box(const int *live, int f) {
    return make_int4(live[4 * f], live[4 * f + 1], live[4 * f + 2], live[4 * f + 3]);
}
static __device__ __forceinline__ bool box_empty(const int4 b) { return b.y <= b.x || b.w <= b.z; }
static __device__ __forceinline__ bool box_mode(const int4 b, const int N) {
    return box_empty(b) || (b.x >= 3 && b.y <= N - 3 && b.z >= 3 && b.w <= N - 3);
}
/* is the pixel e (interior) of the committed field X live */
static __device__ __forceinline__ bool is_live(const float *__restrict__ X, const size_t e, const int N) {
    const float v = X[e];
    return (v != 0.0f && v != 1.0f) || X[e + N] != v || X[e - N] != v || X[e + 1] != v || X[e - 1] != v;
}

/* psi = sum_n hfunc_n (compute_psi_kernels, useful_functions.cpp:305-322) on
 * one 16 x 16 tile per group of 256 threads (ti0 < 0: the group has no tile
 * and only joins the barriers).  The cells whose support box meets the tile
 * are listed 64 at a time in cell order by the group's first wave, so the
 * FP32 sum takes its operands in cell order, as the reference's loop over n
 * does.  With phi != nullptr, hfunc_n inside cell n's window (origin from
 * xc, yc) is taken as hfun(phi_n) (the value the same step's prep stores
 * there), so psi needs no wait for the prep; with phi_win != nullptr as well,
 * phi_n inside its previous window (origin xo, yo) is read from phi_win (the
 * new window not yet copied back).  Per group: a 64-entry list in LDS
 * (ORG_PSI_LDS_GROUP bytes).  Every thread of the workgroup calls this
 * together (it synchronises the workgroup). */
#define ORG_PSI_LDS_GROUP (64 * 4 + 64 * 16 + 64 * 16 + 16)
static __device__ void psi_tile(const DevP &P, const float *__restrict__ hfunc, const float *__restrict__ phi,
                                const int *__restrict__ box, const int *__restrict__ xc, const int *__restrict__ yc,
                                const float *__restrict__ phi_win, const int *__restrict__ xo,
                                const int *__restrict__ yo, const int4 u, const int ti0, const int tj0,
                                unsigned char *lq, float *__restrict__ psi) {
    constexpr int W = ORG_W;
    int *lcell = (int *)lq;
    int4 *lbox = (int4 *)(lq + 64 * 4);
    int4 *lwin = (int4 *)(lq + 64 * 4 + 64 * 16);
    int *lcnt = (int *)(lq + 64 * 4 + 64 * 16 + 64 * 16);
    const int gt = threadIdx.x & 255, N = P.N;
    const size_t NN = (size_t)N * N;
    const int i = ti0 + (gt >> 4), j = tj0 + (gt & 15);
    const bool act = ti0 >= 0 && i < N && j < N && i >= u.x && i < u.y && j >= u.z && j < u.w;
    float acc = 0.0f;
    for (int c0 = 0; c0 < P.nc; c0 += 64) {
        if (gt < 64) {
            const int n = c0 + gt;
            bool hit = false;
            int4 bx = make_int4(0, 0, 0, 0);
            if (ti0 >= 0 && n < P.nc) {
                bx = make_int4(box[4 * n], box[4 * n + 1], box[4 * n + 2], box[4 * n + 3]);
                hit = !(ti0 >= bx.y || ti0 + 16 <= bx.x || tj0 >= bx.w || tj0 + 16 <= bx.z);
            }
            const unsigned long long m = __ballot(hit);
            if (hit) {
                const int pos = __popcll(m & ((1ull << gt) - 1ull));
                lcell[pos] = n;
                lbox[pos] = bx;
                lwin[pos] = make_int4(phi ? xc[n] - P.zs : -1, phi ? yc[n] - P.zs : -1,
                                      phi_win ? xo[n] - P.zs : -1, phi_win ? yo[n] - P.zs : -1);
            }
            if (gt == 0) *lcnt = __popcll(m);
        }
        __syncthreads();
        if (act) {
            /* 4 listed cells at a time: loads unconditional (every listed
             * plane holds the pixel), a cell whose box misses the pixel
             * adds +0, which leaves the sum unchanged bit for bit */
            const int cnt = *lcnt;
            const size_t e = (size_t)i * N + j;
            for (int k0 = 0; k0 < cnt; k0 += 4) {
                float x[4];
                bool cov[4], inw[4];
#pragma unroll
                for (int kk = 0; kk < 4; kk++) {
                    const int k = min(k0 + kk, cnt - 1);
                    const int4 bx = lbox[k];
                    const int4 w = lwin[k];
                    cov[kk] = k0 + kk < cnt && i >= bx.x && i < bx.y && j >= bx.z && j < bx.w;
                    inw[kk] = w.x >= 0 && i >= w.x && i < w.x + W && j >= w.y && j < w.y + W;
                    const bool ino = w.z >= 0 && i >= w.z && i < w.z + W && j >= w.w && j < w.w + W;
                    const bool fw = inw[kk] && ino;
                    const float *pb = fw ? phi_win : (inw[kk] ? phi : hfunc);
                    const size_t po = fw ? (size_t)lcell[k] * (W * W) + (size_t)((i - w.z) * W + (j - w.w))
                                         : (size_t)lcell[k] * NN + e;
                    x[kk] = pb[po];
                }
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int kk = 0; kk < 4; kk++) acc += cov[kk] ? (inw[kk] ? hfun(x[kk]) : x[kk]) : 0.0f;
            }
        }
        __syncthreads();
    }
    if (act) psi[(size_t)i * N + j] = acc;
}

/* k_psi: psi on the 16 x 16 tiles of the grid that meet the union of this
 * step's windows (psi changes nowhere else). */
__global__ __launch_bounds__(256) void k_psi(DevP P, const float *__restrict__ hfunc, const int *__restrict__ box,
                                             const int *__restrict__ uni, float *__restrict__ psi) {
    __shared__ __attribute__((aligned(16))) unsigned char lq[ORG_PSI_LDS_GROUP];
    const int ntx = (P.N + 15) / 16;
    const int ti0 = (blockIdx.x / ntx) * 16, tj0 = (blockIdx.x % ntx) * 16;
    const int4 u = make_int4(uni[0], uni[1], uni[2], uni[3]);
    if (ti0 >= u.y || ti0 + 16 <= u.x || tj0 >= u.w || tj0 + 16 <= u.z) return;
    psi_tile(P, hfunc, nullptr, box, nullptr, nullptr, nullptr, nullptr, nullptr, u, ti0, tj0, lq, psi);
}

/* psi of the batched step: the union of the windows from the centres xc, yc,
 * then every tile meeting it, dealt round-robin to the 256-thread groups of
 * workgroups blk = 0..nblk-1 (1024 threads each), with hfun(phi) inside each
 * cell's window (phi_win, xo, yo: see psi_tile).  scr: 4 * ORG_PSI_LDS_GROUP
 * + 16 bytes of LDS. */
static __device__ void psi_tiles_step(const DevP &P, const int blk, const int nblk, const float *__restrict__ hfunc,
                                      const float *__restrict__ phi, const int *__restrict__ box,
                                      const int *__restrict__ xc, const int *__restrict__ yc,
                                      const float *__restrict__ phi_win, const int *__restrict__ xo,
                                      const int *__restrict__ yo, float *__restrict__ psi, unsigned char *scr) {
    int *su = (int *)(scr + 4 * ORG_PSI_LDS_GROUP);
    const int g = threadIdx.x >> 8, ng = blockDim.x >> 8;
    if (threadIdx.x < 64) {           /* union of the windows, wave 0 */
        int a0 = P.N, a1 = 0, a2 = P.N, a3 = 0;
        for (int n = threadIdx.x; n < P.nc; n += 64) {
            const int x0 = xc[n] - P.zs, y0 = yc[n] - P.zs;
            a0 = min(a0, x0); a1 = max(a1, x0 + ORG_W); a2 = min(a2, y0); a3 = max(a3, y0 + ORG_W);
        }
        for (int o = 32; o > 0; o >>= 1) {
            a0 = min(a0, __shfl_xor(a0, o, 64)); a1 = max(a1, __shfl_xor(a1, o, 64));
            a2 = min(a2, __shfl_xor(a2, o, 64)); a3 = max(a3, __shfl_xor(a3, o, 64));
        }
        if (threadIdx.x == 0) { su[0] = a0; su[1] = a1; su[2] = a2; su[3] = a3; }
    }
    __syncthreads();
    const int4 u = make_int4(su[0], su[1], su[2], su[3]);
    if (u.y <= u.x || u.w <= u.z) return;
    const int tr0 = u.x >> 4, tc0 = u.z >> 4;
    const int ntc = ((u.w - 1) >> 4) - tc0 + 1, ntile = (((u.y - 1) >> 4) - tr0 + 1) * ntc;
    for (int base = blk * ng; base < ntile; base += nblk * ng) {
        const int k = base + g;
        int ti0 = -1, tj0 = -1;
        if (k < ntile) { ti0 = (tr0 + k / ntc) * 16; tj0 = (tc0 + k % ntc) * 16; }
        psi_tile(P, hfunc, phi, box, xc, yc, phi_win, xo, yo, u, ti0, tj0, scr + g * ORG_PSI_LDS_GROUP, psi);
    }
}

/* 9-point Laplacian numerator, cell_phase_functions.cpp:198 and :416 */
#define L9AT(t, g, N) (-12.0f * (t)[g] + 2.0f * ((t)[(g) + (N)] + (t)[(g) - (N)] + (t)[(g) + 1] + (t)[(g) - 1]) \
                       + (t)[(g) + (N) + 1] + (t)[(g) + (N) - 1] + (t)[(g) - (N) - 1] + (t)[(g) - (N) + 1])

/* lumen / ECM reshaping loop state: stop flag, sweeps done, buffer holding
 * the last iterate */
struct LumState { int done, ri, cur, pad; };

/* update_lumen_and_ECM (cell_phase_functions.cpp:395-443) at an interior
 * pixel e: the new lumen (f = 0) or ECM (f = 1) value. */
static __device__ __forceinline__ float lumecm_px(const DevP &P, const int f, const size_t e, const float se,
                                                  const float *__restrict__ lum, const float *__restrict__ ecm,
                                                  const float *__restrict__ psi, const float *__restrict__ hlum,
                                                  const float *__restrict__ hecm) {
    const int N = P.N;
    const float inv4 = 1.0f / (4.0f * P.dx * P.dx);
    const float ps = psi[e];
    const float lap = L9AT(psi, e, (size_t)N) * inv4;
    if (f == 0) {
        const float l = lum[e];
        return l + (1.0f / P.tau_lum) * l * (1.0f - l) * P.dt *
                       (P.xi_lum - P.beta_lc * ps + P.eta_lc * lap - P.beta_le * hecm[e]);
    }
    const float ecm_area = (float)(N * N) * P.dx * P.dx - se;
    const float E = ecm[e];
    return E + (1.0f / P.tau_E) * E * (1.0f - E) * P.dt *
                   (P.alpha_E * ecm_area + P.xi_E - P.beta_ec * ps + P.eta_ec * lap - P.beta_le * hlum[e]);
}

/* update_lumen_and_ECM on row i, then boundary_conditions_lumen_and_ECM
 * (useful_functions.cpp:439-462: lumen edges and corners 0, ECM 1); the
 * first iterate of both reshaping loops goes to both ping-pong buffers. */
static __device__ __forceinline__ void lumecm_row(const DevP &P, const int i, const float se,
                                                  const float *__restrict__ lum, const float *__restrict__ ecm,
                                                  const float *__restrict__ psi, const float *__restrict__ hlum,
                                                  const float *__restrict__ hecm, float *__restrict__ lumA,
                                                  float *__restrict__ lumB, float *__restrict__ ecmA,
                                                  float *__restrict__ ecmB) {
    const int N = P.N;
    const size_t r = (size_t)i * N;
    for (int j = threadIdx.x; j < N; j += blockDim.x) {
        const size_t e = r + j;
        float ln, en;
        if (i == 0 || j == 0 || i == N - 1 || j == N - 1) {
            ln = 0.0f;
            en = 1.0f;
        } else {
            ln = lumecm_px(P, 0, e, se, lum, ecm, psi, hlum, hecm);
            en = lumecm_px(P, 1, e, se, lum, ecm, psi, hlum, hecm);
        }
        lumA[e] = ln; lumB[e] = ln;
        ecmA[e] = en; ecmB[e] = en;
    }
}

/* Loop states of both fields at the start of the reshaping loops. */
static __device__ __forceinline__ void lum_reset_states(const DevP &P, LumState *__restrict__ lst) {
    if (threadIdx.x < 2) {
        const bool run = P.rdev0_lum > P.thr_lum;
        lst[threadIdx.x].done = run ? 0 : 1;
        lst[threadIdx.x].ri = 0;
        lst[threadIdx.x].cur = 0;
    }
}

/* k_lumecm_update: the lumen / ECM update of every row; resets the loop
 * states, the loop barrier and the abort flag. */
__global__ __launch_bounds__(1024) void k_lumecm_update(DevP P, const float *__restrict__ lum,
                                                       const float *__restrict__ ecm, const float *__restrict__ psi,
                                                       const float *__restrict__ hlum, const float *__restrict__ hecm,
                                                       const double *__restrict__ sum_ecm, float *__restrict__ lumA,
                                                       float *__restrict__ lumB, float *__restrict__ ecmA,
                                                       float *__restrict__ ecmB, LumState *__restrict__ lst,
                                                       int *__restrict__ abort_flag, unsigned *__restrict__ bar) {
    if (blockIdx.x == 0 && threadIdx.x == 0) { *abort_flag = 0; bar[0] = 0u; }
    if (blockIdx.x == 0) lum_reset_states(P, lst);
    const float se = (float)(*sum_ecm);
#pragma unroll 2
    for (int i = blockIdx.x; i < P.N; i += gridDim.x)
        lumecm_row(P, i, se, lum, ecm, psi, hlum, hecm, lumA, lumB, ecmA, ecmB);
}

/* periodic ghost-ring source index of boundary_conditions_periodic /
 * boundary_conditions_cells (useful_functions.cpp:391-420, 505-526) */
static __device__ __forceinline__ int ring_src(int i, int N) { return i == 0 ? N - 2 : (i == N - 1 ? 1 : i); }

/* Central difference of the old field for |grad| in the lumen/ECM normals.
 * The reference divides the new field's gradient by the old field's |grad|
 * (cell_phase_functions.cpp:734-763), which is ill-conditioned where the old
 * gradient vanishes: at an extremum whose neighbours are equal in exact
 * arithmetic (an ECM pocket between cells), FP32 rounding leaves a difference
 * of a few ulps, |grad| becomes ~1e-12 and the normal ~1e3, and the reshaping
 * sweeps then overflow to inf/NaN within a few steps.  Two safeguards: a
 * difference of at most 2 FP32 epsilons of its larger operand (2 to 4 ulps,
 * rounding noise rather than a resolved gradient) is taken as zero, and a
 * normal longer than ORG_JMAX (rounding noise that accumulated over steps in
 * such a pocket) is set to zero, the normal the reference gives an exactly
 * flat pixel.  Where the gradient is resolved, the normal's length stays
 * near 1 (the FP64 oracle's largest is 2, at the first step). */
#define ORG_JMAX 4.0f
static __device__ __forceinline__ float old_diff(const float p, const float m) {
    const float d = p - m;
    return fabsf(d) <= 2.0f * 1.1920929e-7f * fmaxf(fabsf(p), fabsf(m)) ? 0.0f : d;
}

/* Normals of one field on row i: |grad| of the old field
 * (cell_phase_functions.cpp:722-741), normals from the new field's gradient
 * (:744-771) with the periodic ring; grows (b0..b3) to the bounding box of
 * the nonzero normals (inclusive).  A sweep changes a pixel only if the
 * normal at it or at one of its four neighbours is nonzero (every flux term
 * carries a face average of the normals). */
static __device__ __forceinline__ void normals_row(const DevP &P, const int i, const float *__restrict__ old,
                                                   const float *__restrict__ nw, float *__restrict__ jx,
                                                   float *__restrict__ jy, int &b0, int &b1, int &b2, int &b3) {
    const int N = P.N;
    const float c = 1.0f / P.dx / P.dx / 4.0f, h = 1.0f / (2.0f * P.dx);
    const int si = ring_src(i, N);
    for (int j = threadIdx.x; j < N; j += blockDim.x) {
        const size_t g = (size_t)si * N + ring_src(j, N);
        float a = old_diff(old[g + N], old[g - N]), b = old_diff(old[g + 1], old[g - 1]);
        float ng = hw_sqrt(c * (a * a + b * b));
        float vx = 0.0f, vy = 0.0f;
        if (ng != 0.0f) {
            const float rng = hw_rcp(ng);
            vx = h * (nw[g + N] - nw[g - N]) * rng;
            vy = h * (nw[g + 1] - nw[g - 1]) * rng;
        }
        const size_t e = (size_t)i * N + j;
        if (vx * vx + vy * vy > ORG_JMAX * ORG_JMAX) { vx = 0.0f; vy = 0.0f; }
        jx[e] = vx; jy[e] = vy;
        if (vx != 0.0f || vy != 0.0f) { b0 = min(b0, i); b1 = max(b1, i); b2 = min(b2, j); b3 = max(b3, j); }
    }
}

/* Normals of field f on rows blk, blk + nblk, ...; the workgroup's box of
 * nonzero normals goes to jbox_part[(f * nblk + blk) * 4 ..]. sb: 4 ints
 * of LDS. */
static __device__ void normals_rows(const DevP &P, const int f, const int blk, const int nblk,
                                    const float *__restrict__ old, const float *__restrict__ nw,
                                    float *__restrict__ jx, float *__restrict__ jy, int *__restrict__ jbox_part,
                                    int *sb) {
    const int N = P.N;
    if (threadIdx.x == 0) { sb[0] = N; sb[1] = -1; sb[2] = N; sb[3] = -1; }
    __syncthreads();
    int b0 = N, b1 = -1, b2 = N, b3 = -1;
#pragma unroll 2
    for (int i = blk; i < N; i += nblk) normals_row(P, i, old, nw, jx, jy, b0, b1, b2, b3);
    box_to_lds(b0, b1, b2, b3, sb);
    __syncthreads();
    if (threadIdx.x < 4) jbox_part[((size_t)f * nblk + blk) * 4 + threadIdx.x] = sb[threadIdx.x];
}

/* Lumen/ECM update and normals of both fields on the row band [r0, r1) of
 * workgroup blk.  The normals of a row read the updated fields on the rows
 * around its ring source (ring_src(i) +- 1), so the workgroup updates those
 * rows itself: the band, one halo row on each side and, for the first and
 * last bands, the rows the periodic ring maps to.  Rows updated by two
 * workgroups get the same values from both.  No synchronisation between
 * workgroups is needed; the boxes of nonzero normals go to jbox_part. */
static __device__ void lumecm_normals_band(const DevP &P, const int blk, const int nblk, const float se,
                                           const float *__restrict__ lum, const float *__restrict__ ecm,
                                           const float *__restrict__ psi, const float *__restrict__ hlum,
                                           const float *__restrict__ hecm, float *__restrict__ L0,
                                           float *__restrict__ L1, float *__restrict__ E0, float *__restrict__ E1,
                                           float *__restrict__ jlx, float *__restrict__ jly, float *__restrict__ jex,
                                           float *__restrict__ jey, int *__restrict__ jbox_part, int *sb) {
    const int N = P.N;
    const int r0 = (int)(((long long)blk * N) / nblk), r1 = (int)(((long long)(blk + 1) * N) / nblk);
#pragma unroll 2
    for (int i = max(r0 - 1, 0); i < min(r1 + 1, N); i++)
        lumecm_row(P, i, se, lum, ecm, psi, hlum, hecm, L0, L1, E0, E1);
    if (r0 == 0)
        for (int i = max(N - 3, 0); i < N; i++) lumecm_row(P, i, se, lum, ecm, psi, hlum, hecm, L0, L1, E0, E1);
    if (r1 == N)
        for (int i = 0; i < min(3, N); i++) lumecm_row(P, i, se, lum, ecm, psi, hlum, hecm, L0, L1, E0, E1);
    __syncthreads();          /* the workgroup reads back its own stores */
    for (int f = 0; f < 2; f++) {
        const float *old = f ? ecm : lum, *nw = f ? E0 : L0;
        float *jx = f ? jex : jlx, *jy = f ? jey : jly;
        if (threadIdx.x == 0) { sb[0] = N; sb[1] = -1; sb[2] = N; sb[3] = -1; }
        __syncthreads();
        int b0 = N, b1 = -1, b2 = N, b3 = -1;
#pragma unroll 2
        for (int i = r0; i < r1; i++) normals_row(P, i, old, nw, jx, jy, b0, b1, b2, b3);
        box_to_lds(b0, b1, b2, b3, sb);
        __syncthreads();
        if (threadIdx.x < 4) jbox_part[((size_t)f * nblk + blk) * 4 + threadIdx.x] = sb[threadIdx.x];
        __syncthreads();
    }
}

/* Box mode: the update (both buffers) and the normals of field f on its
 * live box B, the box's rows split into bands over the workgroups.  The band
 * updates its rows plus one halo row each side and the box's columns plus
 * one each side, which is every updated value its normals read; rows
 * updated by two workgroups get the same values from both.  Outside B grown
 * by one the buffers already hold the (unchanged) field, and the normals
 * outside B are zero and already stored as such (B holds every pixel that
 * was ever live).  The workgroup's box of nonzero normals goes to
 * jbox_part. */
static __device__ void lumecm_normals_box(const DevP &P, const int f, const int4 B, const int blk, const int nblk,
                                          const float se, const float *__restrict__ lum,
                                          const float *__restrict__ ecm, const float *__restrict__ psi,
                                          const float *__restrict__ hlum, const float *__restrict__ hecm,
                                          float *__restrict__ XA, float *__restrict__ XB, float *__restrict__ jx,
                                          float *__restrict__ jy, int *__restrict__ jbox_part, int *sb) {
    const int N = P.N, nt = blockDim.x;
    if (threadIdx.x == 0) { sb[0] = N; sb[1] = -1; sb[2] = N; sb[3] = -1; }
    if (!box_empty(B)) {
        const int nr = B.y - B.x;
        const int r0 = B.x + (int)(((long long)blk * nr) / nblk), r1 = B.x + (int)(((long long)(blk + 1) * nr) / nblk);
        if (r1 > r0) {
            /* update: rows [r0-1, r1+1) within [B.x-1, B.y+1), columns [B.z-1, B.w+1) */
            const int ua = r0 - 1, ub = r1 + 1, ca = B.z - 1, cw = B.w - B.z + 2;
            const int npx = (ub - ua) * cw;
#pragma unroll 4
            for (int q = threadIdx.x; q < npx; q += nt) {
                const int i = ua + q / cw, j = ca + q % cw;
                const size_t e = (size_t)i * N + j;
                const float v = lumecm_px(P, f, e, se, lum, ecm, psi, hlum, hecm);
                XA[e] = v; XB[e] = v;
            }
        }
        __syncthreads();      /* the workgroup reads back its own stores */
        if (r1 > r0) {
            const float *old = f ? ecm : lum;
            const float c = 1.0f / P.dx / P.dx / 4.0f, h = 1.0f / (2.0f * P.dx);
            const int bw = B.w - B.z, npx = (r1 - r0) * bw;
            int b0 = N, b1 = -1, b2 = N, b3 = -1;
#pragma unroll 4
            for (int q = threadIdx.x; q < npx; q += nt) {
                const int i = r0 + q / bw, j = B.z + q % bw;
                const size_t g = (size_t)i * N + j;
                const float a = old_diff(old[g + N], old[g - N]), b = old_diff(old[g + 1], old[g - 1]);
                const float ng = hw_sqrt(c * (a * a + b * b));
                float vx = 0.0f, vy = 0.0f;
                if (ng != 0.0f) {
                    const float rng = hw_rcp(ng);
                    vx = h * (XA[g + N] - XA[g - N]) * rng;
                    vy = h * (XA[g + 1] - XA[g - 1]) * rng;
                }
                if (vx * vx + vy * vy > ORG_JMAX * ORG_JMAX) { vx = 0.0f; vy = 0.0f; }
                jx[g] = vx; jy[g] = vy;
                if (vx != 0.0f || vy != 0.0f) { b0 = min(b0, i); b1 = max(b1, i); b2 = min(b2, j); b3 = max(b3, j); }
            }
            box_to_lds(b0, b1, b2, b3, sb);
        }
    }
    __syncthreads();
    if (threadIdx.x < 4) jbox_part[((size_t)f * nblk + blk) * 4 + threadIdx.x] = sb[threadIdx.x];
    __syncthreads();
}

/* Box mode: the boundary ring of both fields in both buffers (lumen 0,
 * ECM 1, useful_functions.cpp:439-462), by workgroup blk of nblk; its
 * normals are zero. */
static __device__ void lumecm_ring(const DevP &P, float *__restrict__ L0, float *__restrict__ L1,
                                   float *__restrict__ E0, float *__restrict__ E1, const int blk, const int nblk,
                                   const bool corners = true) {
    for (int q = blk * blockDim.x + threadIdx.x; q < 4 * P.N - 4; q += nblk * blockDim.x) {
        int i, j;
        ring_px(q, P.N, i, j);
        if (!corners && (i == 0 || i == P.N - 1) && (j == 0 || j == P.N - 1)) continue;
        const size_t e = (size_t)i * P.N + j;
        L0[e] = 0.0f; L1[e] = 0.0f;
        E0[e] = 1.0f; E1[e] = 1.0f;
    }
}

/* k_lum_normals: normals of field fbase + blockIdx.y (0 lumen, 1 ECM). */
__global__ __launch_bounds__(1024) void k_lum_normals(DevP P, const float *__restrict__ lum_old,
                                                     const float *__restrict__ ecm_old, const float *__restrict__ lumn,
                                                     const float *__restrict__ ecmn, float *__restrict__ jlx,
                                                     float *__restrict__ jly, float *__restrict__ jex,
                                                     float *__restrict__ jey, int *__restrict__ jbox_part,
                                                     int fbase) {
    __shared__ int sb[4];
    const int f = fbase + blockIdx.y;
    normals_rows(P, f, blockIdx.x, gridDim.x, f ? ecm_old : lum_old, f ? ecmn : lumn, f ? jex : jlx, f ? jey : jly,
                 jbox_part, sb);
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
    box_to_lds(b0, b1, b2, b3, sb);
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

/* reshape_flux with the normals already loaded: jx[0..4], jy[0..4] at the
 * pixel and its +row, -row, +column, -column neighbours. */
static __device__ __forceinline__ float reshape_flux_j(float un, const float *u, int su, const float *jx,
                                                       const float *jy, float ka, float kc) {
    const float ja[3 * 3] = {0.0f, jx[2], 0.0f, jx[4], jx[0], jx[3], 0.0f, jx[1], 0.0f};
    const float jb[3 * 3] = {0.0f, jy[2], 0.0f, jy[4], jy[0], jy[3], 0.0f, jy[1], 0.0f};
    return reshape_flux(un, u, su, ja + 4, jb + 4, 3, ka, kc);
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
#define ORG_LUM_MAX 128      /* lumen/ECM workgroups of k_step a tuning may ask for */
#ifndef ORG_LUM_LDS_LOOP
#define ORG_LUM_LDS_LOOP 1   /* k_step's loops with the field in LDS where it fits (box mode) */
#endif
static __device__ __forceinline__ double lum_sweep(const DevP &P, const float *__restrict__ src,
                                                   float *__restrict__ dst, const float *__restrict__ jx,
                                                   const float *__restrict__ jy, const int4 jb,
                                                   float side, int blk, int nblk, bool ring = true) {
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
    /* the boundary ring: corners 0, the rest `side`.  With ring false the
     * caller has put `side` on the whole ring of both buffers, so outside the
     * corners the sweep would rewrite the same value and add exactly 0 to the
     * deviation; only the corners are visited (by the same threads) */
    for (int k = blk * 1024 + (int)threadIdx.x; k < 4 * N - 4; k += stride) {
        int i, j;
        if (k < N) { i = 0; j = k; }
        else if (k < 2 * N) { i = N - 1; j = k - N; }
        else if (k < 3 * N - 2) { i = k - 2 * N + 1; j = 0; }
        else { i = k - (3 * N - 2) + 1; j = N - 1; }
        const size_t e = (size_t)i * N + j;
        const bool corner = (i == 0 || i == N - 1) && (j == 0 || j == N - 1);
        if (!ring && !corner) continue;
        float v = corner ? 0.0f : side;
        dst[e] = v;
        acc += (double)(fabsf(v - src[e]) * dx2);
    }
    return acc;
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

/* Barrier among the `nblk` lumen/ECM workgroups of k_step.  bar[0] counts
 * arrivals over the whole launch (k_post clears it); the e-th barrier of a
 * workgroup releases once it reaches e * nblk, so no reset is needed between
 * barriers.  Every wave drains its stores, thread 0 releases (agent scope),
 * arrives, polls with a sleep and acquires.  The workgroups that wait here
 * wait only on each other and every other workgroup of the launch finishes
 * on its own, so all of them become resident.  A wait longer than `limit`
 * wall-clock ticks, or one that finds the abort flag raised, raises the abort
 * flag and ORG_ETIMEOUT and returns false instead of waiting on. */
static __device__ bool lum_barrier(unsigned *bar, int nblk, int &epoch, int *abort_flag, int *err, long long limit) {
    __shared__ int ok;
    epoch++;
    drain_stores();
    __syncthreads();
    if (threadIdx.x == 0) {
        int good = 1;
        const unsigned target = (unsigned)epoch * (unsigned)nblk;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        unsigned seen = atomicAdd(&bar[0], 1u) + 1u;
        const long long t0 = wall_clock64();
        while (seen < target) {
            __builtin_amdgcn_s_sleep(2);
            seen = __hip_atomic_load(&bar[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (seen >= target) break;
            if (__hip_atomic_load(abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ||
                wall_clock64() - t0 > limit) {
                atomicExch(abort_flag, 1);
                set_err(err, ORG_ETIMEOUT);
                good = 0;
                break;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        ok = good;
    }
    __syncthreads();
    return ok != 0;
}

/* Both reshaping loops (lumen and ECM) run to their stop tests by the loop
 * workgroups of k_step, one barrier per sweep.  Every loop workgroup
 * sums the per-workgroup deviations in the same order, so all take the same
 * stop decision after the same sweep.  part: [2 parities][2 fields][nblk].
 * Returns false if a barrier timed out; cur_out: the buffers holding the
 * last iterates (every workgroup knows them). */
static __device__ bool lum_loop_body(const DevP &P, int blk, int nblk, double *red, float *bufL0, float *bufL1,
                                     float *bufE0, float *bufE1, const float *jlx, const float *jly,
                                     const float *jex, const float *jey, const int *jbox_part, int nparts,
                                     LumState *lst, double *part, unsigned *bar, int &epoch, int *abort_flag,
                                     int *err, long long limit, int *cur_out, bool ring) {
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
                acc = lum_sweep(P, src, dst, f ? jex : jlx, f ? jey : jly, jb[f], f ? 1.0f : 0.0f, blk, nblk,
                                ring);
            }
            double tot = block_sum_d(acc, red);
            if (threadIdx.x == 0) pk[f * nblk + blk] = tot;
        }
        if (!lum_barrier(bar, nblk, epoch, abort_flag, err, limit)) return false;
        for (int f = 0; f < 2; f++) {
            if (s[f].done) continue;
            double rdev = sum_parts(pk + f * nblk, nblk, red);
            lum_stop(P, s[f], rdev, err);
        }
    }
    if (blk == 0 && threadIdx.x < 2) lst[threadIdx.x] = s[threadIdx.x];
    cur_out[0] = s[0].cur;
    cur_out[1] = s[1].cur;
    return true;
}

/* ---- the lumen/ECM role with the fields in LDS ------------------------------
 * The batched step's lumen/ECM update, normals and reshaping loops (box mode)
 * with each loop workgroup keeping one band of rows of each field's sweep
 * domain in LDS from the update to the last sweep.  The sweep domain is the
 * live box grown by one (the update's region): outside the normals' box
 * grown by one a sweep leaves a pixel unchanged exactly (every flux term
 * carries a face average of normals, all zero there), so this equals
 * sweeping only that box.
 *   - Set-up, no exchange: each workgroup computes the update (lumecm_px)
 *     on its band plus ORG_LUM_SPAN + 1 rows on each side and the normals on
 *     its band plus ORG_LUM_SPAN rows (zero outside the live box, from the
 *     old field and the LDS update).
 *   - Spans of ORG_LUM_SPAN sweeps: sweep t of a span covers the band grown
 *     by SPAN - 1 - t rows (the neighbours' rows, recomputed with the same
 *     arithmetic on the same operands, so the same values), so a span needs
 *     the neighbours' rows only once, at its start.  After a span each
 *     workgroup publishes its first and last SPAN rows (at their grid
 *     coordinates, in the exchange plane of the span's parity) and the
 *     deviation of its own rows for each sweep; then, with no barrier, it
 *     waits for the neighbours' rows and every workgroup's deviations of
 *     the span, puts the rows into its halo and takes the sweeps' stop
 *     tests in order (cell_phase_functions.cpp:774, 862-868).  Sweeps after
 *     the one that stops the loop were speculative and are dropped; the
 *     stopping sweep's iterate is still in LDS (its own buffer, or a copy of
 *     the band's rows for the sweeps the span overwrites later).
 *   - Every exchanged value travels in a 64-bit word with the tag of its
 *     exchange (xput / xget / xwait), through relaxed agent-scope accesses
 *     that are coherent across XCDs by themselves: a reader polls the words
 *     it needs until they carry the span's tag, so a span's exchange costs
 *     one store-to-load latency, with no wait for store completion, no
 *     arrival counter and no L2 write-back or invalidation (the other
 *     roles' cached data on the XCD stay valid).  Two planes suffice: a
 *     workgroup overwrites span k's words only in span k + 2, after every
 *     reader has consumed them (it needed their span k + 1 words first).
 *   - At the end the band's last iterate goes to the buffer `cur` (the
 *     other buffer is rewritten before any read by the next step's update),
 *     workgroup 0 writes the four corners of both buffers (the ring writes
 *     of the other workgroups leave them out, so no two workgroups store to
 *     the same address), and one full barrier publishes everything.
 * Each pixel's update, normal and flux is the same arithmetic on the same
 * operands as lumecm_normals_box and lum_sweep, so the fields equal the
 * global-memory path bit for bit; only the summation order of the
 * deviation differs (it feeds the stop test alone). */
#ifndef ORG_LUM_SPAN
#define ORG_LUM_SPAN 2             /* sweeps per exchange (even) */
#endif
/* 64-bit words of the LDS loops' exchange buffer (rows and partial sums) */
#define ORG_XCH_WORDS(NN) (4 * (size_t)(NN) + 2 * 2 * ORG_LUM_SPAN * 2 * ORG_LUM_MAX)
/* LDS of the loop workgroup: [0, 128) red, [128, 160) the prologue's boxes,
 * [256, 1280) per-wave deviations, the bands from ORG_LUM_LDS_BASE */
#define ORG_LUM_LDS_BASE 2048
struct LumBand {
    int r0, r1, c0, c1;           /* sweep domain (the live box grown by one), inclusive */
    int a, b;                     /* this workgroup's rows [a, b) */
    int S, bw;                    /* LDS row stride (bw + 2), domain width */
    float *u[2], *jx, *jy;        /* LDS row 0 = grid row a - SPAN - 1, column 0 = column c0 - 1 */
    float *hist;                  /* the band's rows after sweeps 0 .. SPAN-3 of a span */
};

static __device__ __forceinline__ int lum_own_max(int bh, int nblk) { return (bh + nblk - 1) / nblk; }
static __device__ __forceinline__ int lum_lds_rows(int bh, int nblk) {
    return lum_own_max(bh, nblk) + 2 * ORG_LUM_SPAN + 2;
}

/* LDS bytes of one field's band arrays for live box B (half open) split
 * over nblk workgroups (0 for an empty box). */
static __device__ __forceinline__ int lum_band_bytes(const int4 B, int nblk) {
    if (box_empty(B)) return 0;
    const int bh = B.y - B.x + 2, bw = B.w - B.z + 2;
    return 4 * (4 * lum_lds_rows(bh, nblk) * (bw + 2) + (ORG_LUM_SPAN - 2) * lum_own_max(bh, nblk) * bw);
}

/* Whether the LDS form fits: both fields' bands beside the reduction
 * scratch in k_step's shared memory. */
static __device__ __forceinline__ bool lum_lds_fits(const int4 BL, const int4 BE, int nblk, int lds_bytes) {
    return ORG_LUM_LDS_BASE + lum_band_bytes(BL, nblk) + lum_band_bytes(BE, nblk) <= lds_bytes;
}

/* Tagged exchange words of the LDS loops: 32 bits of payload in the low
 * half, the tag of the exchange in the high half, written and read as one
 * 64-bit agent-scope access (single-copy atomic, coherent across XCDs).  A
 * reader that finds its tag holds the payload of that exchange, so no
 * barrier and no ordering between words is needed.  Tags grow by one per
 * exchange of a context and never repeat within 2^32 exchanges; the
 * buffer is zeroed at creation and no exchange uses tag 0. */
static __device__ __forceinline__ void xput(unsigned long long *p, unsigned v, unsigned tag) {
    __hip_atomic_store(p, ((unsigned long long)tag << 32) | (unsigned long long)v, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}
static __device__ __forceinline__ unsigned long long xget(const unsigned long long *p) {
    return __hip_atomic_load(const_cast<unsigned long long *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
/* Poll *p until it carries tag (w: the word as first read, updated); false
 * once the wait passes `limit` wall-clock ticks or the abort flag is up. */
static __device__ bool xwait(const unsigned long long *p, unsigned tag, unsigned long long &w, int *abort_flag,
                             long long limit) {
    if ((unsigned)(w >> 32) == tag) return true;
    const long long t0 = wall_clock64();
    for (int n = 1;; n++) {
        __builtin_amdgcn_s_sleep(1);
        w = xget(p);
        if ((unsigned)(w >> 32) == tag) return true;
        if ((n & 15) == 0 && (__hip_atomic_load(abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ||
                              wall_clock64() - t0 > limit))
            return false;
    }
}

/* The normal of one pixel (as normals_row and lumecm_normals_box): |grad|
 * of the old field at g, the new field's central differences given. */
static __device__ __forceinline__ void lum_normal_px(const DevP &P, const float *__restrict__ old, const size_t g,
                                                     const int N, const float np, const float nm, const float ep,
                                                     const float em, float &vx, float &vy) {
    const float c = 1.0f / P.dx / P.dx / 4.0f, h = 1.0f / (2.0f * P.dx);
    const float a = old_diff(old[g + N], old[g - N]), b = old_diff(old[g + 1], old[g - 1]);
    const float ng = hw_sqrt(c * (a * a + b * b));
    vx = 0.0f; vy = 0.0f;
    if (ng != 0.0f) {
        const float rng = hw_rcp(ng);
        vx = h * (np - nm) * rng;
        vy = h * (ep - em) * rng;
    }
    if (vx * vx + vy * vy > ORG_JMAX * ORG_JMAX) { vx = 0.0f; vy = 0.0f; }
}

/* The pixels q = r * w + c (0 <= r < rows) of a block of rows taken by the
 * workgroup's threads at q = threadIdx.x, + blockDim.x, ...: (r, c)
 * advance by additions, without a division per pixel. */
struct RowWalk {
    int r, c, dr, dc, w;
    __device__ __forceinline__ explicit RowWalk(int w_) : w(w_) {
        r = (int)threadIdx.x / w; c = (int)threadIdx.x - r * w;
        dr = (int)blockDim.x / w; dc = (int)blockDim.x - dr * w;
    }
    __device__ __forceinline__ void next() {
        r += dr; c += dc;
        if (c >= w) { c -= w; r++; }
    }
};

/* One sweep of grid rows [i0, i1) of a band in LDS (src -> dst); returns
 * this thread's deviation over the rows within [a, b). */
static __device__ __forceinline__ double lum_band_sweep(const LumBand &L, const float *src, float *dst, int i0,
                                                        int i1, float ka, float kc, float dx2) {
    const int base = L.a - ORG_LUM_SPAN - 1, own0 = L.a - i0, own1 = L.b - i0;
    double acc = 0.0;
    for (RowWalk q(L.bw); q.r < i1 - i0; q.next()) {
        const int e = (i0 + q.r - base) * L.S + 1 + q.c;
        const float u = src[e];
        const float v = reshape_flux(u, src + e, L.S, L.jx + e, L.jy + e, L.S, ka, kc);
        dst[e] = v;
        if (q.r >= own0 && q.r < own1) acc += (double)(fabsf(v - u) * dx2);
    }
    return acc;
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

/* Copy-back of field f from its final buffer F (lum <- lumn or ECM <- ECMn)
 * by workgroup blk of nblk.  Full mode: every pixel.  Box mode: the live box
 * B grown by one pixel and the ring (the only pixels that can differ), then
 * the bounding box (inclusive) of the live pixels of the new field within B
 * grown by two (a neighbour of a changed pixel may have become live) goes to
 * live_part[(f * nblk + blk) * 4 ..] (empty: i1 < 0). */
static __device__ void commit_field(const DevP &P, const int f, const int4 B, const int blk, const int nblk,
                                    float *__restrict__ X, const float *__restrict__ F, int *__restrict__ live_part,
                                    int *sb) {
    const int N = P.N, nt = blockDim.x;
    int *lp = live_part + ((size_t)f * nblk + blk) * 4;
    if (threadIdx.x < 4) lp[threadIdx.x] = (threadIdx.x & 1) ? -1 : N;
    if (!box_mode(B, N)) {
#pragma unroll 4
        for (int i = blk; i < N; i += nblk) {
            const size_t r = (size_t)i * N;
            for (int j = threadIdx.x; j < N; j += nt) X[r + j] = F[r + j];
        }
        return;
    }
    for (int q = blk * nt + threadIdx.x; q < 4 * N - 4; q += nblk * nt) {
        int i, j;
        ring_px(q, N, i, j);
        X[(size_t)i * N + j] = F[(size_t)i * N + j];
    }
    if (box_empty(B)) return;
    {
        const int ca = B.z - 1, cw = B.w - B.z + 2, npx = (B.y - B.x + 2) * cw;
#pragma unroll 4
        for (int q = blk * nt + threadIdx.x; q < npx; q += nblk * nt) {
            const size_t e = (size_t)(B.x - 1 + q / cw) * N + (ca + q % cw);
            X[e] = F[e];
        }
    }
    if (threadIdx.x == 0) { sb[0] = N; sb[1] = -1; sb[2] = N; sb[3] = -1; }
    __syncthreads();
    int b0 = N, b1 = -1, b2 = N, b3 = -1;
    {
        const int ra = max(B.x - 2, 1), rb = min(B.y + 2, N - 1), ca = max(B.z - 2, 1), cb = min(B.w + 2, N - 1);
        const int cw = cb - ca, npx = (rb - ra) * cw;
#pragma unroll 4
        for (int q = blk * nt + threadIdx.x; q < npx; q += nblk * nt) {
            const int i = ra + q / cw, j = ca + q % cw;
            if (is_live(F, (size_t)i * N + j, N)) { b0 = min(b0, i); b1 = max(b1, i); b2 = min(b2, j); b3 = max(b3, j); }
        }
    }
    box_to_lds(b0, b1, b2, b3, sb);
    __syncthreads();
    if (threadIdx.x < 4) lp[threadIdx.x] = sb[threadIdx.x];
    __syncthreads();
}

/* k_live_init: the lumen/ECM state of the batched step from scratch (before
 * the first batched step, or after the fields were set or stepped by the
 * per-function entry points): live boxes of lum and ECM (reset, with the
 * copy-back's partial boxes, by k_live_reset), both ping-pong buffers <- the
 * fields, normals <- 0, every part of the h(lum)/h(ECM) pass and both
 * buffers of the lumen/ECM bracket.  Workgroup
 * k < ntiles: tile k; the last: the ring. */
__global__ void k_live_reset(int *live, int *live_part, int nparts, int N, int *ringh) {
    if (threadIdx.x < 8) live[threadIdx.x] = (threadIdx.x & 1) ? 0 : N;
    if (threadIdx.x == 0) *ringh = 0;
    for (int q = threadIdx.x; q < 8 * nparts; q += blockDim.x) live_part[q] = (q & 1) ? -1 : N;
}
__global__ __launch_bounds__(1024) void k_live_init(DevP P, const float *__restrict__ lum,
                                                    const float *__restrict__ ecm, float *__restrict__ hlum,
                                                    float *__restrict__ hecm, double *__restrict__ part,
                                                    float *__restrict__ L0, float *__restrict__ L1,
                                                    float *__restrict__ E0, float *__restrict__ E1,
                                                    float *__restrict__ jlx, float *__restrict__ jly,
                                                    float *__restrict__ jex, float *__restrict__ jey,
                                                    int *__restrict__ live, double *__restrict__ lbrk) {
    __shared__ double red[16];
    __shared__ int sb[8];
    __shared__ __attribute__((aligned(16))) float lds[ORG_LDS_LBRK / 4];
    const int k = blockIdx.x, N = P.N;
    const double tot = block_sum_d(hlum_part(P, k, lum, ecm, hlum, hecm), red);
    if (threadIdx.x == 0) part[k] = tot;
    lbrk_part(P, k, lum, ecm, lbrk, lds);
    if (threadIdx.x < 8) sb[threadIdx.x] = (threadIdx.x & 1) ? -1 : N;
    __syncthreads();
    int bx[2][4] = {{N, -1, N, -1}, {N, -1, N, -1}};
    auto px = [&](int i, int j) {
        const size_t e = (size_t)i * N + j;
        const float a = lum[e], b = ecm[e];
        L0[e] = a; L1[e] = a; E0[e] = b; E1[e] = b;
        jlx[e] = 0.0f; jly[e] = 0.0f; jex[e] = 0.0f; jey[e] = 0.0f;
        if (i >= 1 && i < N - 1 && j >= 1 && j < N - 1) {
            for (int f = 0; f < 2; f++)
                if (is_live(f ? ecm : lum, e, N)) {
                    bx[f][0] = min(bx[f][0], i); bx[f][1] = max(bx[f][1], i);
                    bx[f][2] = min(bx[f][2], j); bx[f][3] = max(bx[f][3], j);
                }
        }
    };
    if (k == hl_ntiles(N)) {
        for (int q = threadIdx.x; q < 4 * N - 4; q += blockDim.x) { int i, j; ring_px(q, N, i, j); px(i, j); }
    } else {
        const int nt = (N + ORG_HT - 1) / ORG_HT, ti = (k / nt) * ORG_HT, tj = (k % nt) * ORG_HT;
        for (int e = threadIdx.x; e < ORG_HT * ORG_HT; e += blockDim.x) {
            const int i = ti + e / ORG_HT, j = tj + e % ORG_HT;
            if (i >= 1 && i < N - 1 && j >= 1 && j < N - 1) px(i, j);
        }
    }
    for (int f = 0; f < 2; f++) box_to_lds(bx[f][0], bx[f][1], bx[f][2], bx[f][3], sb + 4 * f);
    __syncthreads();
    if (threadIdx.x < 2 && sb[4 * threadIdx.x + 1] >= 0) {
        const int f = threadIdx.x;
        atomicMin(&live[4 * f], sb[4 * f]); atomicMax(&live[4 * f + 1], sb[4 * f + 1] + 1);
        atomicMin(&live[4 * f + 2], sb[4 * f + 2]); atomicMax(&live[4 * f + 3], sb[4 * f + 3] + 1);
    }
}

/* Active stress pi of one box pixel (rho_functions.cpp:32-115): label idx,
 * rho r, |grad phi| ng from the central differences (dxp, dyp). */
static __device__ __forceinline__ void active_stress(const DevP &P, const int idx, const float r, const float ng,
                                                     const float dxp, const float dyp, const float inv4,
                                                     float &pxx, float &pxy, float &pyv) {
    const float at = by_label(P.atan_, idx), ao = by_label(P.aort, idx), bb = by_label(P.bb, idx);
    const float r3 = r * r * r;
    const float base = -at * r3 + bb * r3 * r;
    pxx = base; pxy = 0.0f; pyv = base;
    if (ng > 0.0f) {
        const float rng = hw_rcp(ng);
        const float c = (at - ao) * r3 * inv4 * rng * rng;
        pxx += c * dxp * dxp;
        pxy += c * dxp * dyp;
        pyv += c * dyp * dyp;
    }
}

/* ======================================================================== */
/* k_spectral: update_all_cortices (rho_functions.cpp:409-471) for every cell.
 *
 * One workgroup (1024 threads) per cell, one 128 x 128 complex field in LDS,
 * a 128 x 128 complex and two 128 x 128 real scratch planes per cell in
 * global memory (L2 resident): F(pi_xx + i pi_xy), needed at k and -k,
 * pi_yy, and rho on the box.
 * By linearity of the reference's spectral update, Re(F^-1[F(rho) +
 * dt F(react) + D]) / M^2 = rho + dt*react + Re(F^-1 D) / M^2, so the
 * transforms of rho and of the reaction term (rho_functions.cpp:415-422,
 * 462-464) cancel and are not taken.  Real pairs share one complex
 * transform: F(pi_xx + i pi_xy), F(pi_yy), F^-1(H v_x + i H v_y) where
 * H X(k) = (X(k) + conj X(-k)) / 2 is the part whose inverse is real (the
 * reference keeps only Re, rho_functions.cpp:279), F(rho v_x + i rho v_y),
 * and F^-1(H D).  The per-mode solve uses the viscosity of the real-space
 * pixel with the mode's indices (rho_functions.cpp:149-150, 170). */
#ifndef ORG_P1_QUADS
#define ORG_P1_QUADS 2      /* quads of the stress pass whose loads are issued together */
#endif
template <bool GIVEN>
static __device__ __forceinline__ void spectral_body(const DevP &P, const int n, float2 *__restrict__ sm,
                                                     float2 *__restrict__ tw, unsigned char *__restrict__ lab,
                                                     const float *__restrict__ phi,
                                                     const float *__restrict__ rho, const float *__restrict__ lum,
                                                     const float *__restrict__ ecm, const int x0, const int y0,
                                                     float2 *__restrict__ scratch,
                                                     float *__restrict__ rho_box, float *__restrict__ vxbox,
                                                     float *__restrict__ vybox, int write_v,
                                                     const float *__restrict__ vg) {
    /* (x0, y0): FFT box origin = centre - zone_size + 1, useful_functions.cpp:618-619.
     * vg: if not null, this cell's velocity on the box ([2][M*M], vx then vy,
     * from the surrogate; GIVEN): the active stress, its transforms and the
     * per-mode solve are skipped and vg stands for v in the advection. */
    const int t = threadIdx.x, N = P.N;
    constexpr int MM = ORG_MB * ORG_MB;
    constexpr int NPAIR = 63 * ORG_MB + 2 * 65;      /* mode pairs {k, -k} incl. 4 self-paired modes */
    const size_t NN = (size_t)N * N;
    const int n1 = (n == 0) ? P.nc - 1 : n - 1, n2 = (n == P.nc - 1) ? 0 : n + 1;
    if (t < 128) tw[t] = make_float2(kCos128[t], kSin128[t]);
    const float *ph = phi + n * NN, *ph1 = phi + n1 * NN, *ph2 = phi + n2 * NN, *rh = rho + n * NN;
    float2 *SA = scratch + (size_t)n * 2 * MM;
    float *SB = (float *)(SA + MM);
    /* rho on the box, dense (row pitch 128), written by P1 beside pi_yy:
     * the rho * v epilogue reads it from here (recently written, in L2)
     * rather than from the plane */
    float *SR = SB + MM;
    float *rbox = rho_box + (size_t)n * MM;
    const float inv4 = 1.0f / (4.0f * P.dx * P.dx);
    const float isig2 = 1.0f / (2.0f * P.sigma * P.sigma);
    const float dnorm = 1.0f / (2.50662827463f * P.sigma);   /* 1/(sqrt(2 pi) sigma) */
    const float kscale = 6.28318530718f / (ORG_MB * P.dx);
    const float MMf = (float)MM;

    /* P1: active stress (rho_functions.cpp:32-115), reaction (:193-209), in
     * quads of 4 pixels along a row: thread t takes columns 4 (t & 31) ..
     * +3 of rows (t >> 5) + 32 q, q = 0..3, and a quad's loads are 8 four-
     * wide ones plus the two row neighbours at its ends, issued before any
     * use (byte offsets from uniform bases, scheduling barrier) */
    const unsigned rN = 4u * (unsigned)N;
    const int jq = 4 * (t & 31);
    constexpr int QB = ORG_P1_QUADS;
#pragma unroll 1
    for (int qb = 0; qb < 4; qb += QB) {
        f4a fv[QB], fxp[QB], fxm[QB], fl[QB], fe[QB], f1[QB], f2[QB], fr[QB];
        float fym[QB], fyp[QB];
#pragma unroll
        for (int k = 0; k < QB; k++) {
            const int i = (t >> 5) + 32 * (qb + k);
            const unsigned g = 4u * ((unsigned)(x0 + i) * (unsigned)N + (unsigned)(y0 + jq));
            fv[k] = ldb4(ph, g); fxp[k] = ldb4(ph, g + rN); fxm[k] = ldb4(ph, g - rN);
            fym[k] = ldb(ph, g - 4); fyp[k] = ldb(ph, g + 16);
            fl[k] = ldb4(lum, g); fe[k] = ldb4(ecm, g); f1[k] = ldb4(ph1, g); f2[k] = ldb4(ph2, g);
            fr[k] = ldb4(rh, g);
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int k = 0; k < QB; k++) {
            const int i = (t >> 5) + 32 * (qb + k), e0 = i * ORG_MB + jq;
            unsigned labs = 0;
            f4a pyq, rq;
#pragma unroll
            for (int m = 0; m < 4; m++) {
                const float v = fv[k][m];
                const float dxp = fxp[k][m] - fxm[k][m];
                const float dyp = (m < 3 ? fv[k][m + 1] : fyp[k]) - (m > 0 ? fv[k][m - 1] : fym[k]);
                const float ng = hw_sqrt(inv4 * (dxp * dxp + dyp * dyp));
                const int idx = turnover_label(v, fl[k][m], fe[k][m], f1[k][m], f2[k][m]);
                labs |= (unsigned)idx << (8 * m);
                const float r = fr[k][m];
                if (!GIVEN) {
                    float pxx, pxy, pyv;
                    active_stress(P, idx, r, ng, dxp, dyp, inv4, pxx, pxy, pyv);
                    pyq[m] = pyv;
                    sm[i * ORG_LS + jq + m] = make_float2(pxx, pxy);
                }
                const float d = v - 0.5f;
                const float pnd = ng * hw_exp(-d * d * isig2) * dnorm;
                const float react = -by_label(P.deg, idx) * r + by_label(P.gen, idx) * pnd;
                rq[m] = r + P.dt * react;
            }
            *(unsigned *)(lab + e0) = labs;
            if (!GIVEN) {
                *(float4 *)(SB + e0) = make_float4(pyq[0], pyq[1], pyq[2], pyq[3]);
                *(float4 *)(SR + e0) = make_float4(fr[k][0], fr[k][1], fr[k][2], fr[k][3]);
            }
            *(float4 *)(rbox + e0) = make_float4(rq[0], rq[1], rq[2], rq[3]);
        }
    }
    if (GIVEN) {          /* given velocity: rho v straight into the transform buffer */
#pragma unroll 4
        for (int q = 0; q < 16; q++) {
            const int e = t + 1024 * q, i = e >> 7, j = e & 127;
            const float r = ldb(rh, 4u * ((unsigned)(x0 + i) * (unsigned)N + (unsigned)(y0 + j)));
            const float vx = vg[e], vy = vg[MM + e];
            if (write_v) {
                vxbox[(size_t)n * MM + e] = vx;
                vybox[(size_t)n * MM + e] = vy;
            }
            sm[i * ORG_LS + j] = make_float2(r * vx, r * vy);
        }
    } else {
    __syncthreads();
    /* F(pi_xx + i pi_xy) straight from the last stage to SA, and pi_yy (SB,
     * written by other threads in P1, ordered by the transform's barriers)
     * into its place for the next transform */
    fft2d_128<-1>(sm, tw, [&](const float2 *u, const int sub, const int j) {
#pragma unroll
        for (int q0 = 0; q0 < 16; q0 += 8) {      /* in halves: 8 loads in flight */
            float yy[8];
#pragma unroll
            for (int q = 0; q < 8; q++) yy[q] = SB[fft_out_pos(sub, q0 + q) * ORG_MB + j];
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const int i = fft_out_pos(sub, q0 + q);
                SA[i * ORG_MB + j] = u[q0 + q];
                sm[i * ORG_LS + j] = make_float2(yy[q], 0.0f);
            }
        }
    });
    fft2d_128<-1>(sm, tw);
    /* P4: per pair {k, -k}: split F(pi_xx + i pi_xy), solve both modes
     * (rho_functions.cpp:133-190), keep the part of v whose inverse is real,
     * Z = H v_x + i H v_y written in place at k and -k. */
    {
        /* pairs u = t + 1024 m, m = 0..8 (NPAIR = 8 * 1024 + 2), in groups of
         * 4 (the last, m = 8, exists for t < 2 only); a group's spectra are
         * loaded before any use */
        auto pair_ij = [](int u, int &i, int &j) {
            if (u < 63 * ORG_MB) { i = 1 + (u >> 7); j = u & 127; }
            else { const int w = u - 63 * ORG_MB; i = (w < 65) ? 0 : 64; j = (w < 65) ? w : w - 65; }
        };
        auto solve_pair = [&](int u, float2 A, float2 Am) {
            int i, j;
            pair_ij(u, i, j);
            const int im = (ORG_MB - i) & 127, jm = (ORG_MB - j) & 127;
            const int ek = i * ORG_MB + j, em = im * ORG_MB + jm;
            const int lk = i * ORG_LS + j, lm = im * ORG_LS + jm;
            const float2 B = sm[lk], Bm = sm[lm];
            const float2 pxx = make_float2(0.5f * (A.x + Am.x), 0.5f * (A.y - Am.y));
            const float2 pxy = make_float2(0.5f * (A.y + Am.y), 0.5f * (Am.x - A.x));
            const float2 pxxm = make_float2(pxx.x, -pxx.y), pxym = make_float2(pxy.x, -pxy.y);
            const int ir = (i < ORG_MB / 2) ? i : i - ORG_MB, jr = (j < ORG_MB / 2) ? j : j - ORG_MB;
            const int irm = (im < ORG_MB / 2) ? im : im - ORG_MB, jrm = (jm < ORG_MB / 2) ? jm : jm - ORG_MB;
            float2 vxk, vyk, vxm, vym;
            velocity_mode(by_label(P.visc, lab[ek]), P.friction, kscale * ir, kscale * jr, pxx, pxy, B, vxk, vyk);
            velocity_mode(by_label(P.visc, lab[em]), P.friction, kscale * irm, kscale * jrm, pxxm, pxym, Bm, vxm, vym);
            const float2 hx = make_float2(0.5f * (vxk.x + vxm.x), 0.5f * (vxk.y - vxm.y));
            const float2 hy = make_float2(0.5f * (vyk.x + vym.x), 0.5f * (vyk.y - vym.y));
            sm[lk] = make_float2(hx.x - hy.y, hx.y + hy.x);
            if (em != ek) sm[lm] = make_float2(hx.x + hy.y, hy.x - hx.y);
        };
#pragma unroll 1
        for (int m0 = 0; m0 < 8; m0 += 4) {
            float2 A[4], Am[4];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                int i, j;
                pair_ij(t + 1024 * (m0 + k), i, j);
                const int im = (ORG_MB - i) & 127, jm = (ORG_MB - j) & 127;
                A[k] = ldb2(SA, 8u * (unsigned)(i * ORG_MB + j));
                Am[k] = ldb2(SA, 8u * (unsigned)(im * ORG_MB + jm));
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int k = 0; k < 4; k++) solve_pair(t + 1024 * (m0 + k), A[k], Am[k]);
        }
        const int u8 = t + 1024 * 8;
        if (u8 < NPAIR) {
            int i, j;
            pair_ij(u8, i, j);
            const int im = (ORG_MB - i) & 127, jm = (ORG_MB - j) & 127;
            solve_pair(u8, SA[i * ORG_MB + j], SA[im * ORG_MB + jm]);
        }
    }
    __syncthreads();
    /* P5, fused into the inverse transform's last stage: v = Re(.)/M^2 ;
     * rho * v (rho_functions.cpp:267-284), rho from SR (P1) */
    fft2d_128<1>(sm, tw, [&](const float2 *u, const int sub, const int j) {
#pragma unroll
        for (int q0 = 0; q0 < 16; q0 += 8) {      /* in halves: 8 loads in flight */
            float rr[8];
#pragma unroll
            for (int q = 0; q < 8; q++) rr[q] = SR[fft_out_pos(sub, q0 + q) * ORG_MB + j];
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const int i = fft_out_pos(sub, q0 + q), e = i * ORG_MB + j;
                const float2 z = u[q0 + q];
                if (write_v) {
                    vxbox[(size_t)n * MM + e] = z.x / MMf;
                    vybox[(size_t)n * MM + e] = z.y / MMf;
                }
                sm[i * ORG_LS + j] = make_float2(rr[q] * z.x / MMf, rr[q] * z.y / MMf);
            }
        }
    });
    }                     /* the solve (exact velocity) */
    __syncthreads();
    fft2d_128<-1>(sm, tw);
    /* P6: spectral divergence of update_fft_rho (rho_functions.cpp:230-249)
     * per pair, the part whose inverse is real, in place */
    const float cdt = P.dt * kscale;
#pragma unroll 4
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
    /* P7, fused into the inverse transform's last stage: rho^{n+1} on the
     * box (copy_to_rhon, rho_functions.cpp:253-265) */
    fft2d_128<1>(sm, tw, [&](const float2 *u, const int sub, const int j) {
#pragma unroll
        for (int q0 = 0; q0 < 16; q0 += 8) {      /* in halves: 8 loads in flight */
            float rb[8];
#pragma unroll
            for (int q = 0; q < 8; q++) rb[q] = rbox[fft_out_pos(sub, q0 + q) * ORG_MB + j];
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int q = 0; q < 8; q++)
                rbox[fft_out_pos(sub, q0 + q) * ORG_MB + j] = rb[q] + u[q0 + q].x / MMf;
        }
    });
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
    const int n = blockIdx.x;
    spectral_body<false>(P, n, sm, tw, lab, phi, rho, lum, ecm, xc[n] - P.zs + 1, yc[n] - P.zs + 1, scratch,
                         rho_box, vxbox, vybox, write_v, nullptr);
}

/* ---- the velocity surrogate's side of the step --------------------------- */
/* (phi_n, rho_n) on each cell's FFT box (origin centre - zone_size + 1):
 * out [ncell][2][M*M]. */
__global__ __launch_bounds__(1024) void k_crops(DevP P, const float *__restrict__ phi, const float *__restrict__ rho,
                                                const int *__restrict__ xc, const int *__restrict__ yc,
                                                float *__restrict__ out) {
    constexpr int MM = ORG_MB * ORG_MB;
    const int n = blockIdx.x, N = P.N;
    const size_t NN = (size_t)N * N;
    const int x0 = xc[n] - P.zs + 1, y0 = yc[n] - P.zs + 1;
    float *o = out + (size_t)n * 2 * MM;
    for (int e = threadIdx.x; e < MM; e += 1024) {
        const size_t g = n * NN + (size_t)(x0 + (e >> 7)) * N + (y0 + (e & 127));
        o[e] = phi[g];
        o[MM + e] = rho[g];
    }
}

/* The box velocities of the last solve, [ncell][2][M*M] (vx then vy). */
__global__ __launch_bounds__(1024) void k_pack_v(const float *__restrict__ vxbox, const float *__restrict__ vybox,
                                                 float *__restrict__ out) {
    constexpr int MM = ORG_MB * ORG_MB;
    const int n = blockIdx.x;
    for (int e = threadIdx.x; e < MM; e += 1024) {
        out[(size_t)n * 2 * MM + e] = vxbox[(size_t)n * MM + e];
        out[(size_t)n * 2 * MM + MM + e] = vybox[(size_t)n * MM + e];
    }
}

/* use[n] = val for every cell, *count = ncell * val (a fill without
 * hipMemsetAsync). */
__global__ void k_fill_use(int *__restrict__ use, int nc, int val, int *__restrict__ count, float *__restrict__ res) {
    for (int n = threadIdx.x; n < nc; n += blockDim.x) {
        use[n] = val;
        if (res) res[n] = -1.0f;
    }
    if (threadIdx.x == 0 && count) *count = val ? 0 : nc;
}

/* Judge of a given velocity v ([ncell][2][M*M]) on each cell's box: the
 * finite-difference residual of the force balance the per-mode solve
 * inverts (velocity_mode: xi v - eta (lap v + grad div v) = -div pi, i.e.
 * in Fourier space (xi + eta (2 kx^2 + ky^2)) vx + eta kx ky vy =
 * -i (kx pi_xx + ky pi_xy) and its y counterpart), with central differences
 * on the periodic box and the viscosity of each pixel's label,
 *   res = |xi v - eta (lap v + grad div v) + div pi| / |div pi|   (L2 over
 * the box).  use[n] = res <= thr; cells over the threshold add one to
 * *nfall.  pi goes through the cell's spectral scratch. */
__global__ __launch_bounds__(1024) void k_vresid(DevP P, const float *__restrict__ phi, const float *__restrict__ rho,
                                                 const float *__restrict__ lum, const float *__restrict__ ecm,
                                                 const int *__restrict__ xc, const int *__restrict__ yc,
                                                 float2 *__restrict__ scratch, const float *__restrict__ v,
                                                 float thr, float *__restrict__ res, int *__restrict__ use,
                                                 int *__restrict__ nfall) {
    constexpr int MM = ORG_MB * ORG_MB;
    __shared__ unsigned char lab[MM];
    __shared__ double red[16];
    const int n = blockIdx.x, t = threadIdx.x, N = P.N;
    const size_t NN = (size_t)N * N;
    const int n1 = (n == 0) ? P.nc - 1 : n - 1, n2 = (n == P.nc - 1) ? 0 : n + 1;
    const int x0 = xc[n] - P.zs + 1, y0 = yc[n] - P.zs + 1;
    const float *ph = phi + n * NN, *ph1 = phi + n1 * NN, *ph2 = phi + n2 * NN, *rh = rho + n * NN;
    float2 *SA = scratch + (size_t)n * 2 * MM;
    float *SB = (float *)(SA + MM);
    const float inv4 = 1.0f / (4.0f * P.dx * P.dx);
    for (int e = t; e < MM; e += 1024) {
        const int i = e >> 7, j = e & 127;
        const size_t g = (size_t)(x0 + i) * N + (y0 + j);
        const float vv = ph[g], dxp = ph[g + N] - ph[g - N], dyp = ph[g + 1] - ph[g - 1];
        const float ng = hw_sqrt(inv4 * (dxp * dxp + dyp * dyp));
        const int idx = turnover_label(vv, lum[g], ecm[g], ph1[g], ph2[g]);
        lab[e] = (unsigned char)idx;
        float pxx, pxy, pyv;
        active_stress(P, idx, rh[g], ng, dxp, dyp, inv4, pxx, pxy, pyv);
        SA[e] = make_float2(pxx, pxy);
        SB[e] = pyv;
    }
    __syncthreads();      /* the workgroup reads back its own stores */
    const float *vx = v + (size_t)n * 2 * MM, *vy = vx + MM;
    const float h1 = 1.0f / (2.0f * P.dx), h2 = 1.0f / (P.dx * P.dx), h11 = 1.0f / (4.0f * P.dx * P.dx);
    double rr = 0.0, ff = 0.0;
    for (int e = t; e < MM; e += 1024) {
        const int i = e >> 7, j = e & 127;
        const int ip = ((i + 1) & 127) << 7, im = ((i - 1) & 127) << 7, ic = i << 7;
        const int jp = (j + 1) & 127, jm = (j - 1) & 127;
        const float eta = by_label(P.visc, lab[e]);
        const float vxc = vx[e], vyc = vy[e];
        const float dxx_x = (vx[ip + j] - 2.0f * vxc + vx[im + j]) * h2, dyy_x = (vx[ic + jp] - 2.0f * vxc + vx[ic + jm]) * h2;
        const float dxx_y = (vy[ip + j] - 2.0f * vyc + vy[im + j]) * h2, dyy_y = (vy[ic + jp] - 2.0f * vyc + vy[ic + jm]) * h2;
        const float dxy_x = (vx[ip + jp] - vx[ip + jm] - vx[im + jp] + vx[im + jm]) * h11;
        const float dxy_y = (vy[ip + jp] - vy[ip + jm] - vy[im + jp] + vy[im + jm]) * h11;
        const float fx = (SA[ip + j].x - SA[im + j].x) * h1 + (SA[ic + jp].y - SA[ic + jm].y) * h1;
        const float fy = (SA[ip + j].y - SA[im + j].y) * h1 + (SB[ic + jp] - SB[ic + jm]) * h1;
        const float rx = P.friction * vxc - eta * (2.0f * dxx_x + dyy_x + dxy_y) + fx;
        const float ry = P.friction * vyc - eta * (dxx_y + 2.0f * dyy_y + dxy_x) + fy;
        rr += (double)rx * rx + (double)ry * ry;
        ff += (double)fx * fx + (double)fy * fy;
    }
    rr = block_sum_d(rr, red);
    ff = block_sum_d(ff, red);
    if (t == 0) {
        const float q = ff > 0.0 ? (float)sqrt(rr / ff) : (rr > 0.0 ? INFINITY : 0.0f);
        res[n] = q;
        const int ok = q <= thr;
        use[n] = ok;
        if (!ok) atomicAdd(nfall, 1);
    }
}

/* The spectral judge of a given velocity v ([ncell][2][M*M]): the relative
 * L2 distance, in Fourier space, between the candidate and the solution of
 * the per-mode friction-viscosity system update_rho_velocity solves
 * (rho_functions.cpp:133-190), with the viscosity of the mode's pixel
 * (:149-150) and the part whose inverse is real (the reference keeps Re,
 * :279):
 *   res = |F(v) - H(M^-1 b)| / |H(M^-1 b)|,  b = -i k . F(pi),
 * summed over all modes of both components (by Parseval the real-space
 * relative L2 error of v).  The exact solve's own velocity scores the
 * rounding of its transforms (~1e-6), a zero velocity 1.  It takes the
 * active stress pass, F(pi_xx + i pi_xy), F(pi_yy) and F(v_x + i v_y):
 * three of the five transforms of the exact solve.  use[n] = res <= thr;
 * cells over the threshold add one to *nfall. */
__global__ __launch_bounds__(1024) void k_vresid_spec(DevP P, const float *__restrict__ phi,
                                                      const float *__restrict__ rho, const float *__restrict__ lum,
                                                      const float *__restrict__ ecm, const int *__restrict__ xc,
                                                      const int *__restrict__ yc, float2 *__restrict__ scratch,
                                                      const float *__restrict__ v, float thr, float *__restrict__ res,
                                                      int *__restrict__ use, int *__restrict__ nfall) {
    constexpr int MM = ORG_MB * ORG_MB;
    __shared__ float2 sm[ORG_MB * ORG_LS];
    __shared__ float2 tw[128];
    __shared__ unsigned char lab[MM];
    __shared__ double red[16];
    const int n = blockIdx.x, t = threadIdx.x, N = P.N;
    const size_t NN = (size_t)N * N;
    const int n1 = (n == 0) ? P.nc - 1 : n - 1, n2 = (n == P.nc - 1) ? 0 : n + 1;
    const int x0 = xc[n] - P.zs + 1, y0 = yc[n] - P.zs + 1;
    const float *ph = phi + n * NN, *ph1 = phi + n1 * NN, *ph2 = phi + n2 * NN, *rh = rho + n * NN;
    float2 *SA = scratch + (size_t)n * 2 * MM;         /* F(pi_xx + i pi_xy) */
    float *SB = (float *)(SA + MM);                    /* pi_yy, then (as float2) F(pi_yy) */
    float2 *SY = (float2 *)SB;
    const float inv4 = 1.0f / (4.0f * P.dx * P.dx);
    const float kscale = 6.28318530718f / (ORG_MB * P.dx);
    if (t < 128) tw[t] = make_float2(kCos128[t], kSin128[t]);
    for (int e = t; e < MM; e += 1024) {
        const int i = e >> 7, j = e & 127;
        const size_t g = (size_t)(x0 + i) * N + (y0 + j);
        const float vv = ph[g], dxp = ph[g + N] - ph[g - N], dyp = ph[g + 1] - ph[g - 1];
        const float ng = hw_sqrt(inv4 * (dxp * dxp + dyp * dyp));
        const int idx = turnover_label(vv, lum[g], ecm[g], ph1[g], ph2[g]);
        lab[e] = (unsigned char)idx;
        float pxx, pxy, pyv;
        active_stress(P, idx, rh[g], ng, dxp, dyp, inv4, pxx, pxy, pyv);
        sm[i * ORG_LS + j] = make_float2(pxx, pxy);
        SB[e] = pyv;
    }
    __syncthreads();
    /* F(pi_xx + i pi_xy) to SA; pi_yy into the plane for the next transform */
    fft2d_128<-1>(sm, tw, [&](const float2 *u, const int sub, const int j) {
#pragma unroll
        for (int q = 0; q < 16; q++) {
            const int i = fft_out_pos(sub, q);
            SA[i * ORG_MB + j] = u[q];
            sm[i * ORG_LS + j] = make_float2(SB[i * ORG_MB + j], 0.0f);
        }
    });
    fft2d_128<-1>(sm, tw);
    __syncthreads();
    for (int e = t; e < MM; e += 1024) SY[e] = sm[(e >> 7) * ORG_LS + (e & 127)];
    __syncthreads();
    const float *vx = v + (size_t)n * 2 * MM, *vy = vx + MM;
    for (int e = t; e < MM; e += 1024) sm[(e >> 7) * ORG_LS + (e & 127)] = make_float2(vx[e], vy[e]);
    __syncthreads();
    fft2d_128<-1>(sm, tw);
    __syncthreads();
    double err = 0.0, nrm = 0.0;
    auto split = [](float2 a, float2 am, float2 &re_part, float2 &im_part) {
        /* X = F(p + i q) at k and -k -> F(p)(k), F(q)(k) */
        re_part = make_float2(0.5f * (a.x + am.x), 0.5f * (a.y - am.y));
        im_part = make_float2(0.5f * (a.y + am.y), 0.5f * (am.x - a.x));
    };
    for (int e = t; e < MM; e += 1024) {
        const int i = e >> 7, j = e & 127, im = (ORG_MB - i) & 127, jm = (ORG_MB - j) & 127;
        const int em = im * ORG_MB + jm;
        float2 pxx, pxy, pxxm, pxym, cx, cy;
        split(SA[e], SA[em], pxx, pxy);
        split(SA[em], SA[e], pxxm, pxym);
        split(sm[i * ORG_LS + j], sm[im * ORG_LS + jm], cx, cy);
        /* each mode with its own wave numbers (rho_functions.cpp:159-165) */
        const int ir = (i < ORG_MB / 2) ? i : i - ORG_MB, jr = (j < ORG_MB / 2) ? j : j - ORG_MB;
        const int irm = (im < ORG_MB / 2) ? im : im - ORG_MB, jrm = (jm < ORG_MB / 2) ? jm : jm - ORG_MB;
        float2 vxk, vyk, vxm, vym;
        velocity_mode(by_label(P.visc, lab[e]), P.friction, kscale * ir, kscale * jr, pxx, pxy, SY[e], vxk, vyk);
        velocity_mode(by_label(P.visc, lab[em]), P.friction, kscale * irm, kscale * jrm, pxxm, pxym, SY[em], vxm,
                      vym);
        /* the spectrum of the exact solve's real velocity: H v(k) = (v(k) +
         * conj v(-k)) / 2 (its inverse transform / M^2 is the velocity on the
         * box, rho_functions.cpp:279) */
        const float2 hx = make_float2(0.5f * (vxk.x + vxm.x), 0.5f * (vxk.y - vxm.y));
        const float2 hy = make_float2(0.5f * (vyk.x + vym.x), 0.5f * (vyk.y - vym.y));
        const double dxr = (double)cx.x - hx.x, dxi = (double)cx.y - hx.y;
        const double dyr = (double)cy.x - hy.x, dyi = (double)cy.y - hy.y;
        err += dxr * dxr + dxi * dxi + dyr * dyr + dyi * dyi;
        nrm += (double)hx.x * hx.x + (double)hx.y * hx.y + (double)hy.x * hy.x + (double)hy.y * hy.y;
    }
    err = block_sum_d(err, red);
    nrm = block_sum_d(nrm, red);
    if (t == 0) {
        const float q = nrm > 0.0 ? (float)sqrt(err / nrm) : (err > 0.0 ? INFINITY : 0.0f);
        res[n] = q;
        const int ok = q <= thr;
        use[n] = ok;
        if (!ok) atomicAdd(nfall, 1);
    }
}

/* ======================================================================== */
/* Phase field of one cell, one 1024-thread workgroup per cell:
 * update_all_cells (cell_phase_functions.cpp:126-217) and reshaping_cells
 * (cell_phase_functions.cpp:447-580).  The relaxed field (window plus a
 * one-pixel ring) and the iterate live in LDS, the normals (from the
 * start-of-step phi, useful_functions.cpp:356-388) in a per-cell scratch;
 * the stop test runs in the workgroup.  (x0, y0): window origin. */

/* Phase-field update on the window; result to the LDS iterate pnl (row
 * stride W + 2, one-pixel ring) or, with pnl == nullptr, to pw.  Planes are
 * addressed with 32-bit unsigned offsets from their (uniform) base, so the
 * 36 loads of a pixel take one register each and issue together. */
template <bool SHARED>
static __device__ void cells_update(const DevP &P, const int x0, const int y0, const float *__restrict__ ph,
                                    const float *__restrict__ rh, const float *__restrict__ hf,
                                    const float *__restrict__ psi, const float *__restrict__ hlum,
                                    const float *__restrict__ hecm, const double *__restrict__ lbrk,
                                    const float fn, float *__restrict__ pnl, float *__restrict__ pw,
                                    float *__restrict__ jx = nullptr, float *__restrict__ jy = nullptr) {
    constexpr int W = ORG_W, WH = ORG_WH, NPX = ORG_W * ORG_W;
    const unsigned N = (unsigned)P.N;
    const float inv4 = 1.0f / (4.0f * P.dx * P.dx);
    const double inv4d = 1.0 / (4.0 * (double)P.dx * (double)P.dx), itau = 1.0 / (double)P.tau;
    const float isig2 = 1.0f / (2.0f * P.sigma * P.sigma);
    const float cplc = P.cpl / P.tau * (P.dt / P.dx / P.dx / 4.0f) / (2.50662827463f * P.sigma);
#pragma unroll 1
    for (int e = threadIdx.x; e < NPX; e += 1024) {
        const int i = e / W, j = e - i * W;
        const unsigned g = 4u * ((unsigned)(x0 + i) * N + (unsigned)(y0 + j)), r = 4u * N;
        /* all loads first, in one basic block, so they issue together: with
         * SHARED the lumen/ECM part of the bracket comes from lbrk (28
         * loads), otherwise it is formed here from h(lum), h(ECM) (45) */
        const unsigned o[9] = {g, g + r, g - r, g + 4, g - 4, g + r + 4, g + r - 4, g - r - 4, g - r + 4};
        float fp[5], fr[4], fs[9], fh[9], fl[9], fe[9];
        double lb = 0.0;
#pragma unroll
        for (int k = 0; k < 5; k++) fp[k] = ldb(ph, o[k]);
#pragma unroll
        for (int k = 0; k < 4; k++) fr[k] = ldb(rh, o[k + 1]);
#pragma unroll
        for (int k = 0; k < 9; k++) { fs[k] = ldb(psi, o[k]); fh[k] = ldb(hf, o[k]); }
        if (SHARED) {
            lb = *(const double *)((const char *)lbrk + 2u * g);
        } else {
#pragma unroll
            for (int k = 0; k < 9; k++) { fl[k] = ldb(hlum, o[k]); fe[k] = ldb(hecm, o[k]); }
        }
        __builtin_amdgcn_sched_barrier(0);   /* keep the loads ahead of their uses */
        if (!SHARED) lb = lumecm_bracket(P, fl, fe, inv4d);
        const float v = fp[0];
        const float dxp = fp[1] - fp[2], dyp = fp[3] - fp[4];
        const float drx = fr[0] - fr[1], dry = fr[2] - fr[3];
        const float ng = hw_sqrt(inv4 * (dxp * dxp + dyp * dyp));
        const float d = v - 0.5f;
        const float cterm = cplc * hw_exp(-d * d * isig2) * (drx * dxp + dry * dyp) * hw_rcp(ng);
        float out = (ng > 0.0f) ? v - cterm : v;
        /* temp_phase = psi - hfunc_n (cell_phase_functions.cpp:150) on the 3x3 stencil */
        float tt[9];
#pragma unroll
        for (int k = 0; k < 9; k++) tt[k] = fs[k] - fh[k];
        /* the bracket and the increment in double, as in the reference's
         * FP32 build (its double literals promote them); phin is float */
        const double br = (double)(fn - P.beta * tt[0]) + (double)P.eta * l9d(tt) * inv4d + lb;
        out = (float)((double)out + itau * (double)v * (1.0 - (double)v) * (double)P.dt * br);
        if (pnl) pnl[(i + 1) * WH + (j + 1)] = out;
        else pw[e] = out;
        if (jx) {     /* the normals of reshaping_cells at this pixel (see cells_normals) */
            const float h = 1.0f / (2.0f * P.dx), rng = hw_rcp(ng);
            const bool nz = ng > 0.0f;
            jx[(i + 1) * WH + (j + 1)] = nz ? h * dxp * rng : 0.0f;
            jy[(i + 1) * WH + (j + 1)] = nz ? h * dyp * rng : 0.0f;
        }
    }
}

/* cells_normals on the one-pixel ring around the window only (the window's
 * normals come from cells_update). */
static __device__ void cells_normals_ring(const DevP &P, const int x0, const int y0, const float *__restrict__ ph,
                                          float *__restrict__ jx, float *__restrict__ jy,
                                          float *__restrict__ phase) {
    constexpr int WH = ORG_WH, NR = 4 * WH - 4;
    const unsigned N = (unsigned)P.N, r = 4u * N;
    const float inv4 = 1.0f / (4.0f * P.dx * P.dx), h = 1.0f / (2.0f * P.dx);
    for (int q = threadIdx.x; q < NR; q += blockDim.x) {
        int a, b;
        if (q < WH) { a = 0; b = q; }
        else if (q < 2 * WH) { a = WH - 1; b = q - WH; }
        else if (q < 3 * WH - 2) { a = q - 2 * WH + 1; b = 0; }
        else { a = q - (3 * WH - 2) + 1; b = WH - 1; }
        const unsigned g = 4u * ((unsigned)(x0 + a - 1) * N + (unsigned)(y0 + b - 1));
        const float c0 = ldb(ph, g), dxp = ldb(ph, g + r) - ldb(ph, g - r), dyp = ldb(ph, g + 4) - ldb(ph, g - 4);
        const float ng = hw_sqrt(inv4 * (dxp * dxp + dyp * dyp));
        const float rng = hw_rcp(ng);
        const bool nz = ng > 0.0f;
        const int e = a * WH + b;
        jx[e] = nz ? h * dxp * rng : 0.0f;
        jy[e] = nz ? h * dyp * rng : 0.0f;
        phase[e] = c0;
    }
}

/* Normals on the window plus ring from the start-of-step phi (to jx, jy,
 * row stride W + 2) and the ring of the relaxed field (to phase).  Pixels in
 * batches of 4 per thread, their 20 loads issued before any use. */
static __device__ void cells_normals(const DevP &P, const int x0, const int y0, const float *__restrict__ ph,
                                     float *__restrict__ jx, float *__restrict__ jy, float *__restrict__ phase) {
    constexpr int W = ORG_W, WH = ORG_WH, NP = WH * WH, B = 4;
    const unsigned N = (unsigned)P.N, r = 4u * N;
    const float inv4 = 1.0f / (4.0f * P.dx * P.dx), h = 1.0f / (2.0f * P.dx);
#pragma unroll 1
    for (int e0 = threadIdx.x; e0 < NP; e0 += B * 1024) {
        float c0[B], xp[B], xm[B], yp[B], ym[B];
#pragma unroll
        for (int k = 0; k < B; k++) {
            const int e = min(e0 + 1024 * k, NP - 1);
            const int i = e / WH - 1, j = e - (i + 1) * WH - 1;
            const unsigned g = 4u * ((unsigned)(x0 + i) * N + (unsigned)(y0 + j));
            c0[k] = ldb(ph, g); xp[k] = ldb(ph, g + r); xm[k] = ldb(ph, g - r);
            yp[k] = ldb(ph, g + 4); ym[k] = ldb(ph, g - 4);
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int k = 0; k < B; k++) {
            const int e = e0 + 1024 * k;
            if (e >= NP) break;
            const int i = e / WH - 1, j = e - (i + 1) * WH - 1;
            const float dxp = xp[k] - xm[k], dyp = yp[k] - ym[k];
            const float ng = hw_sqrt(inv4 * (dxp * dxp + dyp * dyp));
            const float rng = hw_rcp(ng);
            const bool nz = ng > 0.0f;
            jx[e] = nz ? h * dxp * rng : 0.0f;
            jy[e] = nz ? h * dyp * rng : 0.0f;
            if (i < 0 || j < 0 || i >= W || j >= W) phase[e] = c0[k];
        }
    }
}

/* Reshaping loop until the deviation test passes (cell_phase_functions.cpp:
 * 470-575); phi <- second-to-last iterate (still in phase), phin <- last.
 * With pwin the window goes to a per-cell buffer that the copy-back writes
 * to phi, so other workgroups of the same launch still read the
 * start-of-step phi.  Expects pnl and the phase ring written and a barrier
 * after them. */
static __device__ void cells_relax(const DevP &P, const int n, const int x0, const int y0, float *__restrict__ ph,
                                   float *__restrict__ phase, float *__restrict__ pnl, const float *__restrict__ jx,
                                   const float *__restrict__ jy, double *__restrict__ red, float *__restrict__ pw,
                                   float *__restrict__ pwin, int *__restrict__ counts, int *__restrict__ err) {
    constexpr int W = ORG_W, WH = ORG_WH, NPX = ORG_W * ORG_W;
    const int t = threadIdx.x, N = P.N;
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
        /* flux update (:498-549); 4 pixels per thread at a time, their 40
         * normal loads issued before any use */
#pragma unroll 1
        for (int e0 = t; e0 < NPX; e0 += 4 * 1024) {
            float nx[4][5], ny[4][5];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int e = min(e0 + 1024 * k, NPX - 1);
                const int i = e / W, l = (i + 1) * WH + (e - i * W + 1);
                const unsigned b = 4u * (unsigned)l, r = 4u * WH;
                nx[k][0] = ldb(jx, b); nx[k][1] = ldb(jx, b + r); nx[k][2] = ldb(jx, b - r);
                nx[k][3] = ldb(jx, b + 4); nx[k][4] = ldb(jx, b - 4);
                ny[k][0] = ldb(jy, b); ny[k][1] = ldb(jy, b + r); ny[k][2] = ldb(jy, b - r);
                ny[k][3] = ldb(jy, b + 4); ny[k][4] = ldb(jy, b - 4);
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int e = e0 + 1024 * k;
                if (e < NPX) {
                    const int i = e / W, l = (i + 1) * WH + (e - i * W + 1);
                    const float u = pnl[l];
                    const float nv = reshape_flux_j(u, phase + l, WH, nx[k], ny[k], ka, kc);
                    acc += (double)(fabsf(nv - u) * dscale);
                    pnl[l] = nv;
                }
            }
        }
        rdev = block_sum_d(acc, red);                  /* (:554-562) */
        cnt++;
        if (ri > P.rimax) { if (t == 0) set_err(err, ORG_ERELAX); break; }
        ri++;
    }
    for (int e = t; e < NPX; e += 1024) {
        const int i = e / W, j = e - i * W, l = (i + 1) * WH + (j + 1);
        if (pwin) pwin[e] = phase[l];
        else ph[(size_t)(x0 + i) * N + (y0 + j)] = phase[l];
        pw[e] = pnl[l];
    }
    if (t == 0) counts[n] = cnt;
}

/* k_cells: update_all_cells (mode&1) and / or reshaping_cells (mode&2) of
 * every cell; mode 1 leaves phin in phin_win, mode 2 starts from it, mode
 * 2|4 starts from phi itself. */
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
    constexpr int W = ORG_W, WH = ORG_WH, NPX = ORG_W * ORG_W;
    const int n = blockIdx.x;
    const size_t NN = (size_t)P.N * P.N;
    const int x0 = xc[n] - P.zs, y0 = yc[n] - P.zs;
    float *ph = phi + n * NN;
    float *pw = phin_win + (size_t)n * NPX;
    if (mode & 1) {
        cells_update<false>(P, x0, y0, ph, rho + n * NN, hfunc + n * NN, psi, hlum, hecm, nullptr, f0[n],
                            (mode & 2) ? pnl : nullptr, pw);
        if (!(mode & 2)) return;
    } else if (mode & 4) {  /* phin = phi (the IC program's copy, prepare_initial_files.cpp:87-90) */
        for (int e = threadIdx.x; e < NPX; e += 1024) {
            const int i = e / W, j = e - i * W;
            pnl[(i + 1) * WH + (j + 1)] = ph[(size_t)(x0 + i) * P.N + (y0 + j)];
        }
    } else {
        for (int e = threadIdx.x; e < NPX; e += 1024) {
            const int i = e / W, j = e - i * W;
            pnl[(i + 1) * WH + (j + 1)] = pw[e];
        }
    }
    float *jx = jwin + (size_t)n * 2 * WH * WH, *jy = jx + WH * WH;
    cells_normals(P, x0, y0, ph, jx, jy, phase);
    __syncthreads();
    cells_relax(P, n, x0, y0, ph, phase, pnl, jx, jy, red, pw, nullptr, counts, err);
}

/* Arguments of k_pre, k_step and k_post. */
struct StepArgs {
    /* per-cell planes and scalars; xn, yn, bn: centres and support boxes of
     * the next step, from k_post (or k_centroid before the first step) */
    float *phi, *rho, *hfunc;
    int *xc, *yc, *box, *xn, *yn, *bn, *uni, *counts;
    int *xp, *yp;            /* copy of this step's centres for k_mid (whose preps overwrite xc, yc) */
    const double *s0;
    double *s;
    float *f0;
    /* shared planes */
    float *lum, *ecm;
    float *psi, *hlum, *hecm;
    /* lumen/ECM part of the cell update's bracket: k_step's phase-field role
     * reads it, the copy-back launch after it (k_mid, k_post) writes the
     * next step's */
    double *lbrk;
    int shared_brk;          /* 1: the copy-back launches write lbrk and the phase-field role reads it */
    /* 1 once the ring part of the h(lum)/h(ECM) pass has been taken from a
     * committed batched step (k_live_reset clears it) */
    int *ringh;
    double *part_h, *sum_ecm;
    int nh;                  /* h(lum), h(ECM) workgroups of k_pre */
    int npsi;                /* psi workgroups of k_pre */
    int *live, *live_part;   /* live boxes of lum and ECM; per-workgroup boxes of live pixels from k_post */
    int nlp;                 /* k_post's lumen/ECM workgroups (partial boxes per field) */
    int nlum;                /* lumen/ECM workgroups of k_step (at most ORG_LUM_MAX) */
    /* lumen / ECM loops */
    float *L0, *L1, *E0, *E1, *jlx, *jly, *jex, *jey;
    int *jbox_part;
    LumState *lst;
    double *part_lc;         /* loop deviations: [2 spans][2 fields][ORG_LUM_SPAN sweeps][ORG_LUM_MAX] */
    float *xch;              /* tagged exchange words of the LDS loops (ORG_XCH_WORDS, 64-bit) */
    /* cortex */
    float2 *spec;
    float *rho_box, *vxbox, *vybox;
    int write_v;
    /* surrogate velocities [ncell][2][M*M] and per-cell switch (1: use
     * vgiven instead of the solve); null: every cell solves */
    const float *vgiven;
    const int *vuse;
    /* phase field */
    float *phin_win, *phi_win, *jwin;
    /* barrier of the lumen/ECM workgroups, abort flag */
    unsigned *bar;
    int *abort_flag, *err;
    long long limit;
};

/* Live box of field f merged with the partial boxes of live pixels that the
 * last copy-back found (wave 0 reduces them); valid in every thread.  sb: 4
 * ints of LDS. */
/* The live boxes of lumen (to sb[0..3]) and ECM (sb[4..7]): the stored ones
 * merged with the copy-back's partial boxes, wave f reducing field f.  No
 * barrier inside: the caller's next workgroup barrier publishes sb. */
static __device__ void merged_live_both(const DevP &P, const StepArgs &a, int *sb) {
    const int N = P.N, f = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (f > 1) return;
    int c0 = N, c1 = -1, c2 = N, c3 = -1;
    for (int k = lane; k < a.nlp; k += 64) {
        const int *q = a.live_part + ((size_t)f * a.nlp + k) * 4;
        if (q[1] >= 0) { c0 = min(c0, q[0]); c1 = max(c1, q[1]); c2 = min(c2, q[2]); c3 = max(c3, q[3]); }
    }
    for (int o = 32; o > 0; o >>= 1) {
        c0 = min(c0, __shfl_xor(c0, o, 64)); c1 = max(c1, __shfl_xor(c1, o, 64));
        c2 = min(c2, __shfl_xor(c2, o, 64)); c3 = max(c3, __shfl_xor(c3, o, 64));
    }
    if (lane == 0) {
        int4 r = ld_box(a.live, f);
        if (c1 >= 0) {
            const int4 q = make_int4(c0, c1 + 1, c2, c3 + 1);
            r = box_empty(r) ? q : make_int4(min(r.x, q.x), max(r.y, q.y), min(r.z, q.z), max(r.w, q.w));
        }
        sb[4 * f] = r.x; sb[4 * f + 1] = r.y; sb[4 * f + 2] = r.z; sb[4 * f + 3] = r.w;
    }
}

/* k_pre: the first part of a timestep (run_with_active_gels.cpp:123-142),
 * workgroups in three independent roles:
 *   0 .. ncell-1       cell n: centre and support box from (xn, yn, bn),
 *                      hfunc on the window and the FP64 area
 *                      (compute_hfunc_3d_tables_kernels_loop_n, sum_s_kernels)
 *   next npsi          psi on the union of the windows (compute_psi_kernels),
 *                      with hfun(phi) inside each window
 *   last nh            h(lum), h(ECM) and the parts of the sum of h(ECM)
 *                      (compute_all_hfunc_2d_tables_kernels) that the last
 *                      copy-back may have changed: the ring and, in box mode,
 *                      the tiles meeting a live box grown by one (all tiles in
 *                      full mode). */
#define ORG_LDS_PRE (4 * ORG_PSI_LDS_GROUP + 16)
static __device__ __forceinline__ void pre_cell(DevP P, StepArgs a, const int n, unsigned char *scr) {
    cell_prep(P, n, a.phi, a.hfunc, a.xn, a.yn, a.bn, a.xc, a.yc, a.box, a.s0, a.s, a.f0, (double *)scr);
}
static __device__ __forceinline__ void pre_psi(DevP P, StepArgs a, const int blk, unsigned char *scr) {
    psi_tiles_step(P, blk, a.npsi, a.hfunc, a.phi, a.bn, a.xn, a.yn, nullptr, nullptr, nullptr, a.psi, scr);
}
static __device__ __forceinline__ void pre_hlum(DevP P, StepArgs a, const int hb, unsigned char *scr) {
    const int N = P.N, nt = (N + ORG_HT - 1) / ORG_HT, ntl = nt * nt;
    const int4 BL = ld_box(a.live, 0), BE = ld_box(a.live, 1);
    const bool full = !box_mode(BL, N) || !box_mode(BE, N);
    for (int k = hb; k <= ntl; k += a.nh) {
        bool need = full || k == ntl;
        if (!need) {
            const int ti = (k / nt) * ORG_HT, tj = (k % nt) * ORG_HT;
            for (int f = 0; f < 2 && !need; f++) {
                const int4 B = f ? BE : BL;
                need = !box_empty(B) && ti < B.y + 1 && ti + ORG_HT > B.x - 1 && tj < B.w + 1 &&
                       tj + ORG_HT > B.z - 1;
            }
        }
        if (!need) continue;
        const double tot = block_sum_d(hlum_part(P, k, a.lum, a.ecm, a.hlum, a.hecm), (double *)scr);
        if (threadIdx.x == 0) a.part_h[k] = tot;
    }
}
/* Logical index of block b within a role of C blocks that starts at block
 * B0.  The hardware deals the workgroups of a launch to the 8 XCDs round
 * robin (block b runs on XCD b % 8), each XCD with its own L2; this mapping
 * gives the role's blocks on one XCD a contiguous range of indices, so
 * neighbouring cells (cells are ordered by angle) and neighbouring row bands,
 * which read the same psi, h and lumen/ECM pixels, share an L2.  Identity
 * unless C is a multiple of 8.  Results do not depend on it: every index is
 * still taken by exactly one block. */
#define ORG_NXCD 8
static __device__ __forceinline__ int xcd_index(const int b, const int B0, const int C) {
    if (C % ORG_NXCD) return b - B0;
    return (b % ORG_NXCD) * (C / ORG_NXCD) + (b - B0) / ORG_NXCD;
}

__global__ __launch_bounds__(1024) void k_pre(DevP P, StepArgs a) {
    __shared__ __attribute__((aligned(16))) unsigned char scr[ORG_LDS_PRE];
    const int b = blockIdx.x, nc = P.nc;
    if (b < nc) pre_cell(P, a, xcd_index(b, 0, nc), scr);
    else if (b < nc + a.npsi) pre_psi(P, a, xcd_index(b, nc, a.npsi), scr);
    else pre_hlum(P, a, xcd_index(b, nc + a.npsi, a.nh), scr);
}

/* k_step: the rest of the timestep up to the copy-back
 * (run_with_active_gels.cpp:130-172) in one launch, workgroups in three
 * roles:
 *   0 .. ncell-1       spectral cortex update of cell n (update_all_cortices
 *                      with its turnover labels, normals and prod_norm_delta)
 *   next ORG_LUM_BLOCKS  lumen/ECM: update_lumen_and_ECM and the normals of
 *                      reshaping_lumen_or_ECM per row band, barrier, both
 *                      reshaping loops to their stop tests (one barrier per
 *                      sweep)
 *   last ncell         phase field of cell n: update_all_cells and
 *                      reshaping_cells; the new window goes to phi_win; then
 *                      the next step's centre and support box to (xn, yn, bn)
 * Every role reads only the state at the start of the launch and its own
 * results; only the lumen/ECM workgroups synchronise, among themselves.  The
 * copy-back (k_post) runs after the launch.  The FP sums take the same
 * operands in the same order as the per-function kernels, so both forms
 * agree bit for bit. */
#define ORG_LDS_SPECTRAL (ORG_MB * ORG_LS * 8 + 128 * 8 + ORG_MB * ORG_MB)
#define ORG_LDS_CELLS (2 * ORG_WH * ORG_WH * 4 + 16 * 8 + 48 * 8)
#define ORG_LDS_FUSED (ORG_LDS_SPECTRAL > ORG_LDS_CELLS ? ORG_LDS_SPECTRAL : ORG_LDS_CELLS)
static __device__ __forceinline__ void step_spectral(DevP P, StepArgs a, const int b,
                                                     unsigned char *lds) {
    float2 *sm = (float2 *)lds;
    unsigned char *lab = (unsigned char *)(sm + ORG_MB * ORG_LS + 128);
    const int x0 = a.xc[b] - P.zs + 1, y0 = a.yc[b] - P.zs + 1;
    if (a.vuse && a.vuse[b])
        spectral_body<true>(P, b, sm, sm + ORG_MB * ORG_LS, lab, a.phi, a.rho, a.lum, a.ecm, x0, y0, a.spec,
                            a.rho_box, a.vxbox, a.vybox, a.write_v, a.vgiven + (size_t)b * 2 * ORG_MB * ORG_MB);
    else
        spectral_body<false>(P, b, sm, sm + ORG_MB * ORG_LS, lab, a.phi, a.rho, a.lum, a.ecm, x0, y0, a.spec,
                             a.rho_box, a.vxbox, a.vybox, a.write_v, nullptr);
}
/* The batched step's lumen/ECM role in LDS (see the block comment above
 * ORG_LUM_SPAN); false if a barrier timed out. */
static __device__ bool lum_role_lds(const DevP &P, const StepArgs &A, const int blk, const int nblk, const float se,
                                    const int4 BL, const int4 BE, unsigned char *lds, int &epoch, int *cur_out,
                                    const unsigned tag0) {
    constexpr int K = ORG_LUM_SPAN;
    const int N = P.N;
    const size_t NN = (size_t)N * N;
    const float ka = P.wi * P.wi * (P.rdt / P.dx / P.dx);
    const float kc = P.rdt * P.wi * 1.41421356237f / P.dx;
    const float dx2 = P.dx * P.dx;
    double *red = (double *)lds;
    /* loop states as lum_reset_states sets them (workgroup 0 stores them) */
    LumState s[2];
#pragma unroll
    for (int f = 0; f < 2; f++) {
        s[f].done = P.rdev0_lum > P.thr_lum ? 0 : 1;
        s[f].ri = 0; s[f].cur = 0; s[f].pad = 0;
    }
    /* buffer q of field f, selected without an indexed array (which would
     * live in scratch) */
    auto glob = [=](int f, int q) { return f ? (q ? A.E1 : A.E0) : (q ? A.L1 : A.L0); };
    LumBand B[2];
    float *lp = (float *)(lds + ORG_LUM_LDS_BASE);
#pragma unroll
    for (int f = 0; f < 2; f++) {
        LumBand &L = B[f];
        const int4 Bf = f ? BE : BL;
        if (box_empty(Bf)) { L.a = L.b = 0; L.r0 = 1; L.r1 = 0; L.c0 = 1; L.c1 = 0; L.S = 0; L.bw = 0; continue; }
        L.r0 = Bf.x - 1; L.r1 = Bf.y; L.c0 = Bf.z - 1; L.c1 = Bf.w;
        const int bh = L.r1 - L.r0 + 1;
        L.bw = L.c1 - L.c0 + 1;
        L.a = L.r0 + (int)(((long long)blk * bh) / nblk);
        L.b = L.r0 + (int)(((long long)(blk + 1) * bh) / nblk);
        L.S = L.bw + 2;
        const int rows = lum_lds_rows(bh, nblk);
        L.u[0] = lp; L.u[1] = lp + rows * L.S; L.jx = lp + 2 * rows * L.S; L.jy = lp + 3 * rows * L.S;
        L.hist = lp + 4 * rows * L.S;
        lp += 4 * rows * L.S + (K - 2) * lum_own_max(bh, nblk) * L.bw;
        if (L.b <= L.a) continue;
        const float *X = f ? A.ecm : A.lum;
        const int base = L.a - K - 1, w = L.bw + 2;
        /* the update on grid rows [a-K-1, b+K+1) within the domain and its
         * halo; outside the domain the committed field (both buffers hold it) */
        {
            const int i0 = max(base, L.r0 - 1), i1 = min(L.b + K + 1, L.r1 + 2);
#pragma unroll 4
            for (RowWalk q(w); q.r < i1 - i0; q.next()) {
                const int i = i0 + q.r, j = L.c0 - 1 + q.c;
                const size_t e = (size_t)i * N + j;
                const bool in = i >= L.r0 && i <= L.r1 && j >= L.c0 && j <= L.c1;
                const float v = in ? lumecm_px(P, f, e, se, A.lum, A.ecm, A.psi, A.hlum, A.hecm) : X[e];
                const int l = (i - base) * L.S + q.c;
                L.u[0][l] = v; L.u[1][l] = v;
            }
        }
        __syncthreads();
        /* normals on grid rows [a-K, b+K): inside the live box from the old
         * field and the update, zero outside it */
        {
            const int i0 = max(L.a - K, L.r0 - 1), i1 = min(L.b + K, L.r1 + 2);
            const float *un = L.u[0];
#pragma unroll 4
            for (RowWalk q(w); q.r < i1 - i0; q.next()) {
                const int i = i0 + q.r, j = L.c0 - 1 + q.c;
                const int l = (i - base) * L.S + q.c;
                float vx = 0.0f, vy = 0.0f;
                if (i >= Bf.x && i < Bf.y && j >= Bf.z && j < Bf.w)
                    lum_normal_px(P, X, (size_t)i * N + j, N, un[l + L.S], un[l - L.S], un[l + 1], un[l - 1], vx, vy);
                L.jx[l] = vx; L.jy[l] = vy;
            }
        }
    }
    /* the four corners of the boundary ring: every sweep writes 0 there (also
     * for the ECM, useful_functions.cpp:432-435) and counts the change;
     * workgroup 0, thread k owns corner k of both fields and tracks the value
     * each buffer holds, starting from the ring values (lumen 0, ECM 1) */
    const bool corner_thread = blk == 0 && threadIdx.x < 4;
    size_t ce = 0;
    float cval[2][2] = {{0.0f, 0.0f}, {1.0f, 1.0f}};
    if (corner_thread) ce = (size_t)((threadIdx.x & 1) ? N - 1 : 0) * N + ((threadIdx.x & 2) ? N - 1 : 0);
    int stop_t[2] = {K - 1, K - 1}, span_par[2] = {0, 0};
    /* the exchange words: [2 fields][2 span parities][N][N] rows, then
     * [2 span parities][2K deviations][2 halves][ORG_LUM_MAX] partial sums */
    unsigned long long *xrow = (unsigned long long *)A.xch;
    unsigned long long *xdev = xrow + 4 * NN;
    __shared__ int xfail;
    int k = 0;
    __syncthreads();
    for (; !(s[0].done && s[1].done); k++) {
        const int xp = k & 1;
        const unsigned tag = tag0 + (unsigned)k + 1u;
        unsigned long long *pk = xdev + (size_t)xp * 2 * K * 2 * ORG_LUM_MAX;
        /* per-wave sums of each sweep's deviation, [wave][2K] (the fixed
         * order of the workgroup's sum: waves, then lanes by wave_sum_d) */
        double *wred = (double *)(lds + 256);
        const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
        for (int f = 0; f < 2; f++) {
            if (s[f].done) continue;
            const LumBand &L = B[f];
            const int par = s[f].ri & 1;
            span_par[f] = par;
#pragma unroll
            for (int t = 0; t < K; t++) {
                const int p = (par + t) & 1;
                const float *src = p ? L.u[1] : L.u[0];
                float *dst = p ? L.u[0] : L.u[1];
                double acc = 0.0;
                if (L.b > L.a) {
                    const int i0 = max(L.a - (K - 1 - t), L.r0), i1 = min(L.b + (K - 1 - t), L.r1 + 1);
                    acc = lum_band_sweep(L, src, dst, i0, i1, ka, kc, dx2);
                }
                if (corner_thread) {
                    const float cv = p ? cval[f][1] : cval[f][0];
                    acc += (double)(fabsf(0.0f - cv) * dx2);
                    if (p) cval[f][0] = 0.0f; else cval[f][1] = 0.0f;
                }
                /* the wave's part of this sweep's deviation, before the
                 * sweep's barrier (so the workgroup's sum needs none) */
                const double ws = wave_sum_d(acc);
                if (lane == 0) wred[2 * K * wv + f * K + t] = ws;
                /* every thread read the last span's xfail before this
                 * span's first barrier; the waits below come after the last */
                if (t == K - 1 && threadIdx.x == 0) xfail = 0;
                __syncthreads();
                if (t <= K - 3 && L.b > L.a) {
                    /* keep the band's rows of this iterate: sweep t + 2 overwrites them */
                    float *h = L.hist + (size_t)t * (L.b - L.a) * L.bw;
                    for (RowWalk q(L.bw); q.r < L.b - L.a; q.next())
                        h[q.r * L.bw + q.c] = dst[(q.r + K + 1) * L.S + 1 + q.c];
                }
            }
        }
        /* the 2K deviations of this workgroup (fixed-order sums of the wave
         * parts), published first as tagged halves, then the latest
         * iterate's first K and last K rows, tagged */
        if (threadIdx.x < 2 * K && !((threadIdx.x >= K) ? s[1].done : s[0].done)) {
            double t = 0.0;
            for (int ww = 0; ww < (int)(blockDim.x >> 6); ww++) t += wred[2 * K * ww + threadIdx.x];
            const unsigned long long tb = (unsigned long long)__double_as_longlong(t);
            xput(&pk[(threadIdx.x * 2 + 0) * ORG_LUM_MAX + blk], (unsigned)tb, tag);
            xput(&pk[(threadIdx.x * 2 + 1) * ORG_LUM_MAX + blk], (unsigned)(tb >> 32), tag);
        }
#pragma unroll
        for (int f = 0; f < 2; f++) {
            const LumBand &L = B[f];
            if (s[f].done || !(L.b > L.a)) continue;
            const float *lat = span_par[f] ? L.u[1] : L.u[0];
            const int own = L.b - L.a, nr = min(2 * K, own), base = L.a - K - 1;
            unsigned long long *x = xrow + (size_t)(f * 2 + xp) * NN;
            for (int q = threadIdx.x; q < nr * L.bw; q += blockDim.x) {
                const int r = q / L.bw, j = q % L.bw;
                const int gi = (r < K || own <= 2 * K) ? L.a + r : L.b - (nr - r);
                xput(&x[(size_t)gi * N + L.c0 + j], __float_as_uint(lat[(gi - base) * L.S + 1 + j]), tag);
            }
        }
        /* no barrier: wave 0 waits for every workgroup's deviations of this
         * span and sums them in a fixed order; the other waves wait for the
         * neighbours' rows and put them into the halo.  Every word carries
         * the tag of the exchange it belongs to. */
        bool ok = true;
        if (threadIdx.x < 64) {
            double x[2 * K][2], v[2 * K];
#pragma unroll
            for (int i = 0; i < 2 * K; i++) {
                unsigned long long wv[2][2];
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    const int b2 = (int)threadIdx.x + 64 * h;
                    const bool want = !s[i / K].done && b2 < nblk;
#pragma unroll
                    for (int half = 0; half < 2; half++)
                        wv[h][half] = want ? xget(&pk[(i * 2 + half) * ORG_LUM_MAX + b2]) : 0ull;
                }
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    const int b2 = (int)threadIdx.x + 64 * h;
                    x[i][h] = 0.0;
                    if (!s[i / K].done && b2 < nblk) {
#pragma unroll
                        for (int half = 0; half < 2; half++)
                            ok = ok && xwait(&pk[(i * 2 + half) * ORG_LUM_MAX + b2], tag, wv[h][half], A.abort_flag,
                                             A.limit);
                        x[i][h] = __longlong_as_double((long long)(((wv[h][1] & 0xffffffffull) << 32) |
                                                                   (wv[h][0] & 0xffffffffull)));
                    }
                }
            }
#pragma unroll
            for (int i = 0; i < 2 * K; i++) v[i] = s[i / K].done ? 0.0 : wave_sum_d(x[i][0] + x[i][1]);
            if (threadIdx.x == 0) {
#pragma unroll
                for (int i = 0; i < 2 * K; i++) red[i] = v[i];
            }
        } else {
#pragma unroll
            for (int f = 0; f < 2; f++) {
                const LumBand &L = B[f];
                if (s[f].done || !(L.b > L.a)) continue;
                float *lat = (s[f].ri & 1) ? L.u[1] : L.u[0];
                const unsigned long long *x = xrow + (size_t)(f * 2 + xp) * NN;
                const int base = L.a - K - 1, step = (int)blockDim.x - 64, nq = 2 * K * L.bw;
                /* four words per thread in flight, then the waits */
                for (int q0 = (int)threadIdx.x - 64; q0 < nq; q0 += 4 * step) {
                    unsigned long long wv[4];
                    size_t ga[4];
                    int la[4];
#pragma unroll
                    for (int m = 0; m < 4; m++) {
                        const int q = q0 + m * step;
                        la[m] = -1; ga[m] = 0; wv[m] = 0ull;
                        if (q < nq) {
                            const int r = q / L.bw, j = q % L.bw;
                            const int gi = r < K ? L.a - K + r : L.b + (r - K);
                            if (gi >= L.r0 && gi <= L.r1) {
                                la[m] = (gi - base) * L.S + 1 + j;
                                ga[m] = (size_t)gi * N + L.c0 + j;
                                wv[m] = xget(&x[ga[m]]);
                            }
                        }
                    }
#pragma unroll
                    for (int m = 0; m < 4; m++) {
                        if (la[m] < 0) continue;
                        ok = ok && xwait(&x[ga[m]], tag, wv[m], A.abort_flag, A.limit);
                        lat[la[m]] = __uint_as_float((unsigned)wv[m]);
                    }
                }
            }
        }
        if (!ok) xfail = 1;
        __syncthreads();
        if (xfail) {
            /* a wait passed its limit or another workgroup gave up: stop,
             * and move the tags of later steps past every one used here */
            if (threadIdx.x == 0) {
                atomicExch(A.abort_flag, 1);
                set_err(A.err, ORG_ETIMEOUT);
                atomicAdd(&A.bar[1], 1u << 20);
            }
            return false;
        }
#pragma unroll
        for (int f = 0; f < 2; f++) {
            if (s[f].done) continue;
#pragma unroll
            for (int t = 0; t < K; t++) {
                lum_stop(P, s[f], readfirstlane_d(red[f * K + t]), A.err);
                if (s[f].done) { stop_t[f] = t; break; }
            }
        }
        /* red and xfail are next written after the next span's sweep
         * barriers, so no barrier is needed here */
    }
    /* the next step's exchanges use tags after this step's */
    if (blk == 0 && threadIdx.x == 0) atomicAdd(&A.bar[1], (unsigned)k + 1u);
    /* the band's last iterate into the buffer that holds it; the corners */
#pragma unroll
    for (int f = 0; f < 2; f++) {
        const LumBand &L = B[f];
        float *g = glob(f, s[f].cur);
        if (corner_thread) {
            glob(f, 0)[ce] = cval[f][0];
            glob(f, 1)[ce] = cval[f][1];
        }
        if (!(L.b > L.a)) continue;
        const int own = L.b - L.a, base = L.a - K - 1, t = stop_t[f];
        const float *src = nullptr;
        bool hist = false;
        if (s[f].ri == 0) src = L.u[0];                        /* no sweep ran: the update */
        else if (t <= K - 3) hist = true;
        else src = ((span_par[f] + t + 1) & 1) ? L.u[1] : L.u[0];
        for (RowWalk q(L.bw); q.r < own; q.next()) {
            const float v = hist ? L.hist[(size_t)t * own * L.bw + q.r * L.bw + q.c]
                                 : src[(L.a + q.r - base) * L.S + 1 + q.c];
            g[(size_t)(L.a + q.r) * N + L.c0 + q.c] = v;
        }
    }
    if (blk == 0 && threadIdx.x == 0) { A.lst[0] = s[0]; A.lst[1] = s[1]; }   /* s stays in registers */
    cur_out[0] = s[0].cur;
    cur_out[1] = s[1].cur;
    return true;          /* the copy-back launch reads the bands: no barrier here */
}

static __device__ __forceinline__ void step_lumecm(DevP P, StepArgs a, const int blk, const int nblk,
                                                   unsigned char *lds) {
    double *red = (double *)lds;
    int *sb = (int *)(lds + 128);
    int epoch = 0;
    /* the prologue's three reductions overlap: the parts of the h(ECM) sum
     * (every thread's loads, then the fixed-order block sum) and the live
     * boxes of this step, the stored ones plus what the last copy-back found
     * (waves 0 and 1; every workgroup merges the same parts, the first
     * stores the result, and merging again gives the same boxes) */
    /* the exchanges' tag base (a round trip to memory: read first, used
     * after the update) */
    const unsigned tag0 = __hip_atomic_load(&a.bar[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    double hv = 0.0;
    for (int b = threadIdx.x; b < hl_ntiles(P.N) + 1; b += blockDim.x)
        hv += __hip_atomic_load(&a.part_h[b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    merged_live_both(P, a, sb);
    const double se = block_sum_d(hv, red);      /* its barriers publish sb */
    /* the boxes are the same in every thread: readfirstlane keeps them, and
     * everything the loops derive from them, in scalar registers */
    auto u = [&](int v) { return __builtin_amdgcn_readfirstlane(v); };
    const int4 BL = make_int4(u(sb[0]), u(sb[1]), u(sb[2]), u(sb[3]));
    const int4 BE = make_int4(u(sb[4]), u(sb[5]), u(sb[6]), u(sb[7]));
    __syncthreads();      /* sb is reused below */
    if (blk == 0) {
        if (threadIdx.x == 0) *a.sum_ecm = se;
        lum_reset_states(P, a.lst);
    }
    if (blk == 0 && threadIdx.x == 0) {
        a.live[0] = BL.x; a.live[1] = BL.y; a.live[2] = BL.z; a.live[3] = BL.w;
        a.live[4] = BE.x; a.live[5] = BE.y; a.live[6] = BE.z; a.live[7] = BE.w;
    }
    const bool boxes = box_mode(BL, P.N) && box_mode(BE, P.N);
    int cur[2];
    if (boxes && ORG_LUM_LDS_LOOP && lum_lds_fits(BL, BE, nblk, ORG_LDS_FUSED)) {
        lumecm_ring(P, a.L0, a.L1, a.E0, a.E1, blk, nblk, false);
        if (!lum_role_lds(P, a, blk, nblk, (float)se, BL, BE, lds, epoch, cur, tag0)) return;
    } else {
        if (boxes) {
            lumecm_ring(P, a.L0, a.L1, a.E0, a.E1, blk, nblk);
            lumecm_normals_box(P, 0, BL, blk, nblk, (float)se, a.lum, a.ecm, a.psi, a.hlum, a.hecm, a.L0, a.L1,
                               a.jlx, a.jly, a.jbox_part, sb);
            lumecm_normals_box(P, 1, BE, blk, nblk, (float)se, a.lum, a.ecm, a.psi, a.hlum, a.hecm, a.E0, a.E1,
                               a.jex, a.jey, a.jbox_part, sb);
        } else {
            lumecm_normals_band(P, blk, nblk, (float)se, a.lum, a.ecm, a.psi, a.hlum, a.hecm, a.L0, a.L1, a.E0,
                                a.E1, a.jlx, a.jly, a.jex, a.jey, a.jbox_part, sb);
        }
        if (!lum_barrier(a.bar, nblk, epoch, a.abort_flag, a.err, a.limit)) return;
        if (!lum_loop_body(P, blk, nblk, red, a.L0, a.L1, a.E0, a.E1, a.jlx, a.jly, a.jex, a.jey, a.jbox_part,
                           nblk, a.lst, a.part_lc, a.bar, epoch, a.abort_flag, a.err, a.limit, cur, !boxes))
            return;
    }
}
static __device__ __forceinline__ void step_cells(DevP P, StepArgs a, const int n, unsigned char *lds) {
    constexpr int WH = ORG_WH, NPX = ORG_W * ORG_W;
    const size_t NN = (size_t)P.N * P.N;
    float *phase = (float *)lds, *pnl = phase + WH * WH;
    double *red = (double *)(lds + 2 * WH * WH * 4);
    const int x0 = a.xc[n] - P.zs, y0 = a.yc[n] - P.zs;
    float *ph = a.phi + n * NN;
    float *jx = a.jwin + (size_t)n * 2 * WH * WH, *jy = jx + WH * WH;
    if (a.shared_brk)
        cells_update<true>(P, x0, y0, ph, a.rho + n * NN, a.hfunc + n * NN, a.psi, a.hlum, a.hecm, a.lbrk,
                           a.f0[n], pnl, nullptr, jx, jy);
    else
        cells_update<false>(P, x0, y0, ph, a.rho + n * NN, a.hfunc + n * NN, a.psi, a.hlum, a.hecm, nullptr,
                            a.f0[n], pnl, nullptr, jx, jy);
    cells_normals_ring(P, x0, y0, ph, jx, jy, phase);
    __syncthreads();      /* the workgroup reads back its own stores */
    cells_relax(P, n, x0, y0, ph, phase, pnl, jx, jy, red, a.phin_win + (size_t)n * NPX,
                a.phi_win + (size_t)n * NPX, a.counts, a.err);
    /* the next step's centre and support box (find_centres_cells of the
     * next iteration, useful_functions.cpp:193-227) from the new phi:
     * the window from LDS, the rest of the support box from the plane */
    long long *lred = (long long *)(lds + 2 * WH * WH * 4 + 128);
    const int i0 = a.box[4 * n], i1 = a.box[4 * n + 1], j0 = a.box[4 * n + 2], j1 = a.box[4 * n + 3];
    long long X, Y, C;
    centroid_sums_win(ph, P.N, i0, i1, j0, j1, x0, y0, phase, lred, X, Y, C);
    if (threadIdx.x == 0) {
        place_window(P, n, X, Y, C, a.xc[n], a.yc[n], i0, i1, j0, j1, a.xn, a.yn, a.bn, a.err);
        a.xp[n] = a.xc[n]; a.yp[n] = a.yc[n];
    }
}
__global__ __launch_bounds__(1024) void k_step(DevP P, StepArgs a) {
    __shared__ __attribute__((aligned(16))) unsigned char lds[ORG_LDS_FUSED];
    const int nc = P.nc, b = blockIdx.x;
    if (b < nc) step_spectral(P, a, xcd_index(b, 0, nc), lds);
    else if (b < nc + a.nlum) step_lumecm(P, a, xcd_index(b, nc, a.nlum), a.nlum, lds);
    else step_cells(P, a, xcd_index(b, nc + a.nlum, nc), lds);
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
    if (phi_win) {        /* phi window from k_step (reshaping_cells' phase) */
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

/* The lumen/ECM part of the next step's cell-update bracket into lbrk, from
 * the loops' final iterates (the fields the copy-back commits), on the parts
 * that can have changed: the live boxes grown by two (the bracket's stencil
 * reaches one pixel past the copied region), all of them in full mode, and
 * the ring (its stencil reads the corners, which the sweeps rewrite).  Run by
 * the copy-back launches (k_mid, k_post), after the step's launch, whose
 * phase-field role read the current bracket, has ended: one buffer serves
 * both, and the step's lumen/ECM role ends with its loops (no barrier among
 * its workgroups to read each other's bands).  lds: ORG_LDS_LBRK bytes. */
static __device__ void lbrk_step(const DevP &P, const StepArgs &a, const int blk, const int nblk, float *lds) {
    const LumState sl = a.lst[0], se = a.lst[1];
    const float *FL = sl.cur ? a.L1 : a.L0, *FE = se.cur ? a.E1 : a.E0;
    const int4 BL = ld_box(a.live, 0), BE = ld_box(a.live, 1);
    const int N = P.N, nt = (N + ORG_HT - 1) / ORG_HT, ntl = nt * nt;
    const bool full = !box_mode(BL, N) || !box_mode(BE, N);
    /* parts by residue blk + nblk/2: in k_mid the same workgroups' h parts
     * go by residue blk, so the two loops' busy parts fall on different
     * workgroups */
    for (int k = (blk + nblk / 2) % nblk; k < ntl + ORG_RING_PARTS; k += nblk) {
        bool need = full || k >= ntl;
        if (!need) {
            const int ti = (k / nt) * ORG_HT, tj = (k % nt) * ORG_HT;
            for (int f = 0; f < 2 && !need; f++) {
                const int4 B = f ? BE : BL;
                need = !box_empty(B) && ti < B.y + 2 && ti + ORG_HT > B.x - 2 && tj < B.w + 2 &&
                       tj + ORG_HT > B.z - 2;
            }
        }
        if (need) lbrk_part(P, k, FL, FE, a.lbrk, lds, ORG_RING_PARTS);
    }
}

/* k_post: the copy-back of a timestep (run_with_active_gels.cpp:172-182):
 * rho box and phi window of every cell (workgroups 0..ncell-1); lum <- lumn,
 * ECM <- ECMn (the other workgroups, see commit_field), then the next
 * step's lumen/ECM bracket.  Clears k_step's barrier counter and abort
 * flag. */
static __device__ __forceinline__ void post_cell(DevP P, StepArgs a, const int n) {
    if (n == 0 && threadIdx.x == 0) { a.bar[0] = 0u; *a.abort_flag = 0; }
    commit_cell_body(P, n, a.rho, a.rho_box, a.phi, a.phi_win, a.xc, a.yc);
}
static __device__ __forceinline__ void post_lumecm(DevP P, StepArgs a, const int blk, const int nblk,
                                                   unsigned char *scr) {
    int *sb = (int *)scr;
    const LumState sl = a.lst[0], se = a.lst[1];
    if (blk == 0 && threadIdx.x == 0) {
        a.uni[0] = P.N; a.uni[1] = 0; a.uni[2] = P.N; a.uni[3] = 0;
        if (!sl.done || !se.done) set_err(a.err, ORG_ERELAX);
    }
    commit_field(P, 0, ld_box(a.live, 0), blk, nblk, a.lum, sl.cur ? a.L1 : a.L0, a.live_part, sb);
    commit_field(P, 1, ld_box(a.live, 1), blk, nblk, a.ecm, se.cur ? a.E1 : a.E0, a.live_part, sb);
    if (a.shared_brk) lbrk_step(P, a, blk, nblk, (float *)(scr + 256));
}
#define ORG_LDS_POST (256 + ORG_LDS_LBRK)
__global__ __launch_bounds__(1024) void k_post(DevP P, StepArgs a) {
    __shared__ __attribute__((aligned(16))) unsigned char scr[ORG_LDS_POST];
    const int b = blockIdx.x, nc = P.nc;
    if (b < nc) post_cell(P, a, xcd_index(b, 0, nc));
    else post_lumecm(P, a, xcd_index(b, nc, a.nlp), a.nlp, scr);
}

/* k_mid: the copy-back of one timestep and the first part of the next
 * (k_post then k_pre) in one launch, workgroups in three independent roles:
 *   0 .. ncell-1   cell n: rho box and phi window copy-back, then (same
 *                  workgroup, after its own stores) hfunc on the next window
 *                  and the area, with the centre and support box from
 *                  (xn, yn, bn)
 *   next npsi      psi of the next step: hfun(phi) inside each next window,
 *                  phi inside the previous window read from phi_win (the
 *                  copy-back of the first role may be under way), hfunc
 *                  elsewhere (the preps write hfunc only inside the next
 *                  windows)
 *   last nlp       lum <- lumn, ECM <- ECMn with the live pixels (commit_field),
 *                  h(lum), h(ECM) of the changed tiles from lumn, ECMn, and
 *                  the next step's lumen/ECM bracket (lbrk_step)
 * The previous centres come from (xp, yp), which the preps do not touch. */
#define ORG_LDS_MID (4 * ORG_PSI_LDS_GROUP + 16 > 256 + ORG_LDS_LBRK ? 4 * ORG_PSI_LDS_GROUP + 16 : 256 + ORG_LDS_LBRK)
static __device__ __forceinline__ void mid_cell(DevP P, StepArgs a, const int n, unsigned char *scr) {
    if (n == 0 && threadIdx.x == 0) { a.bar[0] = 0u; *a.abort_flag = 0; }
    commit_cell_body(P, n, a.rho, a.rho_box, a.phi, a.phi_win, a.xc, a.yc);
    __syncthreads();      /* the prep reads back this workgroup's phi window */
    cell_prep(P, n, a.phi, a.hfunc, a.xn, a.yn, a.bn, a.xc, a.yc, a.box, a.s0, a.s, a.f0, (double *)scr);
}
static __device__ __forceinline__ void mid_psi(DevP P, StepArgs a, const int blk, unsigned char *scr) {
    psi_tiles_step(P, blk, a.npsi, a.hfunc, a.phi, a.bn, a.xn, a.yn, a.phi_win, a.xp, a.yp, a.psi, scr);
}
static __device__ __forceinline__ void mid_lumecm(DevP P, StepArgs a, const int blk,
                                                  unsigned char *scr) {
    const int nblk = a.nlp, N = P.N;
    const LumState sl = a.lst[0], se = a.lst[1];
    if (blk == 0 && threadIdx.x == 0) {
        a.uni[0] = N; a.uni[1] = 0; a.uni[2] = N; a.uni[3] = 0;
        if (!sl.done || !se.done) set_err(a.err, ORG_ERELAX);
    }
    const float *FL = sl.cur ? a.L1 : a.L0, *FE = se.cur ? a.E1 : a.E0;
    const int4 BL = ld_box(a.live, 0), BE = ld_box(a.live, 1);
    int *sb = (int *)scr;
    commit_field(P, 0, BL, blk, nblk, a.lum, FL, a.live_part, sb);
    commit_field(P, 1, BE, blk, nblk, a.ecm, FE, a.live_part, sb);
    /* h of the new fields on every part the step may have changed.  The
     * ring: every batched step leaves the same values there (the boundary
     * values, with the corners the sweeps write, all fixed by the
     * parameters; box mode never follows full mode, as boxes only grow), so
     * after the first copy-back since set-up its h values and its part of
     * the sum stand */
    const int nt = (N + ORG_HT - 1) / ORG_HT, ntl = nt * nt;
    const bool full = !box_mode(BL, N) || !box_mode(BE, N);
    for (int k = blk; k <= ntl; k += nblk) {
        bool need = full || (k == ntl && *a.ringh == 0);
        if (!need) {
            const int ti = (k / nt) * ORG_HT, tj = (k % nt) * ORG_HT;
            for (int f = 0; f < 2 && !need; f++) {
                const int4 B = f ? BE : BL;
                need = !box_empty(B) && ti < B.y + 1 && ti + ORG_HT > B.x - 1 && tj < B.w + 1 &&
                       tj + ORG_HT > B.z - 1;
            }
        }
        if (!need) continue;
        const double tot = block_sum_d(hlum_part(P, k, FL, FE, a.hlum, a.hecm), (double *)(scr + 64));
        if (threadIdx.x == 0) {
            a.part_h[k] = tot;
            if (k == ntl) *a.ringh = 1;     /* read above before the sum's barriers */
        }
    }
    if (a.shared_brk) lbrk_step(P, a, blk, nblk, (float *)(scr + 256));
}
__global__ __launch_bounds__(1024) void k_mid(DevP P, StepArgs a) {
    __shared__ __attribute__((aligned(16))) unsigned char scr[ORG_LDS_MID];
    const int b = blockIdx.x, nc = P.nc;
    if (b < nc) mid_cell(P, a, xcd_index(b, 0, nc), scr);
    else if (b < nc + a.npsi) mid_psi(P, a, xcd_index(b, nc, a.npsi), scr);
    else mid_lumecm(P, a, xcd_index(b, nc + a.npsi, a.nlp), scr);
}

/* ---- batches: every organoid of a set in one launch per stage -----------
 * A batch runs k_pre, (k_step, k_mid) x (n - 1), k_step, k_post over several
 * organoids at once (grids, cell counts and constants may all differ).  Each
 * workgroup reads one word of a block map: the organoid (bits 20..27), the
 * role (bits 28..29, in the order of the single-organoid kernel) and the
 * index within the role (bits 0..19), then runs that role with the
 * organoid's constants and buffers.  The roles are the same functions as in
 * k_pre, k_step, k_mid and k_post, so every organoid of a batch evolves bit
 * for bit as it would alone.  In k_step the block map lists the lumen/ECM
 * workgroups of every organoid first: they are the only ones that wait (on
 * workgroups of the same organoid and role), and the host keeps their total
 * within the CU count, so all of them are resident together.  The tables
 * live in constant memory (address space 4): the role arguments load through
 * scalar loads as kernel arguments do. */
struct BatchEnt {
    DevP P;
    StepArgs a;
};
#define ORG_CONST __attribute__((address_space(4)))
#define ORG_BENT(v) (((v) >> 20) & 0xff)
#define ORG_BROLE(v) (((v) >> 28) & 3)
#define ORG_BIDX(v) ((v) & 0xfffff)

__global__ __launch_bounds__(1024) void k_pre_batch(const ORG_CONST BatchEnt *ent, const ORG_CONST int *bmap) {
    __shared__ __attribute__((aligned(16))) unsigned char scr[ORG_LDS_PRE];
    const int v = bmap[blockIdx.x], r = ORG_BROLE(v), i = ORG_BIDX(v);
    const BatchEnt *E = (const BatchEnt *)(ent + ORG_BENT(v));
    if (r == 0) pre_cell(E->P, E->a, i, scr);
    else if (r == 1) pre_psi(E->P, E->a, i, scr);
    else pre_hlum(E->P, E->a, i, scr);
}

__global__ __launch_bounds__(1024) void k_step_batch(const ORG_CONST BatchEnt *ent, const ORG_CONST int *bmap) {
    __shared__ __attribute__((aligned(16))) unsigned char lds[ORG_LDS_FUSED];
    const int v = bmap[blockIdx.x], r = ORG_BROLE(v), i = ORG_BIDX(v);
    const BatchEnt *E = (const BatchEnt *)(ent + ORG_BENT(v));
    if (r == 0) step_spectral(E->P, E->a, i, lds);
    else if (r == 1) step_lumecm(E->P, E->a, i, E->a.nlum, lds);
    else step_cells(E->P, E->a, i, lds);
}

__global__ __launch_bounds__(1024) void k_mid_batch(const ORG_CONST BatchEnt *ent, const ORG_CONST int *bmap) {
    __shared__ __attribute__((aligned(16))) unsigned char scr[ORG_LDS_MID];
    const int v = bmap[blockIdx.x], r = ORG_BROLE(v), i = ORG_BIDX(v);
    const BatchEnt *E = (const BatchEnt *)(ent + ORG_BENT(v));
    if (r == 0) mid_cell(E->P, E->a, i, scr);
    else if (r == 1) mid_psi(E->P, E->a, i, scr);
    else mid_lumecm(E->P, E->a, i, scr);
}

__global__ __launch_bounds__(1024) void k_post_batch(const ORG_CONST BatchEnt *ent, const ORG_CONST int *bmap) {
    __shared__ __attribute__((aligned(16))) unsigned char scr[ORG_LDS_POST];
    const int v = bmap[blockIdx.x], r = ORG_BROLE(v), i = ORG_BIDX(v);
    const BatchEnt *E = (const BatchEnt *)(ent + ORG_BENT(v));
    if (r == 0) post_cell(E->P, E->a, i);
    else post_lumecm(E->P, E->a, i, E->a.nlp, scr);
}

/* update_rho_to_shape (cell_phase_functions.cpp:42-62) on every plane:
 * where phi != 0, rho = rho0 exp(-1/(4 h (1 - h) + tol) + 1/(1 + tol)) with
 * h = phi^2 (3 - 2 phi); then rho = 0 wherever it is below
 * rho0 exp(-1/(2 tol) + 1/(1 + tol)).  Double arithmetic, as the
 * reference's double literals make it. */
__global__ __launch_bounds__(256) void k_rho_to_shape(const float *__restrict__ phi, float *__restrict__ rho,
                                                      size_t count, float rho0, float tol) {
    const double floor_v = (double)rho0 * exp(-1.0 / (2.0 * (double)tol) + 1.0 / (1.0 + (double)tol));
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < count; e += (size_t)gridDim.x * blockDim.x) {
        const float p = phi[e];
        float r = rho[e];
        if (p != 0.0f) {
            const double pd = p, h = pd * pd * (3.0 - 2.0 * pd);
            r = (float)((double)rho0 * exp(-1.0 / (4.0 * h * (1.0 - h) + (double)tol) + 1.0 / (1.0 + (double)tol)));
        }
        if ((double)r < floor_v) r = 0.0f;
        rho[e] = r;
    }
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
