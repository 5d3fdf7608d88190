/*
 * organoid_generic.hip -- the per-function timestep on full planes in global
 * memory, for the forms the LDS kernels of organoid_kernels.hip do not cover:
 * the FP64 drop-in (the reference retyped to double) and any FFT box
 * (M = 2*zone_size - 2; the reference's template has zone_size 200, M = 398).
 *
 * Each kernel is one function of the reference's time loop
 * (run_with_active_gels.cpp:118-184) for all cells in one launch, templated
 * on the stored real type R, with every expression written with the
 * reference's own literal types: for R = float the double literals promote
 * intermediate expressions exactly as the reference's float build does
 * (e.g. rho_functions.cpp:25), for R = double everything is double.  FMA
 * contraction is off, so each operation rounds as written.  Reductions
 * (areas, the sum of h(ECM), reshaping deviations) accumulate in double in
 * a fixed order.  The 2D transforms of update_all_cortices (cufftExecC2C,
 * useful_functions.cpp:646-660: unnormalised, out of place) run in double
 * whatever R is, rows then columns, and store R: radix-2 Cooley-Tukey with
 * the bit-reversed load for powers of two, Bluestein's chirp-z transform
 * (1970) through power-of-two transforms otherwise; twiddles and chirps are
 * computed on the host from the definition in double.
 *
 * Performance is not the point of this path (one launch per reference
 * function, full-plane passes as the reference makes them); the batched
 * LDS form is the fast one.
 */
#include <cmath>
#include <cstdio>
#include <cstring>
#include <type_traits>
#include <vector>

#include "organoid_generic.h"

#pragma clang fp contract(off)

namespace gen {

template <typename R>
struct Cx { R x, y; };

/* ---- FFT plan ----------------------------------------------------------- */
/* Twiddles of the power-of-two transform of length n: tw[k] = (cos, sin) of
 * 2 pi k / n, k < n/2 (the sign is applied by the kernel). */
static void host_twiddles(int n, std::vector<double> &tw) {
    tw.resize(n);
    for (int k = 0; k < n / 2; k++) {
        tw[2 * k] = cos(2.0 * M_PI * k / n);
        tw[2 * k + 1] = sin(2.0 * M_PI * k / n);
    }
}

/* In-place radix-2 DFT on the host (definition X_k = sum x_j e^{s 2 pi i jk/n}),
 * used only to build Bluestein's kernel spectrum. */
static void host_fft_pow2(double *re, double *im, int n, int sign) {
    for (int i = 1, j = 0; i < n; i++) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) { std::swap(re[i], re[j]); std::swap(im[i], im[j]); }
    }
    std::vector<double> cw(n / 2 + 1), sw(n / 2 + 1);
    for (int k = 0; k < n / 2; k++) {
        cw[k] = cos(2.0 * M_PI * k / n);
        sw[k] = sign * sin(2.0 * M_PI * k / n);
    }
    for (int len = 2; len <= n; len <<= 1) {
        const int half = len / 2, step = n / len;
        for (int i = 0; i < n; i += len)
            for (int k = 0; k < half; k++) {
                const double wr = cw[k * step], wi = sw[k * step];
                const int a = i + k, b = i + k + half;
                const double xr = re[b] * wr - im[b] * wi, xi = re[b] * wi + im[b] * wr;
                re[b] = re[a] - xr; im[b] = im[a] - xi;
                re[a] += xr; im[a] += xi;
            }
    }
}

struct FftPlan {
    int M = 0, L = 0;          /* line length; power-of-two length used (L = M, or Bluestein's) */
    bool pow2 = true;
    double *tw = nullptr;      /* device: L/2 (cos, sin) pairs of length L */
    double *chirp[2] = {nullptr, nullptr};   /* device, sign -1 / +1: w_k (M pairs) */
    double *khat[2] = {nullptr, nullptr};    /* device, sign -1 / +1: FFT_L of the conjugate chirp (L pairs) */
    /* schedule: radix-2 stages fused per LDS round trip (1, 2 or 3: radix
     * 2, 4, 8) and lines per workgroup (1, 2 or 4) */
    int fuse = 1, lpw = 1;
};

static bool is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

/* ---- line transforms in LDS ---------------------------------------------- */
/* The oracle's radix-2 transform (oracle/organoid_oracle.c fft_pow2): the
 * input in bit-reversed order, then log2(n) butterfly stages with the
 * twiddle of stage `len` at k * (n / len).  The schedule groups F
 * consecutive stages per LDS round trip: each of nt threads owns groups of
 * 2^F elements (indices base + k + m half0, m < 2^F) and runs the F stages'
 * butterflies among them in registers.  Every butterfly is the same
 * operation on the same operands in any schedule, so the values are the
 * oracle's for F = 1, 2, 3 alike. */
template <int F>
static __device__ __forceinline__ void lds_fft_pow2(double *re, double *im, int n, int sign,
                                                    const double *__restrict__ tw, int t0, int nt) {
    int half0 = 1;
    while (half0 < n) {
        /* stages half0, 2 half0, ... up to F of them (fewer at the end) */
        int f = 0;
        while (f < F && (half0 << f) < n) f++;
        const int g = 1 << f;                          /* elements per group */
        const int span = half0 << f;                   /* indices one group covers */
        for (int t = t0; t < n / g; t += nt) {
            const int blk = t / half0, k = t - blk * half0;
            const int base = blk * span + k;
            double xr[8], xi[8];
#pragma unroll
            for (int m = 0; m < 8; m++)
                if (m < g) { xr[m] = re[base + m * half0]; xi[m] = im[base + m * half0]; }
#pragma unroll
            for (int sft = 0; sft < 3; sft++) {
                if (sft >= f) break;
                const int half = half0 << sft, step = n / (2 * half);
#pragma unroll
                for (int m = 0; m < 8; m++) {
                    if (m >= g || (m & (1 << sft))) continue;
                    const int kk = k + (m & ((1 << sft) - 1)) * half0;     /* index within its len-block */
                    const double wr = tw[2 * (kk * step)], wi = sign * tw[2 * (kk * step) + 1];
                    const int mb = m + (1 << sft);
                    const double br = xr[mb], bi = xi[mb];
                    const double pr = br * wr - bi * wi;
                    const double pi_ = br * wi + bi * wr;
                    xr[mb] = xr[m] - pr; xi[mb] = xi[m] - pi_;
                    xr[m] += pr; xi[m] += pi_;
                }
            }
#pragma unroll
            for (int m = 0; m < 8; m++)
                if (m < g) { re[base + m * half0] = xr[m]; im[base + m * half0] = xi[m]; }
        }
        __syncthreads();
        half0 <<= f;
    }
}

static __device__ __forceinline__ int bitrev(int i, int logn) { return (int)(__brev((unsigned)i) >> (32 - logn)); }

/* nlines transforms of length M, LPW lines per workgroup pass (256 / LPW
 * threads each): line l of plane p = l / lines_per_plane starts at
 * p * plane + (l % lines_per_plane) * line_step, elements elem_step apart,
 * read from src and written to dst (which may be src).  Computed in double;
 * TI, TO: the stored types.  LDS: 2 L doubles per line. */
template <typename TI, typename TO, int F>
__global__ __launch_bounds__(256) void g_fft_lines(const Cx<TI> *src, Cx<TO> *dst,
                                                   int nlines, int lines_per_plane, long plane, int line_step,
                                                   int elem_step, int M, int L, int sign, const double *__restrict__ tw,
                                                   const double *__restrict__ chirp, const double *__restrict__ khat,
                                                   int lpw) {
    extern __shared__ double sh[];
    const int nt = blockDim.x / lpw, sub = threadIdx.x / nt, t0 = threadIdx.x - sub * nt;
    double *re = sh + (size_t)sub * 2 * L, *im = re + L;
    const int logL = 31 - __clz(L);
    for (int line0 = blockIdx.x * lpw; line0 < nlines; line0 += gridDim.x * lpw) {
        const int line = line0 + sub;
        const bool act = line < nlines;
        const long base = act ? (long)(line / lines_per_plane) * plane + (long)(line % lines_per_plane) * line_step : 0;
        if (!chirp) {
            if (act)
                for (int k = t0; k < M; k += nt) {
                    const Cx<TI> v = src[base + (long)k * elem_step];
                    const int r = bitrev(k, logL);
                    re[r] = (double)v.x; im[r] = (double)v.y;
                }
            __syncthreads();
            lds_fft_pow2<F>(re, im, L, sign, tw, t0, nt);
            if (act)
                for (int k = t0; k < M; k += nt) {
                    Cx<TO> o; o.x = (TO)re[k]; o.y = (TO)im[k];
                    dst[base + (long)k * elem_step] = o;
                }
        } else {
            /* Bluestein (the oracle's fft_any): a_k = x_k w_k zero-padded to L,
             * A = F_-(a), C = A * K, c = F_+(C), X_k = w_k c_k / L */
            for (int k = t0; k < L; k += nt) {
                double ar = 0.0, ai = 0.0;
                if (act && k < M) {
                    const Cx<TI> v = src[base + (long)k * elem_step];
                    const double xr = (double)v.x, xi = (double)v.y, wr = chirp[2 * k], wi = chirp[2 * k + 1];
                    ar = xr * wr - xi * wi;
                    ai = xr * wi + xi * wr;
                }
                const int r = bitrev(k, logL);
                re[r] = ar; im[r] = ai;
            }
            __syncthreads();
            lds_fft_pow2<F>(re, im, L, -1, tw, t0, nt);
            double pr[8], pi_[8];      /* this thread's products, then the bit-reversed reload */
            int cnt = 0;
            for (int k = t0; k < L; k += nt, cnt++) {
                const double kr = khat[2 * k], ki = khat[2 * k + 1];
                pr[cnt] = re[k] * kr - im[k] * ki;
                pi_[cnt] = re[k] * ki + im[k] * kr;
            }
            __syncthreads();
            cnt = 0;
            for (int k = t0; k < L; k += nt, cnt++) {
                const int r = bitrev(k, logL);
                re[r] = pr[cnt]; im[r] = pi_[cnt];
            }
            __syncthreads();
            lds_fft_pow2<F>(re, im, L, +1, tw, t0, nt);
            const double rL = 1.0 / L;     /* L a power of two: x * rL is x / L exactly */
            if (act)
                for (int k = t0; k < M; k += nt) {
                    const double cr = re[k] * rL, ci = im[k] * rL, wr = chirp[2 * k], wi = chirp[2 * k + 1];
                    Cx<TO> o;
                    o.x = (TO)(cr * wr - ci * wi);
                    o.y = (TO)(cr * wi + ci * wr);
                    dst[base + (long)k * elem_step] = o;
                }
        }
        __syncthreads();
    }
}


/* ---- the timestep's functions ---------------------------------------------- */
/* Constants as the reference holds them: float globals for R = float
 * (constants_template.cpp), double for the FP64 form. */
template <typename R>
struct GP {
    int N, nc, zs, M, rimax, ncur;
    R dt, dx, sigma, cpl, tau, alpha, beta, eta;
    R tau_lum, beta_lc, eta_lc, xi_lum, alpha_E, tau_E, beta_ec, eta_ec, xi_E, beta_le;
    R wi, rdt, friction;
    R visc[3], aort[3], atan_[3], bb[3], gen[3], deg[3];   /* apical, basal, lateral */
};

#define GPL(n, i, j) ((size_t)(n) * NN + (size_t)(i) * N + (size_t)(j))
#define GG(i, j) ((size_t)(i) * N + (size_t)(j))

/* Fixed-order double sum over the block (blockDim a multiple of 64, <= 1024). */
static __device__ double gblock_sum(double v, double *red) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[w] = v;
    __syncthreads();
    double s = 0.0;
    for (int k = 0; k < (int)(blockDim.x >> 6); k++) s += red[k];
    __syncthreads();
    return s;
}

/* find_centres_cells, useful_functions.cpp:193-227: the integer sums of i,
 * j and the count over the pixels with phi > 0.5, workgroup (b, n) on rows
 * b, b + gridDim.x, ... of cell n, added into acc[n] = {sum i, sum j,
 * count} (integer sums: the same in any order).  acc zero on entry. */
template <typename R>
__global__ __launch_bounds__(256) void g_centroid_part(GP<R> P, const R *__restrict__ phi,
                                                       unsigned long long *__restrict__ acc) {
    __shared__ long long red[3 * 4];
    const int N = P.N, n = blockIdx.y;
    const R *ph = phi + (size_t)n * N * N;
    long long xs = 0, ys = 0, cnt = 0;
    for (int i = blockIdx.x; i < N; i += gridDim.x)
        for (int j = threadIdx.x; j < N; j += blockDim.x)
            if (ph[(size_t)i * N + j] > 0.5) { xs += i; ys += j; cnt += 1; }
    for (int o = 32; o > 0; o >>= 1) {
        xs += __shfl_down(xs, o, 64); ys += __shfl_down(ys, o, 64); cnt += __shfl_down(cnt, o, 64);
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) { red[w] = xs; red[4 + w] = ys; red[8 + w] = cnt; }
    __syncthreads();
    if (threadIdx.x < 3) {
        long long v = 0;
        for (int k = 0; k < (int)(blockDim.x >> 6); k++) v += red[4 * threadIdx.x + k];
        if (v) atomicAdd(&acc[3 * n + threadIdx.x], (unsigned long long)v);
    }
}

/* g_centroid_part's sums restricted to each cell's window (rows and
 * columns [c - zs, c + zs) around the current centre), added to acc (sign
 * > 0) or taken out of it (sign < 0, two's complement: the 64-bit sums wrap
 * and stay exact).  Block (b, n) does rows b, b + gridDim.x, ... of cell
 * n's window. */
template <typename R>
__global__ __launch_bounds__(256) void g_centroid_win(GP<R> P, const R *__restrict__ phi, const int *__restrict__ xc,
                                                      const int *__restrict__ yc, unsigned long long *__restrict__ acc,
                                                      int sign) {
    __shared__ long long red[3 * 4];
    const int N = P.N, n = blockIdx.y, zs = P.zs;
    const R *ph = phi + (size_t)n * N * N;
    long long xs = 0, ys = 0, cnt = 0;
    for (int i = xc[n] - zs + blockIdx.x; i < xc[n] + zs; i += gridDim.x)
        for (int j = yc[n] - zs + threadIdx.x; j < yc[n] + zs; j += blockDim.x)
            if (ph[(size_t)i * N + j] > 0.5) { xs += i; ys += j; cnt += 1; }
    for (int o = 32; o > 0; o >>= 1) {
        xs += __shfl_down(xs, o, 64); ys += __shfl_down(ys, o, 64); cnt += __shfl_down(cnt, o, 64);
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) { red[w] = xs; red[4 + w] = ys; red[8 + w] = cnt; }
    __syncthreads();
    if (threadIdx.x < 3) {
        long long v = 0;
        for (int k = 0; k < (int)(blockDim.x >> 6); k++) v += red[4 * threadIdx.x + k];
        if (v) atomicAdd(&acc[3 * n + threadIdx.x], (unsigned long long)(sign > 0 ? v : -v));
    }
}

/* the centres from g_centroid_part's sums (integer division as the
 * reference's), then the window check (window plus one pixel inside rows /
 * columns 1..N-2) */
template <typename R>
__global__ void g_centroid_fin(GP<R> P, const unsigned long long *__restrict__ acc, int *__restrict__ xc,
                               int *__restrict__ yc, int *__restrict__ err) {
    const int N = P.N;
    for (int n = threadIdx.x; n < P.nc; n += blockDim.x) {
        const long long X = (long long)acc[3 * n], Y = (long long)acc[3 * n + 1], C = (long long)acc[3 * n + 2];
        if (C == 0) { atomicCAS(err, 0, ORG_EEMPTY_CELL); continue; }
        const int x = (int)(X / C), y = (int)(Y / C);
        xc[n] = x; yc[n] = y;
        const int x0 = x - P.zs, y0 = y - P.zs;
        if (x0 - 1 < 1 || x0 + 2 * P.zs > N - 2 || y0 - 1 < 1 || y0 + 2 * P.zs > N - 2) atomicCAS(err, 0, ORG_EWINDOW);
    }
}

/* compute_hfunc_3d_tables_kernels_loop_n, useful_functions.cpp:233-251:
 * block (r, n) does row r of cell n's window. */
template <typename R>
__global__ void g_hfunc_cells(GP<R> P, const R *__restrict__ phi, R *__restrict__ hf, const int *__restrict__ xc,
                              const int *__restrict__ yc, int *__restrict__ hbox) {
    const int N = P.N, n = blockIdx.y, zs = P.zs;
    const size_t NN = (size_t)N * N;
    const int i = xc[n] - zs + blockIdx.x;
    if (blockIdx.x == 0 && threadIdx.x == 0) {             /* the box grows by the window written */
        int *b = hbox + 4 * n;
        const int r0 = xc[n] - zs, c0 = yc[n] - zs, w = 2 * zs;
        const bool empty = b[0] >= b[1] || b[2] >= b[3];
        b[0] = empty ? r0 : min(b[0], r0);
        b[1] = empty ? r0 + w : max(b[1], r0 + w);
        b[2] = empty ? c0 : min(b[2], c0);
        b[3] = empty ? c0 + w : max(b[3], c0 + w);
    }
    for (int j = yc[n] - zs + threadIdx.x; j < yc[n] + zs; j += blockDim.x) {
        const R v = phi[GPL(n, i, j)];
        hf[GPL(n, i, j)] = v * v * (3.0 - 2.0 * v);
    }
}

/* compute_all_hfunc_2d_tables_kernels, useful_functions.cpp:272-285 */
template <typename R>
__global__ void g_hfunc_lumecm(GP<R> P, const R *__restrict__ lum, const R *__restrict__ ecm, R *__restrict__ hl,
                               R *__restrict__ he) {
    const size_t NN = (size_t)P.N * P.N;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < NN; k += (size_t)gridDim.x * blockDim.x) {
        const R a = lum[k], e = ecm[k];
        hl[k] = a * a * (3.0 - 2.0 * a);
        he[k] = e * e * (3.0 - 2.0 * e);
    }
}

/* boundary_conditions_cells, useful_functions.cpp:391-420: periodic ghost
 * ring (period N-2) of `planes` planes; workgroup (p, y) does plane p's
 * ring indices 1 + y blockDim.x + t, 1 + (y + gridDim.y) blockDim.x + t,
 * ... (every value read is an interior pixel, every one written a ring
 * pixel: the workgroups need no order among them). */
template <typename R>
__global__ void g_bc_cells(int N, R *__restrict__ t) {
    const size_t NN = (size_t)N * N;
    R *q = t + (size_t)blockIdx.x * NN;
    for (int i = 1 + blockIdx.y * blockDim.x + threadIdx.x; i < N - 1; i += gridDim.y * blockDim.x) {
        q[(size_t)i * N] = q[(size_t)i * N + N - 2];
        q[(size_t)i * N + N - 1] = q[(size_t)i * N + 1];
        q[i] = q[(size_t)N * (N - 2) + i];
        q[(size_t)N * (N - 1) + i] = q[(size_t)N + i];
    }
    if (blockIdx.y == 0 && threadIdx.x == 0) {
        q[0] = q[(size_t)N * (N - 2) + N - 2];
        q[N - 1] = q[(size_t)N * (N - 2) + 1];
        q[(size_t)N * (N - 1)] = q[(size_t)N + N - 2];
        q[(size_t)N * (N - 1) + N - 1] = q[(size_t)N + 1];
    }
}

/* compute_normal_grad_phi, useful_functions.cpp:326-354, interior of every
 * plane: workgroup (x, n) on rows x + 1, x + 1 + gridDim.x, ... of plane n. */
template <typename R>
__global__ void g_norm_grad(GP<R> P, const R *__restrict__ phi, R *__restrict__ ngp) {
    const int N = P.N, n = blockIdx.y;
    const size_t NN = (size_t)N * N;
    const R dx = P.dx;
    for (int i = 1 + blockIdx.x; i < N - 1; i += gridDim.x)
        for (int j = 1 + threadIdx.x; j < N - 1; j += blockDim.x) {
            const R a = phi[GPL(n, i + 1, j)] - phi[GPL(n, i - 1, j)];
            const R c = phi[GPL(n, i, j + 1)] - phi[GPL(n, i, j - 1)];
            ngp[GPL(n, i, j)] = sqrt((1.0 / dx / dx / 4.0) * (a * a + c * c));
        }
}

/* Each cell's hfunc plane is zero outside hbox[n] = {r0, r1, c0, c1}
 * (rows [r0, r1) x columns [c0, c1); empty when r0 >= r1): g_hbox_scan
 * sets the box from the plane after an upload, g_hfunc_cells widens it by
 * each window it writes.  Adding zero leaves a sum's bits unchanged, so
 * sums over the boxes alone equal the whole-plane sums bit for bit. */
__global__ void g_hbox_reset(int *__restrict__ hbox, int nc, int N) {
    for (int n = threadIdx.x; n < nc; n += blockDim.x) {
        hbox[4 * n] = N; hbox[4 * n + 1] = 0; hbox[4 * n + 2] = N; hbox[4 * n + 3] = 0;
    }
}
template <typename R>
__global__ __launch_bounds__(256) void g_hbox_scan(GP<R> P, const R *__restrict__ hf, int *__restrict__ hbox) {
    __shared__ int bb[4];
    const int N = P.N, n = blockIdx.y;
    const R *h = hf + (size_t)n * N * N;
    int i0 = N, i1 = -1, j0 = N, j1 = -1;
    for (int i = blockIdx.x; i < N; i += gridDim.x)
        for (int j = threadIdx.x; j < N; j += blockDim.x)
            if (h[(size_t)i * N + j] != (R)0.0) { i0 = min(i0, i); i1 = max(i1, i); j0 = min(j0, j); j1 = max(j1, j); }
    if (threadIdx.x == 0) { bb[0] = N; bb[1] = -1; bb[2] = N; bb[3] = -1; }
    for (int o = 32; o > 0; o >>= 1) {
        i0 = min(i0, __shfl_xor(i0, o, 64)); i1 = max(i1, __shfl_xor(i1, o, 64));
        j0 = min(j0, __shfl_xor(j0, o, 64)); j1 = max(j1, __shfl_xor(j1, o, 64));
    }
    __syncthreads();
    if ((threadIdx.x & 63) == 0 && i1 >= 0) {
        atomicMin(&bb[0], i0); atomicMax(&bb[1], i1); atomicMin(&bb[2], j0); atomicMax(&bb[3], j1);
    }
    __syncthreads();
    if (threadIdx.x == 0 && bb[1] >= 0) {
        int *b = hbox + 4 * n;
        atomicMin(&b[0], bb[0]); atomicMax(&b[1], bb[1] + 1); atomicMin(&b[2], bb[2]); atomicMax(&b[3], bb[3] + 1);
    }
}

/* compute_psi_kernels, useful_functions.cpp:305-322: psi = the cells'
 * hfunc added in cell order, each cell's plane read inside its box only
 * (the terms left out are zeros: the same bits as the whole-plane sum).
 * Workgroup (x, y) on columns [256 x, 256 x + 256) of rows y, y +
 * gridDim.y, ...: for each row it first lists in LDS, in cell order, the
 * cells whose box holds the row (nc <= 1024). */
template <typename R>
__global__ __launch_bounds__(256) void g_psi_box(GP<R> P, const R *__restrict__ hf, const int4 *__restrict__ hbox,
                                                 R *__restrict__ psi) {
    __shared__ int4 lst[1024];
    __shared__ int wcnt[4];
    const int N = P.N, nc = P.nc, j = blockIdx.x * 256 + threadIdx.x;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const size_t NN = (size_t)N * N;
    for (int i = blockIdx.y; i < N; i += gridDim.y) {
        int total = 0;
        __syncthreads();                       /* the previous row's list has been read */
        for (int c0 = 0; c0 < nc; c0 += 256) {
            const int n = c0 + threadIdx.x;
            int4 b = make_int4(0, 0, 0, 0);
            bool in = false;
            if (n < nc) { b = hbox[n]; in = i >= b.x && i < b.y; }
            const unsigned long long m = __ballot(in);
            if (lane == 0) wcnt[w] = __popcll(m);
            __syncthreads();
            int off = total;
            for (int k = 0; k < w; k++) off += wcnt[k];
            if (in) lst[off + __popcll(m & ((1ull << lane) - 1ull))] = make_int4(n, b.z, b.w, 0);
            total += wcnt[0] + wcnt[1] + wcnt[2] + wcnt[3];
            __syncthreads();
        }
        if (j < N) {
            const size_t q = (size_t)i * N + j;
            R acc = 0.0;
            for (int k = 0; k < total; k++) {
                const int4 e = lst[k];
                if (j >= e.y && j < e.z) acc += hf[(size_t)e.x * NN + q];
            }
            psi[q] = acc;
        }
    }
}

/* sum_s_kernels, useful_functions.cpp:168-190, one block per cell: the area
 * sum_k h dx^2 in double, f0 = alpha (s0 - s). */
template <typename R>
__global__ __launch_bounds__(1024) void g_sum_s(GP<R> P, const R *__restrict__ hf, const double *__restrict__ s0,
                                                double *__restrict__ s, R *__restrict__ f0) {
    __shared__ double red[16];
    const size_t NN = (size_t)P.N * P.N;
    const int n = blockIdx.x;
    const R dx = P.dx;
    const R *h = hf + (size_t)n * NN;
    double acc = 0.0;
    /* each thread's pixels k = t, t + 1024, ... added in that order, the
     * loads of 32 of them in flight together (one workgroup per cell: its
     * loads in flight set the pace) */
    constexpr int U = 32;
    size_t k = threadIdx.x;
    for (; k + (size_t)(U - 1) * blockDim.x < NN; k += (size_t)U * blockDim.x) {
        R v[U];
#pragma unroll
        for (int u = 0; u < U; u++) v[u] = h[k + (size_t)u * blockDim.x];
#pragma unroll
        for (int u = 0; u < U; u++) acc += (double)(v[u] * dx * dx);
    }
    for (; k < NN; k += blockDim.x) acc += (double)(h[k] * dx * dx);
    const double tot = gblock_sum(acc, red);
    if (threadIdx.x == 0) {
        s[n] = tot;
        f0[n] = (R)(P.alpha * ((R)s0[n] - (R)tot));
    }
}

/* sum_s_kernels for FP64 state in an order that does not depend on the
 * hfunc boxes: each row's sum (lane l of a wave adds the row's columns j =
 * l mod 64 in increasing j, then the wave's fixed tree), then the rows'
 * sums in increasing row.  Every pixel outside the box is zero, so the
 * rows and columns outside it are left out and the bits stay the same.
 * g_area_rows: workgroup (b, n), wave w, on box rows r0 + 4 b + w,
 * stepping 4 gridDim.x; rows[n][i] for the box's rows only. */
__global__ __launch_bounds__(256) void g_area_rows(GP<double> P, const double *__restrict__ hf,
                                                   const int4 *__restrict__ hbox, double *__restrict__ rows) {
    const int N = P.N, n = blockIdx.y, lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int4 b = hbox[n];
    const double dx = P.dx;
    const double *h = hf + (size_t)n * N * N;
    const int c_first = b.z + ((lane - b.z) % 64 + 64) % 64;      /* the lane's first column >= c0 */
    for (int i = b.x + 4 * blockIdx.x + w; i < b.y; i += 4 * gridDim.x) {
        const double *hr = h + (size_t)i * N;
        double acc = 0.0;
        for (int j = c_first; j < b.w; j += 64) acc += (double)(hr[j] * dx * dx);
        for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
        if (lane == 0) rows[(size_t)n * N + i] = acc;
    }
}
/* workgroup n: cell n's row sums in increasing row (staged in LDS), then
 * f0 = alpha (s0 - s) */
__global__ __launch_bounds__(256) void g_area_fin(GP<double> P, const double *__restrict__ rows,
                                                  const int4 *__restrict__ hbox, const double *__restrict__ s0,
                                                  double *__restrict__ s, double *__restrict__ f0) {
    __shared__ double sp[1024];
    const int N = P.N, n = blockIdx.x;
    const int4 b = hbox[n];
    double tot = 0.0;
    for (int c = b.x; c < b.y; c += 1024) {
        const int m = min(1024, b.y - c);
        __syncthreads();
        for (int k = threadIdx.x; k < m; k += blockDim.x) sp[k] = rows[(size_t)n * N + c + k];
        __syncthreads();
        if (threadIdx.x == 0)
            for (int k = 0; k < m; k++) tot += sp[k];
    }
    if (threadIdx.x == 0) {
        s[n] = tot;
        f0[n] = (double)(P.alpha * ((double)s0[n] - (double)tot));
    }
}

/* compute_normal_to_phi, useful_functions.cpp:356-388 (interior; the ring
 * by g_bc_cells): workgroup (x, n) on rows x + 1, x + 1 + gridDim.x, ... of
 * plane n */
template <typename R>
__global__ void g_normals(GP<R> P, const R *__restrict__ phi, const R *__restrict__ ngp, R *__restrict__ jx,
                          R *__restrict__ jy) {
    const int N = P.N, n = blockIdx.y;
    const size_t NN = (size_t)N * N;
    const R dx = P.dx;
    for (int i = 1 + blockIdx.x; i < N - 1; i += gridDim.x)
        for (int j = 1 + threadIdx.x; j < N - 1; j += blockDim.x) {
            const size_t k = GPL(n, i, j);
            const R g = ngp[k];
            if (g > 0.00) {
                jx[k] = (1.0 / (2.0 * dx)) * (phi[GPL(n, i + 1, j)] - phi[GPL(n, i - 1, j)]) / g;
                jy[k] = (1.0 / (2.0 * dx)) * (phi[GPL(n, i, j + 1)] - phi[GPL(n, i, j - 1)]) / g;
            } else {
                jx[k] = 0.0;
                jy[k] = 0.0;
            }
        }
}

/* g_norm_grad and g_normals on each cell's window grown by one pixel only
 * (workgroup (r, n): row xc - zs - 1 + r of cell n, columns yc - zs - 1 ..
 * yc + zs): the same arithmetic at every pixel there.  Inside one step
 * these are the only pixels read (the spectral box and the update on the
 * window, the reshaping's normals on the window and its ring), and the
 * window check keeps them inside rows / columns 1..N-2. */
template <typename R>
__global__ void g_norm_grad_win(GP<R> P, const R *__restrict__ phi, const int *__restrict__ xc,
                                const int *__restrict__ yc, R *__restrict__ ngp) {
    const int N = P.N, n = blockIdx.y, zs = P.zs;
    const size_t NN = (size_t)N * N;
    const R dx = P.dx;
    const int i = xc[n] - zs - 1 + blockIdx.x;
    for (int j = yc[n] - zs - 1 + threadIdx.x; j <= yc[n] + zs; j += blockDim.x) {
        const R a = phi[GPL(n, i + 1, j)] - phi[GPL(n, i - 1, j)];
        const R c = phi[GPL(n, i, j + 1)] - phi[GPL(n, i, j - 1)];
        ngp[GPL(n, i, j)] = sqrt((1.0 / dx / dx / 4.0) * (a * a + c * c));
    }
}
template <typename R>
__global__ void g_normals_win(GP<R> P, const R *__restrict__ phi, const R *__restrict__ ngp,
                              const int *__restrict__ xc, const int *__restrict__ yc, R *__restrict__ jx,
                              R *__restrict__ jy) {
    const int N = P.N, n = blockIdx.y, zs = P.zs;
    const size_t NN = (size_t)N * N;
    const R dx = P.dx;
    const int i = xc[n] - zs - 1 + blockIdx.x;
    for (int j = yc[n] - zs - 1 + threadIdx.x; j <= yc[n] + zs; j += blockDim.x) {
        const size_t k = GPL(n, i, j);
        const R g = ngp[k];
        if (g > 0.00) {
            jx[k] = (1.0 / (2.0 * dx)) * (phi[GPL(n, i + 1, j)] - phi[GPL(n, i - 1, j)]) / g;
            jy[k] = (1.0 / (2.0 * dx)) * (phi[GPL(n, i, j + 1)] - phi[GPL(n, i, j - 1)]) / g;
        } else {
            jx[k] = 0.0;
            jy[k] = 0.0;
        }
    }
}

/* create_turnover_rates_tensions, rho_functions.cpp:286-405, at one pixel:
 * 0 apical, 1 basal, 2 lateral (ring neighbours n-1, n+1, strict '>'). */
template <typename R>
static __device__ __forceinline__ int g_label(const GP<R> &P, const R *__restrict__ phi, const R *__restrict__ lum,
                                              const R *__restrict__ ecm, int n, size_t g) {
    const size_t NN = (size_t)P.N * P.N;
    const int nc = P.nc;
    const int n1 = n == 0 ? nc - 1 : n - 1, n2 = n == nc - 1 ? 0 : n + 1;
    const R ph = phi[(size_t)n * NN + g];
    const R c1l = ph * lum[g], c1e = ph * ecm[g];
    int index;
    R mx;
    if (c1l > c1e) { index = 0; mx = c1l; } else { index = 1; mx = c1e; }
    const R c1c2 = ph * phi[(size_t)n1 * NN + g];
    if (c1c2 > mx) { index = 2; mx = c1c2; }
    const R c1c3 = ph * phi[(size_t)n2 * NN + g];
    if (c1c3 > mx) { index = 3; }
    return index == 3 ? 2 : index;
}

/* update_all_cortices, first part (rho_functions.cpp:413-440): per cell and
 * FFT-box pixel, the five transform inputs load_temp_fft puts in temp_fft
 * (useful_functions.cpp:605-625): rho, the reaction term of create_deg_gen
 * (rho_functions.cpp:193-209) with prod_norm_delta (rho_functions.cpp:13-30),
 * and pi_xx, pi_xy, pi_yy (rho_functions.cpp:32-115); the turnover maps at
 * the pixel from its label.  in5: [5][nc][M][M]. */
template <typename R>
__global__ void g_spec_load(GP<R> P, const R *__restrict__ phi, const R *__restrict__ rho,
                            const R *__restrict__ ngp, const R *__restrict__ lum, const R *__restrict__ ecm,
                            const int *__restrict__ xc, const int *__restrict__ yc, Cx<R> *__restrict__ in5) {
    const int N = P.N, M = P.M, zs = P.zs, n = blockIdx.y;
    const size_t NN = (size_t)N * N, MM = (size_t)M * M, PL5 = MM * P.nc;
    const R dx = P.dx, sigma = P.sigma;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < M * M; k += gridDim.x * blockDim.x) {
        const int i = k / M, j = k % M;
        const int gi = i + 1 + xc[n] - zs, gj = j + 1 + yc[n] - zs;
        const size_t g = GG(gi, gj), q = (size_t)n * NN + g;
        const int lab = g_label(P, phi, lum, ecm, n, g);
        const R ph = phi[q], r = rho[q], gr = ngp[q];
        const R pnd = gr * exp(-(ph - 0.5) * (ph - 0.5) / 2.0 / sigma / sigma) / sqrt(2.0 * M_PI) / sigma;
        const R deg = P.deg[lab], gen = P.gen[lab], at = P.atan_[lab], ao = P.aort[lab], bb = P.bb[lab];
        const R dxp = phi[q + N] - phi[q - N];
        const R dyp = phi[q + 1] - phi[q - 1];
        const R base = -at * r * r * r + bb * r * r * r * r;
        R pxx = base, pxy = 0.0, pyy = base;
        if (gr > 0.0) {
            pxx += (at - ao) * r * r * r * (1.0 / (4.0 * dx * dx)) * dxp * dxp / gr / gr;
            pxy += (at - ao) * r * r * r * (1.0 / (4.0 * dx * dx)) * dxp * dyp / gr / gr;
            pyy += (at - ao) * r * r * r * (1.0 / (4.0 * dx * dx)) * dyp * dyp / gr / gr;
        }
        const size_t o = (size_t)n * MM + k;
        Cx<R> v;
        v.y = 0.0;
        v.x = r; in5[o] = v;
        v.x = -deg * r + gen * pnd; in5[PL5 + o] = v;
        v.x = pxx; in5[2 * PL5 + o] = v;
        v.x = pxy; in5[3 * PL5 + o] = v;
        v.x = pyy; in5[4 * PL5 + o] = v;
    }
}

/* update_rho_velocity, rho_functions.cpp:133-190: per mode, the 2x2 force
 * balance solved for (v_x, v_y); the viscosity is the turnover map's value at
 * the real-space pixel with the mode's indices (:149-150).  f5 as in5
 * transformed; v2: [2][nc][M][M]. */
template <typename R>
__global__ void g_spec_solve(GP<R> P, const R *__restrict__ phi, const R *__restrict__ lum,
                             const R *__restrict__ ecm, const int *__restrict__ xc, const int *__restrict__ yc,
                             const Cx<R> *__restrict__ f5, Cx<R> *__restrict__ v2) {
    const int N = P.N, M = P.M, zs = P.zs, n = blockIdx.y;
    const size_t MM = (size_t)M * M, PL5 = MM * P.nc;
    const R dx = P.dx, xi = P.friction;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < M * M; k += gridDim.x * blockDim.x) {
        const int i = k / M, j = k % M;
        const int ir = (i < M / 2) ? i : i - M, jr = (j < M / 2) ? j : j - M;      /* :159-162 */
        const R kx = (2.0 * M_PI / (M * dx)) * ir;                                  /* :164 */
        const R ky = (2.0 * M_PI / (M * dx)) * jr;                                  /* :165 */
        const size_t g = GG(i + 1 + xc[n] - zs, j + 1 + yc[n] - zs);
        const R eta = P.visc[g_label(P, phi, lum, ecm, n, g)];
        const size_t o = (size_t)n * MM + k;
        const Cx<R> pxx = f5[2 * PL5 + o], pxy = f5[3 * PL5 + o], pyy = f5[4 * PL5 + o];
        const R D = -(4.0 * eta * eta * kx * kx * ky * ky + xi * xi + 3.0 * eta * xi * (kx * kx + ky * ky)
                      + 2.0 * eta * eta * (kx * kx * kx * kx + ky * ky * ky * ky));
        const R A = xi + 2.0 * eta * (kx * kx + ky * ky / 2.0);
        Cx<R> y, x;
        y.x = 0.0; y.y = 0.0; x.x = 0.0; x.y = 0.0;
        y.x += eta * kx * ky * (kx * pxx.y + ky * pxy.y);                     /* :170 */
        y.x -= A * (ky * pyy.y + kx * pxy.y);                                  /* :171 */
        y.x /= D;                                                              /* :172 */
        y.y -= eta * kx * ky * (kx * pxx.x + ky * pxy.x);                     /* :175 */
        y.y += A * (ky * pyy.x + kx * pxy.x);                                  /* :176 */
        y.y /= D;                                                              /* :177 */
        x.x += kx * pxx.y + ky * pxy.y - eta * kx * ky * y.x;                 /* :181 */
        x.x /= A;                                                              /* :182 */
        x.y += -kx * pxx.x - ky * pxy.x - eta * kx * ky * y.y;                /* :185 */
        x.y /= A;                                                              /* :186 */
        v2[o] = x;
        v2[PL5 + o] = y;
    }
}

/* load_fft_rho_v, rho_functions.cpp:267-284, for one component:
 * out = rho Re(F^-1 v)/M^2 (Im 0); the velocity on the box to vbox. */
template <typename R>
__global__ void g_spec_rhov(GP<R> P, const R *__restrict__ rho, const int *__restrict__ xc,
                            const int *__restrict__ yc, const Cx<R> *__restrict__ treal, Cx<R> *__restrict__ out,
                            R *__restrict__ vbox) {
    const int N = P.N, M = P.M, zs = P.zs, n = blockIdx.y;
    const size_t NN = (size_t)N * N, MM = (size_t)M * M;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < M * M; k += gridDim.x * blockDim.x) {
        const int i = k / M, j = k % M;
        const size_t q = (size_t)n * NN + GG(i + 1 + xc[n] - zs, j + 1 + yc[n] - zs);
        const size_t o = (size_t)n * MM + k;
        Cx<R> t;
        t.x = rho[q] * treal[o].x / (M * M);
        t.y = 0.0;
        out[o] = t;
        if (vbox) vbox[o] = treal[o].x / (M * M);
    }
}

/* update_fft_rho, rho_functions.cpp:213-251 (forward Euler, no diffusion) */
template <typename R>
__global__ void g_spec_update(GP<R> P, Cx<R> *__restrict__ frho, const Cx<R> *__restrict__ freact,
                              const Cx<R> *__restrict__ frvx, const Cx<R> *__restrict__ frvy) {
    const int M = P.M;
    const size_t MM = (size_t)M * M, tot = MM * P.nc;
    const R dt = P.dt, dx = P.dx;
    for (size_t o = (size_t)blockIdx.x * blockDim.x + threadIdx.x; o < tot; o += (size_t)gridDim.x * blockDim.x) {
        const int k = (int)(o % MM), i = k / M, j = k % M;
        Cx<R> f = frho[o];
        f.x += dt * freact[o].x;                                               /* :222 */
        f.y += dt * freact[o].y;                                               /* :223 */
        const int ir = (i < M / 2) ? i : i - M, jr = (j < M / 2) ? j : j - M;
        f.x += dt * (2.0 * M_PI / (M * dx)) * (ir * frvx[o].y + jr * frvy[o].y);   /* :243 */
        f.y -= dt * (2.0 * M_PI / (M * dx)) * (ir * frvx[o].x + jr * frvy[o].x);   /* :247 */
        frho[o] = f;
    }
}

/* copy_to_rhon, rho_functions.cpp:253-265: rhon = Re(F^-1 rho^)/M^2 on the box */
template <typename R>
__global__ void g_spec_commit(GP<R> P, const int *__restrict__ xc, const int *__restrict__ yc,
                              const Cx<R> *__restrict__ t, R *__restrict__ rhon) {
    const int N = P.N, M = P.M, zs = P.zs, n = blockIdx.y;
    const size_t NN = (size_t)N * N, MM = (size_t)M * M;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < M * M; k += gridDim.x * blockDim.x) {
        const int i = k / M, j = k % M;
        rhon[(size_t)n * NN + GG(i + 1 + xc[n] - zs, j + 1 + yc[n] - zs)] = t[(size_t)n * MM + k].x / (M * M);
    }
}

/* copy_cell_tables_1_by_1's rho <- rhon (useful_functions.cpp:679-693) on
 * the spectral boxes g_spec_commit writes (workgroup (i, n): box row i of
 * cell n), for a step whose rhon equalled rho elsewhere */
template <typename R>
__global__ void g_copy_boxes(GP<R> P, const int *__restrict__ xc, const int *__restrict__ yc,
                             const R *__restrict__ rhon, R *__restrict__ rho) {
    const int N = P.N, M = P.M, zs = P.zs, n = blockIdx.y;
    const size_t NN = (size_t)N * N;
    const size_t row = (size_t)n * NN + GG(blockIdx.x + 1 + xc[n] - zs, 1 + yc[n] - zs);
    for (int j = threadIdx.x; j < M; j += blockDim.x) rho[row + j] = rhon[row + j];
}

/* 9-point Laplacian numerator of cell_phase_functions.cpp (e.g. :198) */
#define GL9(t, g) (-12.0 * (t)[(g)] + 2.0 * ((t)[(g) + N] + (t)[(g) - N] + (t)[(g) + 1] + (t)[(g) - 1]) \
                   + (t)[(g) + N + 1] + (t)[(g) + N - 1] + (t)[(g) - N - 1] + (t)[(g) - N + 1])

/* Fixed-order double partial sums of f(k) over k < n: block b sums the
 * contiguous range [b * per, (b + 1) * per), part[b]. */
template <typename R>
__global__ __launch_bounds__(1024) void g_sum_parts(const R *__restrict__ a, const R *__restrict__ b, R dx,
                                                    size_t n, size_t per, int mode, double *__restrict__ part) {
    __shared__ double red[16];
    double acc = 0.0;
    const size_t k0 = (size_t)blockIdx.x * per, k1 = k0 + per < n ? k0 + per : n;
    for (size_t k = k0 + threadIdx.x; k < k1; k += blockDim.x) {
        if (mode == 0) acc += (double)(a[k] * dx * dx);                      /* sum of h(ECM) dx^2 */
        else acc += fabs((double)(a[k] - b[k])) * dx * dx;                   /* reshaping deviation */
    }
    const double t = gblock_sum(acc, red);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}
/* the parts added in order (nparts <= 1024): all loaded into LDS at once,
 * then one thread's sum there */
__global__ void g_sum_final(const double *__restrict__ part, int nparts, double *__restrict__ out) {
    __shared__ double sp[1024];
    for (int k = threadIdx.x; k < nparts; k += blockDim.x) sp[k] = part[k];
    __syncthreads();
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double s = 0.0;
        for (int k = 0; k < nparts; k++) s += sp[k];
        *out = s;
    }
}

/* update_lumen_and_ECM, cell_phase_functions.cpp:372-444 (interior), then
 * boundary_conditions_lumen_and_ECM, useful_functions.cpp:439-462. */
template <typename R>
__global__ void g_lumecm_update(GP<R> P, const R *__restrict__ lum, const R *__restrict__ ecm,
                                const R *__restrict__ psi, const R *__restrict__ hl, const R *__restrict__ he,
                                const double *__restrict__ sum_ecm, R *__restrict__ lumn, R *__restrict__ ecmn) {
    const int N = P.N;
    const size_t NN = (size_t)N * N;
    const R dx = P.dx, dt = P.dt;
    const R sum_e = (R)(*sum_ecm);
    (void)NN;
    for (int i = blockIdx.x; i < N; i += gridDim.x)                        /* workgroup x on rows x, x + gridDim.x, ... */
    for (int j = threadIdx.x; j < N; j += blockDim.x) {
        const size_t q = (size_t)i * N + j;
        if (i < 1 || i > N - 2 || j < 1 || j > N - 2) {
            lumn[q] = 0.0;                                                  /* edges and corners */
            ecmn[q] = 1.0;
            continue;
        }
        const R l = lum[q], e = ecm[q];
        const R lap = GL9(psi, q);
        R ln = l;                                                          /* :408 */
        ln += (1.0 / P.tau_lum) * l * (1.0 - l) * dt * (                   /* :409 */
            P.xi_lum - P.beta_lc * psi[q] + P.eta_lc * lap / (4.0 * dx * dx) - P.beta_le * he[q]);
        R en = e;                                                          /* :424 */
        en += (1.0 / P.tau_E) * e * (1.0 - e) * dt * (                     /* :425 */
            P.alpha_E * (N * N * dx * dx - sum_e) + P.xi_E - P.beta_ec * psi[q]
            + P.eta_ec * lap / (4.0 * dx * dx) - P.beta_le * hl[q]);
        lumn[q] = ln;
        ecmn[q] = en;
    }
}

/* reshaping_lumen_or_ECM's |grad| of the old field (cell_phase_functions.cpp:
 * 734-736) and normals from the new one over it (:757-763), interior; the
 * rings by g_bc_cells. */
template <typename R>
__global__ void g_lum_grad(GP<R> P, const R *__restrict__ old, R *__restrict__ ng) {
    const int N = P.N;
    const R dx = P.dx;
    for (int i = 1 + blockIdx.x; i < N - 1; i += gridDim.x)
        for (int j = 1 + threadIdx.x; j < N - 1; j += blockDim.x) {
            const size_t q = (size_t)i * N + j;
            const R a = old[q + N] - old[q - N], c = old[q + 1] - old[q - 1];
            ng[q] = sqrt((1.0 / dx / dx / 4.0) * (a * a + c * c));
        }
}
template <typename R>
__global__ void g_lum_normals(GP<R> P, const R *__restrict__ nw, const R *__restrict__ ng, R *__restrict__ jx,
                              R *__restrict__ jy) {
    const int N = P.N;
    const R dx = P.dx;
    for (int i = 1 + blockIdx.x; i < N - 1; i += gridDim.x)
        for (int j = 1 + threadIdx.x; j < N - 1; j += blockDim.x) {
            const size_t q = (size_t)i * N + j;
            const R g = ng[q];
            if (g != 0.0) {
                jx[q] = (1.0 / (2.0 * dx)) * (nw[q + N] - nw[q - N]) / g;
                jy[q] = (1.0 / (2.0 * dx)) * (nw[q + 1] - nw[q - 1]) / g;
            } else {
                jx[q] = 0.0;
                jy[q] = 0.0;
            }
        }
}

/* The loop-invariant factors of the reshaping flux (cell_phase_functions.cpp:
 * 511-547 and :802-838), each formed as the reference's expression forms it
 * (left to right, FP contraction off): a = wi wi (rdt / dx / dx) in R, and
 * in double a4 = a 0.25 and c = rdt wi sqrt(2) / dx 0.5.  Formed once per
 * kernel: within a loop the compiler cannot hoist them itself (they are read
 * through a reference the loop's stores might alias), and c alone is a
 * double division. */
template <typename R>
struct GFluxK {
    R a;
    double a4, c;
};
template <typename R>
static __device__ __forceinline__ GFluxK<R> g_flux_k(R wi, R rdt, R dx) {
    GFluxK<R> k;
    k.a = wi * wi * (rdt / dx / dx);
    k.a4 = k.a * 0.25;
    k.c = rdt * wi * sqrt(2.0) / dx * 0.5;
    return k;
}

/* The flux through the face between a pixel and its neighbour at row
 * offset +1 (g_face_x) or column offset +1 (g_face_y), with the relaxed
 * field read through U(a, b) (offsets from the pixel) and the normals at
 * jx / jy + qj (row stride sj): one of the four face terms of the reshaping
 * update (cell_phase_functions.cpp:511-547, :802-838).  The neighbour's
 * opposite term takes the same operands in the same order, so it is the
 * same value: the faces can be computed once and shared. */
#define JX(a, b) jx[qj + (long)(a) * sj + (b)]
#define JY(a, b) jy[qj + (long)(a) * sj + (b)]
template <typename R, typename FU>
static __device__ __forceinline__ double g_face_x(FU U, const R *__restrict__ jx, const R *__restrict__ jy, size_t qj,
                                                  int sj, const GFluxK<R> &K) {
    return 0.5 * (JX(1, 0) + JX(0, 0)) * (
        K.a * (U(1, 0) - U(0, 0)) * 0.5 * (JX(1, 0) + JX(0, 0))
        + K.a4 * (U(1, 1) + U(0, 1) - U(1, -1) - U(0, -1)) * 0.5 * (JY(1, 0) + JY(0, 0))
        - K.c * (U(1, 0) * (1.0 - U(1, 0)) + U(0, 0) * (1.0 - U(0, 0))));
}
template <typename R, typename FU>
static __device__ __forceinline__ double g_face_y(FU U, const R *__restrict__ jx, const R *__restrict__ jy, size_t qj,
                                                  int sj, const GFluxK<R> &K) {
    return 0.5 * (JY(0, 1) + JY(0, 0)) * (
        K.a * (U(0, 1) - U(0, 0)) * 0.5 * (JY(0, 1) + JY(0, 0))
        + K.a4 * (U(1, 1) + U(1, 0) - U(-1, 1) - U(-1, 0)) * 0.5 * (JX(0, 1) + JX(0, 0))
        - K.c * (U(0, 1) * (1.0 - U(0, 1)) + U(0, 0) * (1.0 - U(0, 0))));
}
#undef JX
#undef JY

/* The interface flux at a pixel, the relaxed field read through U(a, b):
 * the faces below and right added, above and left subtracted, each rounded
 * to R as the reference's `+=` / `-=` do. */
template <typename R, typename FU>
static __device__ __forceinline__ R g_flux(R un, FU U, const R *__restrict__ jx, const R *__restrict__ jy, size_t qj,
                                           int sj, const GFluxK<R> &K) {
    un += g_face_x(U, jx, jy, qj, sj, K);
    un += g_face_y(U, jx, jy, qj, sj, K);
    un -= g_face_x([&](int a, int b) { return U(a - 1, b); }, jx, jy, qj - sj, sj, K);
    un -= g_face_y([&](int a, int b) { return U(a, b - 1); }, jx, jy, qj - 1, sj, K);
    return un;
}

/* The conservative interface flux of both reshaping loops at grid index q
 * (row stride N): the new value of un given the relaxed field u. */
template <typename R>
static __device__ __forceinline__ R g_reshape(R un, const R *__restrict__ u, const R *__restrict__ jx,
                                              const R *__restrict__ jy, int N, size_t q, const GFluxK<R> &K) {
    return g_flux(un, [&](int a, int b) { return u[q + (long)a * N + b]; }, jx, jy, q, N, K);
}

/* One sweep of reshaping_lumen_or_ECM's loop (cell_phase_functions.cpp:
 * 777-844): the reference copies lum <- lumn and updates lumn from its own
 * value; here u holds the previous iterate and un the buffer of the one
 * before, so the update starts from u[q] (the value the copy would leave in
 * lumn).  Then the Dirichlet ring of boundary_conditions_dirichlet
 * (useful_functions.cpp:422-437: edges `side`, corners 0). */
template <typename R>
__global__ void g_lum_sweep(GP<R> P, const R *__restrict__ u, R *__restrict__ un, const R *__restrict__ jx,
                            const R *__restrict__ jy, R side) {
    const int N = P.N;
    const GFluxK<R> K = g_flux_k(P.wi, P.rdt, P.dx);
    for (int i = blockIdx.x; i < N; i += gridDim.x)                        /* workgroup x on rows x, x + gridDim.x, ... */
    for (int j = threadIdx.x; j < N; j += blockDim.x) {
        const size_t q = (size_t)i * N + j;
        if (i < 1 || i > N - 2 || j < 1 || j > N - 2) {
            const bool corner = (i == 0 || i == N - 1) && (j == 0 || j == N - 1);
            un[q] = corner ? (R)0.0 : side;
            continue;
        }
        un[q] = g_reshape(u[q], u, jx, jy, N, q, K);
    }
}

/* update_all_cells, cell_phase_functions.cpp:126-217, on every cell's window
 * (block (r, n): window row r of cell n); temp_phase = psi - hfunc_n
 * (:141-153) is formed where the Laplacian reads it. */
template <typename R>
__global__ void g_update_cells(GP<R> P, const R *__restrict__ phi, const R *__restrict__ rho,
                               const R *__restrict__ ngp, const R *__restrict__ psi, const R *__restrict__ hf,
                               const R *__restrict__ hl, const R *__restrict__ he, const R *__restrict__ f0,
                               const int *__restrict__ xc, const int *__restrict__ yc, R *__restrict__ phin) {
    const int N = P.N, zs = P.zs, n = blockIdx.y;
    const size_t NN = (size_t)N * N;
    const R dx = P.dx, dt = P.dt, tau = P.tau, sigma = P.sigma;
    const R *ph = phi + (size_t)n * NN, *rh = rho + (size_t)n * NN, *ng = ngp + (size_t)n * NN;
    const R *hn = hf + (size_t)n * NN;
    const int i = xc[n] - zs + blockIdx.x;
    const R f0n = f0[n];
    for (int j = yc[n] - zs + threadIdx.x; j < yc[n] + zs; j += blockDim.x) {
        const size_t q = GG(i, j);
        R tp[9];                                   /* temp_phase on the 3 x 3 stencil */
        {
            int k = 0;
            for (int a = -1; a <= 1; a++)
                for (int b = -1; b <= 1; b++, k++) {
                    const size_t g = q + (long)a * N + b;
                    tp[k] = psi[g] - hn[g];
                }
        }
        const double l9tp = -12.0 * tp[4] + 2.0 * (tp[7] + tp[1] + tp[5] + tp[3]) + tp[8] + tp[6] + tp[0] + tp[2];
        const R v = ph[q];
        R pn = v;                                                                      /* :170 */
        if (ng[q] > 0.0) {                                                             /* :177-186 */
            pn -= P.cpl / tau * (dt / dx / dx / 4.0) * exp(-(v - 0.5) * (v - 0.5) / 2.0 / sigma / sigma)
                  / sqrt(2.0 * M_PI) / sigma
                  * ((rh[q + N] - rh[q - N]) * (ph[q + N] - ph[q - N]) + (rh[q + 1] - rh[q - 1]) * (ph[q + 1] - ph[q - 1]))
                  / ng[q];
        }
        pn += (1.0 / tau) * v * (1.0 - v) * dt * (                                     /* :190-211 */
            f0n
            - P.beta * tp[4]
            + P.eta * l9tp / (4.0 * dx * dx)
            - P.beta_lc * hl[q]
            + P.eta_lc * GL9(hl, q) / (4.0 * dx * dx)
            - P.beta_ec * he[q]
            + P.eta_ec * GL9(he, q) / (4.0 * dx * dx));
        phin[(size_t)n * NN + q] = pn;
    }
}

/* reshaping_cells, cell_phase_functions.cpp:447-580: one block per cell runs
 * the cell's whole loop on its window (copy phase <- phasen, flux update,
 * deviation sum_window |phasen - phase| dx^2 / N_current in double, stop
 * when <= rtol 2 pi R wi rdt); phase ends at the second-to-last iterate. */
template <typename R>
__global__ __launch_bounds__(1024) void g_relax_cells(GP<R> P, R *__restrict__ phi, R *__restrict__ phin,
                                                      const R *__restrict__ jx, const R *__restrict__ jy,
                                                      const int *__restrict__ xc, const int *__restrict__ yc,
                                                      double thr, double rdev0, int *__restrict__ counts,
                                                      int *__restrict__ relax_flag) {
    __shared__ double red[16];
    const int N = P.N, zs = P.zs, n = blockIdx.x, W = 2 * zs;
    const size_t NN = (size_t)N * N;
    R *ph = phi + (size_t)n * NN, *pn = phin + (size_t)n * NN;
    const R *jxn = jx + (size_t)n * NN, *jyn = jy + (size_t)n * NN;
    const int x0 = xc[n] - zs, y0 = yc[n] - zs;
    const R dx = P.dx;
    const GFluxK<R> K = g_flux_k(P.wi, P.rdt, dx);
    double rdev = rdev0;
    int ri = 0, count = 0;
    while (rdev > thr) {
        for (int k = threadIdx.x; k < W * W; k += blockDim.x) {
            const size_t q = GG(x0 + k / W, y0 + k % W);
            ph[q] = pn[q];
        }
        __syncthreads();
        for (int k = threadIdx.x; k < W * W; k += blockDim.x) {
            const size_t q = GG(x0 + k / W, y0 + k % W);
            pn[q] = g_reshape(pn[q], ph, jxn, jyn, N, q, K);
        }
        __syncthreads();
        double acc = 0.0;
        for (int k = threadIdx.x; k < W * W; k += blockDim.x) {
            const size_t q = GG(x0 + k / W, y0 + k % W);
            acc += fabs((double)(pn[q] - ph[q])) * dx * dx / P.ncur;
        }
        rdev = gblock_sum(acc, red);
        count++;
        if (ri > P.rimax) { if (threadIdx.x == 0) *relax_flag = 1; break; }   /* :565-568 */
        ri++;
    }
    if (threadIdx.x == 0) counts[n] = count;
}

/* ---- loops split over several workgroups ------------------------------------
 * The reshaping loops on big windows and on the whole grid run on the
 * device in one launch each: a group of workgroups shares every sweep, each
 * on a band of rows, and meets at a barrier after it.  Only the band's
 * first and last rows (and the grid's boundary ring) are read by other
 * workgroups: they are stored and loaded with agent-scope atomics, which
 * reach every XCD's view without a cache flush, everything else with plain
 * stores and loads that only the same workgroup reads back.  So the barrier
 * needs no release or acquire: every thread waits for its own stores to
 * complete, then thread 0 arrives on a monotonic counter (the e-th barrier
 * of the launch releases at e * members) and polls it with a sleep.  All
 * workgroups of the launch must be resident together: the host launches at
 * most one per compute unit.  A wait longer than `limit` wall-clock ticks,
 * or one that finds the abort flag raised, raises the flag and every
 * workgroup leaves its loop (ORG_ETIMEOUT). */
static __device__ bool g_barrier(unsigned *ctr, unsigned target, int *abort_flag, long long limit) {
    __shared__ int ok;
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) {
        int good = 1;
        unsigned seen = __hip_atomic_fetch_add(ctr, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1u;
        const long long t0 = wall_clock64();
        while (seen < target) {
            __builtin_amdgcn_s_sleep(1);
            seen = __hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (seen >= target) break;
            if (__hip_atomic_load(abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ||
                wall_clock64() - t0 > limit) {
                atomicExch(abort_flag, 1);
                good = 0;
                break;
            }
        }
        ok = good;
    }
    __syncthreads();
    return ok != 0;
}

/* The group partials at p (n <= 1024, agent-scope loads, one per thread)
 * into LDS; after a __syncthreads, g_lds_sum adds them in order: the same
 * value in every workgroup of the group. */
static __device__ __forceinline__ void g_group_load(const double *p, int n, double *lds) {
    if ((int)threadIdx.x < n)
        lds[threadIdx.x] = __hip_atomic_load(const_cast<double *>(&p[threadIdx.x]), __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT);
}
static __device__ __forceinline__ double g_lds_sum(const double *lds, int n) {
    double r = 0.0;
    for (int q = 0; q < n; q++) r += lds[q];
    return r;
}

template <typename T>
static __device__ __forceinline__ void g_xs(T *p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
static __device__ __forceinline__ T g_xl(const T *p) {
    return __hip_atomic_load(const_cast<T *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

/* One pixel of a banded sweep: the relaxed field's rows above, at and below
 * the pixel (up, mid, dn: each the pixel's column in that row; a
 * neighbour's row comes from its copy in LDS) and the normals on the grid;
 * u0 returns the centre value. */
template <typename R>
static __device__ __forceinline__ R g_flux_rows(const R *up, const R *mid, const R *dn, const R *__restrict__ jx,
                                                const R *__restrict__ jy, size_t qj, int sj, const GFluxK<R> &K,
                                                R &u0) {
    auto U = [&](int a, int b) { return a < 0 ? up[b] : (a > 0 ? dn[b] : mid[b]); };
    u0 = mid[0];
    return g_flux(u0, U, jx, jy, qj, sj, K);
}

/* reshaping_cells (cell_phase_functions.cpp:447-580) with K workgroups per
 * cell (blockIdx.x = n K + b), workgroup b on window rows [b W / K,
 * (b + 1) W / K).  The iterates live in two (W + 2)^2 window buffers per
 * cell whose one-pixel ring holds phi around the window (what the
 * reference's stencil reads there); the first holds phin's window.  Sweep
 * k reads one buffer and writes the other, so the reference's copy
 * phase <- phasen before each sweep (:480-492) is the change of roles, and
 * one barrier per sweep suffices.  A workgroup reads and writes the window
 * buffers only on its own rows (plain accesses); the first and last rows
 * of every sweep's result also go to an exchange buffer with agent-scope
 * stores, from which the neighbours load them into LDS with agent-scope
 * loads after the barrier (the two rows next to the first and last band
 * are the static ring rows).  The deviation is each workgroup's FP64 sum
 * over its band, the K parts added in order by every workgroup (so all take
 * the same stop decision).  At the end phin's window gets the last iterate
 * and phi's the second-to-last.  S: [nc][2][W + 2][W + 2]; X: [nc][2
 * parities][K][2][W + 2]; part: [nc][2][K]; ctr: [nc] arrival counters,
 * zero at launch.  Dynamic LDS: 2 (W + 2) reals. */
template <typename R>
__global__ __launch_bounds__(1024) void g_relax_cells_mb(GP<R> P, R *__restrict__ phi, R *__restrict__ phin,
                                                         const R *__restrict__ jx, const R *__restrict__ jy,
                                                         const int *__restrict__ xc, const int *__restrict__ yc,
                                                         double thr, double rdev0, R *__restrict__ S,
                                                         R *__restrict__ X, double *__restrict__ part,
                                                         unsigned *__restrict__ ctr, int *__restrict__ counts,
                                                         int *__restrict__ relax_flag, int *__restrict__ abort_flag,
                                                         long long limit, int K) {
    __shared__ double red[16];
    __shared__ double sums[1024];
    extern __shared__ char g_dyn[];
    const int N = P.N, zs = P.zs, W = 2 * zs, W2 = W + 2;
    R *hu = (R *)g_dyn, *hd = hu + W2;                       /* the rows above and below the band */
    const int n = blockIdx.x / K, b = blockIdx.x % K;
    const size_t NN = (size_t)N * N, SS = (size_t)W2 * W2;
    R *ph = phi + (size_t)n * NN, *pn = phin + (size_t)n * NN;
    const R *jxn = jx + (size_t)n * NN, *jyn = jy + (size_t)n * NN;
    R *S0 = S + (size_t)n * 2 * SS, *S1 = S0 + SS;
    R *Xn = X + (size_t)n * 2 * K * 2 * W2;                  /* [2][K][2][W2] */
    auto xrow = [&](int par, int blk, int e) { return Xn + (((size_t)par * K + blk) * 2 + e) * W2; };
    const int x0 = xc[n] - zs, y0 = yc[n] - zs;              /* window origin on the grid */
    const int r0 = b * W / K, r1 = (b + 1) * W / K;          /* this workgroup's window rows */
    const int rows = r1 - r0;
    /* buffer rows r0+1 .. r1 (window rows r0 .. r1-1) and, for the first and
     * last workgroup, the ring rows 0 and W+1: ring columns from phi, the
     * window from phin; the band's first and last rows also to the exchange */
    const int f0 = (b == 0) ? 0 : r0 + 1, f1 = (b == K - 1) ? W2 : r1 + 1;
    for (int k = threadIdx.x; k < (f1 - f0) * W2; k += blockDim.x) {
        const int sr = f0 + k / W2, sc = k % W2;
        const size_t g = (size_t)(x0 - 1 + sr) * N + (y0 - 1 + sc);
        const bool ring = sr == 0 || sr == W + 1 || sc == 0 || sc == W + 1;
        const R v = ring ? ph[g] : pn[g];
        const size_t q = (size_t)sr * W2 + sc;
        S0[q] = v;
        if (ring) S1[q] = v;
        if (sr == r0 + 1) g_xs(&xrow(0, b, 0)[sc], v);
        if (sr == r1) g_xs(&xrow(0, b, 1)[sc], v);
    }
    unsigned epoch = 1;
    if (!g_barrier(&ctr[n], epoch * K, abort_flag, limit)) return;
    const R dx = P.dx;
    const GFluxK<R> FK = g_flux_k(P.wi, P.rdt, dx);
    double rdev = rdev0;
    int ri = 0, count = 0, cur = 0;
    /* halo rows of the sweep to come (and, from the second sweep on, the
     * previous sweep's partials) are loaded together after each barrier */
    auto load_halo = [&](const R *src, int par) {
        for (int k = threadIdx.x; k < 2 * W2; k += blockDim.x) {
            const int e = k / W2, c = k % W2;
            R v;
            if (e == 0) v = b == 0 ? src[c] : g_xl(&xrow(par, b - 1, 1)[c]);
            else v = b == K - 1 ? src[(size_t)(W + 1) * W2 + c] : g_xl(&xrow(par, b + 1, 0)[c]);
            hu[k] = v;
        }
    };
    load_halo(S0, 0);
    __syncthreads();
    while (rdev > thr) {
        const R *src = cur ? S1 : S0;
        R *dst = cur ? S0 : S1;
        const int par = count & 1;
        double acc = 0.0;
        for (int k = threadIdx.x; k < rows * W; k += blockDim.x) {
            const int wr = r0 + k / W, wc = k % W;
            const size_t qs = (size_t)(wr + 1) * W2 + (wc + 1);
            const size_t qj = (size_t)(x0 + wr) * N + (y0 + wc);
            const R *up = wr == r0 ? hu + (wc + 1) : src + (qs - W2);
            const R *dn = wr == r1 - 1 ? hd + (wc + 1) : src + (qs + W2);
            R o;
            const R v = g_flux_rows(up, src + qs, dn, jxn, jyn, qj, N, FK, o);
            dst[qs] = v;
            if (wr == r0) g_xs(&xrow(par ^ 1, b, 0)[wc + 1], v);
            if (wr == r1 - 1) g_xs(&xrow(par ^ 1, b, 1)[wc + 1], v);
            acc += fabs((double)(v - o)) * dx * dx / P.ncur;
        }
        /* the exchanged rows' ring columns (static) */
        if (threadIdx.x < 4) {
            const int e = threadIdx.x >> 1, c = (threadIdx.x & 1) ? W + 1 : 0;
            const size_t qs = (size_t)(e ? r1 : r0 + 1) * W2 + c;
            g_xs(&xrow(par ^ 1, b, e)[c], src[qs]);
        }
        const double t = gblock_sum(acc, red);
        if (threadIdx.x == 0) g_xs(&part[((size_t)n * 2 + par) * K + b], t);
        epoch++;
        if (!g_barrier(&ctr[n], epoch * K, abort_flag, limit)) return;
        g_group_load(&part[((size_t)n * 2 + par) * K], K, sums);
        load_halo(dst, par ^ 1);                     /* for the next sweep (unused if the loop stops) */
        __syncthreads();
        rdev = g_lds_sum(sums, K);
        count++;
        cur ^= 1;
        if (ri > P.rimax) { if (threadIdx.x == 0 && b == 0) *relax_flag = 1; break; }   /* :565-568 */
        ri++;
    }
    /* phin <- the last iterate, phi <- the one before (the reference's copy
     * at the start of the last sweep); own rows only */
    const R *last = cur ? S1 : S0, *prev = cur ? S0 : S1;
    for (int k = threadIdx.x; k < rows * W; k += blockDim.x) {
        const int wr = r0 + k / W, wc = k % W;
        const size_t qs = (size_t)(wr + 1) * W2 + (wc + 1);
        const size_t g = (size_t)(x0 + wr) * N + (y0 + wc);
        pn[g] = last[qs];
        if (count > 0) ph[g] = prev[qs];
    }
    if (threadIdx.x == 0 && b == 0) counts[n] = count;
}

/* Tagged exchange words: 32 bits of payload in the low half,
 * the tag of the exchange in the high half, stored and loaded as one 64-bit
 * agent-scope access (single-copy atomic and coherent across the XCDs'
 * L2s): a reader that finds its tag holds that exchange's payload, so the
 * workgroups of a loop meet without a barrier, a wait for store completion
 * or a cache flush.  Tags grow by one per sweep; the host passes each
 * launch a base above every tag used before (the buffers are zeroed at
 * set-up and no exchange uses tag 0). */
static __device__ __forceinline__ void g_tput(unsigned long long *p, unsigned v, unsigned tag) {
    __hip_atomic_store(p, ((unsigned long long)tag << 32) | (unsigned long long)v, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}
static __device__ __forceinline__ unsigned long long g_tget(const unsigned long long *p) {
    return __hip_atomic_load(const_cast<unsigned long long *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
/* poll *p until it carries tag (w: the word as first read); false once the
 * wait passes `limit` wall-clock ticks or the abort flag is up */
static __device__ bool g_twait(const unsigned long long *p, unsigned tag, unsigned long long &w, int *abort_flag,
                               long long limit) {
    if ((unsigned)(w >> 32) == tag) return true;
    const long long t0 = wall_clock64();
    for (int k = 1;; k++) {
        __builtin_amdgcn_s_sleep(1);
        w = g_tget(p);
        if ((unsigned)(w >> 32) == tag) return true;
        if ((k & 15) == 0 && (__hip_atomic_load(abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ||
                              wall_clock64() - t0 > limit))
            return false;
    }
}

/* The banded reshaping loops (g_relax_cells_lds, g_lum_loop_lds; FP32 or
 * FP64 state) run S = 1 or 2 sweeps per exchange (ORG_GEN_SPAN_CELLS,
 * ORG_GEN_SPAN_LUM) and take at most ORG_GEN_LDS_MAX bytes of dynamic LDS
 * (the static part adds up to 5 KB; a compute unit has 160 KB). */
#ifndef ORG_GEN_SPAN_CELLS
#define ORG_GEN_SPAN_CELLS 1
#endif
#ifndef ORG_GEN_SPAN_LUM
#define ORG_GEN_SPAN_LUM 1
#endif
#define ORG_GEN_LDS_MAX (150 * 1024)

/* One workgroup's band of a banded reshaping loop with spans of S sweeps
 * and values of type V: own rows [r0, r1) of a domain (the cell's window,
 * the normals' box) `bw` columns wide whose one-pixel surround is static.
 * LDS row l holds domain row r0 - S + l, LDS column c domain column c - 1
 * (L = bw + 2).  X0 holds the span's input, X1 the other iterate (both with
 * the static surround), JX / JY the normals, SV (S = 2; from row S on) the
 * span input's own rows.  Every band has S rows or more.  A value travels
 * as sizeof(V) / 4 tagged words (its 32-bit halves, low first). */
template <typename V>
struct GBand {
    V *X0, *X1, *JX, *JY, *SV;
    int L, bw, rows, b, K;
    unsigned long long *xr;   /* [2 parities][K][2 ends][S][bw][words per value] tagged rows */
    unsigned long long *xd;   /* [2 parities][K][S][2 halves] tagged deviations */
};
/* dynamic LDS of a band of up to `rows` rows, `bw` columns, values of vbytes */
static __host__ __device__ inline size_t g_band_lds(int S, int rows, int bw, size_t vbytes) {
    return ((size_t)4 * (rows + 2 * S) + (S > 1 ? rows : 0)) * (bw + 2) * vbytes;
}
/* tagged words of a group of K bands `bw` columns wide, values of vbytes */
static __host__ __device__ inline size_t g_band_words(int S, int K, int bw, size_t vbytes) {
    return (size_t)2 * K * 2 * S * bw * (vbytes / 4) + (size_t)2 * K * S * 2;
}
/* the dynamic LDS of the launch (16-byte aligned: FP64 arrays follow the
 * static part) */
static __device__ inline char *g_dyn_base() {
    extern __shared__ __attribute__((aligned(16))) char g_band_dyn[];
    return g_band_dyn;
}
/* G's LDS arrays, for bands of up to rmax rows (G.L set) */
template <typename V>
static __device__ inline void g_band_lds_split(GBand<V> &G, int S, int rmax) {
    const size_t plane = (size_t)(rmax + 2 * S) * G.L;
    G.X0 = (V *)g_dyn_base(); G.X1 = G.X0 + plane; G.JX = G.X1 + plane; G.JY = G.JX + plane;
    G.SV = G.JY + plane - (size_t)S * G.L;          /* own rows only: from row S on */
}
/* word h of a value and a value from its words */
static __device__ __forceinline__ unsigned g_word(float v, int) { return __float_as_uint(v); }
static __device__ __forceinline__ unsigned g_word(double v, int h) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    return h ? (unsigned)(b >> 32) : (unsigned)b;
}
template <typename V>
static __device__ __forceinline__ V g_join(const unsigned long long *w);
template <>
__device__ __forceinline__ float g_join<float>(const unsigned long long *w) { return __uint_as_float((unsigned)w[0]); }
template <>
__device__ __forceinline__ double g_join<double>(const unsigned long long *w) {
    return __longlong_as_double((long long)(((w[1] & 0xffffffffull) << 32) | (w[0] & 0xffffffffull)));
}

/* a / b rounded to nearest from y = RN(1 / b) (Markstein: q = RN(a y) is
 * within an ulp of a / b, r = a - b q is exact in one fma, RN(q + r y) is
 * RN(a / b) for normal operands): the IEEE quotient with three FP64
 * operations instead of a division sequence */
static __device__ __forceinline__ double g_div_rn(double a, double b, double y) {
    const double q = a * y;
    return fma(fma(-q, b, a), y, q);
}

/* pixel k = t, t + step, ... of a rows x bw block as (row, column), with no
 * integer division per pixel */
struct GRowCol {
    int r, c, dr, dc, bw;
    __device__ GRowCol(int k, int step, int bw_) : r(k / bw_), c(k % bw_), dr(step / bw_), dc(step % bw_), bw(bw_) {}
    __device__ void next() {
        r += dr;
        c += dc;
        if (c >= bw) { c -= bw; r++; }
    }
};

/* The loop `while (rdev > thr) { sweep; rdev = deviation; count++; if (ri >
 * rimax) { relax flag; break; } ri++; }` of reshaping_cells
 * (cell_phase_functions.cpp:475-568) and reshaping_lumen_or_ECM (:774-866)
 * on one band.  Sweep t of a span covers the band grown by S - 1 - t rows
 * on each side that has a neighbouring band (those rows are the
 * neighbour's, recomputed with the same arithmetic on the same operands,
 * so the same values), so a span needs the neighbours' rows once, at its
 * start.  A sweep's deviation is over the own rows, with the same
 * pixel-to-thread split and block sum (each wave's tree sum, then the
 * waves in order) as the exchange-buffer kernels, so the same value, over
 * ncur for the cells (DIV).  After a span each workgroup publishes its
 * first and last S rows and its S deviations as tagged words (tag tag0 +
 * span number), then, with no barrier, waits for its neighbours' rows
 * (into the span's last iterate's halo rows) and the K deviations of each
 * sweep, and takes the sweeps' stop tests in order; sweeps after the one
 * that stops the loop were speculative and are dropped.  Two planes
 * suffice: a workgroup overwrites span k's words in span k + 2, after every
 * reader has consumed them (it needed their span k + 1 deviations first).
 * Returns the sweep t of the last span at which the loop stopped, with X0
 * and X1 as at that span's start (the iterate: X1 for t = 0, X0 for t = 1;
 * the one before: X0 (S = 1) or SV (S = 2) for t = 0, X1 for t = 1), -1 if
 * the loop did not run, -2 if a wait timed out or another workgroup gave
 * up (the abort flag is then up).  KMAX: the largest K. */
template <int S, int KMAX, bool DIV, typename V>
static __device__ int g_band_loop(const GP<V> &P, GBand<V> &G, double thr, double rdev, int &count, int &ri,
                                  unsigned tag0, int *relax_flag, int *abort_flag, long long limit, double *red,
                                  double *sums, int *xfail) {
    static_assert(S == 1 || S == 2, "spans of one or two sweeps");
    constexpr int MD = (S * KMAX + 63) / 64;        /* deviations per lane of the polling wave */
    constexpr int WPV = (int)(sizeof(V) / 4);       /* tagged words per value */
    if (!(rdev > thr)) return -1;
    const int L = G.L, bw = G.bw, rows = G.rows, b = G.b, K = G.K;
    const bool up = b > 0, dn = b < K - 1;          /* neighbouring bands above / below */
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const V dx = P.dx;
    const GFluxK<V> FK = g_flux_k(P.wi, P.rdt, dx);
    const double ncd = (double)P.ncur, rnc = 1.0 / ncd;   /* the deviation's divisor and its reciprocal */
    for (unsigned span = 1;; span++) {
        const unsigned tag = tag0 + span;
        const int par = span & 1;
        for (int t = 0; t < S; t++) {
            const V *u = (t & 1) ? G.X1 : G.X0;
            V *un = (t & 1) ? G.X0 : G.X1;
            double acc = 0.0;
            GRowCol rc(threadIdx.x, blockDim.x, bw);
            for (int k = threadIdx.x; k < rows * bw; k += blockDim.x, rc.next()) {
                const int q = (rc.r + S) * L + (rc.c + 1);
                V o;
                const V v = g_flux_rows(u + q - L, u + q, u + q + L, (const V *)G.JX, (const V *)G.JY, (size_t)q, L,
                                        FK, o);
                un[q] = v;
                if (S > 1 && t == 0) G.SV[q] = o;
                if constexpr (DIV) acc += g_div_rn(fabs((double)(v - o)) * dx * dx, ncd, rnc);
                else acc += fabs((double)(v - o)) * dx * dx;
            }
            if (S > 1) {
                const int ga = up ? S - 1 - t : 0, gb = dn ? S - 1 - t : 0;
                for (int k = threadIdx.x; k < (ga + gb) * bw; k += blockDim.x) {
                    const int r = k / bw, l = r < ga ? S - ga + r : S + rows + (r - ga);
                    const int q = l * L + (k % bw + 1);
                    V o;
                    un[q] = g_flux_rows(u + q - L, u + q, u + q + L, (const V *)G.JX, (const V *)G.JY, (size_t)q, L,
                                        FK, o);
                }
            }
            for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
            if (lane == 0) red[t * 16 + wv] = acc;
            __syncthreads();
        }
        V *out = (S & 1) ? G.X1 : G.X0;             /* the span's last iterate */
        for (int k = threadIdx.x; k < 2 * S * bw * WPV; k += blockDim.x) {
            const int h = k % WPV, kv = k / WPV;
            const int e = kv / (S * bw), s = (kv / bw) % S, c = kv % bw;
            const int l = e ? rows + s : S + s;     /* the first / last S own rows */
            g_tput(&G.xr[(((((size_t)par * K + b) * 2 + e) * S + s) * bw + c) * WPV + h],
                   g_word(out[l * L + c + 1], h), tag);
        }
        if ((int)threadIdx.x < S) {
            const int t = threadIdx.x;
            double d = 0.0;
            for (int k = 0; k < nw; k++) d += red[t * 16 + k];
            const unsigned long long tb = (unsigned long long)__double_as_longlong(d);
            g_tput(&G.xd[(((size_t)par * K + b) * S + t) * 2 + 0], (unsigned)tb, tag);
            g_tput(&G.xd[(((size_t)par * K + b) * S + t) * 2 + 1], (unsigned)(tb >> 32), tag);
        }
        bool ok = true;
        if (threadIdx.x < 64) {
            /* the group's deviations, every first load in flight before the first wait */
            unsigned long long lo[MD], hi[MD];
#pragma unroll
            for (int m = 0; m < MD; m++) {
                const int j = lane + 64 * m;
                lo[m] = hi[m] = 0ull;
                if (j < S * K) {
                    const unsigned long long *a = &G.xd[(((size_t)par * K + j % K) * S + j / K) * 2];
                    lo[m] = g_tget(a);
                    hi[m] = g_tget(a + 1);
                }
            }
#pragma unroll
            for (int m = 0; m < MD; m++) {
                const int j = lane + 64 * m;
                if (j < S * K) {
                    const unsigned long long *a = &G.xd[(((size_t)par * K + j % K) * S + j / K) * 2];
                    ok = ok && g_twait(a, tag, lo[m], abort_flag, limit) && g_twait(a + 1, tag, hi[m], abort_flag, limit);
                    sums[(j / K) * K + j % K] =
                        __longlong_as_double((long long)(((hi[m] & 0xffffffffull) << 32) | (lo[m] & 0xffffffffull)));
                }
            }
        } else {
            /* the neighbours' rows into the halo rows of `out`, four values
             * per thread in flight, then the waits */
            const int nq = 2 * S * bw, step = (int)blockDim.x - 64;
            for (int q0 = (int)threadIdx.x - 64; q0 < nq; q0 += 4 * step) {
                unsigned long long w[4][WPV];
                const unsigned long long *ap[4];
                int la[4];
#pragma unroll
                for (int m = 0; m < 4; m++) {
                    const int q = q0 + m * step;
                    la[m] = -1; ap[m] = nullptr;
#pragma unroll
                    for (int h = 0; h < WPV; h++) w[m][h] = 0ull;
                    if (q < nq) {
                        const int e = q / (S * bw), s = (q / bw) % S, c = q % bw;
                        if (e ? dn : up) {          /* else static rows there */
                            /* above: b - 1's last S rows; below: b + 1's first */
                            ap[m] = &G.xr[(((((size_t)par * K + (e ? b + 1 : b - 1)) * 2 + (e ? 0 : 1)) * S + s) * bw +
                                           c) * WPV];
                            la[m] = (e ? S + rows + s : s) * L + c + 1;
#pragma unroll
                            for (int h = 0; h < WPV; h++) w[m][h] = g_tget(ap[m] + h);
                        }
                    }
                }
#pragma unroll
                for (int m = 0; m < 4; m++) {
                    if (la[m] < 0) continue;
#pragma unroll
                    for (int h = 0; h < WPV; h++) ok = ok && g_twait(ap[m] + h, tag, w[m][h], abort_flag, limit);
                    out[la[m]] = g_join<V>(w[m]);
                }
            }
        }
        if (!ok) *xfail = 1;
        __syncthreads();
        if (*xfail) {
            if (threadIdx.x == 0) atomicExch(abort_flag, 1);
            return -2;
        }
        for (int t = 0; t < S; t++) {
            rdev = g_lds_sum(sums + t * K, K);
            count++;
            if (ri > P.rimax) { if (threadIdx.x == 0 && b == 0) *relax_flag = 1; return t; }
            ri++;
            if (!(rdev > thr)) return t;
        }
        if (S & 1) { V *x = G.X0; G.X0 = G.X1; G.X1 = x; }
    }
}
/* the iterate the loop stopped at and the one before (g_band_loop's return t >= 0) */
template <int S, typename V>
static __device__ inline const V *g_band_last(const GBand<V> &G, int t) { return (t & 1) ? G.X0 : G.X1; }
template <int S, typename V>
static __device__ inline const V *g_band_prev(const GBand<V> &G, int t) {
    return t == 0 ? (S == 1 ? G.X0 : G.SV) : G.X1;
}

/* reshaping_cells (cell_phase_functions.cpp:447-580) with K workgroups per
 * cell (blockIdx.x = n K + b), workgroup b on window rows
 * [b W / K, (b + 1) W / K) held in LDS (g_band_loop; the window's one-pixel
 * ring, phi around the window, is the static surround).  The iterates
 * equal g_relax_cells_mb's bit for bit, and with the same K so do the
 * per-workgroup deviations and their in-order sum.  At the end phin's
 * window gets the last iterate and phi's the second-to-last (own rows).
 * xw: [nc] groups of g_band_words(S, K, W, sizeof(V)) words.  Dynamic LDS: g_band_lds(S, rmax, W, sizeof(V)), rmax >= the
 * largest band; K <= 64. */
template <int S, typename V>
__global__ __launch_bounds__(1024) void g_relax_cells_lds(GP<V> P, V *__restrict__ phi, V *__restrict__ phin,
                                                          const V *__restrict__ jx, const V *__restrict__ jy,
                                                          const int *__restrict__ xc,
                                                          const int *__restrict__ yc, double thr, double rdev0,
                                                          unsigned long long *__restrict__ xw, unsigned tag0,
                                                          int *__restrict__ counts, int *__restrict__ cmax,
                                                          int *__restrict__ relax_flag, int *__restrict__ abort_flag,
                                                          long long limit, int K, int rmax) {
    __shared__ double red[S * 16];
    __shared__ double sums[S * 64];
    __shared__ int xfail;
    const int N = P.N, zs = P.zs, W = 2 * zs, L = W + 2;
    const int n = blockIdx.x / K, b = blockIdx.x % K;
    const size_t NN = (size_t)N * N;
    V *ph = phi + (size_t)n * NN, *pn = phin + (size_t)n * NN;
    const V *jxn = jx + (size_t)n * NN, *jyn = jy + (size_t)n * NN;
    const int x0 = xc[n] - zs, y0 = yc[n] - zs;              /* window origin on the grid */
    const int r0 = b * W / K, r1 = (b + 1) * W / K, rows = r1 - r0;
    GBand<V> G;
    G.L = L; G.bw = W; G.rows = rows; G.b = b; G.K = K;
    g_band_lds_split(G, S, rmax);
    G.xr = xw + (size_t)n * g_band_words(S, K, W, sizeof(V));
    G.xd = G.xr + (size_t)2 * K * 2 * S * W * (sizeof(V) / 4);
    /* the window rows r0 - S .. r1 + S - 1 there are, and the ring rows -1
     * and W: ring pixels from phi, the window from phin */
    const int l0 = b == 0 ? S - 1 : 0, l1 = b == K - 1 ? S + rows + 1 : rows + 2 * S;
    for (int k = threadIdx.x; k < (l1 - l0) * L; k += blockDim.x) {
        const int l = l0 + k / L, c = k % L, wr = r0 - S + l, wc = c - 1;
        const size_t g = (size_t)(x0 + wr) * N + (y0 + wc);
        const bool ring = wr < 0 || wr >= W || wc < 0 || wc >= W;
        const V v = ring ? ph[g] : pn[g];
        const int q = l * L + c;
        G.X0[q] = v; G.X1[q] = v;
        G.JX[q] = jxn[g]; G.JY[q] = jyn[g];
    }
    if (threadIdx.x == 0) xfail = 0;      /* set once: after it goes to 1 every thread returns */
    __syncthreads();
    int ri = 0, count = 0;
    const int last = g_band_loop<S, 64, true>(P, G, thr, rdev0, count, ri, tag0, relax_flag, abort_flag, limit, red,
                                              sums, &xfail);
    if (last == -2) return;
    if (last >= 0) {
        /* phin <- the last iterate, phi <- the one before; own rows only */
        const V *vl = g_band_last<S>(G, last), *vp = g_band_prev<S>(G, last);
        for (int k = threadIdx.x; k < rows * W; k += blockDim.x) {
            const int q = (k / W + S) * L + (k % W + 1);
            const size_t g = (size_t)(x0 + r0 + k / W) * N + (y0 + k % W);
            pn[g] = vl[q];
            ph[g] = vp[q];
        }
    }
    if (threadIdx.x == 0 && b == 0) {
        counts[n] = count;
        atomicMax(cmax, count);
    }
}

/* Sweeps 2, 3, ... of reshaping_lumen_or_ECM's loop
 * (cell_phase_functions.cpp:774-866) with K workgroups; the caller ran
 * sweep 1 on the whole grid (lumn <- F(lum), the reference's copy lum <-
 * lumn being the change of roles: X_1 is in bufA, X_0 in bufB) and wrote
 * the Dirichlet ring of boundary_conditions_dirichlet into bufB too, which
 * sweep 2 would write, so both buffers now hold the same constant ring and
 * the ring adds 0 to every later deviation.  Sweep k + 1 reads X_k (bufA
 * for odd k) and writes the other buffer.  A sweep leaves a pixel whose
 * 3 x 3 neighbourhood has zero normals unchanged exactly (every flux term
 * carries a face average of the normals), and after sweep 1 both buffers
 * agree outside box = [box[0], box[1]) x [box[2], box[3]) (the nonzero
 * normals grown by one, inside rows / columns 1..N-2), so only the box is
 * swept, workgroup b on box rows [r0, r1).  A workgroup reads and writes
 * the buffers only on its rows and on static pixels; its first and last
 * rows also go to an exchange buffer (agent-scope stores), from which the
 * neighbours load them into LDS (agent-scope loads) after each barrier.
 * count / ri: sweeps done and the rimax counter on entry; rdev: the last
 * deviation.  out[0]: sweeps done in all (the last iterate is in bufA when
 * odd, in bufB when even).  X: [2][K][2][box width + 2]; part: [2][K]; ctr:
 * one arrival counter, zero at launch.  Dynamic LDS: 2 (box width + 2)
 * reals. */
template <typename R>
__global__ __launch_bounds__(1024) void g_lum_loop_mb(GP<R> P, R *__restrict__ bufA, R *__restrict__ bufB,
                                                      const R *__restrict__ jx, const R *__restrict__ jy,
                                                      const int *__restrict__ box, double thr, double rdev,
                                                      int count, int ri, R *__restrict__ X,
                                                      double *__restrict__ part, unsigned *__restrict__ ctr,
                                                      int *__restrict__ out, int *__restrict__ relax_flag,
                                                      int *__restrict__ abort_flag, long long limit, int K) {
    __shared__ double red[16];
    __shared__ double sums[1024];
    extern __shared__ char g_dyn[];
    const int N = P.N, b = blockIdx.x;
    const int bi0 = box[0], bi1 = box[1], bj0 = box[2], bj1 = box[3];
    const int bw = bj1 - bj0, bh = bi1 - bi0, L = bw + 2;    /* the caller launches only for a nonempty box */
    R *hu = (R *)g_dyn, *hd = hu + L;                        /* grid rows r0-1 and r1, columns bj0-1 .. bj1 */
    auto xrow = [&](int par, int blk, int e) { return X + (((size_t)par * K + blk) * 2 + e) * L; };
    const int r0 = bi0 + b * bh / K, r1 = bi0 + (b + 1) * bh / K;     /* this workgroup's box rows */
    const R dx = P.dx;
    const GFluxK<R> FK = g_flux_k(P.wi, P.rdt, dx);
    /* the exchange rows of X_count (in the buffer the next sweep reads) */
    {
        const R *u = (count & 1) ? bufA : bufB;
        for (int k = threadIdx.x; k < 2 * L; k += blockDim.x) {
            const int e = k / L, c = k % L;
            g_xs(&xrow(count & 1, b, e)[c], u[(size_t)(e ? r1 - 1 : r0) * N + (bj0 - 1 + c)]);
        }
    }
    unsigned epoch = 1;
    if (!g_barrier(ctr, epoch * K, abort_flag, limit)) return;
    auto load_halo = [&](const R *u, int par) {
        for (int k = threadIdx.x; k < 2 * L; k += blockDim.x) {
            const int e = k / L, c = k % L;
            R v;
            if (e == 0) v = b == 0 ? u[(size_t)(r0 - 1) * N + (bj0 - 1 + c)] : g_xl(&xrow(par, b - 1, 1)[c]);
            else v = b == K - 1 ? u[(size_t)r1 * N + (bj0 - 1 + c)] : g_xl(&xrow(par, b + 1, 0)[c]);
            hu[k] = v;
        }
    };
    load_halo((count & 1) ? bufA : bufB, count & 1);
    __syncthreads();
    while (rdev > thr) {
        const R *u = (count & 1) ? bufA : bufB;
        R *un = (count & 1) ? bufB : bufA;
        const int par = count & 1;
        double acc = 0.0;
        for (int k = threadIdx.x; k < (r1 - r0) * bw; k += blockDim.x) {
            const int i = r0 + k / bw, j = bj0 + k % bw;
            const size_t q = (size_t)i * N + j;
            const int c = j - bj0 + 1;                            /* column in the exchange rows */
            const R *up = i == r0 ? hu + c : u + (q - N);
            const R *dn = i == r1 - 1 ? hd + c : u + (q + N);
            R o;
            const R v = g_flux_rows(up, u + q, dn, jx, jy, q, N, FK, o);
            un[q] = v;
            if (i == r0) g_xs(&xrow(par ^ 1, b, 0)[c], v);
            if (i == r1 - 1) g_xs(&xrow(par ^ 1, b, 1)[c], v);
            acc += fabs((double)(v - o)) * dx * dx;
        }
        /* the exchanged rows' columns bj0-1 and bj1 (static) */
        if (threadIdx.x < 4) {
            const int e = threadIdx.x >> 1, c = (threadIdx.x & 1) ? L - 1 : 0;
            g_xs(&xrow(par ^ 1, b, e)[c], u[(size_t)(e ? r1 - 1 : r0) * N + (bj0 - 1 + c)]);
        }
        const double t = gblock_sum(acc, red);
        if (threadIdx.x == 0) g_xs(&part[par * K + b], t);
        epoch++;
        if (!g_barrier(ctr, epoch * K, abort_flag, limit)) return;
        g_group_load(&part[par * K], K, sums);
        load_halo(un, par ^ 1);                      /* for the next sweep (unused if the loop stops) */
        __syncthreads();
        rdev = g_lds_sum(sums, K);
        count++;
        if (ri > P.rimax) { if (threadIdx.x == 0 && b == 0) *relax_flag = 1; break; }   /* :862-865 */
        ri++;
    }
    if (threadIdx.x == 0 && b == 0) out[0] = count;
}

/* Sweeps 2, 3, ... of reshaping_lumen_or_ECM's loop as g_lum_loop_mb, with
 * workgroup b's band of box rows [r0, r1) held in LDS
 * (g_band_loop; the rows and columns just outside the box are the static
 * surround, the same in both buffers).  The iterates equal g_lum_loop_mb's
 * bit for bit, and with the same K so do the per-workgroup deviations and
 * their in-order sum.  At
 * the end both buffers get the band's rows of the iterates they hold as
 * the sweeps of g_lum_loop_mb leave them (X_count in bufA for odd count).
 * xw: g_band_words(S, K, box width, sizeof(V)) words.  Dynamic LDS:
 * g_band_lds(S, rmax, box width, sizeof(V)), rmax >= the largest band;
 * every band has S rows or more; K <= 128. */
template <int S, typename V>
__global__ __launch_bounds__(1024) void g_lum_loop_lds(GP<V> P, V *__restrict__ bufA, V *__restrict__ bufB,
                                                       const V *__restrict__ jx, const V *__restrict__ jy,
                                                       const int *__restrict__ box,
                                                       double thr, double rdev, int count, int ri,
                                                       unsigned long long *__restrict__ xw, unsigned tag0,
                                                       int *__restrict__ out, int *__restrict__ relax_flag,
                                                       int *__restrict__ abort_flag, long long limit, int K, int rmax) {
    __shared__ double red[S * 16];
    __shared__ double sums[S * 128];
    __shared__ int xfail;
    const int N = P.N, b = blockIdx.x;
    const int bi0 = box[0], bi1 = box[1], bj0 = box[2], bj1 = box[3];
    const int bw = bj1 - bj0, bh = bi1 - bi0, L = bw + 2;
    const int r0 = bi0 + b * bh / K, r1 = bi0 + (b + 1) * bh / K, rows = r1 - r0;
    GBand<V> G;
    G.L = L; G.bw = bw; G.rows = rows; G.b = b; G.K = K;
    g_band_lds_split(G, S, rmax);
    G.xr = xw;
    G.xd = xw + (size_t)2 * K * 2 * S * bw * (sizeof(V) / 4);
    /* X0: X_count (bufA for odd count), X1: the other buffer; grid row r0 -
     * S + l, column bj0 - 1 + c, for the box rows there are and the static
     * rows bi0 - 1 and bi1 */
    const V *cur = (count & 1) ? bufA : bufB, *oth = (count & 1) ? bufB : bufA;
    const int l0 = b == 0 ? S - 1 : 0, l1 = b == K - 1 ? S + rows + 1 : rows + 2 * S;
    for (int k = threadIdx.x; k < (l1 - l0) * L; k += blockDim.x) {
        const int l = l0 + k / L, c = k % L;
        const size_t g = (size_t)(r0 - S + l) * N + (bj0 - 1 + c);
        const int q = l * L + c;
        G.X0[q] = cur[g]; G.X1[q] = oth[g];
        G.JX[q] = jx[g]; G.JY[q] = jy[g];
    }
    if (threadIdx.x == 0) xfail = 0;      /* set once: after it goes to 1 every thread returns */
    __syncthreads();
    const int last = g_band_loop<S, 128, false>(P, G, thr, rdev, count, ri, tag0, relax_flag, abort_flag, limit, red,
                                                sums, &xfail);
    if (last == -2) return;
    if (last >= 0) {
        const V *vl = g_band_last<S>(G, last), *vp = g_band_prev<S>(G, last);
        V *dl = (count & 1) ? bufA : bufB, *dp = (count & 1) ? bufB : bufA;
        for (int k = threadIdx.x; k < rows * bw; k += blockDim.x) {
            const int q = (k / bw + S) * L + (k % bw + 1);
            const size_t g = (size_t)(r0 + k / bw) * N + (bj0 + k % bw);
            dl[g] = vl[q];
            dp[g] = vp[q];
        }
    }
    if (threadIdx.x == 0 && b == 0) out[0] = count;
}

/* the Dirichlet ring of boundary_conditions_dirichlet (useful_functions.cpp:
 * 422-437): edges `side`, corners 0 */
template <typename R>
__global__ void g_ring_dirichlet(int N, R *__restrict__ t, R side) {
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < 4 * N; k += gridDim.x * blockDim.x) {
        const int e = k / N, c = k % N;
        size_t q;
        int i, j;
        if (e == 0) { i = 0; j = c; } else if (e == 1) { i = N - 1; j = c; } else if (e == 2) { i = c; j = 0; }
        else { i = c; j = N - 1; }
        q = (size_t)i * N + j;
        t[q] = ((i == 0 || i == N - 1) && (j == 0 || j == N - 1)) ? (R)0.0 : side;
    }
}

#define ORG_NZ_PARTS 1024

/* The box of nonzero normals (interior): workgroup b's bounds {rows min,
 * max, columns min, max} over its rows 1 + b, 1 + b + gridDim.x, ... into
 * part[b] ({N, -1, N, -1} when it saw none); g_nz_box_fin combines them. */
template <typename R>
__global__ void g_nz_box(int N, const R *__restrict__ jx, const R *__restrict__ jy, int4 *__restrict__ part) {
    int i0 = N, i1 = -1, j0 = N, j1 = -1;
    for (int i = 1 + blockIdx.x; i < N - 1; i += gridDim.x)
        for (int j = 1 + threadIdx.x; j < N - 1; j += blockDim.x) {
            const size_t q = (size_t)i * N + j;
            if (jx[q] != (R)0.0 || jy[q] != (R)0.0) {
                i0 = min(i0, i); i1 = max(i1, i); j0 = min(j0, j); j1 = max(j1, j);
            }
        }
    /* the workgroup's bounds (min / max in any order) into its part */
    __shared__ int bb[4];
    if (threadIdx.x == 0) { bb[0] = N; bb[1] = -1; bb[2] = N; bb[3] = -1; }
    for (int o = 32; o > 0; o >>= 1) {
        i0 = min(i0, __shfl_xor(i0, o, 64)); i1 = max(i1, __shfl_xor(i1, o, 64));
        j0 = min(j0, __shfl_xor(j0, o, 64)); j1 = max(j1, __shfl_xor(j1, o, 64));
    }
    __syncthreads();
    if ((threadIdx.x & 63) == 0 && i1 >= 0) {
        atomicMin(&bb[0], i0); atomicMax(&bb[1], i1); atomicMin(&bb[2], j0); atomicMax(&bb[3], j1);
    }
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = make_int4(bb[0], bb[1], bb[2], bb[3]);
}
/* g_nz_box's nb parts into box (grown by one, inside rows / columns
 * 1..N-2; {N, 0, N, 0} when no normal is nonzero).  One workgroup of 256. */
__global__ __launch_bounds__(256) void g_nz_box_fin(int N, const int4 *__restrict__ part, int nb,
                                                    int *__restrict__ box) {
    __shared__ int bb[4];
    int i0 = N, i1 = -1, j0 = N, j1 = -1;
    for (int k = threadIdx.x; k < nb; k += blockDim.x) {
        const int4 v = part[k];
        i0 = min(i0, v.x); i1 = max(i1, v.y); j0 = min(j0, v.z); j1 = max(j1, v.w);
    }
    if (threadIdx.x == 0) { bb[0] = N; bb[1] = -1; bb[2] = N; bb[3] = -1; }
    for (int o = 32; o > 0; o >>= 1) {
        i0 = min(i0, __shfl_xor(i0, o, 64)); i1 = max(i1, __shfl_xor(i1, o, 64));
        j0 = min(j0, __shfl_xor(j0, o, 64)); j1 = max(j1, __shfl_xor(j1, o, 64));
    }
    __syncthreads();
    if ((threadIdx.x & 63) == 0 && i1 >= 0) {
        atomicMin(&bb[0], i0); atomicMax(&bb[1], i1); atomicMin(&bb[2], j0); atomicMax(&bb[3], j1);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        if (bb[1] >= 0) {
            box[0] = max(1, bb[0] - 1); box[1] = min(N - 1, bb[1] + 2);
            box[2] = max(1, bb[2] - 1); box[3] = min(N - 1, bb[3] + 2);
        } else {
            box[0] = N; box[1] = 0; box[2] = N; box[3] = 0;
        }
    }
}

/* a launch's arrival counters and abort flag start at zero (a fill
 * kernel, not a memset, so the sequence can be captured) */
__global__ void g_zero_u32(unsigned *p, int n) {
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) p[k] = 0u;
}

template <typename R>
__global__ void g_copy(const R *__restrict__ src, R *__restrict__ dst, size_t n) {
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x)
        dst[k] = src[k];
}

#undef GPL
#undef GG
}  // namespace gen

namespace gen {

/* update_rho_to_shape (cell_phase_functions.cpp:42-62) on every plane, in
 * double as the reference's double literals make it */
template <typename R>
__global__ void g_rho_to_shape(const R *__restrict__ phi, R *__restrict__ rho, size_t count, double rho0,
                               double tol) {
    const double floor_v = rho0 * exp(-1.0 / (2.0 * tol) + 1.0 / (1.0 + tol));
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < count; e += (size_t)gridDim.x * blockDim.x) {
        const R p = phi[e];
        R r = rho[e];
        if (p != 0.0) {
            const double pd = p, h = pd * pd * (3.0 - 2.0 * pd);
            r = (R)(rho0 * exp(-1.0 / (4.0 * h * (1.0 - h) + tol) + 1.0 / (1.0 + tol)));
        }
        if ((double)r < floor_v) r = 0.0;
        rho[e] = r;
    }
}

/* side x side squares of every cell's planes, anchored at (xc + off, yc + off) */
template <typename R>
__global__ void g_gather(int N, const R *__restrict__ src, const int *__restrict__ xc, const int *__restrict__ yc,
                         int off, int side, R *__restrict__ out) {
    const int n = blockIdx.y;
    const size_t NN = (size_t)N * N;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < side * side; k += gridDim.x * blockDim.x)
        out[(size_t)n * side * side + k] =
            src[(size_t)n * NN + (size_t)(xc[n] + off + k / side) * N + (yc[n] + off + k % side)];
}

/* ---- host side ----------------------------------------------------------- */

#define GHIP(expr)                                                                                        \
    do {                                                                                                  \
        hipError_t _e = (expr);                                                                           \
        if (_e != hipSuccess) {                                                                           \
            fprintf(stderr, "organoid (generic path): %s failed: %s (%s:%d)\n", #expr, hipGetErrorString(_e), \
                    __FILE__, __LINE__);                                                                  \
            return ORG_EHIP;                                                                              \
        }                                                                                                 \
    } while (0)
#define GTRY(x) do { int _r = (x); if (_r) return _r; } while (0)
#define GLAUNCH(kern, grid, block, shm, ...)                                                             \
    do {                                                                                                  \
        hipLaunchKernelGGL(kern, grid, block, shm, s, __VA_ARGS__);                                       \
        GHIP(hipGetLastError());                                                                          \
    } while (0)

static int dev_alloc(std::vector<void *> &allocs, void **p, size_t bytes) {
    GHIP(hipMalloc(p, bytes ? bytes : 1));
    allocs.push_back(*p);
    return ORG_OK;
}

/* Bluestein tables of one sign: the chirp w_k = exp(i sign pi k^2 / M) with
 * k^2 reduced mod 2M first (exact), and the L-point transform (sign -1) of
 * the conjugate chirp wrapped around (b_k = b_{L-k} = conj w_k). */
static void bluestein_tables(int M, int L, int sign, std::vector<double> &chirp, std::vector<double> &khat) {
    chirp.assign(2 * M, 0.0);
    for (int k = 0; k < M; k++) {
        const long long k2 = ((long long)k * k) % (2LL * M);
        const double ang = sign * M_PI * (double)k2 / M;
        chirp[2 * k] = cos(ang);
        chirp[2 * k + 1] = sin(ang);
    }
    std::vector<double> br(L, 0.0), bi(L, 0.0);
    br[0] = chirp[0];
    bi[0] = -chirp[1];
    for (int k = 1; k < M; k++) {
        br[k] = chirp[2 * k]; bi[k] = -chirp[2 * k + 1];
        br[L - k] = chirp[2 * k]; bi[L - k] = -chirp[2 * k + 1];
    }
    host_fft_pow2(br.data(), bi.data(), L, -1);
    khat.resize(2 * L);
    for (int k = 0; k < L; k++) { khat[2 * k] = br[k]; khat[2 * k + 1] = bi[k]; }
}

static int make_plan(int M, FftPlan &pl, std::vector<void *> &allocs) {
    pl.M = M;
    pl.pow2 = is_pow2(M);
    pl.L = M;
    if (!pl.pow2) {
        pl.L = 1;
        while (pl.L < 2 * M - 1) pl.L <<= 1;
    }
    if (pl.L > 2048) {
        fprintf(stderr, "organoid (generic path): FFT box %d too large (line transforms up to 2048 points)\n", M);
        return ORG_EINVAL;
    }
    std::vector<double> tw;
    host_twiddles(pl.L, tw);
    GTRY(dev_alloc(allocs, (void **)&pl.tw, sizeof(double) * tw.size()));
    GHIP(hipMemcpy(pl.tw, tw.data(), sizeof(double) * tw.size(), hipMemcpyHostToDevice));
    if (!pl.pow2) {
        for (int si = 0; si < 2; si++) {
            std::vector<double> ch, kh;
            bluestein_tables(M, pl.L, si ? +1 : -1, ch, kh);
            GTRY(dev_alloc(allocs, (void **)&pl.chirp[si], sizeof(double) * ch.size()));
            GTRY(dev_alloc(allocs, (void **)&pl.khat[si], sizeof(double) * kh.size()));
            GHIP(hipMemcpy(pl.chirp[si], ch.data(), sizeof(double) * ch.size(), hipMemcpyHostToDevice));
            GHIP(hipMemcpy(pl.khat[si], kh.data(), sizeof(double) * kh.size(), hipMemcpyHostToDevice));
        }
    }
    return ORG_OK;
}

/* nplanes M x M transforms src -> dst (unnormalised, sign -1 or +1): rows
 * into the double work buffer, then columns out of it (cufftExecC2C on each
 * plane, useful_functions.cpp:646-660, with the double intermediate the
 * oracle's transform keeps). */
template <typename TI, typename TO>
static int fft_lines_launch(hipStream_t s, const FftPlan &pl, const Cx<TI> *src, Cx<TO> *dst, int nlines,
                            long plane, int line_step, int elem_step, int sign) {
    const int M = pl.M, L = pl.L, si = sign > 0 ? 1 : 0;
    const dim3 grid(std::max(1, std::min((nlines + pl.lpw - 1) / pl.lpw, 8192))), block(256);
    const size_t shm = 2 * sizeof(double) * L * pl.lpw;
    const double *ch = pl.pow2 ? nullptr : pl.chirp[si], *kh = pl.pow2 ? nullptr : pl.khat[si];
    switch (pl.fuse) {
    case 1:
        GLAUNCH((g_fft_lines<TI, TO, 1>), grid, block, shm, src, dst, nlines, M, plane, line_step, elem_step, M, L,
                sign, pl.tw, ch, kh, pl.lpw);
        break;
    case 2:
        GLAUNCH((g_fft_lines<TI, TO, 2>), grid, block, shm, src, dst, nlines, M, plane, line_step, elem_step, M, L,
                sign, pl.tw, ch, kh, pl.lpw);
        break;
    default:
        GLAUNCH((g_fft_lines<TI, TO, 3>), grid, block, shm, src, dst, nlines, M, plane, line_step, elem_step, M, L,
                sign, pl.tw, ch, kh, pl.lpw);
        break;
    }
    return ORG_OK;
}

/* nplanes M x M transforms src -> dst (unnormalised, sign -1 or +1): rows
 * into the double work buffer, then columns out of it (cufftExecC2C on each
 * plane, useful_functions.cpp:646-660, with the double intermediate the
 * oracle's transform keeps). */
template <typename R>
static int fft2d_planes(hipStream_t s, const FftPlan &pl, const Cx<R> *src, Cx<R> *dst, Cx<double> *work,
                        int nplanes, int sign) {
    const int M = pl.M;
    const long plane = (long)M * M;
    GTRY((fft_lines_launch<R, double>(s, pl, src, work, nplanes * M, plane, M, 1, sign)));
    return fft_lines_launch<double, R>(s, pl, (const Cx<double> *)work, dst, nplanes * M, plane, 1, M, sign);
}

/* A schedule is valid when every thread holds at most 8 elements of a
 * Bluestein line and the lines fit the workgroup's LDS. */
static bool schedule_ok(const FftPlan &pl, int fuse, int lpw) {
    if (fuse < 1 || fuse > 3 || (lpw != 1 && lpw != 2 && lpw != 4)) return false;
    if ((size_t)pl.L * lpw > 2048) return false;                 /* 256 / lpw threads, <= 8 elements each */
    return 2 * sizeof(double) * (size_t)pl.L * lpw <= 65536;
}

/* The autotuner of the transform schedule: times every valid (fuse, lines
 * per workgroup) on nplanes forward transforms of the given buffers (two
 * runs each, the faster kept) and leaves the fastest in the plan.  All
 * schedules give the same values. */
template <typename R>
static int tune_fft(hipStream_t s, FftPlan &pl, Cx<R> *a, Cx<R> *b, Cx<double> *work, int nplanes,
                    std::vector<float> *times) {
    hipEvent_t e0, e1;
    GHIP(hipEventCreate(&e0));
    GHIP(hipEventCreate(&e1));
    float best = 1e30f;
    int bf = 1, bl = 1;
    for (int fuse = 1; fuse <= 3; fuse++)
        for (int lpw = 1; lpw <= 4; lpw *= 2) {
            if (!schedule_ok(pl, fuse, lpw)) continue;
            pl.fuse = fuse;
            pl.lpw = lpw;
            float tmin = 1e30f;
            for (int rep = 0; rep < 2; rep++) {
                GHIP(hipEventRecord(e0, s));
                GTRY(fft2d_planes<R>(s, pl, a, b, work, nplanes, -1));
                GHIP(hipEventRecord(e1, s));
                GHIP(hipEventSynchronize(e1));
                float ms = 0.f;
                GHIP(hipEventElapsedTime(&ms, e0, e1));
                tmin = std::min(tmin, ms);
            }
            if (times) times->push_back(tmin);
            if (tmin < best) { best = tmin; bf = fuse; bl = lpw; }
        }
    pl.fuse = bf;
    pl.lpw = bl;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return ORG_OK;
}

/* The values the host reads back at a sync point of the step, in one
 * device block so that one copy (into pinned host memory) brings them all:
 * each separate small copy costs ~25 us of idle GPU. */
struct GLoopCtl {
    int relax, abrt, cmax, count;   /* a loop launch: relax flag, abort flag, largest count, count */
};
struct GCtl {
    double rdev[2];            /* reshape_lumecm: each field's first-sweep deviation (0 lumen, 1 ECM) */
    int box[2][4];             /* reshape_lumecm: each field's box of nonzero normals */
    GLoopCtl loop[2];          /* [0]: the cell loop's or the lumen loop's, [1]: the ECM loop's */
    int err;                   /* find_centres: the window check */
    int pad[3];
};

template <typename R>
struct GenCtx final : Generic {
    org_params p{};
    GP<R> P{};
    hipStream_t s = nullptr;
    int N = 0, nc = 0, zs = 0, W = 0, M = 0;
    size_t NN = 0, PP = 0, MM = 0;
    std::vector<void *> allocs;
    FftPlan plan;
    /* per-cell planes (declare_tables.cpp:97-213) */
    R *phi = nullptr, *phin = nullptr, *rho = nullptr, *rhon = nullptr, *hf = nullptr, *ngp = nullptr;
    R *jx = nullptr, *jy = nullptr;
    /* shared planes; the loop's pairs swap roles between sweeps */
    R *lum = nullptr, *lumn = nullptr, *ecm = nullptr, *ecmn = nullptr, *psi_ = nullptr, *hl = nullptr, *he = nullptr;
    R *ngl = nullptr, *jlx = nullptr, *jly = nullptr, *jex = nullptr, *jey = nullptr;   /* lumen's, ECM's normals */
    /* FFT boxes: five transforms' inputs / spectra ([5][nc][M][M]), the velocity modes ([2][nc][M][M]) */
    Cx<R> *A = nullptr, *B = nullptr, *V = nullptr;
    Cx<double> *work = nullptr;
    R *vxb = nullptr, *vyb = nullptr, *win = nullptr;
    R *f0 = nullptr;
    double *s_ = nullptr, *s0 = nullptr, *sum_e = nullptr, *part = nullptr;
    int *xc = nullptr, *yc = nullptr, *err = nullptr, *counts = nullptr, *relax = nullptr;
    int nparts = 1;
    /* the loops split over several workgroups: workgroups per cell / for
     * the lumen-ECM loop (1: the cell loop runs one workgroup per cell),
     * window buffers, per-workgroup deviations, arrival counters + abort flag */
    int kc = 1, kl = 1;
    R *swin = nullptr;
    double *mb_part = nullptr;
    unsigned *mb_ctr = nullptr;      /* arrival counters: [nc] cell groups, [nc], [nc + 1] the lumen, ECM groups */
    R *mb_xch = nullptr;             /* exchanged band rows: cells [nc][2][kc][2][W + 2]; lumen/ECM [2][kl][2][N] */
    /* the cell loop with its bands in LDS and tagged exchanges
     * (g_relax_cells_lds): workgroups per cell (0: not used), the largest
     * band, its LDS bytes, the tagged words, the launch's largest sweep
     * count, and the next launch's tag base */
    int kcl = 0, rmax_l = 0;
    size_t lds_l = 0;
    unsigned long long *mb_tw = nullptr;
    /* and the lumen/ECM loop's (g_lum_loop_lds): the most workgroups (0: not
     * used) and its tagged words (g_band_words for klm bands of N columns) */
    int klm = 0;
    unsigned long long *mb_lw[2] = {nullptr, nullptr};   /* the lumen loop's, the ECM loop's */
    unsigned tag_lum[2] = {1, 1};                        /* each one's next tag base */
    int *mb_cmax = nullptr;
    int4 *nz_part = nullptr;                 /* g_nz_box's per-workgroup bounds, [2][ORG_NZ_PARTS] */
    unsigned long long *cen_acc = nullptr;   /* the centroid sums, [nc][3] */
    int *hbox = nullptr;                     /* [nc][4]: where each cell's hfunc may be nonzero */
    bool hbox_stale = true;                  /* hfunc uploaded since the boxes were last scanned */
    double *srows = nullptr;                 /* FP64 areas' row sums, [nc][N] */
    unsigned tag_next = 1;                   /* the cell loop's next tag base */
    long long wait_limit = 0;
    int count_lum = 0, count_ecm = 0;
    bool in_step = false, ngp_fresh = false;   /* step_scope: ngp formed in this step from the current phi */
    /* step_chain: cen_acc holds the centroid sums of the current phi (set by
     * find_centres inside a chain, kept by reshape_cells' window deltas) */
    bool chain = false, cen_live = false;
    /* rho equals rhon on every pixel (both planes' writers keep this up to
     * date), and this step's spectral update started from that state: its
     * copy-back then needs only the boxes it wrote */
    bool rho_synced = true, boxes_only = false;
    int win_err = 0;          /* the last centroid's window check (the oracle's sticky check) */

    GCtl *ctl = nullptr, *hctl = nullptr;    /* on the device; its pinned host mirror */
    int *abrt = nullptr;                     /* &ctl->abrt: the loops' abort flag */

    ~GenCtx() override {
        if (s) (void)hipStreamSynchronize(s);
        for (void *q : allocs) (void)hipFree(q);
        if (hctl) (void)hipHostFree(hctl);
    }
    /* ctl -> hctl and wait: the one readback of a sync point */
    int read_ctl() {
        GHIP(hipMemcpyAsync(hctl, ctl, sizeof(GCtl), hipMemcpyDeviceToHost, s));
        GHIP(hipStreamSynchronize(s));
        return ORG_OK;
    }
    /* a loop launch's relax flag, abort flag, largest count and count to zero (one fill kernel) */
    int zero_loop_ctl(int slot) {
        GLAUNCH(g_zero_u32, dim3(1), dim3(64), 0, (unsigned *)&ctl->loop[slot], 4);
        return ORG_OK;
    }
    int real_bytes() const override { return (int)sizeof(R); }

    template <typename T>
    int A_(T **ptr, size_t count) { return dev_alloc(allocs, (void **)ptr, count * sizeof(T)); }

    int init(const org_params &pp, const double *s0_cell, hipStream_t stream) {
        p = pp;
        s = stream;
        N = p.N; nc = p.ncell; zs = p.zone_size; W = 2 * zs; M = 2 * zs - 2;
        NN = (size_t)N * N; PP = NN * nc; MM = (size_t)M * M;
        P.N = N; P.nc = nc; P.zs = zs; P.M = M; P.rimax = p.rimax; P.ncur = p.n_current;
        P.dt = (R)p.dt; P.dx = (R)p.dx; P.sigma = (R)p.sigma; P.cpl = (R)p.coupling_PF_AG;
        P.tau = (R)p.tau; P.alpha = (R)p.alpha; P.beta = (R)p.beta; P.eta = (R)p.eta;
        P.tau_lum = (R)p.tau_lum; P.beta_lc = (R)p.beta_lum_cells; P.eta_lc = (R)p.eta_lum_cells;
        P.xi_lum = (R)p.xi_lum; P.alpha_E = (R)p.alpha_ECM; P.tau_E = (R)p.tau_ECM;
        P.beta_ec = (R)p.beta_ECM_cells; P.eta_ec = (R)p.eta_ECM_cells; P.xi_E = (R)p.xi_ECM;
        P.beta_le = (R)p.beta_lum_ECM; P.wi = (R)p.wi; P.rdt = (R)p.rdt; P.friction = (R)p.friction;
        const double tab[6][3] = {{p.visc_ap, p.visc_bas, p.visc_lat}, {p.a_ort_ap, p.a_ort_bas, p.a_ort_lat},
                                  {p.a_tan_ap, p.a_tan_bas, p.a_tan_lat}, {p.b_ap, p.b_bas, p.b_lat},
                                  {p.gen_ap, p.gen_bas, p.gen_lat}, {p.deg_ap, p.deg_bas, p.deg_lat}};
        for (int l = 0; l < 3; l++) {
            P.visc[l] = (R)tab[0][l]; P.aort[l] = (R)tab[1][l]; P.atan_[l] = (R)tab[2][l];
            P.bb[l] = (R)tab[3][l]; P.gen[l] = (R)tab[4][l]; P.deg[l] = (R)tab[5][l];
        }
        GTRY(make_plan(M, plan, allocs));
        for (R **f : {&phi, &phin, &rho, &rhon, &hf, &ngp, &jx, &jy}) GTRY(A_(f, PP));
        for (R **f : {&lum, &lumn, &ecm, &ecmn, &psi_, &hl, &he, &ngl, &jlx, &jly, &jex, &jey}) GTRY(A_(f, NN));
        GTRY(A_(&A, 5 * nc * MM)); GTRY(A_(&B, 5 * nc * MM)); GTRY(A_(&V, 2 * nc * MM));
        GTRY(A_(&work, 5 * nc * MM));
        GTRY(A_(&vxb, nc * MM)); GTRY(A_(&vyb, nc * MM));
        GTRY(A_(&win, (size_t)nc * std::max(W * W, M * M)));
        nparts = (int)std::min<size_t>(256, (NN + 4095) / 4096);
        {
            int dev = 0, ncu = 0, khz = 0;
            GHIP(hipGetDevice(&dev));
            GHIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
            GHIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
            if (ncu <= 0 || khz <= 0) return ORG_EHIP;
            wait_limit = 2000LL * khz;                 /* 2 s */
            /* one workgroup per compute unit at most (all must be resident),
             * about 8192 pixels or more each: a sweep's share then takes
             * ~13 us against ~5 us for the barrier (measured on the 400^2
             * windows of the template configuration: 4096 pixels per
             * workgroup left the FP64-bound flux at ~12 us and doubled the
             * barrier's wait for the slowest workgroup) */
            kc = std::min(std::min(ncu / nc, (W * W + 8191) / 8192), W / 8);
            if (kc < 2) kc = 1;
            kl = (int)std::min<size_t>(std::min<size_t>((size_t)ncu, (NN + 4095) / 4096), (size_t)(N / 8));
            if (kl < 1) kl = 1;
            if (2 * sizeof(R) * (size_t)(N + 2) > 65536 || 2 * sizeof(R) * (size_t)(W + 2) > 65536) {
                fprintf(stderr, "organoid (generic path): grid side %d / window %d too large for the loops' LDS rows\n",
                        N, W);
                return ORG_EINVAL;
            }
        }
        GTRY(A_(&f0, nc)); GTRY(A_(&s_, nc)); GTRY(A_(&s0, nc)); GTRY(A_(&sum_e, 1));
        GTRY(A_(&part, nparts));
        GTRY(A_(&xc, nc)); GTRY(A_(&yc, nc)); GTRY(A_(&counts, nc));
        GTRY(A_(&ctl, 1));
        GHIP(hipHostMalloc((void **)&hctl, sizeof(GCtl)));
        err = &ctl->err; relax = &ctl->loop[0].relax; abrt = &ctl->loop[0].abrt; mb_cmax = &ctl->loop[0].cmax;
        GTRY(A_(&cen_acc, (size_t)3 * nc));
        GTRY(A_(&hbox, (size_t)4 * nc));
        GTRY(A_(&nz_part, (size_t)2 * ORG_NZ_PARTS));
        if (std::is_same<R, double>::value) GTRY(A_(&srows, (size_t)N * nc));
        if (kc > 1) GTRY(A_(&swin, (size_t)nc * 2 * (W + 2) * (W + 2)));
        GTRY(A_(&mb_part, (size_t)2 * std::max(nc * kc, kl)));
        GTRY(A_(&mb_ctr, (size_t)nc + 2));

        GTRY(A_(&mb_xch, std::max((size_t)nc * 2 * kc * 2 * (W + 2), (size_t)2 * kl * 2 * N)));
        {
            int ncu = 0, dev = 0;
            GHIP(hipGetDevice(&dev));
            GHIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
            /* one workgroup per compute unit at most (all resident), at
             * least 4 window rows each, 64 deviations a sweep at most */
            kcl = std::min(std::min(ncu / nc, W / 4), 64);
            if (kcl >= 2) {
                rmax_l = (W + kcl - 1) / kcl;
                lds_l = g_band_lds(ORG_GEN_SPAN_CELLS, rmax_l, W, sizeof(R));
                if (lds_l > ORG_GEN_LDS_MAX) kcl = 0;
            } else {
                kcl = 0;
            }
            if (kcl) {
                const size_t words = (size_t)nc * g_band_words(ORG_GEN_SPAN_CELLS, kcl, W, sizeof(R));
                GTRY(A_(&mb_tw, words));
                GHIP(hipMemset(mb_tw, 0, sizeof(unsigned long long) * words));
                GHIP(hipFuncSetAttribute((const void *)g_relax_cells_lds<ORG_GEN_SPAN_CELLS, R>,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_l));
            }
            /* the lumen/ECM loop: one band of box rows per workgroup, up to
             * one workgroup per compute unit and 128 in all (with 256, the
             * polls of twice the deviations cost more than the smaller
             * bands save: 1.38 against 1.17 ms per template step); each
             * launch checks that its bands fit */
            klm = std::min(std::min(ncu, 128), N - 2);
            if (klm >= 2) {
                const size_t words = g_band_words(ORG_GEN_SPAN_LUM, klm, N, sizeof(R));
                for (int f = 0; f < 2; f++) {
                    GTRY(A_(&mb_lw[f], words));
                    GHIP(hipMemset(mb_lw[f], 0, sizeof(unsigned long long) * words));
                }
                GHIP(hipFuncSetAttribute((const void *)g_lum_loop_lds<ORG_GEN_SPAN_LUM, R>,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, ORG_GEN_LDS_MAX));
            } else {
                klm = 0;
            }
        }
        std::vector<double> s0v(nc);
        for (int n = 0; n < nc; n++) s0v[n] = (p.use_s0_cell && s0_cell) ? s0_cell[n] : p.s0;
        std::vector<int> cen(nc, N / 2);
        GHIP(hipMemcpy(s0, s0v.data(), sizeof(double) * nc, hipMemcpyHostToDevice));
        GHIP(hipMemcpy(xc, cen.data(), sizeof(int) * nc, hipMemcpyHostToDevice));
        GHIP(hipMemcpy(yc, cen.data(), sizeof(int) * nc, hipMemcpyHostToDevice));
        for (R *f : {phi, phin, rho, rhon, hf, ngp, jx, jy}) GHIP(hipMemset(f, 0, sizeof(R) * PP));
        for (R *f : {lum, lumn, ecm, ecmn, psi_, hl, he, ngl, jlx, jly, jex, jey})
            GHIP(hipMemset(f, 0, sizeof(R) * NN));
        GHIP(hipMemset(vxb, 0, sizeof(R) * nc * MM)); GHIP(hipMemset(vyb, 0, sizeof(R) * nc * MM));
        GHIP(hipMemset(f0, 0, sizeof(R) * nc)); GHIP(hipMemset(s_, 0, sizeof(double) * nc));
        GHIP(hipMemset(ctl, 0, sizeof(GCtl))); GHIP(hipMemset(counts, 0, sizeof(int) * nc));
        GHIP(hipMemset(A, 0, sizeof(Cx<R>) * 5 * nc * MM));
        GHIP(hipMemset(B, 0, sizeof(Cx<R>) * 5 * nc * MM));
        GHIP(hipDeviceSynchronize());          /* the memsets above ran on the null stream */
        /* the transform schedule, timed on the step's largest batch (5 ncell planes) */
        GTRY(tune_fft<R>(s, plan, A, B, work, 5 * nc, &fft_times));
        return ORG_OK;
    }
    std::vector<float> fft_times;
    int get_fft_schedule(int32_t *fuse, int32_t *lpw) override {
        *fuse = plan.fuse;
        *lpw = plan.lpw;
        return ORG_OK;
    }
    int set_fft_schedule(int fuse, int lpw) override {
        if (!schedule_ok(plan, fuse, lpw)) return ORG_EINVAL;
        GHIP(hipStreamSynchronize(s));
        plan.fuse = fuse;
        plan.lpw = lpw;
        return ORG_OK;
    }

    R *field(int f, size_t *n) {
        switch (f) {
        case ORG_F_PHI: *n = PP; return phi;
        case ORG_F_RHO: *n = PP; return rho;
        case ORG_F_HFUNC: *n = PP; return hf;
        case ORG_F_LUM: *n = NN; return lum;
        case ORG_F_ECM: *n = NN; return ecm;
        case ORG_F_PSI: *n = NN; return psi_;
        case ORG_F_HLUM: *n = NN; return hl;
        case ORG_F_HECM: *n = NN; return he;
        case ORG_F_LUMN: *n = NN; return lumn;
        case ORG_F_ECMN: *n = NN; return ecmn;
        default: *n = 0; return nullptr;
        }
    }

    /* host <-> device in the caller's precision */
    int upload(R *d, const void *host, size_t n, int f64) {
        if ((sizeof(R) == 8) == (f64 != 0)) {
            GHIP(hipMemcpyAsync(d, host, n * sizeof(R), hipMemcpyHostToDevice, s));
            GHIP(hipStreamSynchronize(s));
            return ORG_OK;
        }
        std::vector<R> t(n);
        if (f64) for (size_t k = 0; k < n; k++) t[k] = (R)((const double *)host)[k];
        else for (size_t k = 0; k < n; k++) t[k] = (R)((const float *)host)[k];
        GHIP(hipMemcpyAsync(d, t.data(), n * sizeof(R), hipMemcpyHostToDevice, s));
        GHIP(hipStreamSynchronize(s));
        return ORG_OK;
    }
    int download(const R *d, void *host, size_t n, int f64) {
        if ((sizeof(R) == 8) == (f64 != 0)) {
            GHIP(hipMemcpyAsync(host, d, n * sizeof(R), hipMemcpyDeviceToHost, s));
            GHIP(hipStreamSynchronize(s));
            return ORG_OK;
        }
        std::vector<R> t(n);
        GHIP(hipMemcpyAsync(t.data(), d, n * sizeof(R), hipMemcpyDeviceToHost, s));
        GHIP(hipStreamSynchronize(s));
        if (f64) for (size_t k = 0; k < n; k++) ((double *)host)[k] = (double)t[k];
        else for (size_t k = 0; k < n; k++) ((float *)host)[k] = (float)t[k];
        return ORG_OK;
    }

    int set_field(int f, const void *host, int f64) override {
        size_t n;
        R *d = field(f, &n);
        if (!d || !host) return ORG_EINVAL;
        if (d == rho) rho_synced = false;
        if (d == hf) hbox_stale = true;
        return upload(d, host, n, f64);
    }
    int get_field(int f, void *host, int f64) override {
        size_t n;
        R *d = field(f, &n);
        if (!d || !host) return ORG_EINVAL;
        return download(d, host, n, f64);
    }
    int get_window(int which, void *host, int f64) override {
        if (!host) return ORG_EINVAL;
        const dim3 blk(256);
        switch (which) {
        case ORG_W_PHIN:
            GLAUNCH((g_gather<R>), dim3((W * W + 255) / 256, nc), blk, 0, N, (const R *)phin, xc, yc, -zs, W, win);
            return download(win, host, (size_t)nc * W * W, f64);
        case ORG_W_RHON:
            GLAUNCH((g_gather<R>), dim3((M * M + 255) / 256, nc), blk, 0, N, (const R *)rhon, xc, yc, 1 - zs, M, win);
            return download(win, host, (size_t)nc * MM, f64);
        case ORG_W_VX: return download(vxb, host, (size_t)nc * MM, f64);
        case ORG_W_VY: return download(vyb, host, (size_t)nc * MM, f64);
        default: return ORG_EINVAL;
        }
    }
    int get_centres(int32_t *x, int32_t *y) override {
        GHIP(hipMemcpyAsync(x, xc, sizeof(int) * nc, hipMemcpyDeviceToHost, s));
        GHIP(hipMemcpyAsync(y, yc, sizeof(int) * nc, hipMemcpyDeviceToHost, s));
        GHIP(hipStreamSynchronize(s));
        return ORG_OK;
    }
    int get_areas(double *sv, double *f0v) override {
        std::vector<R> f(nc);
        GHIP(hipMemcpyAsync(sv, s_, sizeof(double) * nc, hipMemcpyDeviceToHost, s));
        GHIP(hipMemcpyAsync(f.data(), f0, sizeof(R) * nc, hipMemcpyDeviceToHost, s));
        GHIP(hipStreamSynchronize(s));
        for (int n = 0; n < nc; n++) f0v[n] = (double)f[n];
        return ORG_OK;
    }
    int get_iterations(int32_t *cells, int32_t *l, int32_t *e) override {
        GHIP(hipMemcpyAsync(cells, counts, sizeof(int) * nc, hipMemcpyDeviceToHost, s));
        GHIP(hipStreamSynchronize(s));
        *l = count_lum;
        *e = count_ecm;
        return ORG_OK;
    }

    int copy(R *dst, const R *src, size_t n) {
        GHIP(hipMemcpyAsync(dst, src, n * sizeof(R), hipMemcpyDeviceToDevice, s));
        return ORG_OK;
    }
    int grid_for(size_t n) const { return (int)std::min<size_t>(4096, (n + 255) / 256); }
    /* workgroups for a launch of one row per workgroup (at most 65535) */
    static int row_blocks(int rows) { return std::max(1, std::min(rows, 65535)); }
    int bc(R *planes, int nplanes) {
        GLAUNCH((g_bc_cells<R>), dim3(nplanes, std::max(1, (N - 2 + 63) / 64)), dim3(64), 0, N, planes);
        return ORG_OK;
    }
    int window_check() const { return win_err; }

    /* run_with_active_gels.cpp:92-98: lumn <- lum, ECMn <- ECM, phin <- phi, rhon <- rho */
    int prepare() override {
        GTRY(copy(ecmn, ecm, NN));
        GTRY(copy(lumn, lum, NN));
        GTRY(copy(phin, phi, PP));
        GTRY(copy(rhon, rho, PP));
        rho_synced = true;
        GHIP(hipMemsetAsync(psi_, 0, sizeof(R) * NN, s));
        GHIP(hipStreamSynchronize(s));
        count_lum = count_ecm = 0;
        return ORG_OK;
    }
    int find_centres() override {
        boxes_only = false;          /* the boxes move with the centres */
        GLAUNCH(g_zero_u32, dim3(1), dim3(64), 0, (unsigned *)err, 1);
        if (!(chain && cen_live)) {
            GLAUNCH(g_zero_u32, dim3(1), dim3(256), 0, (unsigned *)cen_acc, 6 * nc);
            GLAUNCH((g_centroid_part<R>), dim3(std::min(N, std::max(1, 4096 / nc)), nc), dim3(256), 0, P,
                    (const R *)phi, cen_acc);
        }
        cen_live = chain;
        GLAUNCH((g_centroid_fin<R>), dim3(1), dim3(64), 0, P, (const unsigned long long *)cen_acc, xc, yc, err);
        int e = 0;
        GTRY(read_ctl());
        e = hctl->err;
        if (e == ORG_EEMPTY_CELL) return e;
        win_err = e;
        return e;
    }
    int centroid_win(int sign) {
        /* about 2048 workgroups in all, at least 16 rows' worth per cell */
        GLAUNCH((g_centroid_win<R>), dim3(std::min(W, std::max(16, 2048 / nc)), nc), dim3(256), 0, P, (const R *)phi,
                (const int *)xc, (const int *)yc, cen_acc, sign);
        return ORG_OK;
    }
    int hfunc_cells() override {
        GTRY(window_check());
        GLAUNCH((g_hfunc_cells<R>), dim3(W, nc), dim3(256), 0, P, (const R *)phi, hf, (const int *)xc, (const int *)yc,
                hbox);
        return ORG_OK;
    }
    int hfunc_lum_ecm() override {
        GLAUNCH((g_hfunc_lumecm<R>), dim3(grid_for(NN)), dim3(256), 0, P, (const R *)lum, (const R *)ecm, hl, he);
        return ORG_OK;
    }
    /* the hfunc boxes from the planes, after an upload of hfunc */
    int hbox_fresh() {
        if (!hbox_stale) return ORG_OK;
        GLAUNCH(g_hbox_reset, dim3(1), dim3(256), 0, hbox, nc, N);
        GLAUNCH((g_hbox_scan<R>), dim3(std::min(N, std::max(1, 4096 / nc)), nc), dim3(256), 0, P, (const R *)hf, hbox);
        hbox_stale = false;
        return ORG_OK;
    }
    int psi() override {
        GTRY(hbox_fresh());
        GLAUNCH((g_psi_box<R>), dim3((N + 255) / 256, row_blocks(N)), dim3(256), 0, P, (const R *)hf,
                (const int4 *)hbox, psi_);
        return ORG_OK;
    }
    int sum_s() override {
        if constexpr (std::is_same<R, double>::value) {
            /* FP64 state: row sums over each cell's hfunc box, on 16
             * workgroups per cell; FP32 keeps g_sum_s's order, whose sum the
             * float build's f0 rounds the same way */
            GTRY(hbox_fresh());
            GLAUNCH(g_area_rows, dim3(16, nc), dim3(256), 0, P, (const double *)hf, (const int4 *)hbox, srows);
            GLAUNCH(g_area_fin, dim3(nc), dim3(256), 0, P, (const double *)srows, (const int4 *)hbox,
                    (const double *)s0, s_, f0);
        } else {
            GLAUNCH((g_sum_s<R>), dim3(nc), dim3(1024), 0, P, (const R *)hf, (const double *)s0, s_, f0);
        }
        return ORG_OK;
    }
    /* compute_normal_grad_phi + boundary_conditions_cells (run_with_active_gels.cpp:135-136) */
    int norm_grad() {
        if (in_step && ngp_fresh) return ORG_OK;
        if (windows_only()) {
            GLAUNCH((g_norm_grad_win<R>), dim3(W + 2, nc), dim3(256), 0, P, (const R *)phi, (const int *)xc,
                    (const int *)yc, ngp);
        } else {
            GLAUNCH((g_norm_grad<R>), dim3(row_blocks(N - 2), nc), dim3(256), 0, P, (const R *)phi, ngp);
            GTRY(bc(ngp, nc));
        }
        ngp_fresh = in_step;
        return ORG_OK;
    }
    /* inside one step (step_scope) |grad phi| and the normals are read only
     * on each window grown by one (the spectral box, M = W - 2, lies inside):
     * they are formed there alone, without the planes' periodic rings; the
     * per-function entry points form them on every cell's whole plane */
    bool windows_only() const { return in_step && M <= W; }
    void step_scope(bool on) override {
        in_step = on;
        ngp_fresh = false;
    }
    void step_chain(bool on) override {
        chain = on;
        cen_live = false;
    }
    /* run_with_active_gels.cpp:130-154: turnover maps, prod_norm_delta and
     * the whole per-cell spectral loop of update_all_cortices
     * (rho_functions.cpp:409-471) for all cells at once */
    int cortices() override {
        GTRY(window_check());
        /* rhon is written on the boxes (and its ring, a copy of interior
         * pixels no box reaches): inside a step the copy-back that follows
         * restores rho == rhon from the boxes alone; otherwise (and if the
         * step ends before its copy-back) a whole-plane copy-back does */
        boxes_only = in_step && rho_synced;
        rho_synced = false;
        GTRY(norm_grad());
        const size_t pl = (size_t)nc * MM;
        const dim3 g(std::max(1, std::min(64, (int)((MM + 255) / 256))), nc), b(256);
        GLAUNCH((g_spec_load<R>), g, b, 0, P, (const R *)phi, (const R *)rho, (const R *)ngp, (const R *)lum,
                (const R *)ecm, (const int *)xc, (const int *)yc, A);
        GTRY(fft2d_planes<R>(s, plan, A, B, work, 5 * nc, -1));                       /* :417-440 */
        GLAUNCH((g_spec_solve<R>), g, b, 0, P, (const R *)phi, (const R *)lum, (const R *)ecm, (const int *)xc,
                (const int *)yc, (const Cx<R> *)B, V);                                 /* :445 */
        GTRY(fft2d_planes<R>(s, plan, V, A, work, 2 * nc, +1));                       /* :449, :455 */
        GLAUNCH((g_spec_rhov<R>), g, b, 0, P, (const R *)rho, (const int *)xc, (const int *)yc, (const Cx<R> *)A,
                A + 2 * pl, vxb);
        GLAUNCH((g_spec_rhov<R>), g, b, 0, P, (const R *)rho, (const int *)xc, (const int *)yc,
                (const Cx<R> *)(A + pl), A + 3 * pl, vyb);
        GTRY(fft2d_planes<R>(s, plan, A + 2 * pl, A, work, 2 * nc, -1));              /* :451, :457 */
        GLAUNCH((g_spec_update<R>), dim3(grid_for(pl)), b, 0, P, B, (const Cx<R> *)(B + pl), (const Cx<R> *)A,
                (const Cx<R> *)(A + pl));                                              /* :462 */
        GTRY(fft2d_planes<R>(s, plan, B, A + 4 * pl, work, nc, +1));                  /* :464 */
        GLAUNCH((g_spec_commit<R>), g, b, 0, P, (const int *)xc, (const int *)yc, (const Cx<R> *)(A + 4 * pl), rhon);
        return bc(rhon, nc);                                                           /* :469 */
    }
    /* fixed-order double sum (mode 0: h dx^2 over a; 1: |a - b| dx^2) into *out */
    int plane_sum(const R *a, const R *b2, int mode, double *out) {
        const size_t per = (NN + nparts - 1) / nparts;
        GLAUNCH((g_sum_parts<R>), dim3(nparts), dim3(1024), 0, a, b2, P.dx, NN, per, mode, part);
        GLAUNCH(g_sum_final, dim3(1), dim3(64), 0, (const double *)part, nparts, out);
        return ORG_OK;
    }
    int lumecm_update() override {
        GTRY(plane_sum(he, nullptr, 0, sum_e));                                        /* :376-386 */
        GLAUNCH((g_lumecm_update<R>), dim3(row_blocks(N)), dim3(256), 0, P, (const R *)lum, (const R *)ecm,
                (const R *)psi_, (const R *)hl, (const R *)he, (const double *)sum_e, lumn, ecmn);
        return ORG_OK;
    }
    /* workgroups of g_lum_loop_lds for a bh x bw box: as many as allowed,
     * ORG_GEN_SPAN_LUM rows or more each; 0 (the exchange-buffer loop runs)
     * when the bands do not fit in LDS */
    int lum_lds_blocks(int bh, int bw) const {
        if (klm < 2) return 0;
        const int kb = std::min(klm, bh / ORG_GEN_SPAN_LUM);
        if (kb < 2 || g_band_lds(ORG_GEN_SPAN_LUM, (bh + kb - 1) / kb, bw, sizeof(R)) > ORG_GEN_LDS_MAX) return 0;
        return kb;
    }
    /* One lumen/ECM loop (reshaping_lumen_or_ECM, cell_phase_functions.cpp:
     * 709-871) in three phases around two readbacks of the control block:
     * lum_pre (normals, sweep 1 on the whole grid, its deviation and the box
     * of nonzero normals), lum_mid (the host's stop and rimax decisions, the
     * ring, the loop launch), lum_post (the loop's count, relax and abort
     * flags; the buffers' roles).  Each field has its own normals, control
     * slots, exchange words and tags, so the lumen's and the ECM's loops can
     * share their readbacks (reshape_lumecm_pair): neither reads the other's
     * field. */
    struct LumJob {
        int f = 0;                 /* 0 lumen, 1 ECM */
        bool started = false;      /* the start deviation was above the threshold */
        double rdev = 0.0, thr = 0.0;
        int count = 0, ri = 0, failed = 0;
        int launched = 0;          /* 0 no loop launch, 1 banded (LDS), 2 exchange-buffer */
    };
    R *&fld(int f) { return f ? ecm : lum; }
    R *&fldn(int f) { return f ? ecmn : lumn; }
    int lum_pre(LumJob &j) {
        const int f = j.f;
        R *u = fld(f), *un = fldn(f), *nx = f ? jex : jlx, *ny = f ? jey : jly;
        const R side = f ? (R)1.0 : (R)0.0;
        const R rtol = (R)p.rtol, rdt = (R)p.rdt;
        GLAUNCH((g_lum_grad<R>), dim3(row_blocks(N - 2)), dim3(256), 0, P, (const R *)u, ngl);   /* :734-736 */
        GTRY(bc(ngl, 1));                                                                     /* :741 */
        GLAUNCH((g_lum_normals<R>), dim3(row_blocks(N - 2)), dim3(256), 0, P, (const R *)un, (const R *)ngl, nx, ny);
        GTRY(bc(nx, 1));                                                                      /* :770-771 */
        GTRY(bc(ny, 1));
        j.rdev = (double)(p.rdev_ini * rtol * rdt);                                           /* :720 */
        j.thr = (double)(p.rdev_threshold * rtol * rdt);                                      /* :774 */
        j.count = 0; j.ri = 0; j.failed = 0; j.launched = 0;
        j.started = j.rdev > j.thr;
        if (j.started) {
            /* sweep 1 on the whole grid: lum <- lumn (the role change), lumn <- F(lum) into lum's buffer */
            GLAUNCH((g_lum_sweep<R>), dim3(row_blocks(N)), dim3(256), 0, P, (const R *)un, u, (const R *)nx,
                    (const R *)ny, side);
            GTRY(plane_sum(u, un, 1, &ctl->rdev[f]));                                         /* :848-855 */
            /* per-workgroup bounds, then one workgroup combines them */
            const int nb = std::max(1, std::min(N - 2, ORG_NZ_PARTS));
            int4 *part = nz_part + (size_t)f * ORG_NZ_PARTS;
            GLAUNCH((g_nz_box<R>), dim3(nb), dim3(256), 0, N, (const R *)nx, (const R *)ny, part);
            GLAUNCH(g_nz_box_fin, dim3(1), dim3(256), 0, N, (const int4 *)part, nb, ctl->box[f]);
        }
        return ORG_OK;
    }
    /* after a readback of the control block */
    int lum_mid(LumJob &j) {
        if (!j.started) return ORG_OK;
        const int f = j.f;
        R *u = fld(f), *un = fldn(f), *nx = f ? jex : jlx, *ny = f ? jey : jly;
        const R side = f ? (R)1.0 : (R)0.0;
        j.rdev = hctl->rdev[f];
        const int *bx = hctl->box[f];
        j.count = 1;
        if (j.ri > p.rimax) j.failed = 1;                                                     /* :862-865 */
        else j.ri++;
        const int bh = bx[1] - bx[0], bw = bx[3] - bx[2];
        if (j.failed || !(j.rdev > j.thr)) return ORG_OK;
        /* the ring sweep 2 writes into lumn's buffer, now: it is then the
         * same constant ring in both buffers */
        GLAUNCH((g_ring_dirichlet<R>), dim3((4 * N + 255) / 256), dim3(256), 0, N, un, side);
        if (bh <= 0 || bw <= 0) {
            /* no nonzero normals: sweep 2 changes nothing (both buffers already agree) */
            j.count = 2;
            if (j.ri > p.rimax) j.failed = 1;
            return ORG_OK;
        }
        GLoopCtl *lc = &ctl->loop[f];
        GTRY(zero_loop_ctl(f));
        if (int kb = lum_lds_blocks(bh, bw)) {
            /* sweeps 2, 3, ... with the bands in LDS */
            const int rows_max = (bh + kb - 1) / kb;
            GLAUNCH((g_lum_loop_lds<ORG_GEN_SPAN_LUM, R>), dim3(kb), dim3(1024),
                    g_band_lds(ORG_GEN_SPAN_LUM, rows_max, bw, sizeof(R)), P, u, un, (const R *)nx, (const R *)ny,
                    (const int *)ctl->box[f], j.thr, j.rdev, j.count, j.ri, mb_lw[f], tag_lum[f], &lc->count,
                    &lc->relax, &lc->abrt, wait_limit, kb, rows_max);
            j.launched = 1;
        } else {
            /* sweeps 2, 3, ... on the device, bands of at least 8 box rows
             * and about 8192 pixels */
            const int k = std::max(1, std::min(std::min(kl, (bh + 7) / 8), (int)(((size_t)bh * bw + 8191) / 8192)));
            const size_t shm = 2 * sizeof(R) * (size_t)(bw + 2);
            GLAUNCH(g_zero_u32, dim3(1), dim3(64), 0, mb_ctr + nc + f, 1);
            GLAUNCH((g_lum_loop_mb<R>), dim3(k), dim3(1024), shm, P, u, un, (const R *)nx, (const R *)ny,
                    (const int *)ctl->box[f], j.thr, j.rdev, j.count, j.ri, mb_xch, mb_part, mb_ctr + nc + f,
                    &lc->count, &lc->relax, &lc->abrt, wait_limit, k);
            j.launched = 2;
        }
        return ORG_OK;
    }
    /* after a readback of the control block, if lum_mid launched a loop */
    int lum_post(LumJob &j) {
        const int f = j.f;
        if (j.launched) {
            const GLoopCtl &h = hctl->loop[f];
            if (h.abrt) {
                if (j.launched == 1) tag_lum[f] += 1u << 20;   /* past every tag the aborted launch may have used */
                fprintf(stderr, "organoid (generic path): the lumen/ECM loop's workgroups waited on each "
                                "other for more than 2 s (not all resident?)\n");
                return ORG_ETIMEOUT;
            }
            if (j.launched == 1) tag_lum[f] += (unsigned)h.count + 2u;
            j.count = h.count;
            j.failed = h.relax;
        }
        if (j.count & 1) std::swap(fld(f), fldn(f));   /* an odd count leaves the last iterate in lum's buffer */
        (f ? count_ecm : count_lum) = j.count;
        return ORG_OK;
    }
    int reshape_lumecm(int which, int *failed) override {
        LumJob j;
        j.f = which ? 1 : 0;
        GTRY(lum_pre(j));
        if (j.started) {
            GTRY(read_ctl());
            GTRY(lum_mid(j));
            if (j.launched) GTRY(read_ctl());
        }
        GTRY(lum_post(j));
        *failed = j.failed;
        return ORG_OK;
    }
    int reshape_lumecm_pair(int *failed) override {
        LumJob j[2];
        j[0].f = 0;
        j[1].f = 1;
        GTRY(lum_pre(j[0]));
        GTRY(lum_pre(j[1]));
        if (j[0].started || j[1].started) GTRY(read_ctl());
        GTRY(lum_mid(j[0]));
        GTRY(lum_mid(j[1]));
        if (j[0].launched || j[1].launched) GTRY(read_ctl());
        GTRY(lum_post(j[0]));
        GTRY(lum_post(j[1]));
        failed[0] = j[0].failed;
        failed[1] = j[1].failed;
        return ORG_OK;
    }
    int update_cells() override {
        GTRY(window_check());
        GTRY(norm_grad());
        GLAUNCH((g_update_cells<R>), dim3(W, nc), dim3(256), 0, P, (const R *)phi, (const R *)rho, (const R *)ngp,
                (const R *)psi_, (const R *)hf, (const R *)hl, (const R *)he, (const R *)f0, (const int *)xc,
                (const int *)yc, phin);
        return bc(phin, nc);                                          /* run_with_active_gels.cpp:170 */
    }
    int reshape_cells(int from_phi, int *failed) override {
        GTRY(window_check());
        if (from_phi) GTRY(copy(phin, phi, PP));
        GTRY(norm_grad());
        if (windows_only()) {
            GLAUNCH((g_normals_win<R>), dim3(W + 2, nc), dim3(256), 0, P, (const R *)phi, (const R *)ngp,
                    (const int *)xc, (const int *)yc, jx, jy);
        } else {
            GLAUNCH((g_normals<R>), dim3(row_blocks(N - 2), nc), dim3(256), 0, P, (const R *)phi, (const R *)ngp, jx,
                    jy);
            GTRY(bc(jx, nc));                                         /* run_with_active_gels.cpp:164-166 */
            GTRY(bc(jy, nc));
        }
        const R rtol = (R)p.rtol, Rr = (R)p.R, wi = (R)p.wi, rdt = (R)p.rdt;
        const double thr = (double)(rtol * 2.0 * M_PI * Rr * wi * rdt);                       /* :475 */
        const double rdev0 = (double)(2.0 * rtol * 2.0 * M_PI * Rr * wi * rdt);               /* :473 */
        GTRY(zero_loop_ctl(0));
        /* the loops write phi on the windows alone: carried centroid sums
         * lose the windows' old pixels here and get their new ones after */
        if (cen_live) GTRY(centroid_win(-1));
        if (kcl) {
            GLAUNCH((g_relax_cells_lds<ORG_GEN_SPAN_CELLS, R>), dim3(nc * kcl), dim3(1024), lds_l, P, phi, phin,
                    (const R *)jx, (const R *)jy, (const int *)xc, (const int *)yc, thr, rdev0, mb_tw, tag_next,
                    counts, mb_cmax, relax, abrt, wait_limit, kcl, rmax_l);
            if (cen_live) GTRY(centroid_win(1));
            GTRY(read_ctl());
            const int f[3] = {hctl->loop[0].relax, hctl->loop[0].abrt, hctl->loop[0].cmax};
            if (f[1]) {
                tag_next += 1u << 20;      /* past every tag the aborted launch may have used */
                fprintf(stderr, "organoid (generic path): a cell's reshaping workgroups waited on each other "
                                "for more than 2 s (not all resident?)\n");
                return ORG_ETIMEOUT;
            }
            tag_next += (unsigned)f[2] + 2u;
            *failed = f[0];
            return ORG_OK;
        }
        if (kc > 1) {
            GLAUNCH(g_zero_u32, dim3(1), dim3(64), 0, mb_ctr, nc + 1);
            GLAUNCH((g_relax_cells_mb<R>), dim3(nc * kc), dim3(1024), 2 * sizeof(R) * (W + 2), P, phi, phin,
                    (const R *)jx, (const R *)jy, (const int *)xc, (const int *)yc, thr, rdev0, swin, mb_xch,
                    mb_part, mb_ctr, counts, relax, abrt, wait_limit, kc);
        } else {
            GLAUNCH((g_relax_cells<R>), dim3(nc), dim3(1024), 0, P, phi, phin, (const R *)jx, (const R *)jy,
                    (const int *)xc, (const int *)yc, thr, rdev0, counts, relax);
        }
        if (cen_live) GTRY(centroid_win(1));
        GTRY(read_ctl());
        const int f[2] = {hctl->loop[0].relax, hctl->loop[0].abrt};
        if (f[1]) {
            fprintf(stderr, "organoid (generic path): a cell's reshaping workgroups waited on each other "
                            "for more than 2 s (not all resident?)\n");
            return ORG_ETIMEOUT;
        }
        *failed = f[0];
        return ORG_OK;
    }
    int commit_rho() override {
        if (boxes_only) {
            GLAUNCH((g_copy_boxes<R>), dim3(M, nc), dim3(256), 0, P, (const int *)xc, (const int *)yc,
                    (const R *)rhon, rho);
        } else {
            GTRY(copy(rho, rhon, PP));
        }
        boxes_only = false;
        rho_synced = true;
        return ORG_OK;
    }
    int commit_lumecm() override {
        GTRY(copy(ecm, ecmn, NN));
        return copy(lum, lumn, NN);
    }
    int rho_to_shape(double rho0, double tol) override {
        rho_synced = false;
        boxes_only = false;
        GLAUNCH((g_rho_to_shape<R>), dim3(grid_for(PP)), dim3(256), 0, (const R *)phi, rho, PP, rho0, tol);
        GHIP(hipStreamSynchronize(s));
        return ORG_OK;
    }
    int zero_gel_inputs() override {
        GHIP(hipMemsetAsync(psi_, 0, sizeof(R) * NN, s));
        GHIP(hipMemsetAsync(hl, 0, sizeof(R) * NN, s));
        GHIP(hipMemsetAsync(he, 0, sizeof(R) * NN, s));
        return ORG_OK;
    }
};

int create_generic(const org_params *p, const double *s0_cell, int f64, hipStream_t stream, Generic **out) {
    *out = nullptr;
    if (2 * p->zone_size - 2 < 2 || p->N < 2 * p->zone_size + 6) return ORG_EINVAL;
    int rc;
    if (f64) {
        auto *g = new GenCtx<double>();
        rc = g->init(*p, s0_cell, stream);
        if (rc) { delete g; return rc; }
        *out = g;
    } else {
        auto *g = new GenCtx<float>();
        rc = g->init(*p, s0_cell, stream);
        if (rc) { delete g; return rc; }
        *out = g;
    }
    return ORG_OK;
}

template <typename R>
static int debug_fft2d_t(int M, double *data, int sign, int fuse, int lpw) {
    std::vector<void *> allocs;
    FftPlan pl;
    Cx<R> *buf = nullptr;
    Cx<double> *work = nullptr;
    hipStream_t s = nullptr;
    int rc = make_plan(M, pl, allocs);
    if (!rc && !schedule_ok(pl, fuse, lpw)) rc = ORG_EINVAL;
    pl.fuse = fuse;
    pl.lpw = lpw;
    const size_t MM = (size_t)M * M;
    std::vector<Cx<R>> h(MM);
    for (size_t k = 0; k < MM; k++) { h[k].x = (R)data[2 * k]; h[k].y = (R)data[2 * k + 1]; }
    if (!rc) rc = dev_alloc(allocs, (void **)&buf, sizeof(Cx<R>) * MM);
    if (!rc) rc = dev_alloc(allocs, (void **)&work, sizeof(Cx<double>) * MM);
    if (!rc && hipMemcpy(buf, h.data(), sizeof(Cx<R>) * MM, hipMemcpyHostToDevice) != hipSuccess) rc = ORG_EHIP;
    if (!rc) rc = fft2d_planes<R>(s, pl, buf, buf, work, 1, sign);
    if (!rc && hipMemcpy(h.data(), buf, sizeof(Cx<R>) * MM, hipMemcpyDeviceToHost) != hipSuccess) rc = ORG_EHIP;
    for (void *q : allocs) (void)hipFree(q);
    if (rc) return rc;
    for (size_t k = 0; k < MM; k++) { data[2 * k] = (double)h[k].x; data[2 * k + 1] = (double)h[k].y; }
    return ORG_OK;
}

int debug_fft2d(int M, double *data, int sign, int f64, int fuse, int lpw) {
    if (M < 2 || M > 1024 || !data || (sign != 1 && sign != -1)) return ORG_EINVAL;
    return f64 ? debug_fft2d_t<double>(M, data, sign, fuse, lpw) : debug_fft2d_t<float>(M, data, sign, fuse, lpw);
}

}  // namespace gen
