/*
 * fno.hip -- the Fourier-neural-operator velocity surrogate on gfx950
 * (include/organoid_fno.h): one batched BF16 matrix-core GEMM (MFMA
 * 16x16x32, FP32 accumulation) that every layer, transform and gradient is
 * expressed in, a few element-wise kernels, and the forward / backward /
 * Adam orchestration.
 *
 * The truncated spectral convolution runs as four GEMMs per direction:
 * along y a real GEMM with the 2*modes (cos, -sin) rows of the real DFT,
 * along x a real GEMM over (x, re/im) with the complex twiddles of the 2*modes
 * retained kx, the per-mode complex channel mix as a real 2w x 2w GEMM per
 * mode, and the transposed pair for the inverse.  Layouts (B batch, w width,
 * P = M*M pixels, m = modes):
 *   v, u, dv      [B][w][M][M]          (channel planes)
 *   Z, dZ         [B][w][M][2][m]        y-transform: row (b,c,x), (re|im, ky)
 *   Zh, dZh       [B][w][2m][2][m]       x-transform: (kx_idx, re|im, ky)
 *   Xh, Yh        [nmodes][B][w][2]      per-mode channel vectors
 *   Zp, dZp       [B][w][M][2][m]        inverse x-transform
 * Every GEMM call below states its operands in those layouts.
 */
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "organoid.h"
#include "organoid_fno.h"

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

#define FG_BK 32
#define FG_LD (FG_BK + 8)     /* LDS row stride (bf16): 80 B, keeps 16-B fragment reads aligned */

static __device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }
static __device__ __forceinline__ float gelu_d(float x) {
    return 0.5f * (1.0f + erff(x * 0.70710678118654752f)) + x * 0.39894228040143268f * expf(-0.5f * x * x);
}

template <bool AF>
static __device__ __forceinline__ float ld_op(const void *p, int64_t o) {
    return AF ? ((const float *)p)[o] : (float)((const bf16 *)p)[o];
}

/* C = epilogue(alpha A B), one BM x BN tile of C (64 or 32 each) per
 * 256-thread workgroup, the 4 waves laid out by TileShape (64 x 64: 2 x 2
 * waves of 32 x 32, i.e. 2 x 2 MFMA 16x16x32 tiles), K in steps of 32 through
 * LDS.  Tiles are staged with the contiguous dimension across threads
 * (k if the operand's k-stride is 1, else m or n), converted to BF16, and
 * stored k-contiguous so each lane reads its 8-element fragment as one
 * 16-byte LDS load: A fragment rows = lane % 16, k = 8 (lane / 16) .. +7;
 * B fragment columns = lane % 16, same k; result rows 4 (lane / 16) .. +3,
 * column lane % 16. */
/* One K segment of a GEMM: C += A B over k < K (strides as fno_gemm_args). */
struct GemmSeg {
    const void *A;
    int64_t sAm, sAk, sA1, sA2;
    const void *B;
    int64_t sBk, sBn, sB1, sB2;
    int K;
};

/* Eight consecutive operand elements from a 16-byte aligned offset. */
template <bool F32>
static __device__ __forceinline__ void ld8(const void *p, int64_t o, float (&v)[8]) {
    if (F32) {
        const float4 a = *(const float4 *)((const float *)p + o), b = *(const float4 *)((const float *)p + o + 4);
        v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w; v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
    } else {
        const bf16x8 a = *(const bf16x8 *)((const bf16 *)p + o);
#pragma unroll
        for (int i = 0; i < 8; i++) v[i] = (float)a[i];
    }
}

/* Stage one ROWS x 32 operand tile (rows r0.., k0..) into LDS as [row][k]:
 * thread t < ROWS * 4 takes 8 elements consecutive along the operand's
 * unit-stride dimension (k if kc, else the row), as one vector load when they
 * are in range and 16-byte aligned (vec: the offsets are multiples of 8 BF16
 * / 4 FP32 elements), and stores them k-contiguous. */
template <bool F32, int ROWS>
static __device__ __forceinline__ void stage_tile(const void *p, const int64_t o0, const int64_t s_row,
                                                  const int64_t s_k, const bool kc, const bool vec, const int r0,
                                                  const int nrow, const int k0, const int K, bf16 *S) {
    const int e0 = threadIdx.x * 8;
    if (e0 >= ROWS * FG_BK) return;
    const int r = kc ? (e0 >> 5) : (e0 & (ROWS - 1)), k = kc ? (e0 & 31) : (e0 / ROWS);
    const int gr = r0 + r, gk = k0 + k;
    float v[8];
    const bool inb = kc ? (gr < nrow && gk + 8 <= K) : (gk < K && gr + 8 <= nrow);
    if (vec && (kc || s_row == 1) && inb) {
        ld8<F32>(p, o0 + (int64_t)gr * s_row + (int64_t)gk * s_k, v);
    } else {
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int rr = kc ? gr : gr + j, kk = kc ? gk + j : gk;
            v[j] = (rr < nrow && kk < K) ? ld_op<F32>(p, o0 + (int64_t)rr * s_row + (int64_t)kk * s_k) : 0.0f;
        }
    }
    if (kc) {
        bf16x8 b;
#pragma unroll
        for (int j = 0; j < 8; j++) b[j] = (bf16)v[j];
        *(bf16x8 *)&S[r * FG_LD + k] = b;
    } else {
#pragma unroll
        for (int j = 0; j < 8; j++) S[(r + j) * FG_LD + k] = (bf16)v[j];
    }
}

/* Wave layout of a BM x BN tile over the 4 waves of a workgroup: WGM x WGN
 * waves of (BM / WGM) x (BN / WGN), each MI x NI MFMA 16 x 16 tiles. */
template <int BM, int BN>
struct TileShape {
    static constexpr int WGM = (BM == 64 && BN == 32) ? 4 : ((BM == 32 && BN == 64) ? 1 : 2);
    static constexpr int WGN = 4 / WGM;
    static constexpr int TM = BM / WGM, TN = BN / WGN;
    static constexpr int MI = TM / 16, NI = TN / 16;
};

template <bool AF, bool BF, int BM, int BN>
static __device__ __forceinline__ void gemm_seg(const GemmSeg &g, const int M, const int N, const int z1,
                                                const int z2, const int m0, const int n0, bf16 *As, bf16 *Bs,
                                                f32x4 (&acc)[TileShape<BM, BN>::MI][TileShape<BM, BN>::NI]) {
    using T = TileShape<BM, BN>;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int64_t oA = (int64_t)z1 * g.sA1 + (int64_t)z2 * g.sA2;
    const int64_t oB = (int64_t)z1 * g.sB1 + (int64_t)z2 * g.sB2;
    const bool akc = g.sAk == 1, bkc = g.sBk == 1;
    const int va = AF ? 4 : 8, vb = BF ? 4 : 8;
    const bool veca = oA % va == 0 && (akc ? g.sAm : g.sAk) % va == 0;
    const bool vecb = oB % vb == 0 && (bkc ? g.sBn : g.sBk) % vb == 0;
    const int wm = (w % T::WGM) * T::TM, wn = (w / T::WGM) * T::TN;
    for (int k0 = 0; k0 < g.K; k0 += FG_BK) {
        stage_tile<AF, BM>(g.A, oA, g.sAm, g.sAk, akc, veca, m0, M, k0, g.K, As);
        stage_tile<BF, BN>(g.B, oB, g.sBn, g.sBk, bkc, vecb, n0, N, k0, g.K, Bs);
        __syncthreads();
        bf16x8 a[T::MI], b[T::NI];
#pragma unroll
        for (int i = 0; i < T::MI; i++)
            a[i] = *(const bf16x8 *)&As[(wm + i * 16 + (lane & 15)) * FG_LD + (lane >> 4) * 8];
#pragma unroll
        for (int j = 0; j < T::NI; j++)
            b[j] = *(const bf16x8 *)&Bs[(wn + j * 16 + (lane & 15)) * FG_LD + (lane >> 4) * 8];
#pragma unroll
        for (int i = 0; i < T::MI; i++)
#pragma unroll
            for (int j = 0; j < T::NI; j++)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        __syncthreads();
    }
}

/* C = epilogue(alpha A B), one BM x BN tile of C (64 or 32 each) per
 * 256-thread workgroup, the 4 waves laid out by TileShape (64 x 64: 2 x 2
 * waves of 32 x 32, i.e. 2 x 2 MFMA 16x16x32 tiles), K in steps of 32 through
 * LDS.  Tiles are staged with the contiguous dimension across threads
 * (k if the operand's k-stride is 1, else m or n), converted to BF16, and
 * stored k-contiguous so each lane reads its 8-element fragment as one
 * 16-byte LDS load: A fragment rows = lane % 16, k = 8 (lane / 16) .. +7;
 * B fragment columns = lane % 16, same k; result rows 4 (lane / 16) .. +3,
 * column lane % 16.  SEG2: a second K segment (BF16 operands, s2) adds into
 * the same accumulators before the epilogue; bias_m: the bias is indexed by
 * the row instead of the column. */
template <bool AF, bool BF, bool SEG2, int BM, int BN>
__global__ __launch_bounds__(256) void k_fno_gemm(fno_gemm_args p, GemmSeg s2, int bias_m) {
    using T = TileShape<BM, BN>;
    __shared__ __attribute__((aligned(16))) bf16 As[BM * FG_LD];
    __shared__ __attribute__((aligned(16))) bf16 Bs[BN * FG_LD];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int z = blockIdx.z, z1 = z / p.nz2, z2 = z - z1 * p.nz2;
    const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BN;
    const int wm = (w % T::WGM) * T::TM, wn = (w / T::WGM) * T::TN;
    f32x4 acc[T::MI][T::NI];
#pragma unroll
    for (int i = 0; i < T::MI; i++)
#pragma unroll
        for (int j = 0; j < T::NI; j++) acc[i][j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    const GemmSeg s1 = {p.A, p.sAm, p.sAk, p.sA1, p.sA2, p.B, p.sBk, p.sBn, p.sB1, p.sB2, p.K};
    gemm_seg<AF, BF, BM, BN>(s1, p.M, p.N, z1, z2, m0, n0, As, Bs, acc);
    if (SEG2) gemm_seg<false, false, BM, BN>(s2, p.M, p.N, z1, z2, m0, n0, As, Bs, acc);
    const int64_t oC = (int64_t)z1 * p.sC1 + (int64_t)z2 * p.sC2;
#pragma unroll
    for (int i = 0; i < T::MI; i++)
#pragma unroll
        for (int j = 0; j < T::NI; j++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int m = m0 + wm + i * 16 + (lane >> 4) * 4 + r, n = n0 + wn + j * 16 + (lane & 15);
                if (m >= p.M || n >= p.N) continue;
                const int64_t o = oC + (int64_t)m * p.sCm + (int64_t)n * p.sCn;
                float val = p.alpha * acc[i][j][r];
                if (p.c_mode == 1) val += ((const float *)p.C)[o];
                if (p.bias) val += p.bias[bias_m ? m : n];
                const float av = p.act ? gelu(val) : val;
                if (p.c_mode == 2) ((bf16 *)p.C)[o] = (bf16)av;
                else ((float *)p.C)[o] = val;
                if (p.C2) ((bf16 *)p.C2)[o] = (bf16)av;
            }
}

/* Strided 5-D copy with conversion: out[i0..i4] (+)= in[i0..i4] (i4 fastest). */
struct Copy5 {
    const void *in;
    void *out;
    int d[5];
    int64_t si[5], so[5];
    int in_f32, out_bf16, accumulate;
};
__global__ __launch_bounds__(256) void k_copy5(Copy5 c, int64_t total) {
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < total; q += (int64_t)gridDim.x * blockDim.x) {
        int64_t r = q, oi = 0, oo = 0;
#pragma unroll
        for (int a = 4; a >= 0; a--) {
            const int64_t x = r % c.d[a];
            r /= c.d[a];
            oi += x * c.si[a];
            oo += x * c.so[a];
        }
        const float v = c.in_f32 ? ((const float *)c.in)[oi] : (float)((const bf16 *)c.in)[oi];
        if (c.out_bf16) ((bf16 *)c.out)[oo] = (bf16)v;
        else if (c.accumulate) ((float *)c.out)[oo] += v;
        else ((float *)c.out)[oo] = v;
    }
}

__global__ void k_cvt_bf16(const float *__restrict__ x, bf16 *__restrict__ y, int64_t n) {
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
        y[q] = (bf16)x[q];
}

/* Rr, Ri [nm][w][w] -> the real 2w x 2w matrix of each mode's complex mix,
 * rows (i, re|im), columns (o, re|im):  [Rr Ri; -Ri Rr] per (i, o). */
__global__ void k_expand_R(const float *__restrict__ Rr, const float *__restrict__ Ri, bf16 *__restrict__ Wx,
                           int nm, int w) {
    const int64_t n = (int64_t)nm * w * w;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        const int mode = (int)(q / (w * w)), i = (int)(q / w % w), o = (int)(q % w);
        const float rr = Rr[q], ri = Ri[q];
        bf16 *W = Wx + (int64_t)mode * 4 * w * w;
        const int W2 = 2 * w;
        W[(2 * i) * W2 + 2 * o] = (bf16)rr;
        W[(2 * i + 1) * W2 + 2 * o] = (bf16)(-ri);
        W[(2 * i) * W2 + 2 * o + 1] = (bf16)ri;
        W[(2 * i + 1) * W2 + 2 * o + 1] = (bf16)rr;
    }
}

/* Gradient of the expanded matrices back onto (Rr, Ri), accumulated. */
__global__ void k_fold_dR(const float *__restrict__ G, float *__restrict__ dRr, float *__restrict__ dRi, int nm,
                          int w) {
    const int64_t n = (int64_t)nm * w * w;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        const int mode = (int)(q / (w * w)), i = (int)(q / w % w), o = (int)(q % w);
        const float *g = G + (int64_t)mode * 4 * w * w;
        const int W2 = 2 * w;
        dRr[q] += g[(2 * i) * W2 + 2 * o] + g[(2 * i + 1) * W2 + 2 * o + 1];
        dRi[q] += g[(2 * i) * W2 + 2 * o + 1] - g[(2 * i + 1) * W2 + 2 * o];
    }
}

/* du = dv * GELU'(u) (in place allowed). */
__global__ void k_gelu_bwd(const float *dv, const float *__restrict__ u, float *du, int64_t n) {
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
        du[q] = dv[q] * gelu_d(u[q]);
}

/* g[o] += sum over b, p of X[b][o][p]  (one workgroup per o, FP64 sum in a
 * fixed order). */
__global__ __launch_bounds__(256) void k_rowsum(const float *__restrict__ X, float *__restrict__ g, int B, int O,
                                                int64_t P) {
    __shared__ double red[256];
    const int o = blockIdx.x;
    double acc = 0.0;
    for (int b = 0; b < B; b++) {
        const float *x = X + ((int64_t)b * O + o) * P;
        for (int64_t p = threadIdx.x; p < P; p += 256) acc += (double)x[p];
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) g[o] += (float)red[0];
}

/* g[i] += sum over z of part[z][i]. */
__global__ void k_zsum(const float *__restrict__ part, float *__restrict__ g, int nz, int64_t n) {
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        float s = 0.0f;
        for (int z = 0; z < nz; z++) s += part[(int64_t)z * n + q];
        g[q] += s;
    }
}

/* The lift v0[b][c][p] = sum_i P[c][i] in_i + bP[c] with the input channels
 * in = (phi s0, rho s1, i / M, j / M) from the crops [B][2][P]; FP32
 * arithmetic, v0 stored BF16.  xin (training: the BF16 input the weight
 * gradient of P needs) may be null. */
__global__ __launch_bounds__(256) void k_lift(const float *__restrict__ in, const float *__restrict__ Pw,
                                              const float *__restrict__ bP, bf16 *__restrict__ v0,
                                              bf16 *__restrict__ xin, int B, int M, int w, float s0, float s1) {
    const int64_t P = (int64_t)M * M, n = (int64_t)B * P;
    const float inv = 1.0f / (float)M;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        const int64_t b = q / P, p = q % P;
        const float *ib = in + b * 2 * P;
        const float x0 = ib[p] * s0, x1 = ib[P + p] * s1, x2 = (float)(p / M) * inv, x3 = (float)(p % M) * inv;
        if (xin) {
            bf16 *xb = xin + b * 4 * P;
            xb[p] = (bf16)x0; xb[P + p] = (bf16)x1; xb[2 * P + p] = (bf16)x2; xb[3 * P + p] = (bf16)x3;
        }
        bf16 *vb = v0 + b * w * P + p;
        for (int c = 0; c < w; c++)
            vb[(int64_t)c * P] = (bf16)(Pw[4 * c] * x0 + Pw[4 * c + 1] * x1 + Pw[4 * c + 2] * x2 + Pw[4 * c + 3] * x3 + bP[c]);
    }
}

/* Inference projection for width 32 and 128 hidden channels:
 * out[b][o][p] = scale_o (sum_j Q2[o][j] GELU(sum_c Q1[j][c] v[b][c][p] + c1[j]) + c2[o]).
 * One 64-pixel tile per 256-thread workgroup: the 32-channel tile staged in
 * LDS as [pixel][channel], Q1 v on the matrix cores (one MFMA 16x16x32 per
 * 16 hidden channels, wave w on pixels 16w..16w+15), GELU and the 128 -> 2
 * contraction in registers, the sum over the 16 lanes holding one pixel's
 * hidden channels by lane shuffles.  Nothing of the hidden layer reaches
 * memory. */
__global__ __launch_bounds__(256) void k_proj_fused(const bf16 *__restrict__ v, const bf16 *__restrict__ Q1,
                                                    const float *__restrict__ c1, const float *__restrict__ Q2,
                                                    const float *__restrict__ c2, float *__restrict__ out,
                                                    int64_t P, float s0, float s1) {
    constexpr int W = 32, LD = W + 8;
    __shared__ __attribute__((aligned(16))) bf16 As[64 * LD];
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int64_t b = blockIdx.y, p0 = (int64_t)blockIdx.x * 64;
    const bf16 *vb = v + b * W * P;
#pragma unroll
    for (int j = 0; j < 8; j++) {
        const int e = t * 8 + j, c = e >> 6, px = e & 63;
        As[px * LD + c] = (p0 + px < P) ? vb[(int64_t)c * P + p0 + px] : (bf16)0.0f;
    }
    __syncthreads();
    const bf16x8 a = *(const bf16x8 *)&As[(wv * 16 + (lane & 15)) * LD + (lane >> 4) * 8];
    float o0[4] = {0.0f, 0.0f, 0.0f, 0.0f}, o1[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int jb = 0; jb < 8; jb++) {
        const int j = jb * 16 + (lane & 15);
        const bf16x8 bq = *(const bf16x8 *)&Q1[j * W + (lane >> 4) * 8];
        const f32x4 acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bq, f32x4{0.0f, 0.0f, 0.0f, 0.0f}, 0, 0, 0);
        const float cj = c1[j], q0 = Q2[j], q1 = Q2[128 + j];
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const float h = gelu(acc[r] + cj);
            o0[r] += q0 * h;
            o1[r] += q1 * h;
        }
    }
#pragma unroll
    for (int r = 0; r < 4; r++)
#pragma unroll
        for (int m = 1; m < 16; m <<= 1) {
            o0[r] += __shfl_xor(o0[r], m, 64);
            o1[r] += __shfl_xor(o1[r], m, 64);
        }
    if ((lane & 15) == 0) {
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int64_t p = p0 + wv * 16 + (lane >> 4) * 4 + r;
            if (p < P) {
                out[(b * 2) * P + p] = (o0[r] + c2[0]) * s0;
                out[(b * 2 + 1) * P + p] = (o1[r] + c2[1]) * s1;
            }
        }
    }
}

/* Loss terms: d = out - target / out_scale; dout = 2 d / count; per-block
 * FP64 partial sums of d^2. */
__global__ __launch_bounds__(256) void k_loss(const float *__restrict__ out, const float *__restrict__ tgt,
                                              float *__restrict__ dout, double *__restrict__ part, int B,
                                              int64_t P, float is0, float is1) {
    __shared__ double red[256];
    const int64_t n = (int64_t)B * 2 * P;
    const float sc = 2.0f / (float)n;
    double acc = 0.0;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        const int c = (int)(q / P % 2);
        const float d = out[q] - tgt[q] * (c ? is1 : is0);
        dout[q] = sc * d;
        acc += (double)d * (double)d;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

/* out[b][c][p] *= scale[c] */
__global__ void k_scale_out(float *__restrict__ out, int B, int64_t P, float s0, float s1) {
    const int64_t n = (int64_t)B * 2 * P;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
        out[q] *= (q / P % 2) ? s1 : s0;
}

__global__ void k_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, int64_t n, float lr, float b1, float b2, float eps, float c1,
                       float c2) {
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        const float gq = g[q];
        const float mq = b1 * m[q] + (1.0f - b1) * gq, vq = b2 * v[q] + (1.0f - b2) * gq * gq;
        m[q] = mq;
        v[q] = vq;
        p[q] -= lr * (mq / c1) / (sqrtf(vq / c2) + eps);
    }
}

/* ======================================================================== */
struct fno_net {
    fno_config cfg{};
    int dev = 0;
    hipStream_t st = nullptr;
    int64_t P = 0;
    int nm = 0, L = 0, w = 0, m = 0, M = 0, proj = 0, Bmax = 0;
    size_t npar = 0;
    float in_scale[2] = {1.0f, 1.0f}, out_scale[2] = {1.0f, 1.0f};
    /* parameters: offsets into the flat array */
    size_t oP = 0, obP = 0, oQ1 = 0, oc1 = 0, oQ2 = 0, oc2 = 0;
    std::vector<size_t> oRr, oRi, oW, ob;
    float *par = nullptr, *grad = nullptr, *am = nullptr, *av = nullptr;
    /* BF16 copies of the weights */
    bf16 *wQ1 = nullptr, *wQ2 = nullptr;
    std::vector<bf16 *> wW, wX;
    /* transform tables (BF16) */
    bf16 *Fy = nullptr, *T = nullptr, *Ti = nullptr, *Fi = nullptr;
    /* saved activations */
    bf16 *xin = nullptr, *q = nullptr;
    std::vector<bf16 *> v, xh;
    std::vector<float *> u;
    float *qpre = nullptr, *out = nullptr;
    /* workspace */
    float *Z = nullptr, *Zh = nullptr, *Yh = nullptr, *Zp = nullptr;
    bf16 *Yh2 = nullptr;
    float *dout = nullptr, *dq = nullptr, *dv0 = nullptr, *dv1 = nullptr, *dXh = nullptr, *gexp = nullptr,
          *part = nullptr;
    bf16 *dYh = nullptr, *dZh = nullptr;
    double *lpart = nullptr, *lpart_h = nullptr;
    std::vector<void *> allocs;
    bool weights_fresh = false;
    int tile[FNO_STAGE_COUNT] = {};   /* GEMM tile variant per stage (FNO_TILE_*) */
};

namespace {

constexpr int KC = 2048;     /* K chunk of the weight-gradient GEMMs */
constexpr int LOSS_BLOCKS = 1024;

int halloc(fno_net *n, void **p, size_t bytes) {
    if (hipMalloc(p, bytes ? bytes : 1) != hipSuccess) return ORG_EHIP;
    n->allocs.push_back(*p);
    return ORG_OK;
}
template <typename T>
int dal(fno_net *n, T **p, size_t count) {
    return halloc(n, (void **)p, count * sizeof(T));
}

int grid_for(int64_t n) { return (int)std::min<int64_t>((n + 255) / 256, 8192); }

int launch_ok() { return hipGetLastError() == hipSuccess ? ORG_OK : ORG_EHIP; }

/* One GEMM launch with tile variant t (FNO_TILE_*). */
template <bool AF, bool BF, bool SEG2>
void launch_tiled(fno_net *n, int t, const fno_gemm_args &a, const GemmSeg &s2, int bias_m, int64_t nz) {
    const int bm = (t == FNO_TILE_32x64 || t == FNO_TILE_32x32) ? 32 : 64;
    const int bn = (t == FNO_TILE_64x32 || t == FNO_TILE_32x32) ? 32 : 64;
    dim3 g((a.M + bm - 1) / bm, (a.N + bn - 1) / bn, (unsigned)nz);
    switch (t) {
    case FNO_TILE_32x64: k_fno_gemm<AF, BF, SEG2, 32, 64><<<g, dim3(256), 0, n->st>>>(a, s2, bias_m); break;
    case FNO_TILE_64x32: k_fno_gemm<AF, BF, SEG2, 64, 32><<<g, dim3(256), 0, n->st>>>(a, s2, bias_m); break;
    case FNO_TILE_32x32: k_fno_gemm<AF, BF, SEG2, 32, 32><<<g, dim3(256), 0, n->st>>>(a, s2, bias_m); break;
    default: k_fno_gemm<AF, BF, SEG2, 64, 64><<<g, dim3(256), 0, n->st>>>(a, s2, bias_m); break;
    }
}

int gemm(fno_net *n, const fno_gemm_args &a, int stage = FNO_STAGE_API) {
    if (a.M <= 0 || a.N <= 0 || a.K <= 0 || a.nz1 <= 0 || a.nz2 <= 0 || !a.A || !a.B || !a.C) return ORG_EINVAL;
    if (a.c_mode < 0 || a.c_mode > 2) return ORG_EINVAL;
    const int64_t nz = (int64_t)a.nz1 * a.nz2;
    if (nz > 65535) return ORG_EINVAL;
    const GemmSeg none{};
    const int t = n->tile[stage];
    if (a.a_f32 && a.b_f32) launch_tiled<true, true, false>(n, t, a, none, 0, nz);
    else if (a.a_f32) launch_tiled<true, false, false>(n, t, a, none, 0, nz);
    else if (a.b_f32) launch_tiled<false, true, false>(n, t, a, none, 0, nz);
    else launch_tiled<false, false, false>(n, t, a, none, 0, nz);
    return launch_ok();
}

/* A GEMM with a second K segment (FP32 A, BF16 B in the first; BF16 both in
 * the second) and the bias per row. */
int gemm2(fno_net *n, const fno_gemm_args &a, const GemmSeg &s2, int stage) {
    if (a.M <= 0 || a.N <= 0 || a.K <= 0 || s2.K <= 0 || a.nz1 <= 0 || a.nz2 <= 0 || !a.A || !a.B || !a.C ||
        !s2.A || !s2.B || !a.a_f32 || a.b_f32 || a.c_mode == 1)
        return ORG_EINVAL;
    const int64_t nz = (int64_t)a.nz1 * a.nz2;
    if (nz > 65535) return ORG_EINVAL;
    launch_tiled<true, false, true>(n, n->tile[stage], a, s2, 1, nz);
    return launch_ok();
}

fno_gemm_args G0() {
    fno_gemm_args a;
    std::memset(&a, 0, sizeof(a));
    a.nz1 = a.nz2 = 1;
    a.alpha = 1.0f;
    return a;
}

int copy5(fno_net *n, const void *in, int in_f32, void *out, int out_bf16, int accumulate, const int d[5],
          const int64_t si[5], const int64_t so[5]) {
    Copy5 c;
    c.in = in; c.out = out; c.in_f32 = in_f32; c.out_bf16 = out_bf16; c.accumulate = accumulate;
    int64_t tot = 1;
    for (int a = 0; a < 5; a++) { c.d[a] = d[a]; c.si[a] = si[a]; c.so[a] = so[a]; tot *= d[a]; }
    hipLaunchKernelGGL(k_copy5, dim3(grid_for(tot)), dim3(256), 0, n->st, c, tot);
    return launch_ok();
}

/* [B][w][2m][2][m] (kx_idx, re|im, ky) <-> [nm][B][w][2] (mode = kx_idx*m + ky) */
int to_modes(fno_net *n, int B, const void *in, int in_f32, void *out, int out_bf16) {
    const int w = n->w, m = n->m;
    const int d[5] = {B, w, 2 * m, 2, m};
    const int64_t si[5] = {(int64_t)w * 4 * m * m, 4 * m * m, 2 * m, m, 1};
    const int64_t so[5] = {2 * w, 2, (int64_t)m * B * 2 * w, 1, (int64_t)B * 2 * w};
    return copy5(n, in, in_f32, out, out_bf16, 0, d, si, so);
}
int from_modes(fno_net *n, int B, const void *in, int in_f32, void *out, int out_bf16) {
    const int w = n->w, m = n->m;
    const int d[5] = {B, w, 2 * m, 2, m};
    const int64_t si[5] = {2 * w, 2, (int64_t)m * B * 2 * w, 1, (int64_t)B * 2 * w};
    const int64_t so[5] = {(int64_t)w * 4 * m * m, 4 * m * m, 2 * m, m, 1};
    return copy5(n, in, in_f32, out, out_bf16, 0, d, si, so);
}

#define TRY(x) do { int rc_ = (x); if (rc_) return rc_; } while (0)

/* BF16 copies of the weights and the expanded mode matrices. */
int prep_weights(fno_net *n) {
    auto cvt = [&](size_t off, size_t cnt, bf16 *dst) {
        hipLaunchKernelGGL(k_cvt_bf16, dim3(grid_for(cnt)), dim3(256), 0, n->st, n->par + off, dst, (int64_t)cnt);
        return launch_ok();
    };
    TRY(cvt(n->oQ1, (size_t)n->proj * n->w, n->wQ1));
    TRY(cvt(n->oQ2, (size_t)2 * n->proj, n->wQ2));
    for (int l = 0; l < n->L; l++) {
        TRY(cvt(n->oW[l], (size_t)n->w * n->w, n->wW[l]));
        const int64_t cnt = (int64_t)n->nm * n->w * n->w;
        hipLaunchKernelGGL(k_expand_R, dim3(grid_for(cnt)), dim3(256), 0, n->st, n->par + n->oRr[l],
                           n->par + n->oRi[l], n->wX[l], n->nm, n->w);
        TRY(launch_ok());
    }
    return ORG_OK;
}

/* Pointwise GEMM over the pixels of each batch member:
 * C[b][n][p] (=, +=) sum_k A[b][k][p] Wt[k][n]   (A: [B][K][P]; Wt with
 * strides (sk, sn)). */
fno_gemm_args pointwise(fno_net *n, int B, const void *A, int a_f32, int K, const void *Wt, int64_t sk,
                        int64_t sn, void *C, int N) {
    fno_gemm_args a = G0();
    a.A = A; a.sAm = 1; a.sAk = n->P; a.sA1 = (int64_t)K * n->P; a.a_f32 = a_f32;
    a.B = Wt; a.sBk = sk; a.sBn = sn;
    a.C = C; a.sCm = 1; a.sCn = n->P; a.sC1 = (int64_t)N * n->P;
    a.M = (int)n->P; a.N = N; a.K = K; a.nz1 = B;
    return a;
}

/* Weight gradient g[mo][no] += sum_{b,p} X[b][mo][p] Y[b][no][p]
 * (X: [B][Mo][P], Y: [B][No][P]) through K-chunked partials. */
int wgrad(fno_net *n, int B, const void *X, int x_f32, int Mo, const void *Y, int y_f32, int No, float *g) {
    const int nch = (int)((n->P + KC - 1) / KC);
    fno_gemm_args a = G0();
    a.A = X; a.sAm = n->P; a.sAk = 1; a.sA1 = (int64_t)Mo * n->P; a.sA2 = KC; a.a_f32 = x_f32;
    a.B = Y; a.sBk = 1; a.sBn = n->P; a.sB1 = (int64_t)No * n->P; a.sB2 = KC; a.b_f32 = y_f32;
    a.C = n->part; a.sCm = No; a.sCn = 1; a.sC1 = (int64_t)nch * Mo * No; a.sC2 = (int64_t)Mo * No;
    a.M = Mo; a.N = No; a.K = (int)std::min<int64_t>(KC, n->P); a.nz1 = B; a.nz2 = nch;
    if (n->P % KC != 0 && n->P > KC) return ORG_EINVAL;
    TRY(gemm(n, a, FNO_STAGE_WGRAD));
    const int64_t cnt = (int64_t)Mo * No;
    hipLaunchKernelGGL(k_zsum, dim3(grid_for(cnt)), dim3(256), 0, n->st, n->part, g, B * nch, cnt);
    return launch_ok();
}

int rowsum(fno_net *n, int B, const float *X, int O, float *g) {
    hipLaunchKernelGGL(k_rowsum, dim3(O), dim3(256), 0, n->st, X, g, B, O, n->P);
    return launch_ok();
}

/* Spectral part of layer l: u[l] = K_l(v[l]) (stored, overwriting). */
int spectral_fwd(fno_net *n, int B, int l) {
    const int w = n->w, m = n->m, M = n->M;
    const int64_t rows = (int64_t)B * w * M;
    if (rows > INT32_MAX) return ORG_EINVAL;
    fno_gemm_args a = G0();
    /* y: Z[row][(ri,ky)] = sum_y v[row][y] Fy[(ri,ky)][y] */
    a.A = n->v[l]; a.sAm = M; a.sAk = 1;
    a.B = n->Fy; a.sBk = 1; a.sBn = M;
    a.C = n->Z; a.sCm = 2 * m; a.sCn = 1;
    a.M = (int)rows; a.N = 2 * m; a.K = M;
    TRY(gemm(n, a, FNO_STAGE_Y));
    /* x: Zh[b,c][(kx,ri')][ky] = sum_(x,ri) T[(kx,ri')][(x,ri)] Z[b,c][(x,ri)][ky] */
    a = G0();
    a.A = n->T; a.sAm = 2 * M; a.sAk = 1;
    a.B = n->Z; a.sBk = m; a.sBn = 1; a.sB1 = (int64_t)M * 2 * m; a.b_f32 = 1;
    a.C = n->Zh; a.sCm = m; a.sCn = 1; a.sC1 = (int64_t)4 * m * m;
    a.M = 4 * m; a.N = m; a.K = 2 * M; a.nz1 = B * w;
    TRY(gemm(n, a, FNO_STAGE_X));
    TRY(to_modes(n, B, n->Zh, 1, n->xh[l], 1));
    /* per mode: Yh[mode][b][(o,ri)] = sum_(c,ri) Xh[mode][b][(c,ri)] Wx[mode][(c,ri)][(o,ri)] */
    a = G0();
    a.A = n->xh[l]; a.sAm = 2 * w; a.sAk = 1; a.sA1 = (int64_t)B * 2 * w;
    a.B = n->wX[l]; a.sBk = 2 * w; a.sBn = 1; a.sB1 = (int64_t)4 * w * w;
    a.C = n->Yh; a.sCm = 2 * w; a.sCn = 1; a.sC1 = (int64_t)B * 2 * w;
    a.M = B; a.N = 2 * w; a.K = 2 * w; a.nz1 = n->nm;
    TRY(gemm(n, a, FNO_STAGE_MIX));
    TRY(from_modes(n, B, n->Yh, 1, n->Yh2, 1));
    /* inverse x: Zp[b,o][(x,ri)][ky] = sum_(kx,ri') Ti[(x,ri)][(kx,ri')] Yh2[b,o][(kx,ri')][ky] */
    a = G0();
    a.A = n->Ti; a.sAm = 4 * m; a.sAk = 1;
    a.B = n->Yh2; a.sBk = m; a.sBn = 1; a.sB1 = (int64_t)4 * m * m;
    a.C = n->Zp; a.sCm = m; a.sCn = 1; a.sC1 = (int64_t)M * 2 * m;
    a.M = 2 * M; a.N = m; a.K = 4 * m; a.nz1 = B * w;
    return gemm(n, a, FNO_STAGE_IX);
}

/* The rest of layer l, one GEMM per (b, x) row with two K segments:
 *   u[b][o][x][y] = sum_(ri,ky) Zp[b][o][x][(ri,ky)] Fi[(ri,ky)][y]      (inverse y)
 *                 + sum_c W[o][c] v[b][c][x][y] + b[o]                  (pointwise)
 * v[l+1] = GELU(u) (identity after the last layer); train: u kept in FP32. */
int layer_out(fno_net *n, int B, int l, bool train) {
    const int w = n->w, m = n->m, M = n->M;
    const int64_t P = n->P;
    fno_gemm_args a = G0();
    a.A = n->Zp; a.sAm = (int64_t)M * 2 * m; a.sAk = 1; a.sA1 = (int64_t)w * M * 2 * m; a.sA2 = 2 * m; a.a_f32 = 1;
    a.B = n->Fi; a.sBk = M; a.sBn = 1;
    a.M = w; a.N = M; a.K = 2 * m; a.nz1 = B; a.nz2 = M;
    a.sCm = P; a.sCn = 1; a.sC1 = (int64_t)w * P; a.sC2 = M;
    a.bias = n->par + n->ob[l];
    a.act = (l < n->L - 1);
    if (train) { a.C = n->u[l]; a.c_mode = 0; a.C2 = n->v[l + 1]; }
    else { a.C = n->v[l + 1]; a.c_mode = 2; }
    GemmSeg s2{};
    s2.A = n->wW[l]; s2.sAm = w; s2.sAk = 1;
    s2.B = n->v[l]; s2.sBk = P; s2.sBn = 1; s2.sB1 = (int64_t)w * P; s2.sB2 = M;
    s2.K = w;
    return gemm2(n, a, s2, FNO_STAGE_LAYER);
}

/* Forward pass.  train: keep what backward needs (the BF16 input, the FP32
 * projection pre-activations) and leave the unscaled output in n->out;
 * otherwise write the scaled velocity to vout (fused projection when the
 * network has width 32 and 128 hidden channels). */
int forward(fno_net *n, const float *in, int B, bool train, float *vout) {
    if (!n->weights_fresh) {
        TRY(prep_weights(n));
        n->weights_fresh = true;
    }
    const int w = n->w;
    hipLaunchKernelGGL(k_lift, dim3(grid_for((int64_t)B * n->P)), dim3(256), 0, n->st, in, n->par + n->oP,
                       n->par + n->obP, n->v[0], train ? n->xin : (bf16 *)nullptr, B, n->M, w, n->in_scale[0],
                       n->in_scale[1]);
    TRY(launch_ok());
    fno_gemm_args a;
    for (int l = 0; l < n->L; l++) {
        TRY(spectral_fwd(n, B, l));
        TRY(layer_out(n, B, l, train));
    }
    /* projection */
    if (!train && w == 32 && n->proj == 128) {
        hipLaunchKernelGGL(k_proj_fused, dim3((unsigned)((n->P + 63) / 64), B), dim3(256), 0, n->st, n->v[n->L],
                           n->wQ1, n->par + n->oc1, n->par + n->oQ2, n->par + n->oc2, vout, n->P, n->out_scale[0],
                           n->out_scale[1]);
        return launch_ok();
    }
    a = pointwise(n, B, n->v[n->L], 0, w, n->wQ1, 1, w, n->qpre, n->proj);
    a.bias = n->par + n->oc1; a.C2 = n->q; a.act = 1;
    TRY(gemm(n, a, FNO_STAGE_PROJ));
    a = pointwise(n, B, n->q, 0, n->proj, n->wQ2, 1, n->proj, train ? n->out : vout, 2);
    a.bias = n->par + n->oc2;
    TRY(gemm(n, a, FNO_STAGE_PROJ));
    if (!train) {
        hipLaunchKernelGGL(k_scale_out, dim3(grid_for((int64_t)B * 2 * n->P)), dim3(256), 0, n->st, vout, B, n->P,
                           n->out_scale[0], n->out_scale[1]);
        return launch_ok();
    }
    return ORG_OK;
}

/* Backward of the spectral part of layer l: du (f32 [B][w][P]) -> gradient
 * of (Rr, Ri) accumulated, dv += K_l^T du. */
int spectral_bwd(fno_net *n, int B, int l, const float *du, float *dv) {
    const int w = n->w, m = n->m, M = n->M;
    const int64_t rows = (int64_t)B * w * M;
    fno_gemm_args a = G0();
    /* dZp[row][(ri,ky)] = sum_y du[row][y] Fi[(ri,ky)][y] */
    a.A = du; a.sAm = M; a.sAk = 1; a.a_f32 = 1;
    a.B = n->Fi; a.sBk = 1; a.sBn = M;
    a.C = n->Zp; a.sCm = 2 * m; a.sCn = 1;
    a.M = (int)rows; a.N = 2 * m; a.K = M;
    TRY(gemm(n, a, FNO_STAGE_BWD_SPECTRAL));
    /* dYh2[b,o][(kx,ri')][ky] = sum_(x,ri) Ti[(x,ri)][(kx,ri')] dZp[b,o][(x,ri)][ky] */
    a = G0();
    a.A = n->Ti; a.sAm = 1; a.sAk = 4 * m;
    a.B = n->Zp; a.sBk = m; a.sBn = 1; a.sB1 = (int64_t)M * 2 * m; a.b_f32 = 1;
    a.C = n->Zh; a.sCm = m; a.sCn = 1; a.sC1 = (int64_t)4 * m * m;
    a.M = 4 * m; a.N = m; a.K = 2 * M; a.nz1 = B * w;
    TRY(gemm(n, a, FNO_STAGE_BWD_SPECTRAL));
    TRY(to_modes(n, B, n->Zh, 1, n->dYh, 1));
    /* dWx[mode] = Xh[mode]^T dYh[mode] */
    a = G0();
    a.A = n->xh[l]; a.sAm = 1; a.sAk = 2 * w; a.sA1 = (int64_t)B * 2 * w;
    a.B = n->dYh; a.sBk = 2 * w; a.sBn = 1; a.sB1 = (int64_t)B * 2 * w;
    a.C = n->gexp; a.sCm = 2 * w; a.sCn = 1; a.sC1 = (int64_t)4 * w * w;
    a.M = 2 * w; a.N = 2 * w; a.K = B; a.nz1 = n->nm;
    TRY(gemm(n, a, FNO_STAGE_BWD_SPECTRAL));
    const int64_t cnt = (int64_t)n->nm * w * w;
    hipLaunchKernelGGL(k_fold_dR, dim3(grid_for(cnt)), dim3(256), 0, n->st, n->gexp, n->grad + n->oRr[l],
                       n->grad + n->oRi[l], n->nm, w);
    TRY(launch_ok());
    /* dXh[mode] = dYh[mode] Wx[mode]^T */
    a = G0();
    a.A = n->dYh; a.sAm = 2 * w; a.sAk = 1; a.sA1 = (int64_t)B * 2 * w;
    a.B = n->wX[l]; a.sBk = 1; a.sBn = 2 * w; a.sB1 = (int64_t)4 * w * w;
    a.C = n->dXh; a.sCm = 2 * w; a.sCn = 1; a.sC1 = (int64_t)B * 2 * w;
    a.M = B; a.N = 2 * w; a.K = 2 * w; a.nz1 = n->nm;
    TRY(gemm(n, a, FNO_STAGE_BWD_SPECTRAL));
    TRY(from_modes(n, B, n->dXh, 1, n->dZh, 1));
    /* dZ[b,c][(x,ri)][ky] = sum_(kx,ri') T[(kx,ri')][(x,ri)] dZh[b,c][(kx,ri')][ky] */
    a = G0();
    a.A = n->T; a.sAm = 1; a.sAk = 2 * M;
    a.B = n->dZh; a.sBk = m; a.sBn = 1; a.sB1 = (int64_t)4 * m * m;
    a.C = n->Z; a.sCm = m; a.sCn = 1; a.sC1 = (int64_t)M * 2 * m;
    a.M = 2 * M; a.N = m; a.K = 4 * m; a.nz1 = B * w;
    TRY(gemm(n, a, FNO_STAGE_BWD_SPECTRAL));
    /* dv[row][y] += sum_(ri,ky) dZ[row][(ri,ky)] Fy[(ri,ky)][y] */
    a = G0();
    a.A = n->Z; a.sAm = 2 * m; a.sAk = 1; a.a_f32 = 1;
    a.B = n->Fy; a.sBk = M; a.sBn = 1;
    a.C = dv; a.sCm = M; a.sCn = 1; a.c_mode = 1;
    a.M = (int)rows; a.N = M; a.K = 2 * m;
    return gemm(n, a, FNO_STAGE_BWD_SPECTRAL);
}

int backward(fno_net *n, int B) {
    const int w = n->w, proj = n->proj;
    const int64_t P = n->P;
    if (hipMemsetAsync(n->grad, 0, n->npar * sizeof(float), n->st) != hipSuccess) return ORG_EHIP;
    /* projection Q2 */
    TRY(wgrad(n, B, n->dout, 1, 2, n->q, 0, proj, n->grad + n->oQ2));
    TRY(rowsum(n, B, n->dout, 2, n->grad + n->oc2));
    fno_gemm_args a = pointwise(n, B, n->dout, 1, 2, n->wQ2, proj, 1, n->dq, proj);
    TRY(gemm(n, a, FNO_STAGE_BWD_POINTWISE));
    const int64_t nq = (int64_t)B * proj * P;
    hipLaunchKernelGGL(k_gelu_bwd, dim3(grid_for(nq)), dim3(256), 0, n->st, n->dq, n->qpre, n->dq, nq);
    TRY(launch_ok());
    /* projection Q1 */
    TRY(wgrad(n, B, n->dq, 1, proj, n->v[n->L], 0, w, n->grad + n->oQ1));
    TRY(rowsum(n, B, n->dq, proj, n->grad + n->oc1));
    float *dv = n->dv0, *dvn = n->dv1;
    a = pointwise(n, B, n->dq, 1, proj, n->wQ1, w, 1, dv, w);
    TRY(gemm(n, a, FNO_STAGE_BWD_POINTWISE));
    const int64_t nv = (int64_t)B * w * P;
    for (int l = n->L - 1; l >= 0; l--) {
        if (l < n->L - 1) {
            hipLaunchKernelGGL(k_gelu_bwd, dim3(grid_for(nv)), dim3(256), 0, n->st, dv, n->u[l], dv, nv);
            TRY(launch_ok());
        }
        const float *du = dv;
        TRY(wgrad(n, B, du, 1, w, n->v[l], 0, w, n->grad + n->oW[l]));
        TRY(rowsum(n, B, du, w, n->grad + n->ob[l]));
        /* dv_l = W_l^T du (pointwise), then the spectral part's transpose */
        a = pointwise(n, B, du, 1, w, n->wW[l], w, 1, dvn, w);
        TRY(gemm(n, a, FNO_STAGE_BWD_POINTWISE));
        TRY(spectral_bwd(n, B, l, du, dvn));
        std::swap(dv, dvn);
    }
    /* lift */
    TRY(wgrad(n, B, dv, 1, w, n->xin, 0, 4, n->grad + n->oP));
    return rowsum(n, B, dv, w, n->grad + n->obP);
}

/* Transform tables in double, rounded to BF16 once. */
int make_tables(fno_net *n) {
    const int M = n->M, m = n->m;
    const double tw = 2.0 * M_PI / M;
    auto kx_of = [&](int idx) { return idx < m ? idx : M - 2 * m + idx; };
    std::vector<float> fy((size_t)2 * m * M), fi((size_t)2 * m * M), T((size_t)4 * m * 2 * M),
        Ti((size_t)2 * M * 4 * m);
    for (int ri = 0; ri < 2; ri++)
        for (int ky = 0; ky < m; ky++)
            for (int y = 0; y < M; y++) {
                const double th = tw * ((ky * y) % M), c = (ky == 0) ? 1.0 : 2.0;
                fy[(size_t)(ri * m + ky) * M + y] = (float)(ri ? -std::sin(th) : std::cos(th));
                fi[(size_t)(ri * m + ky) * M + y] = (float)((ri ? -std::sin(th) : std::cos(th)) * c / ((double)M * M));
            }
    for (int k = 0; k < 2 * m; k++)
        for (int x = 0; x < M; x++) {
            const double th = tw * (((int64_t)kx_of(k) * x) % M), c = std::cos(th), s = std::sin(th);
            /* forward (e^{-i th}): re' = c re + s im, im' = -s re + c im */
            T[(size_t)(2 * k) * 2 * M + 2 * x] = (float)c;
            T[(size_t)(2 * k) * 2 * M + 2 * x + 1] = (float)s;
            T[(size_t)(2 * k + 1) * 2 * M + 2 * x] = (float)(-s);
            T[(size_t)(2 * k + 1) * 2 * M + 2 * x + 1] = (float)c;
            /* inverse (e^{+i th}): re = c re' - s im', im = s re' + c im' */
            Ti[(size_t)(2 * x) * 4 * m + 2 * k] = (float)c;
            Ti[(size_t)(2 * x) * 4 * m + 2 * k + 1] = (float)(-s);
            Ti[(size_t)(2 * x + 1) * 4 * m + 2 * k] = (float)s;
            Ti[(size_t)(2 * x + 1) * 4 * m + 2 * k + 1] = (float)c;
        }
    auto up = [&](const std::vector<float> &h, bf16 **d) -> int {
        std::vector<bf16> b(h.size());
        for (size_t i = 0; i < h.size(); i++) b[i] = (bf16)h[i];
        TRY(dal(n, d, h.size()));
        return hipMemcpy(*d, b.data(), b.size() * sizeof(bf16), hipMemcpyHostToDevice) == hipSuccess ? ORG_OK
                                                                                                  : ORG_EHIP;
    };
    TRY(up(fy, &n->Fy));
    TRY(up(fi, &n->Fi));
    TRY(up(T, &n->T));
    TRY(up(Ti, &n->Ti));
    return ORG_OK;
}

int setdev(fno_net *n) { return hipSetDevice(n->dev) == hipSuccess ? ORG_OK : ORG_EHIP; }

}  // namespace

extern "C" {

int fno_create(const fno_config *cfg, int device, fno_net **out) {
    if (!cfg || !out) return ORG_EINVAL;
    const fno_config c = *cfg;
    if (c.M < 8 || c.M > 1024 || (c.M & 1) || c.width < 1 || c.width > 256 || c.modes < 1 || 2 * c.modes > c.M ||
        c.layers < 1 || c.layers > 16 || c.proj < 1 || c.proj > 1024 || c.batch_max < 1)
        return ORG_EINVAL;
    if ((int64_t)c.M * c.M > KC && ((int64_t)c.M * c.M) % KC != 0) return ORG_EINVAL;
    if ((int64_t)c.batch_max * c.width * c.M > INT32_MAX / 2) return ORG_EINVAL;
    if ((int64_t)c.batch_max * c.width > 65535) return ORG_EINVAL;
    fno_net *n = new fno_net();
    n->cfg = c; n->dev = device;
    n->M = c.M; n->w = c.width; n->m = c.modes; n->L = c.layers; n->proj = c.proj; n->Bmax = c.batch_max;
    n->P = (int64_t)c.M * c.M;
    n->nm = 2 * c.modes * c.modes;
    if (setdev(n)) { delete n; return ORG_EHIP; }
    const int w = n->w, L = n->L, B = n->Bmax;
    size_t off = 0;
    n->oP = off; off += (size_t)w * 4;
    n->obP = off; off += w;
    for (int l = 0; l < L; l++) {
        n->oRr.push_back(off); off += (size_t)n->nm * w * w;
        n->oRi.push_back(off); off += (size_t)n->nm * w * w;
        n->oW.push_back(off); off += (size_t)w * w;
        n->ob.push_back(off); off += w;
    }
    n->oQ1 = off; off += (size_t)n->proj * w;
    n->oc1 = off; off += n->proj;
    n->oQ2 = off; off += (size_t)2 * n->proj;
    n->oc2 = off; off += 2;
    n->npar = off;
    int rc = ORG_OK;
    auto A = [&](auto **p, size_t cnt) { if (!rc) rc = dal(n, p, cnt); };
    A(&n->par, n->npar); A(&n->grad, n->npar); A(&n->am, n->npar); A(&n->av, n->npar);
    A(&n->wQ1, (size_t)n->proj * w); A(&n->wQ2, (size_t)2 * n->proj);
    n->wW.assign(L, nullptr); n->wX.assign(L, nullptr); n->u.assign(L, nullptr); n->xh.assign(L, nullptr);
    n->v.assign(L + 1, nullptr);
    for (int l = 0; l < L; l++) {
        A(&n->wW[l], (size_t)w * w);
        A(&n->wX[l], (size_t)n->nm * 4 * w * w);
        A(&n->u[l], (size_t)B * w * n->P);
        A(&n->xh[l], (size_t)n->nm * B * 2 * w);
    }
    for (int l = 0; l <= L; l++) A(&n->v[l], (size_t)B * w * n->P);
    A(&n->xin, (size_t)B * 4 * n->P);
    A(&n->qpre, (size_t)B * n->proj * n->P); A(&n->q, (size_t)B * n->proj * n->P);
    A(&n->out, (size_t)B * 2 * n->P);
    const size_t zsz = (size_t)B * w * n->M * 2 * n->m;
    A(&n->Z, zsz); A(&n->Zp, zsz);
    A(&n->Zh, (size_t)B * w * 4 * n->m * n->m); A(&n->Yh2, (size_t)B * w * 4 * n->m * n->m);
    A(&n->Yh, (size_t)n->nm * B * 2 * w);
    A(&n->dout, (size_t)B * 2 * n->P); A(&n->dq, (size_t)B * n->proj * n->P);
    A(&n->dv0, (size_t)B * w * n->P); A(&n->dv1, (size_t)B * w * n->P);
    A(&n->dXh, (size_t)n->nm * B * 2 * w); A(&n->dYh, (size_t)n->nm * B * 2 * w);
    A(&n->dZh, (size_t)B * w * 4 * n->m * n->m);
    A(&n->gexp, (size_t)n->nm * 4 * w * w);
    const int nch = (int)((n->P + KC - 1) / KC);
    A(&n->part, (size_t)B * nch * (size_t)std::max(n->proj, 4) * std::max(w, 4) + (size_t)B * nch * 2 * n->proj);
    A(&n->lpart, LOSS_BLOCKS);
    if (!rc && hipHostMalloc((void **)&n->lpart_h, LOSS_BLOCKS * sizeof(double)) != hipSuccess) rc = ORG_EHIP;
    if (!rc) rc = make_tables(n);
    if (!rc) {
        if (hipMemset(n->par, 0, n->npar * sizeof(float)) != hipSuccess ||
            hipMemset(n->am, 0, n->npar * sizeof(float)) != hipSuccess ||
            hipMemset(n->av, 0, n->npar * sizeof(float)) != hipSuccess)
            rc = ORG_EHIP;
    }
    if (rc) { fno_destroy(n); return rc; }
    *out = n;
    return ORG_OK;
}

int fno_destroy(fno_net *n) {
    if (!n) return ORG_EINVAL;
    (void)hipSetDevice(n->dev);
    for (void *p : n->allocs) (void)hipFree(p);
    if (n->lpart_h) (void)hipHostFree(n->lpart_h);
    delete n;
    return ORG_OK;
}

size_t fno_param_count(const fno_net *n) { return n ? n->npar : 0; }

int fno_set_params(fno_net *n, const float *host, size_t cnt) {
    if (!n || !host || cnt != n->npar) return ORG_EINVAL;
    TRY(setdev(n));
    if (hipStreamSynchronize(n->st) != hipSuccess) return ORG_EHIP;
    n->weights_fresh = false;
    return hipMemcpy(n->par, host, cnt * sizeof(float), hipMemcpyHostToDevice) == hipSuccess ? ORG_OK : ORG_EHIP;
}

int fno_get_params(fno_net *n, float *host, size_t cnt) {
    if (!n || !host || cnt != n->npar) return ORG_EINVAL;
    TRY(setdev(n));
    if (hipStreamSynchronize(n->st) != hipSuccess) return ORG_EHIP;
    return hipMemcpy(host, n->par, cnt * sizeof(float), hipMemcpyDeviceToHost) == hipSuccess ? ORG_OK : ORG_EHIP;
}

int fno_set_scales(fno_net *n, const float in_scale[2], const float out_scale[2]) {
    if (!n || !in_scale || !out_scale) return ORG_EINVAL;
    for (int c = 0; c < 2; c++) {
        if (!(std::isfinite(in_scale[c]) && std::isfinite(out_scale[c]) && out_scale[c] != 0.0f)) return ORG_EINVAL;
        n->in_scale[c] = in_scale[c];
        n->out_scale[c] = out_scale[c];
    }
    return ORG_OK;
}

int fno_set_stream(fno_net *n, void *stream) {
    if (!n) return ORG_EINVAL;
    n->st = (hipStream_t)stream;
    return ORG_OK;
}

int fno_get_stream(const fno_net *n, void **stream) {
    if (!n || !stream) return ORG_EINVAL;
    *stream = (void *)n->st;
    return ORG_OK;
}

int fno_get_config(const fno_net *n, fno_config *out) {
    if (!n || !out) return ORG_EINVAL;
    *out = n->cfg;
    return ORG_OK;
}

int fno_forward(fno_net *n, const float *in, int B, float *out) {
    if (!n || !in || !out || B < 1 || B > n->Bmax) return ORG_EINVAL;
    TRY(setdev(n));
    return forward(n, in, B, false, out);
}

int fno_loss_grad(fno_net *n, const float *in, const float *target, int B, double *loss, float *grad_host) {
    if (!n || !in || !target || B < 1 || B > n->Bmax) return ORG_EINVAL;
    TRY(setdev(n));
    TRY(forward(n, in, B, true, nullptr));
    hipLaunchKernelGGL(k_loss, dim3(LOSS_BLOCKS), dim3(256), 0, n->st, n->out, target, n->dout, n->lpart, B, n->P,
                       1.0f / n->out_scale[0], 1.0f / n->out_scale[1]);
    TRY(launch_ok());
    TRY(backward(n, B));
    if (hipMemcpyAsync(n->lpart_h, n->lpart, LOSS_BLOCKS * sizeof(double), hipMemcpyDeviceToHost, n->st) !=
            hipSuccess ||
        hipStreamSynchronize(n->st) != hipSuccess)
        return ORG_EHIP;
    if (loss) {
        double s = 0.0;
        for (int i = 0; i < LOSS_BLOCKS; i++) s += n->lpart_h[i];
        *loss = s / ((double)B * 2 * n->P);
    }
    if (grad_host &&
        hipMemcpy(grad_host, n->grad, n->npar * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
        return ORG_EHIP;
    return ORG_OK;
}

int fno_adam_step(fno_net *n, float lr, float b1, float b2, float eps, int t) {
    if (!n || t < 1 || !(lr >= 0.0f) || !(b1 >= 0.0f && b1 < 1.0f) || !(b2 >= 0.0f && b2 < 1.0f)) return ORG_EINVAL;
    TRY(setdev(n));
    const float c1 = 1.0f - std::pow(b1, (float)t), c2 = 1.0f - std::pow(b2, (float)t);
    hipLaunchKernelGGL(k_adam, dim3(grid_for((int64_t)n->npar)), dim3(256), 0, n->st, n->par, n->grad, n->am, n->av,
                       (int64_t)n->npar, lr, b1, b2, eps, c1, c2);
    n->weights_fresh = false;
    return launch_ok();
}

int fno_set_tiles(fno_net *n, const int32_t tiles[FNO_STAGE_COUNT]) {
    if (!n || !tiles) return ORG_EINVAL;
    for (int k = 0; k < FNO_STAGE_COUNT; k++)
        if (tiles[k] < 0 || tiles[k] >= FNO_TILE_COUNT) return ORG_EINVAL;
    for (int k = 0; k < FNO_STAGE_COUNT; k++) n->tile[k] = tiles[k];
    return ORG_OK;
}

int fno_get_tiles(const fno_net *n, int32_t tiles[FNO_STAGE_COUNT]) {
    if (!n || !tiles) return ORG_EINVAL;
    for (int k = 0; k < FNO_STAGE_COUNT; k++) tiles[k] = n->tile[k];
    return ORG_OK;
}

int fno_gemm(fno_net *n, const fno_gemm_args *a) {
    if (!n || !a) return ORG_EINVAL;
    TRY(setdev(n));
    return gemm(n, *a);
}

}  // extern "C"
