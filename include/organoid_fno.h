/*
 * organoid_fno.h -- C-ABI of the Fourier-neural-operator velocity surrogate
 * (gfx950, BF16 MFMA with FP32 accumulation).
 *
 * The reference has no learned part: its cortex velocity comes from the
 * exact spectral friction-viscosity solve inside update_all_cortices
 * (rho_functions.cpp:409-471, the velocity itself at rho_functions.cpp:
 * 134-190).  This library is a surrogate for that solve, (phi_n, rho_n) on a
 * cell's M x M FFT box -> (vx, vy) on the same box, trained on pairs the
 * exact solve produces, with the exact solve kept as judge and fallback
 * (org_step_surrogate in organoid.h).
 *
 * Network (Li et al., "Fourier neural operator for parametric PDEs", 2020):
 *   input      4 channels per pixel: phi * in_scale[0], rho * in_scale[1],
 *              i / M, j / M
 *   lift       v0 = P in + bP                               (4 -> width)
 *   layer l    u_l = K_l(v_l) + W_l v_l + b_l,
 *              v_{l+1} = GELU(u_l) for l < layers-1, v_layers = u_{layers-1}
 *   K_l        truncated spectral convolution: the transform of v over the
 *              modes kx in {0..modes-1, M-modes..M-1}, ky in {0..modes-1},
 *              a complex width x width matrix per mode, the inverse of the
 *              real transform (modes ky > 0 counted twice, Im of ky = 0
 *              dropped: numpy.fft.irfft2 of the truncated spectrum)
 *   projection out = Q2 GELU(Q1 v_layers + c1) + c2        (width -> proj -> 2)
 *   velocity   (vx, vy) = out * out_scale
 * GELU(x) = x/2 (1 + erf(x / sqrt 2)).  Activations and weights enter the
 * matrix cores as BF16, accumulation and the stored pre-activations are
 * FP32, master weights and the Adam state FP32.
 *
 * Parameters are one flat FP32 array, in this order (row-major):
 *   P [width][4], bP [width],
 *   per layer: Rr [nmodes][width][width], Ri [same], W [width][width],
 *              b [width]       (nmodes = 2*modes*modes, mode = kx_idx*modes
 *                               + ky, kx_idx < modes: kx = kx_idx, else
 *                               kx = M - 2*modes + kx_idx; R[mode][i][o]
 *                               maps input channel i to output o)
 *   Q1 [proj][width], c1 [proj], Q2 [2][proj], c2 [2].
 *
 * Device pointers are plain pointers on the library's device; every entry
 * point returns 0 (ORG_OK) or an ORG_E* code from organoid.h.
 */
#ifndef ORGANOID_FNO_H
#define ORGANOID_FNO_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fno_net fno_net;

typedef struct fno_config {
    int32_t M;           /* box side (128)                                   */
    int32_t width;       /* channels of the Fourier layers (32)              */
    int32_t modes;       /* retained modes per axis (12)                     */
    int32_t layers;      /* Fourier layers (4)                               */
    int32_t proj;        /* hidden channels of the projection (128)          */
    int32_t batch_max;   /* largest batch forward / training calls take      */
} fno_config;

/* Create a network (parameters zero, scales 1) on device `device`. */
int fno_create(const fno_config *cfg, int device, fno_net **out);
int fno_destroy(fno_net *net);
/* Number of FP32 parameters (the flat layout above). */
size_t fno_param_count(const fno_net *net);
/* Host <-> device copies of the flat parameter array. */
int fno_set_params(fno_net *net, const float *host, size_t n);
int fno_get_params(fno_net *net, float *host, size_t n);
/* Input and output scales (see the network above). */
int fno_set_scales(fno_net *net, const float in_scale[2], const float out_scale[2]);
/* The stream the library launches on (0: the null stream). */
int fno_set_stream(fno_net *net, void *hip_stream);
int fno_get_stream(const fno_net *net, void **hip_stream);
/* The configuration the network was created with. */
int fno_get_config(const fno_net *net, fno_config *out);

/* Inference.  in: [batch][2][M][M] FP32 (phi, rho) on the device,
 * out: [batch][2][M][M] FP32 (vx, vy). */
int fno_forward(fno_net *net, const float *in, int batch, float *out);

/* One forward and backward pass of the loss
 *   L = mean over batch, component, pixel of (out - target / out_scale)^2
 * (out before the output scale).  in, target: [batch][2][M][M] FP32 on the
 * device.  Gradients are kept on the device for fno_adam_step; loss and, if
 * grad_host is not NULL, the gradient (flat layout, fno_param_count floats)
 * are returned on the host. */
int fno_loss_grad(fno_net *net, const float *in, const float *target, int batch, double *loss,
                  float *grad_host);

/* Adam (Kingma & Ba 2015) on the gradients of the last fno_loss_grad;
 * step t >= 1 enters the bias corrections. */
int fno_adam_step(fno_net *net, float lr, float beta1, float beta2, float eps, int t);

/* GEMM tile shapes (rows x columns of C per 256-thread workgroup) and the
 * stages of the network whose GEMMs share one tile choice.  The default is
 * 64 x 64 everywhere; autotune.py times the variants per stage and sets
 * them.  The tile changes no result: every variant accumulates the same
 * products in the same K order. */
enum {
    FNO_TILE_64x64 = 0, FNO_TILE_32x64 = 1, FNO_TILE_64x32 = 2, FNO_TILE_32x32 = 3, FNO_TILE_COUNT = 4
};
enum {
    FNO_STAGE_Y = 0,             /* forward transform along y                        */
    FNO_STAGE_X = 1,             /* forward transform along x                        */
    FNO_STAGE_MIX = 2,           /* per-mode channel mix                             */
    FNO_STAGE_IX = 3,            /* inverse transform along x                        */
    FNO_STAGE_LAYER = 4,         /* inverse y transform + pointwise mix (2 segments) */
    FNO_STAGE_PROJ = 5,          /* projection (training form)                       */
    FNO_STAGE_BWD_SPECTRAL = 6,  /* backward of the spectral convolution             */
    FNO_STAGE_BWD_POINTWISE = 7, /* backward of the pointwise products               */
    FNO_STAGE_WGRAD = 8,         /* weight gradients (K-chunked partials)            */
    FNO_STAGE_API = 9,           /* fno_gemm                                         */
    FNO_STAGE_COUNT = 10
};
int fno_set_tiles(fno_net *net, const int32_t tiles[FNO_STAGE_COUNT]);
int fno_get_tiles(const fno_net *net, int32_t tiles[FNO_STAGE_COUNT]);

/* The batched BF16 matrix-core GEMM every layer is built from, exposed for
 * its unit tests (tile: the FNO_STAGE_API entry):
 *   C[z][m][n] = epilogue(alpha * sum_k A[z][m][k] B[z][k][n])
 * with z = z1 * nz2 + z2 and element offsets
 *   A: z1*sA1 + z2*sA2 + m*sAm + k*sAk   (FP32 if a_f32, else BF16)
 *   B: z1*sB1 + z2*sB2 + k*sBk + n*sBn   (FP32 if b_f32, else BF16)
 *   C: z1*sC1 + z2*sC2 + m*sCm + n*sCn
 * FP32 operands are rounded to BF16 (nearest even) on the way in.  Epilogue:
 * val = alpha*acc (+ C, c_mode 1) (+ bias[n]); c_mode 0 / 1 store val in
 * FP32 C, c_mode 2 stores act(val) in BF16 C; C2 (BF16, C's offsets), if
 * not NULL, receives act(val); act 1 is GELU, 0 the identity. */
typedef struct fno_gemm_args {
    const void *A; int64_t sAm, sAk, sA1, sA2;
    const void *B; int64_t sBk, sBn, sB1, sB2;
    void *C; int64_t sCm, sCn, sC1, sC2;
    void *C2;
    const float *bias;
    int32_t M, N, K, nz1, nz2;
    int32_t a_f32, b_f32, c_mode, act;
    float alpha;
} fno_gemm_args;
int fno_gemm(fno_net *net, const fno_gemm_args *args);

#ifdef __cplusplus
}
#endif

#endif
