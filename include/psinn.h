/*
 * psinn.h -- C ABI of the learned-preconditioner kernels (libpsisolve.so).
 *
 * These entry points have no counterpart in the reference (pyOM2 has no
 * learned preconditioner): they implement the encoder-decoder network that
 * the flexible-CG path of psi_fcg_* applies to residuals.  Tensors are
 * caller-owned device arrays; activations are NHWC bfloat16 (uint16_t bits)
 * with channel counts that are multiples of 16; master weights are fp32 in
 * [tap][cin][cout] order (tap = 3*ky + kx).  Every function returns 0 on
 * success, 5 on a bad argument, 6 on a HIP launch error.
 */
#ifndef PSINN_H
#define PSINN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* D = A B for bf16 A (16 x 32) and B (32 x 16), row-major, fp32 D (16 x 16),
 * via one v_mfma_f32_16x16x32_bf16, the convolutions' MFMA: lane-layout check. */
int psinn_mfma_probe(const uint16_t *A, const uint16_t *B, float *D, void *stream);

/* 3x3 conv, zero padding, input = [in1 (C1, nearest x2 upsampled if up1) |
 * in2 (C2)], output cout_store channels of cout_pad computed (bf16 out16,
 * cout_store a multiple of 16, or fp32 out32), + bias, optional ReLU.  wpack
 * from psinn_pack_weights. */
int psinn_conv3x3(const uint16_t *in1, int C1, int up1, const uint16_t *in2, int C2, int B, int H,
                  int W, const uint16_t *wpack, int cout_pad, const float *bias, int relu,
                  uint16_t *out16, float *out32, int cout_store, void *stream);

/* Accumulates dW[tap][ci][co] (fp32, CIN = C1+C2, COUT = cout_pad) and
 * dbias[co] from the conv input and the pre-activation gradient g (bf16,
 * cout_pad channels). */
int psinn_conv3x3_wgrad(const uint16_t *in1, int C1, int up1, const uint16_t *in2, int C2, int B,
                        int H, int W, const uint16_t *g, int cout_pad, float *dW, float *dbias,
                        void *stream);

/* Pack fp32 w[tap][cin][cout] into MFMA B fragments (transposed = 1: the
 * flipped, transposed weights of the data-gradient convolution); k16 and n16
 * are the 16-channel blocks along K and N, and pack holds k16 * n16 * 2560
 * bf16 (per 16 input channels, 5 K steps of two taps, the tenth zero). */
int psinn_pack_weights(const float *w, int cin, int cout, int transposed, int k16, int n16,
                       uint16_t *pack, void *stream);

int psinn_avgpool2(const uint16_t *in, int B, int H, int W, int C, uint16_t *out, void *stream);
int psinn_sumpool2(const float *fine, int B, int H, int W, int CF, int c0, int C, float *coarse,
                   void *stream);
/* out = (act > 0) * (A[:, cA0:cA0+C] + 0.25 * upsample(Bc)), bf16. */
int psinn_grad_combine(const uint16_t *act, int B, int H, int W, int C, const float *A, int CA,
                       int cA0, const float *Bc, uint16_t *out, void *stream);
int psinn_expand16(const float *g, long n, uint16_t *out, void *stream);
int psinn_adam(float *w, const float *g, float *m, float *v, long n, float lr, float b1, float b2,
               float eps, int step, float gscale, void *stream);
/* Network input from a residual in solver layout (sh x sw, fp64 or fp32):
 * channel 0 = Z r * scale[0], 1 = inverse depth, 2 = active mask (Z != 0). */
int psinn_net_input(const void *r, int r_is_f64, const float *Z, const float *hinv, int sh, int sw,
                    int H, int W, const double *scale, uint16_t *out, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* PSINN_H */
