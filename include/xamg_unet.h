/*
 * Learned initial guess for the solve (part of libxamg_hip.so).  No reference
 * counterpart: the reference starts every solve from the x the caller passes
 * (XAMG_solver_solve2, api/c/xamg_c_api.cpp:522-531).  A 3D U-Net maps each right-hand
 * side b (one column of an nv-interleaved multivector on an nx x ny x nz grid) to x0; the
 * unchanged solver then runs from x0 to the same tolerance.
 *
 * Network: `levels` levels with channels c[0..levels-1]; level l: two 3x3x3 convolutions
 * + ReLU (the first from c[l-1], or 1 input channel at l = 0), 2x2x2 average pooling
 * between levels; decoder level l: nearest x2 upsampling of level l+1, a 3x3x3
 * convolution c[l+1] -> c[l] added to the level-l encoder output, ReLU, a 3x3x3
 * convolution + ReLU; a 1x1 head c[0] -> 1.  Activations bf16, convolutions on BF16 MFMA
 * with fp32 accumulation, fp32 master weights trained with Adam on the loss
 * 0.5 * mean((net(b/s) - x/(s*o))^2), s the RMS of b, o a fixed output scale.
 *
 * Parameter order (xh_unet_get_params / set_params / loss_grad): encoder convolutions
 * (level 0 first, first then second), decoder convolutions (level 0 first, upsampling
 * convolution then second), head; each convolution w[27][cin][cout] (tap t = 9*(dz+1) +
 * 3*(dy+1) + (dx+1)) then b[cout]; the head w[c0] then b.
 * All functions return 0 on success, a negative code on failure (xh_last_error()).
 */
#ifndef XAMG_UNET_H_
#define XAMG_UNET_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct xh_unet_s *xh_unet;

/* channels: `levels` values, each a multiple of 16 and <= 64; weights drawn from `seed` */
int xh_unet_create(int levels, const int32_t *channels, uint64_t seed, xh_unet *out);
int xh_unet_destroy(xh_unet U);
int xh_unet_nparams(xh_unet U, int64_t *n);
int xh_unet_get_params(xh_unet U, float *params);
int xh_unet_set_params(xh_unet U, const float *params);
int xh_unet_get_output_scale(xh_unet U, float *o);
int xh_unet_set_output_scale(xh_unet U, float o);

/* Loss and gradients for a batch of n single-channel volumes d x h x w (host fp32, volume
 * index x fastest), the network applied to `in` as given (no normalisation): loss =
 * 0.5 * mean((net(in) - tgt)^2); grads receives nparams values; y (nullable) the output.
 * d, h, w must be multiples of 2^(levels-1). */
int xh_unet_loss_grad(xh_unet U, int n, int d, int h, int w, const float *in,
                      const float *tgt, float *loss, float *grads, float *y);

/* Training on solver pairs: nsets device multivector pairs (b, x), each nx*ny*nz rows by nv
 * columns (fp64, interleaved); `steps` Adam steps (learning rate lr) on batches of
 * `batch` random crops of crop^3 cells; sets the output scale from the data first.
 * loss_hist (nullable) receives the loss of every step. */
int xh_unet_train(xh_unet U, int nx, int ny, int nz, int nsets, const double *const *db,
                  const double *const *dx, int nv, int crop, int batch, int steps, float lr,
                  uint64_t seed, float *loss_hist);

/* x0 = net(b) for every column of a device multivector b (nx*ny*nz rows by nv, fp64
 * interleaved) into the device multivector x0 of the same shape */
int xh_unet_infer(xh_unet U, int nx, int ny, int nz, int nv, const double *db, double *dx0);

/* d = a b on one MFMA (16x32 by 32x16, bf16 operands, fp32 result), host row-major arrays:
 * a check of the operand layout the convolution kernels assume */
int xh_unet_mfma_selftest(const float *a, const float *b, float *d);

#ifdef __cplusplus
}
#endif

#endif
