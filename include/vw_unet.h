/*
 * vw_unet.h -- C-ABI of the learned deconvolver (U-Net-lite) on gfx950.
 *
 * No counterpart in the reference (SSshuishui/viss_optimize_ws stops at the
 * dirty map).  The network maps a dirty lm image (img_n x img_n, the
 * reconstruction's grid image) to an estimate of the sky on the same grid:
 * four levels of 16, 32, 64 and 128 channels, two 3x3 convolutions with ReLU
 * per level, 2x2 average pooling down, nearest-neighbour upsampling and a
 * 3x3 convolution up with an additive skip from the encoder, a 1x1 output
 * convolution added to the input.  Every layer runs in hand-written HIP
 * kernels (forward, backward data, backward weights) with FP32 accumulation;
 * activations are stored in FP16 or FP32, weights and the Adam state in FP32.
 * Loss: mean squared error over the pixels inside the unit disk.
 * All calls are synchronous at return and return 0 on success (VW_ERR_* of
 * vissws.h otherwise; vw_unet_last_error gives the message).
 */
#ifndef VW_UNET_H
#define VW_UNET_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* img_n: image side, a multiple of 8 (three 2x2 poolings); batch_max:
 * largest batch of any call; act_fp16: 1 stores activations in FP16;
 * seed: He-normal initialisation (mt19937_64).  Device memory: activations
 * and their gradients for batch_max images (0.88 GB at 256 x 256, batch 16,
 * FP16), weight-gradient partials (0.12 GB there), the training set of
 * vw_unet_set_data. */
int vw_unet_create(int32_t img_n, int32_t batch_max, int32_t act_fp16, uint64_t seed, void** ctx_out);
void vw_unet_destroy(void* ctx);
const char* vw_unet_last_error(void* ctx);

/* parameters as one FP32 vector (conv weights [co][ci][3][3] then biases, layer
 * by layer in forward order) */
int64_t vw_unet_param_count(void* ctx);
int vw_unet_get_params(void* ctx, float* out);
int vw_unet_set_params(void* ctx, const float* in);

/* Loss and its gradient for n images X, targets Y ([n][img_n][img_n] FP32,
 * host); grad (may be NULL) receives param_count values.  No update. */
int vw_unet_loss_grad(void* ctx, const float* X, const float* Y, int32_t n, double* loss, float* grad);

/* Training set resident on the device (stored in the activation type). */
int vw_unet_set_data(void* ctx, const float* X, const float* Y, int64_t count);
/* steps of Adam (beta1 0.9, beta2 0.999, eps 1e-8) on batches drawn without
 * replacement per epoch from the resident set (mt19937_64 seeded shuffles);
 * losses (may be NULL) receives the batch loss of every step. */
int vw_unet_train(void* ctx, int64_t steps, int32_t batch, float lr, uint64_t seed, float* losses);
/* Network output for n images (host in, host out). */
int vw_unet_predict(void* ctx, const float* X, int32_t n, float* out);
/* vw_unet_predict that allocates no device memory: the FP32 staging of the
 * images in and out goes through a static device array of the library
 * (16 MiB: 4,194,304 values, e.g. 64 images of 256 x 256), whose address on
 * the device is resolved by vw_unet_create.  The largest batch,
 * min(batch_max, n) x img_n^2 values, must fit it (VW_ERR_RANGE otherwise).
 * No counterpart in the reference: for callers that cannot allocate device
 * memory at call time (tight serving loops, harnesses that forbid hipMalloc).
 * Calls on one device share the array; calls that overlap there race. */
int vw_unet_predict_noalloc(void* ctx, const float* X, int32_t n, float* out);
/* Device time of the last vw_unet_train call (ms, HIP events). */
int vw_unet_train_ms(void* ctx, double* ms);

#ifdef __cplusplus
}
#endif

#endif /* VW_UNET_H */
