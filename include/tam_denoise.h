/*
 * tam_denoise.h -- C-ABI of the learned fluence denoiser (part of libtam.so).
 *
 * No reference counterpart: Tissue-Ablation-MC has no learned component.
 * The denoiser maps a low-packet-count fluence map plus the optical property
 * maps of a tissue grid to the high-count fluence, so a coupling cycle can
 * run few packets (tam_dn_apply on a tam_ctx after tam_mc_run).
 *
 * Network (fixed): a 3-level 3D U-Net, 16/32/64 channels, two 3x3x3
 * convolutions + ReLU per level, 2x2x2 average pooling, nearest 2x
 * upsampling with skip concatenation, a 1x1x1 head added to the input's
 * log-fluence channel.  Weights and activations are BF16 with FP32
 * accumulation (MFMA implicit-GEMM kernels for the forward, data-gradient
 * and weight-gradient convolutions); master weights and Adam state FP32.
 *
 * Features per voxel (4): log(1 + phi_lo), log10(mu_a), log10(mu_s) - 2,
 * 5 (g - 0.8), with phi = path length / (packets * voxel volume in cm^3)
 * (fluence per launched packet, cm^-2) and mu_a, mu_s clamped below at
 * 1e-3 and 1e-2 cm^-1 (ablated voxels).  Target: log(1 + phi_hi).
 * Grid edges must be multiples of 32 (two poolings, 32-voxel rows).
 */
#ifndef TAM_DENOISE_H
#define TAM_DENOISE_H

#include <stdint.h>

#include "tam.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tam_dn tam_dn;

/* A network for nx*ny*nz grids on `device`, with `capacity` dataset slots
 * held in device memory (12 bytes per voxel each: BF16 features, FP32
 * target) and He-initialised
 * weights drawn from `seed`. */
int  tam_dn_create(int32_t device, int32_t nx, int32_t ny, int32_t nz, int32_t capacity, uint64_t seed,
                   tam_dn** out);
void tam_dn_destroy(tam_dn* dn);
const char* tam_dn_last_error(const tam_dn* dn);

/* Parameters as one FP32 vector (layer by layer: weights [27][cin][cout],
 * bias [cout]; the head's 16 weights and bias last). */
int64_t tam_dn_num_params(const tam_dn* dn);
int tam_dn_get_params(tam_dn* dn, float* out);
int tam_dn_set_params(tam_dn* dn, const float* in);

/* Dataset slot from two transport runs of the same tissue on the same
 * device and grid: `lo` (features) and `hi` (target), with their packet
 * counts.  Device to device, no host copies. */
int tam_dn_store(tam_dn* dn, int32_t slot, const tam_ctx* lo, double npk_lo, const tam_ctx* hi,
                 double npk_hi);
/* Dataset slot from host arrays: features [4][nvox], target [nvox]
 * (log(1 + phi_hi)), interior order x fastest. */
int tam_dn_store_host(tam_dn* dn, int32_t slot, const float* features, const float* target);

/* `steps` Adam steps (learning rate lr), one random slot of [0, nslots) per
 * step, drawn from `seed`; losses[s] = mean |pred - target| of step s. */
int tam_dn_train(tam_dn* dn, int32_t nslots, int32_t steps, double lr, uint64_t seed, float* losses);
/* Mean |pred - target| and mean |input - target| (the noisy input's own
 * error) on a stored slot. */
int tam_dn_eval(tam_dn* dn, int32_t slot, double* l1_pred, double* l1_input);

/* Prediction, target and the noisy input channel (all log(1 + phi), [nvox])
 * of a stored slot, to the host (any pointer may be NULL). */
int tam_dn_predict_slot(tam_dn* dn, int32_t slot, float* pred, float* target, float* input);

/* Host-array forms for tests: prediction log(1 + phi) [nvox] of features
 * [4][nvox]; loss and the gradient of the mean-L1 loss w.r.t. every
 * parameter (layout of tam_dn_get_params). */
int tam_dn_forward_host(tam_dn* dn, const float* features, float* out);
int tam_dn_gradients_host(tam_dn* dn, const float* features, const float* target, float* grads, double* loss);

/* Denoise a context's transport tally in place: features from its tally
 * (npk packets) and optical maps, prediction written back as path length,
 * so tam_normalise turns it into jmean as usual.  Same device and grid. */
int tam_dn_apply(tam_dn* dn, tam_ctx* ctx, double npk);

/* Device time of the last call, ms. */
double tam_dn_last_ms(const tam_dn* dn);

#ifdef __cplusplus
}
#endif
#endif
