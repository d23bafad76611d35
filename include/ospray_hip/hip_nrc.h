/* Neural radiance cache of the hip path tracer: a fully fused MLP on the
 * gfx950 bf16 matrix cores (v_mfma_f32_16x16x32_bf16), fp32 accumulation,
 * trained online with Adam on the path tracer's own path suffixes.
 *
 * No counterpart in the reference renderer: the cache terminates paths
 * early (after a fixed number of bounces) with the network's estimate of
 * the radiance leaving the vertex.  Network: 160 inputs (frequency
 * encoding, 12 bands x sin/cos of the position in the scene box and of the
 * direction, then albedo, normal, position, direction, a constant 1),
 * 4 hidden ReLU layers of 64, 3 linear outputs; relative-L2 loss
 * (y - t)^2 / (y^2 + 0.01) with the denominator held constant.
 *
 * Every function enqueues on the given stream, returns the launch's
 * hipError_t, never throws and never allocates; buffers are device memory.
 */
#pragma once

#include <hip/hip_runtime_api.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HPT_NRC_INPUTS 160
#define HPT_NRC_HIDDEN 64
#define HPT_NRC_OUTPUTS 3
/* weights, row-major (out, in): W0 64x160, W1..W3 64x64, W4 16x64 (rows 3..15 zero) */
#define HPT_NRC_PARAMS 23552
/* transposed copies used by the backward pass: W1^T..W3^T 64x64, W4^T 64x32 */
#define HPT_NRC_TPARAMS 14336
/* queries per tile: one wave of the training kernel */
#define HPT_NRC_TILE 64

#ifndef HPT_API
#define HPT_API __attribute__((visibility("default")))
#endif

typedef struct HptNrc
{
  uint16_t *w; /* bf16 weights, HPT_NRC_PARAMS */
  uint16_t *wT; /* bf16 transposes, HPT_NRC_TPARAMS */
  float *master; /* fp32 weights, HPT_NRC_PARAMS */
  float *adamM, *adamV; /* Adam moments, HPT_NRC_PARAMS each */
  float *partial; /* per-tile gradients, maxTiles x HPT_NRC_PARAMS */
  float *tileLoss; /* per-tile loss, maxTiles */
  float *loss; /* 1: the last training step's loss */
  int32_t maxTiles;
} HptNrc;

/* He-uniform initial weights from `seed`, zero Adam moments. */
HPT_API hipError_t hpt_nrc_init(const HptNrc *N, uint32_t seed, hipStream_t s);
/* raw12 = n x (position, direction, normal, albedo) floats -> n x 160 bf16
 * features; positions mapped over the box [lo, hi]. */
HPT_API hipError_t hpt_nrc_encode(const float *raw12, int n, const float *lo3, const float *hi3, uint16_t *features,
    hipStream_t s);
/* out3 = n x 3 network outputs for n x 160 bf16 features. */
HPT_API hipError_t hpt_nrc_infer(const HptNrc *N, const uint16_t *features, int n, float *out3, hipStream_t s);
/* The path tracer's cache queries fused in one kernel: for the n path
 * slots in queue, read the slot's vertex record rec16 (position,
 * throughput.x | direction, throughput.y | normal, throughput.z | albedo,
 * unused), encode it, run the network, and add throughput * max(output, 0)
 * to contrib4[slot].xyz. */
HPT_API hipError_t hpt_nrc_query(const HptNrc *N, const float *rec16, const uint32_t *queue, int n, const float *lo3,
    const float *hi3, float *contrib4, hipStream_t s);
/* Forward + backward of the relative-L2 loss over n <= maxTiles * 64
 * (features, targets) pairs: per-tile gradients (scaled by 1 / (3 n)) and
 * losses into N->partial and N->tileLoss. */
HPT_API hipError_t hpt_nrc_train(
    const HptNrc *N, const uint16_t *features, const float *target3, int n, hipStream_t s);
/* Sums the gradients of ceil(n / 64) tiles and takes Adam step `step`
 * (1-based; beta1 0.9, beta2 0.99, eps 1e-8) with learning rate lr:
 * updates the fp32 weights, the moments, the bf16 weights and transposes,
 * and N->loss. */
HPT_API hipError_t hpt_nrc_adam(const HptNrc *N, int n, float lr, int step, hipStream_t s);

#ifdef __cplusplus
}
#endif
