/*
 * howmhd_mlp.h — hand-written HIP kernels of the learned smoothness indicator:
 * an 18 -> 32 -> 32 -> 1 MLP (ReLU, ReLU, sigmoid) with bf16 weights and FP32
 * accumulation on the MFMA units, its training (forward, backward, Adam) and
 * the sample collection from the FP64 solver.  No counterpart in the
 * reference; the indicator and its label follow the WENO5 nonlinear weights of
 * src/ssprk.cpp:903-908 and 926-931.
 *
 * Parameter vector (FP32 master copy, HOWMHD_MLP_NPARAM floats):
 *   W1[32][18], b1[32], W2[32][32], b2[32], w3[32], b3
 * Features (18 floats per interface): the 6-cell WENO stencils of rho, p and
 * |B|, each mapped to [0, 1] by its min and max over the 6 cells (all 0 when
 * max - min <= 1e-5 max(|min|, |max|)).
 * Label: 1 when max over waves, sides and r of |w_r - d_r| > 0.05, with
 * d = (0.1, 0.6, 0.3).
 *
 * Pointers marked [dev] are device memory; `stream` is a hipStream_t (NULL:
 * the default stream).  Calls are asynchronous on that stream.
 */
#ifndef HOWMHD_MLP_H
#define HOWMHD_MLP_H

#include <stdint.h>
#include "howmhd_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

#define HOWMHD_MLP_NPARAM 1697
#define HOWMHD_MLP_NFEAT 18

/* logits[i] = z3 of sample i (X: n x 18 [dev], params [dev], logits [dev]). */
HOWMHD_API int howmhd_mlp_infer(const float* X, int64_t n, const float* params, float* logits, void* stream);

/* Adds the gradient of sum_i w_i * BCE(sigmoid(z3_i), y_i) * scale to grad
 * [dev] (w_i = pos_weight for y_i = 1, else 1) and that loss to *loss [dev].
 * y: n floats in {0, 1} [dev]. */
HOWMHD_API int howmhd_mlp_grad(const float* X, const float* y, int64_t n, const float* params, float pos_weight,
                    float scale, float* grad, float* loss, void* stream);

/* One Adam step on params [dev] with moments m, v [dev] (t >= 1). */
HOWMHD_API int howmhd_mlp_adam(float* params, float* m, float* v, const float* grad, int64_t t, float lr, float beta1,
                    float beta2, float eps, void* stream);

/* Samples of the current stage-0 state: for every interface of the sweep
 * along `dir` (0 x, 1 y, 2 z; n_d + 1 interfaces per line) writes the 18
 * features to X [dev] and the label to y [dev] (float 0/1), computed by the
 * FP64 characteristic WENO5 of the plain method.  *count receives the number
 * of samples written (host).  capacity is the number of rows X and y hold. */
HOWMHD_API int howmhd_gpu_collect_samples(howmhd_gpu_ctx* ctx, int dir, float* X, float* y, int64_t capacity,
                               int64_t* count);

/* The network's logit z3 for every interface of the sweep along `dir` of the
 * current stage-0 state, as the routed sweep computes it (features from the
 * sweep's LDS tile, MFMA forward pass), in the sample order of
 * howmhd_gpu_collect_samples; logits [dev].  Needs weights loaded
 * (set_indicator) and precision fp64.  Synchronises. */
HOWMHD_API int howmhd_gpu_indicator_logits(howmhd_gpu_ctx* ctx, int dir, float* logits, int64_t capacity,
                                           int64_t* count);

/* Loads indicator weights (params on the host, HOWMHD_MLP_NPARAM floats) and
 * the flag threshold on the output probability (0 < threshold < 1).
 *   enable 1: routing with the plain-WENO fallback: each routed step counts
 *             its wavefronts; when at least 90% of them ran the full WENO5
 *             (the network found little to skip), the next 8 steps run the
 *             plain sweep, then one step routes again.  The counts of a
 *             routed step travel back asynchronously; the next ssprk waits
 *             for them before it chooses its sweep.
 *   enable 2: routing on every step (no fallback), counted.
 *   enable 0: off (params may be NULL).
 * A context created with indicator HOWMHD_INDICATOR_MLP refuses to step until
 * weights are loaded.  Resets the counters.  Synchronises. */
HOWMHD_API int howmhd_gpu_set_indicator(howmhd_gpu_ctx* ctx, const float* params, float threshold, int enable);

/* Interfaces the indicator flagged / interfaces evaluated since the last
 * call (or set_indicator), and the same count at wavefront granularity (a
 * wavefront runs the full WENO5 when any of its interfaces is flagged).
 * Steps that ran the plain sweep (fallback) are not counted.  Synchronises
 * and resets the counters. */
HOWMHD_API int howmhd_gpu_indicator_stats(howmhd_gpu_ctx* ctx, int64_t* flagged, int64_t* total,
                               int64_t* waves_flagged, int64_t* waves_total);

/* Steps that routed and steps that ran the plain sweep as the fallback since
 * set_indicator (enable 1; enable 2 counts every step as routed). */
HOWMHD_API int howmhd_gpu_indicator_routing(howmhd_gpu_ctx* ctx, int64_t* routed_steps, int64_t* fallback_steps);

#ifdef __cplusplus
}
#endif

#endif
