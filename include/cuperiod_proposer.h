/*
 * C ABI of the learned candidate proposer (gfx950): a 1-D convolutional network that
 * reads a coarse BLS spectrum and a cadence histogram of each light curve and scores
 * every coarse frequency cell as the location of the full-grid BLS best period.
 * cuPeriod has no counterpart: the proposer narrows a batch's fine search to a few
 * frequency windows per curve, and the plain full-grid search (cpb_bls_search_batch)
 * stays its fallback and the reference its recall is measured against.
 *
 * Network, per curve, on L = 4096 positions (ascending log frequency):
 *   h1 = GELU(conv7(x, W1) + b1)          x: 2 channels (FP32), W1: [64][7][2] FP32
 *   h2..h4 = GELU(conv7(h, Wl) + bl)       64 -> 64 channels, Wl: [64][7][64] BF16
 *   logit = h4 . wh + bh                   one logit per position
 * conv7(h, W)[x][o] = sum_{k<7, c} W[o][k][c] * h[x + k - 3][c] (zero padding).
 * Activations are channels-last BF16 [B][L][64]; the 64 -> 64 layers and their
 * gradients run on BF16 MFMA (v_mfma_f32_16x16x32_bf16) with FP32 accumulation.
 *
 * All arrays are HIP device pointers; calls enqueue on `stream` and return. Status codes
 * are those of cuperiod_bls.h (CPB_OK, CPB_E_INVALID for a bad size, CPB_E_DEVICE for a
 * launch failure).
 */
#ifndef CUPERIOD_PROPOSER_H
#define CUPERIOD_PROPOSER_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CPP_CHANNELS 64
#define CPP_TAPS 7
#define CPP_POSITIONS 4096
#define CPP_HIST_BINS 512

/* Features of `n_curves` curves into x[B][L][2] (FP32): channel 0 the coarse power
 * standardised per curve, (p - mean) / std (0 when std = 0), clipped to [-clip, clip],
 * read from `power` at out_base[c] + (L-1-x) (the coarse grid in ascending-period
 * order); channel 1 the curve's cadence histogram, 512 bins of tau / baseline
 * (tau = t - t_min, from `tau` at point_base[c], n_points[c] points), scaled so the
 * mean bin is 1, at bin x / 8. mean/std come from summary[c][0..1]. */
int cpp_features(void* stream, const double* power, const int64_t* out_base,
                 const double* summary, const double* tau, const int64_t* point_base,
                 const int32_t* n_points, const double* baseline, int32_t n_curves,
                 float clip, float* x);

/* Stacked features of `n_curves` curves into x[B][L][2] (FP32). Curve c's time chunks
 * are chunks chunk_start[c] .. chunk_start[c+1]-1, each searched on the curve's coarse
 * grid with its outputs at out_base[k] (ascending period, so cell x reads
 * power[out_base[k] + L-1-x]). Channel 0: S[x] = sum over the chunks of power^2 (a power
 * that is not finite and positive counts 0), standardised over the curve's L cells and
 * clipped to [-clip, clip]; channel 1: the cadence histogram, as cpp_features. */
int cpp_stack_features(void* stream, const double* power, const int32_t* chunk_start,
                       const int64_t* out_base, const double* tau, const int64_t* point_base,
                       const int32_t* n_points, const double* baseline, int32_t n_curves,
                       float clip, float* x);

/* h1 = GELU(conv7(x, W1) + b1) (and z1 = the pre-activation, when z1 is not NULL). */
int cpp_layer1_forward(void* stream, const float* x, const float* w1, const float* b1,
                       int32_t n_curves, uint16_t* z1, uint16_t* h1);

/* 64 -> 64 convolution on MFMA. epilogue 0: out = GELU(acc + bias), zsave = acc + bias
 * (when not NULL); epilogue 1: out = acc * GELU'(zsrc) (backward through the previous
 * layer's activation; bias unused); epilogue 2: out = acc + bias. W is [64][7][64] BF16
 * (BF16 values are passed as their 16-bit patterns). */
int cpp_conv_forward(void* stream, const uint16_t* in, const uint16_t* w, const float* bias,
                     const uint16_t* zsrc, uint16_t* zsave, uint16_t* out, int32_t n_curves,
                     int32_t epilogue);

/* Weight and bias gradients of a 64 -> 64 layer, accumulated (added) into FP32
 * dw[64][7][64] and db[64]: dw[o][k][c] += sum_x dz[x][o] * in[x + k - 3][c]. */
int cpp_conv_wgrad(void* stream, const uint16_t* dz, const uint16_t* in, int32_t n_curves,
                   float* dw, float* db);

/* Layer-1 gradients, accumulated: dw1[o][k][c] += sum_x dz1[x][o] * x[x + k - 3][c]. */
int cpp_layer1_wgrad(void* stream, const uint16_t* dz1, const float* x, int32_t n_curves,
                     float* dw1, float* db1);

/* wt[c][k][o] = w[o][6 - k][c]: the weights of the transposed convolution that carries
 * a gradient back through a 64 -> 64 layer. */
int cpp_flip_weights(void* stream, const uint16_t* w, uint16_t* wt);

/* Head, training: logits, softmax over the L positions, cross-entropy against a uniform
 * target on label[c] = (lo, hi), the cells of the curve's refinement window around its
 * full-grid best period; the loss is summed into *loss (scaled by `scale`), dlogit =
 * scale * (softmax - target); writes dz4 = dlogit * wh * GELU'(z4) and accumulates dwh,
 * dbh. */
int cpp_head_train(void* stream, const uint16_t* h4, const uint16_t* z4, const float* wh,
                   const float* bh, const int32_t* label, int32_t n_curves, float scale,
                   uint16_t* dz4, float* dwh, float* dbh, float* loss);

/* Head, inference: softmax probabilities of each curve's positions, then `n_windows`
 * windows by greedy selection. Cell x has frequency f_lo[c] * exp(x * dlog[c]); a window
 * centred on the most probable unclaimed cell covers the cells within half_width[c]
 * (frequency units) of it, and at least `radius` cells each side. Writes centre[c][w]
 * (-1 when no cell is left), mass[c][w], the probability of the cells it covers that
 * no earlier window claimed, and (when span is not NULL) span[c][w][0..1], the first and
 * last cell it covers: a search that is to earn that probability must cover every
 * frequency within half_width of every cell in the span. */
int cpp_head_windows(void* stream, const uint16_t* h4, const float* wh, const float* bh,
                     const double* f_lo, const double* dlog, const double* half_width,
                     int32_t n_curves, int32_t n_windows, int32_t radius, int32_t* centre,
                     float* mass, int32_t* span);

/* Adam step on n FP32 parameters p with gradients g (then zeroes g); also writes the
 * BF16 copy pb (when not NULL). t is the 1-based step number. */
int cpp_adam(void* stream, float* p, float* g, float* m, float* v, uint16_t* pb, int64_t n,
             float lr, float beta1, float beta2, float eps, int32_t t);

#ifdef __cplusplus
}
#endif

#endif /* CUPERIOD_PROPOSER_H */
