/*
 * flsgpu learned ALP parameter selector, C ABI (device pointers, async on
 * `stream`). No reference counterpart: it proposes the per-vector
 * (exponent, factor) and the ALP / ALP-RD choice that the reference finds by
 * search (src/alp/encoder.hpp:134-300); the search remains the fallback and
 * the source of training labels. Decoding stays lossless whatever the
 * proposal (values that do not round-trip become exceptions).
 *
 * Network: 32 features -> 64 ReLU -> 64 ReLU -> 40 logits
 * (19 exponent + 19 factor + 2 scheme), fp32. Parameter layout (floats):
 * W1[32][64], b1[64], W2[64][64], b2[64], W3[64][40], b3[40].
 */
#ifndef FLSGPU_SELECTOR_H
#define FLSGPU_SELECTOR_H

#include <stdint.h>

#include "flsgpu.h"

#ifdef __cplusplus
extern "C" {
#endif

#define FLSGPU_SEL_INPUTS 32
#define FLSGPU_SEL_HIDDEN 64
#define FLSGPU_SEL_OUTPUTS 40
#define FLSGPU_SEL_PARAMS 8872

/* Features of each vector from 32 samples, values[v * vector_stride +
 * i * sample_stride] for i < 32: (1024, 32) reads the samples at slots
 * 0, 32, ..., 992 of a column's vectors (the samples of
 * src/alp/encoder.hpp:248-249); (32, 1) reads a [n_vecs][32] sample matrix.
 * features[n_vecs][32]. */
int flsgpu_sel_features_f32(const float *values, uint64_t n_vecs, uint32_t vector_stride,
                            uint32_t sample_stride, float *features, flsgpu_stream_t stream);
int flsgpu_sel_features_f64(const double *values, uint64_t n_vecs, uint32_t vector_stride,
                            uint32_t sample_stride, float *features, flsgpu_stream_t stream);
/* Forward pass; h1/h2 ([n][64], post-ReLU) may be NULL when no backward follows. */
int flsgpu_sel_forward(const float *params, const float *x, uint32_t n, float *h1, float *h2,
                       float *logits, flsgpu_stream_t stream);
/* Softmax cross-entropy over the three heads (labels[n][3] = exponent,
 * factor, scheme), gradients of the mean loss w.r.t. every parameter
 * (grad[8872]); dlogits/dh1/dh2 are caller-provided workspace. */
int flsgpu_sel_backward(const float *params, const float *x, const float *h1, const float *h2,
                        const float *logits, const int32_t *labels, uint32_t n, float *dlogits,
                        float *dh1, float *dh2, float *grad, float *loss_per_row,
                        flsgpu_stream_t stream);
/* One Adam step (step counts from 1). */
int flsgpu_sel_adam(float *params, const float *grad, float *m, float *v, float lr, float beta1,
                    float beta2, float eps, uint32_t step, flsgpu_stream_t stream);
/* Proposal per row: exponent = argmax of its head, factor = argmax over
 * factors <= exponent, scheme (0 = ALP, 1 = ALP-RD; scheme may be NULL).
 * Type-agnostic: for float columns use flsgpu_sel_select_f32, which never
 * proposes factor 10 (outside the float factor table). */
int flsgpu_sel_predict(const float *params, const float *x, uint32_t n, uint8_t *exps,
                       uint8_t *facs, uint8_t *scheme, flsgpu_stream_t stream);
/* Learned proposal + verification. For each vector, the n_candidates
 * (exponent, factor) pairs with the largest logit sums are scored on the
 * vector's 32 samples (same addressing as flsgpu_sel_features_*) with the
 * search's size estimate and tie rules (src/alp/encoder.hpp:153-190); the
 * best one is kept. With n_candidates >= the number of pairs (190 for
 * doubles, 66 for floats) this is the exhaustive search itself.
 * scheme = 1 (ALP-RD) when the kept estimate reaches the ALP-RD threshold
 * (48 x 32 bits for doubles, 22 x 32 for floats, encoder.hpp:208-211).
 * scheme and est_bits may be NULL. */
int flsgpu_sel_select_f32(const float *params, const float *features, const float *values,
                          uint64_t n_vecs, uint32_t vector_stride, uint32_t sample_stride,
                          uint32_t n_candidates, uint8_t *exps, uint8_t *facs, uint8_t *scheme,
                          uint64_t *est_bits, flsgpu_stream_t stream);
int flsgpu_sel_select_f64(const float *params, const float *features, const double *values,
                          uint64_t n_vecs, uint32_t vector_stride, uint32_t sample_stride,
                          uint32_t n_candidates, uint8_t *exps, uint8_t *facs, uint8_t *scheme,
                          uint64_t *est_bits, flsgpu_stream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* FLSGPU_SELECTOR_H */
