/*
 * CPU oracle for the FastLanes / ALP / ALP-RD decode path.
 *
 * TEST INFRASTRUCTURE ONLY. This is a plain-C restatement of the reference
 * CPU codec (SvenHepkema/msc-thesis-project: the src/fls sources and src/alp).
 * Only tests/, the smoke check and the CPU-baseline leg of the benchmark may
 * load it, and only as the checker or the timed CPU baseline. The product
 * (the HIP library) never links or calls it.
 *
 * Pinning (what each function is checked against, see tests/test_oracle.py):
 *   fls_*          : compiled reference src/fls/{pack,unpack,ffor,unffor}.cpp
 *                    (oracle/_ref), golden fixtures, brute-force bit model,
 *                    pack/unpack identity, vbw=0 / vbw=W closed forms.
 *   alp_decode_*   : compiled reference falp.cpp + decoder.hpp (oracle/_ref)
 *                    for bit widths < W, golden fixtures, 128-bit integer model.
 *   alp_encode_*   : compiled reference encoder.hpp/sampler.hpp (oracle/_ref),
 *                    golden fixtures, decode(encode(x)) == x bitwise.
 *   galp_convert_* : invariants (stable per-lane partition, headers
 *                    recomputable, both layouts decode identically) and golden
 *                    fixtures; structs.cuh needs CUDA headers and does not
 *                    compile here, so it has no oracle/_ref counterpart.
 *   alprd_*        : decode(encode(x)) == x bitwise; dictionary choice pinned
 *                    against rd.hpp via oracle/_ref on tie-free inputs.
 */
#ifndef FLSGPU_ORACLE_H
#define FLSGPU_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- FastLanes bit-packing: one 1024-value vector per call --------------
 * Layout (src/fls/unffor.cpp:2982-3030): W = element bits, N_LANES = 1024/W.
 * Value at logical index l + N_LANES*j (lane l, slot j) occupies bits
 * [j*bw, (j+1)*bw) of lane l's stream in[l + N_LANES*r], r = 0..bw-1. */
void orc_fls_pack_u8(const uint8_t *in, uint8_t *out, unsigned bw);
void orc_fls_pack_u16(const uint16_t *in, uint16_t *out, unsigned bw);
void orc_fls_pack_u32(const uint32_t *in, uint32_t *out, unsigned bw);
void orc_fls_pack_u64(const uint64_t *in, uint64_t *out, unsigned bw);
void orc_fls_unpack_u8(const uint8_t *in, uint8_t *out, unsigned bw);
void orc_fls_unpack_u16(const uint16_t *in, uint16_t *out, unsigned bw);
void orc_fls_unpack_u32(const uint32_t *in, uint32_t *out, unsigned bw);
void orc_fls_unpack_u64(const uint64_t *in, uint64_t *out, unsigned bw);
void orc_fls_ffor_u8(const uint8_t *in, uint8_t *out, unsigned bw, uint8_t base);
void orc_fls_ffor_u16(const uint16_t *in, uint16_t *out, unsigned bw, uint16_t base);
void orc_fls_ffor_u32(const uint32_t *in, uint32_t *out, unsigned bw, uint32_t base);
void orc_fls_ffor_u64(const uint64_t *in, uint64_t *out, unsigned bw, uint64_t base);
void orc_fls_unffor_u8(const uint8_t *in, uint8_t *out, unsigned bw, uint8_t base);
void orc_fls_unffor_u16(const uint16_t *in, uint16_t *out, unsigned bw, uint16_t base);
void orc_fls_unffor_u32(const uint32_t *in, uint32_t *out, unsigned bw, uint32_t base);
void orc_fls_unffor_u64(const uint64_t *in, uint64_t *out, unsigned bw, uint64_t base);

/* ---- Column-level FFOR decode (src/engine/data.cuh:300-313) -------------
 * bases may be NULL (plain bit-packing, data.cuh:287-298). */
void orc_ffor_column_decode_u32(size_t n_vecs, const uint32_t *packed, const uint8_t *bit_widths,
                                const uint64_t *vector_offsets, const uint32_t *bases, uint32_t *out);
void orc_ffor_column_decode_u64(size_t n_vecs, const uint64_t *packed, const uint8_t *bit_widths,
                                const uint64_t *vector_offsets, const uint64_t *bases, uint64_t *out);
/* Multi-threaded variant (OpenMP-free: pthreads), used only as CPU baseline. */
void orc_ffor_column_decode_u32_mt(size_t n_vecs, const uint32_t *packed, const uint8_t *bit_widths,
                                   const uint64_t *vector_offsets, const uint32_t *bases, uint32_t *out,
                                   int n_threads);

/* ---- ALP constants (src/alp/constants.hpp:49-55, 76-144) ---------------- */
extern const float orc_alp_frac_f32[11];
extern const float orc_alp_exp_f32[11];
extern const int32_t orc_alp_fact_f32[10];
extern const double orc_alp_frac_f64[21];
extern const double orc_alp_exp_f64[24];
extern const int64_t orc_alp_fact_f64[19];

/* ---- ALP decode of one vector: falp + patch ----------------------------
 * out[i] = (T)(INT)((UINT)(unpack(i) + base) * (UINT)FACT[fac]) * FRAC[exp]
 * (src/flsgpu/alp.cuh:40-45, src/alp/falp.cpp bodies such as :95-130).
 * patch: out[positions[k]] = exceptions[k] (src/alp/decoder.hpp:137-145). */
void orc_alp_decode_vec_f32(const uint32_t *in, float *out, unsigned bw, uint32_t base,
                            unsigned fac, unsigned exp);
void orc_alp_decode_vec_f64(const uint64_t *in, double *out, unsigned bw, uint64_t base,
                            unsigned fac, unsigned exp);

/* Column-level ALP decode (src/alp/alp-bindings.cu:208-227, 229-255).
 * Standard layout: counts != NULL. G-ALP layout: offsets_counts != NULL
 * (count = sum of per-lane counts, alp-bindings.cu:241-245). */
void orc_alp_column_decode_f32(size_t n_vecs, const uint32_t *packed, const uint8_t *bit_widths,
                               const uint64_t *vector_offsets, const uint32_t *bases,
                               const uint8_t *factor_indices, const uint8_t *fraction_indices,
                               const uint64_t *exceptions_offsets, const float *exceptions,
                               const uint16_t *positions, const uint16_t *counts,
                               const uint16_t *offsets_counts, float *out);
void orc_alp_column_decode_f64(size_t n_vecs, const uint64_t *packed, const uint8_t *bit_widths,
                               const uint64_t *vector_offsets, const uint64_t *bases,
                               const uint8_t *factor_indices, const uint8_t *fraction_indices,
                               const uint64_t *exceptions_offsets, const double *exceptions,
                               const uint16_t *positions, const uint16_t *counts,
                               const uint16_t *offsets_counts, double *out);
void orc_alp_column_decode_f64_mt(size_t n_vecs, const uint64_t *packed, const uint8_t *bit_widths,
                                  const uint64_t *vector_offsets, const uint64_t *bases,
                                  const uint8_t *factor_indices, const uint8_t *fraction_indices,
                                  const uint64_t *exceptions_offsets, const double *exceptions,
                                  const uint16_t *positions, const uint16_t *counts, double *out,
                                  int n_threads);

/* ---- G-ALP lane-divided exception layout (src/flsgpu/structs.cuh:231-306) */
void orc_galp_convert_f32(size_t n_vecs, const float *exceptions, const uint16_t *positions,
                          const uint16_t *counts, float *out_exceptions, uint16_t *out_positions,
                          uint16_t *out_offsets_counts);
void orc_galp_convert_f64(size_t n_vecs, const double *exceptions, const uint16_t *positions,
                          const uint16_t *counts, double *out_exceptions, uint16_t *out_positions,
                          uint16_t *out_offsets_counts);

/* ---- ALP encoder (src/alp/encoder.hpp, sampler.hpp) --------------------- */
/* first_level_sample (sampler.hpp:11-48); returns number of samples. */
size_t orc_alp_sample_f32(const float *data, size_t offset, size_t n, float *sample);
size_t orc_alp_sample_f64(const double *data, size_t offset, size_t n, double *sample);
/* encode_value (encoder.hpp:76-84), SAFE variant when safe != 0. */
int32_t orc_alp_encode_value_f32(float v, unsigned fac, unsigned exp, int safe);
int64_t orc_alp_encode_value_f64(double v, unsigned fac, unsigned exp, int safe);
/* find_top_k_combinations (encoder.hpp:134-230). Writes up to 5 (exp,fac)
 * pairs; returns k, or 0 if the scheme switches to ALP_RD. */
int orc_alp_top_k_f32(const float *sample, size_t n_samples, uint8_t *exps, uint8_t *facs);
int orc_alp_top_k_f64(const double *sample, size_t n_samples, uint8_t *exps, uint8_t *facs);
/* Exhaustive search over one sample vector: the inner loop of
 * find_top_k_combinations for a single smp_n (encoder.hpp:145-197).
 * Returns the estimated size in bits; writes the best (exp, fac). */
uint64_t orc_alp_search_one_f32(const float *sample, size_t n_samples, uint8_t *exp, uint8_t *fac);
uint64_t orc_alp_search_one_f64(const double *sample, size_t n_samples, uint8_t *exp, uint8_t *fac);
/* Size estimate (bits) of one (exponent, factor) on the samples, as the
 * search computes it; UINT64_MAX when fewer than two values round-trip. */
uint64_t orc_alp_estimate_f32(const float *sample, size_t n_samples, unsigned exp, unsigned fac);
uint64_t orc_alp_estimate_f64(const double *sample, size_t n_samples, unsigned exp, unsigned fac);
/* find_best_exponent_factor_from_combinations (encoder.hpp:236-300). */
void orc_alp_best_combination_f32(const uint8_t *exps, const uint8_t *facs, int k, const float *vec,
                                  uint8_t *exp, uint8_t *fac);
void orc_alp_best_combination_f64(const uint8_t *exps, const uint8_t *facs, int k, const double *vec,
                                  uint8_t *exp, uint8_t *fac);
/* encode_simdized (encoder.hpp:302-379): returns the exception count. */
unsigned orc_alp_encode_vec_f32(const float *vec, unsigned fac, unsigned exp, int32_t *encoded,
                                float *exceptions, uint16_t *positions);
unsigned orc_alp_encode_vec_f64(const double *vec, unsigned fac, unsigned exp, int64_t *encoded,
                                double *exceptions, uint16_t *positions);
/* analyze_ffor (encoder.hpp:104-115). */
void orc_alp_analyze_ffor_f32(const int32_t *encoded, uint8_t *bw, int32_t *base);
void orc_alp_analyze_ffor_f64(const int64_t *encoded, uint8_t *bw, int64_t *base);

/* Column-level encode (src/alp/alp-bindings.cu:83-206). The caller provides
 * output arrays sized for the worst case: packed n_vecs*1024 words,
 * exceptions/positions n_vecs*1024. Returns 0 on success, 1 when the column
 * sampling chooses ALP_RD (the reference throws, alp-bindings.cu:76-78).
 * sizes[0] = packed words, sizes[1] = exceptions, sizes[2] = compressed bytes
 * ALP, sizes[3] = compressed bytes G-ALP (alp-bindings.cu:19-43). */
int orc_alp_column_encode_f32(const float *in, size_t n_vecs, uint32_t *packed, uint8_t *bit_widths,
                              uint64_t *vector_offsets, uint32_t *bases, uint8_t *factor_indices,
                              uint8_t *fraction_indices, uint64_t *exceptions_offsets,
                              float *exceptions, uint16_t *positions, uint16_t *counts,
                              uint64_t *sizes);
int orc_alp_column_encode_f64(const double *in, size_t n_vecs, uint64_t *packed, uint8_t *bit_widths,
                              uint64_t *vector_offsets, uint64_t *bases, uint8_t *factor_indices,
                              uint8_t *fraction_indices, uint64_t *exceptions_offsets,
                              double *exceptions, uint16_t *positions, uint16_t *counts,
                              uint64_t *sizes);

/* ---- ALP-RD (src/alp/rd.hpp) ---------------------------------------------
 * Dictionary search over a sample (rd.hpp:27-98). Ties in left-part counts
 * are broken by ascending left value (the reference leaves them to
 * unordered_map order + std::sort). Returns right_bw; writes dict, dict_size,
 * left_bw. */
unsigned orc_alprd_find_dictionary_f64(const double *sample, size_t n_samples, uint16_t *dict,
                                       unsigned *dict_size, unsigned *left_bw);
unsigned orc_alprd_find_dictionary_f32(const float *sample, size_t n_samples, uint16_t *dict,
                                       unsigned *dict_size, unsigned *left_bw);
/* Column encode/decode. Right parts: FFOR base 0 at right_bw (u64/u32
 * FastLanes); left parts: dictionary index at left_bw (u16 FastLanes);
 * exceptions: original left part (u16) + position (u16) per vector.
 * At exception positions the stored left index is 0. */
void orc_alprd_column_encode_f64(const double *in, size_t n_vecs, unsigned right_bw,
                                 const uint16_t *dict, unsigned dict_size, unsigned left_bw,
                                 uint64_t *right_packed, uint16_t *left_packed,
                                 uint64_t *exceptions_offsets, uint16_t *exceptions,
                                 uint16_t *positions, uint16_t *counts);
void orc_alprd_column_decode_f64(size_t n_vecs, unsigned right_bw, const uint16_t *dict,
                                 unsigned left_bw, const uint64_t *right_packed,
                                 const uint16_t *left_packed, const uint64_t *exceptions_offsets,
                                 const uint16_t *exceptions, const uint16_t *positions,
                                 const uint16_t *counts, double *out);
void orc_alprd_column_encode_f32(const float *in, size_t n_vecs, unsigned right_bw,
                                 const uint16_t *dict, unsigned dict_size, unsigned left_bw,
                                 uint32_t *right_packed, uint16_t *left_packed,
                                 uint64_t *exceptions_offsets, uint16_t *exceptions,
                                 uint16_t *positions, uint16_t *counts);
void orc_alprd_column_decode_f32(size_t n_vecs, unsigned right_bw, const uint16_t *dict,
                                 unsigned left_bw, const uint32_t *right_packed,
                                 const uint16_t *left_packed, const uint64_t *exceptions_offsets,
                                 const uint16_t *exceptions, const uint16_t *positions,
                                 const uint16_t *counts, float *out);

#ifdef __cplusplus
}
#endif
#endif
