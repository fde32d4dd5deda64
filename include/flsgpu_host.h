/*
 * flsgpu host-side encoder, C ABI (CPU, multi-threaded). Produces the column
 * arrays the decoders in flsgpu.h read; byte-identical to the reference
 * encoder's output for the same input.
 *
 * All buffers are caller-allocated host memory sized for the worst case
 * (documented per field); the encoder fills them and reports the used sizes.
 * n_threads <= 0 uses every hardware thread.
 */
#ifndef FLSGPU_HOST_H
#define FLSGPU_HOST_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Positive status: column sampling chose ALP_RD (the reference throws
 * alp::EncodingException here, src/alp/alp-bindings.cu:63-81). */
#define FLSGPU_SCHEME_ALPRD 1

/* host::ALPColumn<T> (src/flsgpu/structs.cuh:192-230). */
typedef struct flsgpu_alp_host_column {
  uint64_t n_values;              /* out */
  uint64_t n_vecs;                /* out */
  void *packed;                   /* >= n_values words of the float's width */
  uint8_t *bit_widths;            /* [n_vecs] */
  uint64_t *vector_offsets;       /* [n_vecs] */
  void *bases;                    /* [n_vecs] */
  uint8_t *factor_indices;        /* [n_vecs] */
  uint8_t *fraction_indices;      /* [n_vecs] */
  uint64_t *exceptions_offsets;   /* [n_vecs] */
  void *exceptions;               /* >= n_values */
  uint16_t *positions;            /* >= n_values */
  uint16_t *counts;               /* [n_vecs] */
  uint64_t n_packed_words;        /* out */
  uint64_t n_exceptions;          /* out */
  uint64_t compressed_bytes_alp;  /* out: alp-bindings.cu:19-43 size model */
  uint64_t compressed_bytes_galp; /* out */
} flsgpu_alp_host_column;

/* ALP-RD column (format of flsgpu_alprd_column in flsgpu.h). */
typedef struct flsgpu_alprd_host_column {
  uint64_t n_values;              /* out */
  uint64_t n_vecs;                /* out */
  uint32_t right_bit_width;       /* out */
  uint32_t left_bit_width;        /* out */
  uint32_t dictionary_size;       /* out */
  uint16_t dictionary[8];         /* out */
  void *right_packed;             /* >= n_vecs * (W-1) * N_LANES words */
  uint16_t *left_packed;          /* >= n_vecs * 3 * 64 */
  uint64_t *exceptions_offsets;   /* [n_vecs] */
  uint16_t *exceptions;           /* >= n_values */
  uint16_t *positions;            /* >= n_values */
  uint16_t *counts;               /* [n_vecs] */
  uint64_t n_exceptions;          /* out */
  uint64_t compressed_bytes;      /* out */
} flsgpu_alprd_host_column;

const char *flsgpu_encoder_last_error(void);

/* alp::encode (src/alp/alp-bindings.cu:83-206): first-rowgroup sampling and
 * top-k (exponent, factor) search, per-vector choice among the top k,
 * exceptions, FFOR at base = vector minimum. n_values % 1024 == 0. */
int flsgpu_alp_encode_f32(const float *in, uint64_t n_values, flsgpu_alp_host_column *col,
                          int n_threads);
int flsgpu_alp_encode_f64(const double *in, uint64_t n_values, flsgpu_alp_host_column *col,
                          int n_threads);
/* Same column assembly with caller-chosen per-vector (exponent, factor)
 * (e.g. from the learned selector); exponent <= 10 / 18, factor <= exponent. */
int flsgpu_alp_encode_with_params_f32(const float *in, uint64_t n_values, const uint8_t *exps,
                                      const uint8_t *facs, flsgpu_alp_host_column *col,
                                      int n_threads);
int flsgpu_alp_encode_with_params_f64(const double *in, uint64_t n_values, const uint8_t *exps,
                                      const uint8_t *facs, flsgpu_alp_host_column *col,
                                      int n_threads);
/* Exhaustive (exponent, factor) search on 32 samples per vector (the inner
 * search of find_top_k_combinations, src/alp/encoder.hpp:145-197).
 * samples: [n_vecs * 32]; est_bits may be NULL. */
int flsgpu_alp_search_f32(const float *samples, uint64_t n_vecs, uint8_t *exps, uint8_t *facs,
                          uint64_t *est_bits, int n_threads);
int flsgpu_alp_search_f64(const double *samples, uint64_t n_vecs, uint8_t *exps, uint8_t *facs,
                          uint64_t *est_bits, int n_threads);
/* ALPColumn -> ALPExtendedColumn exception layout (src/flsgpu/structs.cuh:231-306):
 * per vector, stable partition by lane; offsets_counts[v*N_LANES + l] =
 * (count << 10) | offset. Fails (FLSGPU_EINVAL) when a lane holds more than
 * 63 exceptions, which the 16-bit header cannot encode. */
int flsgpu_galp_convert_f32(uint64_t n_vecs, const float *exceptions, const uint16_t *positions,
                            const uint16_t *counts, const uint64_t *exceptions_offsets,
                            float *out_exceptions, uint16_t *out_positions,
                            uint16_t *out_offsets_counts);
int flsgpu_galp_convert_f64(uint64_t n_vecs, const double *exceptions, const uint16_t *positions,
                            const uint16_t *counts, const uint64_t *exceptions_offsets,
                            double *out_exceptions, uint16_t *out_positions,
                            uint16_t *out_offsets_counts);
/* ALP-RD encode (src/alp/rd.hpp:27-141), one dictionary per column from the
 * first-rowgroup sample; equal left-part counts are ordered by value. */
int flsgpu_alprd_encode_f32(const float *in, uint64_t n_values, flsgpu_alprd_host_column *col,
                            int n_threads);
int flsgpu_alprd_encode_f64(const double *in, uint64_t n_values, flsgpu_alprd_host_column *col,
                            int n_threads);
/* FastLanes FFOR packing (src/fls/ffor.cpp:29771-29966) of n_vecs vectors at
 * the given per-vector bit widths; bases may be NULL (plain bit-packing).
 * packed must hold sum(bit_widths) * N_LANES words. */
int flsgpu_ffor_pack_u32(const uint32_t *values, uint64_t n_vecs, const uint8_t *bit_widths,
                         const uint32_t *bases, uint32_t *packed, uint64_t *vector_offsets);
int flsgpu_ffor_pack_u64(const uint64_t *values, uint64_t n_vecs, const uint8_t *bit_widths,
                         const uint64_t *bases, uint64_t *packed, uint64_t *vector_offsets);

#ifdef __cplusplus
}
#endif
#endif /* FLSGPU_HOST_H */
