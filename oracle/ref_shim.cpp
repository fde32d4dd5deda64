// Test infrastructure only: an extern "C" face over the reference's own CPU
// codec, compiled from the sources where they lie under the reference
// checkout (src/fls/*.cpp, src/alp/falp.cpp and the header-only
// src/alp/{encoder,decoder,sampler,rd}.hpp). Built by oracle/Makefile into
// oracle/_ref/libflsref.so; used by tests to pin the C oracle and, optionally,
// as the CPU baseline. No reference source is copied into this repository.
//
// The column assembly of alp::encode (src/alp/alp-bindings.cu:83-206) lives
// in a CUDA translation unit that does not compile without CUDA headers, so
// ref_alp_column_encode_* below drives the reference's own per-vector
// encoder calls in the same order that file does.
#include <cstdint>
#include <cstring>
#include <vector>

#include "alp/encoder.hpp"
#include "alp/falp.hpp"
#include "alp/rd.hpp"
#include "fls/fls-bindings.hpp"

extern "C" {

#define FLS_WRAP(UT, SUF)                                                                          \
  void ref_fls_pack_##SUF(const UT *in, UT *out, unsigned bw) { fls::pack(in, out, (uint8_t)bw); } \
  void ref_fls_unpack_##SUF(const UT *in, UT *out, unsigned bw) {                                  \
    fls::unpack(in, out, (uint8_t)bw);                                                             \
  }                                                                                                \
  void ref_fls_ffor_##SUF(const UT *in, UT *out, unsigned bw, UT base) {                           \
    fls::ffor(in, out, (uint8_t)bw, &base);                                                        \
  }                                                                                                \
  void ref_fls_unffor_##SUF(const UT *in, UT *out, unsigned bw, UT base) {                         \
    fls::unffor(in, out, (uint8_t)bw, &base);                                                      \
  }

FLS_WRAP(uint8_t, u8)
FLS_WRAP(uint16_t, u16)
FLS_WRAP(uint32_t, u32)
FLS_WRAP(uint64_t, u64)

void ref_ffor_column_decode_u32(size_t n_vecs, const uint32_t *packed, const uint8_t *bit_widths,
                                const uint64_t *vector_offsets, const uint32_t *bases,
                                uint32_t *out) {
  for (size_t v = 0; v < n_vecs; ++v) {
    fls::unffor(packed + vector_offsets[v], out + v * 1024, bit_widths[v], &bases[v]);
  }
}

void ref_ffor_column_decode_u64(size_t n_vecs, const uint64_t *packed, const uint8_t *bit_widths,
                                const uint64_t *vector_offsets, const uint64_t *bases,
                                uint64_t *out) {
  for (size_t v = 0; v < n_vecs; ++v) {
    fls::unffor(packed + vector_offsets[v], out + v * 1024, bit_widths[v], &bases[v]);
  }
}

void ref_falp_f32(const uint32_t *in, float *out, unsigned bw, uint32_t base, unsigned fac,
                  unsigned exp) {
  generated::falp::fallback::scalar::falp(in, out, (uint8_t)bw, &base, (uint8_t)fac, (uint8_t)exp);
}

void ref_falp_f64(const uint64_t *in, double *out, unsigned bw, uint64_t base, unsigned fac,
                  unsigned exp) {
  generated::falp::fallback::scalar::falp(in, out, (uint8_t)bw, &base, (uint8_t)fac, (uint8_t)exp);
}

void ref_patch_f32(float *out, const float *exc, const uint16_t *pos, uint16_t count) {
  alp::decoder<float>::patch_exceptions(out, exc, pos, &count);
}
void ref_patch_f64(double *out, const double *exc, const uint16_t *pos, uint16_t count) {
  alp::decoder<double>::patch_exceptions(out, exc, pos, &count);
}

// The reference's CPU decode of an ALPColumn (src/alp/alp-bindings.cu:208-228):
// per vector, falp then patch_exceptions, for vectors [v0, v1); vector v goes
// to out + v * 1024. A range lets a caller spread one column over threads.
#define ALP_COLUMN_DECODE(T, UT, SUF)                                                              \
  void ref_alp_column_decode_##SUF(size_t v0, size_t v1, const UT *packed,                         \
                                   const uint8_t *bit_widths, const uint64_t *vector_offsets,     \
                                   const UT *bases, const uint8_t *factor_indices,               \
                                   const uint8_t *fraction_indices, const T *exceptions,         \
                                   const uint16_t *positions, const uint64_t *exceptions_offsets, \
                                   const uint16_t *counts, T *out) {                             \
    for (size_t v = v0; v < v1; ++v) {                                                             \
      generated::falp::fallback::scalar::falp(packed + vector_offsets[v], out + v * 1024,          \
                                              bit_widths[v], &bases[v], factor_indices[v],         \
                                              fraction_indices[v]);                                \
      alp::decoder<T>::patch_exceptions(out + v * 1024, exceptions + exceptions_offsets[v],        \
                                        positions + exceptions_offsets[v], &counts[v]);           \
    }                                                                                              \
  }
ALP_COLUMN_DECODE(float, uint32_t, f32)
ALP_COLUMN_DECODE(double, uint64_t, f64)

size_t ref_alp_sample_f32(const float *data, size_t offset, size_t n, float *sample) {
  return alp::sampler::first_level_sample<float>(data, offset, n, sample);
}
size_t ref_alp_sample_f64(const double *data, size_t offset, size_t n, double *sample) {
  return alp::sampler::first_level_sample<double>(data, offset, n, sample);
}

int32_t ref_alp_encode_value_f32(float v, unsigned fac, unsigned exp) {
  return alp::encoder<float>::encode_value<true>(v, (uint8_t)fac, (uint8_t)exp);
}
int64_t ref_alp_encode_value_f64(double v, unsigned fac, unsigned exp) {
  return alp::encoder<double>::encode_value<true>(v, (uint8_t)fac, (uint8_t)exp);
}

// Runs encoder<T>::init on a column; returns k (0 when the scheme is ALP_RD).
#define TOPK_WRAP(T, SUF)                                                                          \
  int ref_alp_top_k_##SUF(const T *data, size_t n, uint8_t *exps, uint8_t *facs) {                 \
    std::vector<T> sample(n + 64);                                                                 \
    alp::state<T> st;                                                                              \
    alp::encoder<T>::init(data, 0, n, sample.data(), st);                                          \
    if (st.scheme != alp::Scheme::ALP) return 0;                                                   \
    for (size_t i = 0; i < st.k_combinations; ++i) {                                               \
      exps[i] = (uint8_t)st.best_k_combinations[i].first;                                          \
      facs[i] = (uint8_t)st.best_k_combinations[i].second;                                         \
    }                                                                                              \
    return (int)st.k_combinations;                                                                 \
  }
TOPK_WRAP(float, f32)
TOPK_WRAP(double, f64)

#define BEST_WRAP(T, SUF)                                                                          \
  void ref_alp_best_combination_##SUF(const uint8_t *exps, const uint8_t *facs, int k,             \
                                      const T *vec, uint8_t *exp, uint8_t *fac) {                  \
    std::vector<std::pair<int, int>> comb;                                                         \
    for (int i = 0; i < k; ++i) comb.emplace_back(exps[i], facs[i]);                               \
    alp::encoder<T>::find_best_exponent_factor_from_combinations(comb, (uint8_t)k, vec, 1024,      \
                                                                 *fac, *exp);                      \
  }
BEST_WRAP(float, f32)
BEST_WRAP(double, f64)

#define ENCVEC_WRAP(T, ST, SUF)                                                                    \
  unsigned ref_alp_encode_vec_##SUF(const T *vec, unsigned fac, unsigned exp, ST *encoded,         \
                                    T *exceptions, uint16_t *positions) {                          \
    uint16_t count = 0;                                                                            \
    alp::encoder<T>::encode_simdized(vec, exceptions, positions, &count, encoded, (uint8_t)fac,    \
                                     (uint8_t)exp);                                                \
    return count;                                                                                  \
  }                                                                                                \
  void ref_alp_analyze_ffor_##SUF(const ST *encoded, uint8_t *bw, ST *base) {                      \
    alp::encoder<T>::analyze_ffor(encoded, *bw, base);                                             \
  }
ENCVEC_WRAP(float, int32_t, f32)
ENCVEC_WRAP(double, int64_t, f64)

// Column encode in the order of alp-bindings.cu:83-206 (configure_alpstate,
// then per vector encode + analyze_ffor + fls::ffor). Returns 1 on ALP_RD.
#define COLENC_WRAP(T, ST, UT, SUF, N_LANES)                                                       \
  int ref_alp_column_encode_##SUF(const T *in, size_t n_vecs, UT *packed, uint8_t *bit_widths,     \
                                  uint64_t *vector_offsets, UT *bases, uint8_t *factor_indices,    \
                                  uint8_t *fraction_indices, uint64_t *exceptions_offsets,         \
                                  T *exceptions, uint16_t *positions, uint16_t *counts) {          \
    const size_t n = n_vecs * 1024;                                                                \
    std::vector<T> sample(n + 64);                                                                 \
    alp::state<T> st;                                                                              \
    alp::encoder<T>::init(in, 0, n, sample.data(), st);                                            \
    if (st.scheme != alp::Scheme::ALP) return 1;                                                   \
    std::vector<ST> enc(1024);                                                                     \
    size_t w_off = 0, e_off = 0;                                                                   \
    for (size_t v = 0; v < n_vecs; ++v) {                                                          \
      alp::encoder<T>::encode(in + v * 1024, exceptions + e_off, positions + e_off, &counts[v],    \
                              enc.data(), st);                                                     \
      ST base;                                                                                     \
      alp::encoder<T>::analyze_ffor(enc.data(), bit_widths[v], &base);                             \
      bases[v] = (UT)base;                                                                         \
      fls::ffor(reinterpret_cast<const UT *>(enc.data()), packed + w_off, bit_widths[v],           \
                &bases[v]);                                                                        \
      factor_indices[v] = st.fac;                                                                  \
      fraction_indices[v] = st.exp;                                                                \
      vector_offsets[v] = w_off;                                                                   \
      exceptions_offsets[v] = e_off;                                                               \
      w_off += (size_t)bit_widths[v] * (N_LANES);                                                  \
      e_off += counts[v];                                                                          \
    }                                                                                              \
    return 0;                                                                                      \
  }
COLENC_WRAP(float, int32_t, uint32_t, f32, 32)
COLENC_WRAP(double, int64_t, uint64_t, f64, 16)

// ALP-RD dictionary search (rd.hpp:83-98) on a sample; returns right_bw and
// the dictionary in the reference's order.
#define RD_WRAP(T, SUF)                                                                            \
  unsigned ref_alprd_find_dictionary_##SUF(const T *sample, size_t n, uint16_t *dict,              \
                                           unsigned *dict_size, unsigned *left_bw) {               \
    alp::state<T> st;                                                                              \
    st.sampled_values_n = n;                                                                       \
    std::vector<T> s(sample, sample + n);                                                          \
    alp::rd_encoder<T>::find_best_dictionary(s.data(), st);                                        \
    for (unsigned i = 0; i < st.actual_dictionary_size; ++i) dict[i] = st.left_parts_dict[i];      \
    *dict_size = st.actual_dictionary_size;                                                        \
    *left_bw = st.left_bit_width;                                                                  \
    return st.right_bit_width;                                                                     \
  }
RD_WRAP(double, f64)
RD_WRAP(float, f32)

// ALP-RD decode of one vector (rd.hpp:146-172) from unpacked parts.
void ref_alprd_decode_vec_f64(double *out, uint64_t *right, uint16_t *left, uint16_t *exceptions,
                              uint16_t *positions, uint16_t count, const uint16_t *dict,
                              unsigned dict_size, unsigned right_bw) {
  alp::state<double> st;
  st.right_bit_width = (uint8_t)right_bw;
  st.actual_dictionary_size = (uint8_t)dict_size;
  for (unsigned i = 0; i < dict_size; ++i) st.left_parts_dict[i] = dict[i];
  alp::rd_encoder<double>::decode(out, right, left, exceptions, positions, &count, st);
}

}  // extern "C"
