// Declarations shared by the ALP-RD kernel instantiation unit and its C ABI.
#pragma once

#include "common.hpp"

namespace flsgpu_hip {

template <typename T>
struct AlprdArgs {
  using UT = typename uint_of<T>::type;
  uint64_t n_vecs;
  unsigned right_bw, left_bw, dict_size;
  uint16_t dict[8];
  const UT *right;
  const uint16_t *left;
  const uint64_t *exceptions_offsets;
  const uint16_t *exceptions;
  const uint16_t *positions;
  const uint16_t *counts;
};

// Checks an ALP-RD descriptor and converts it to kernel arguments. The
// exception arrays are read only for vectors whose count is nonzero, which
// the host cannot see, so they must be valid pointers whenever n_vecs > 0.
template <typename T>
int make_rd_args(const flsgpu_alprd_column *c, AlprdArgs<T> *a) {
  constexpr unsigned W = sizeof(T) * 8;
  if (!c) return set_error(FLSGPU_EINVAL, "null column");
  if (c->n_values != c->n_vecs * 1024)
    return set_error(FLSGPU_EINVAL, "n_values must equal n_vecs * 1024");
  if (c->right_bit_width < W - 16 || c->right_bit_width > W - 1)
    return set_error(FLSGPU_EINVAL, "right_bit_width %u outside %u..%u", c->right_bit_width,
                     W - 16, W - 1);
  if (c->left_bit_width < 1 || c->left_bit_width > 3)
    return set_error(FLSGPU_EINVAL, "left_bit_width %u outside 1..3", c->left_bit_width);
  if (c->dictionary_size < 1 || c->dictionary_size > 8 ||
      c->dictionary_size > (1u << c->left_bit_width))
    return set_error(FLSGPU_EINVAL, "dictionary_size %u invalid", c->dictionary_size);
  if (c->n_vecs && (!c->right_packed || !c->left_packed || !c->exceptions_offsets ||
                    !c->counts || !c->exceptions || !c->positions))
    return set_error(FLSGPU_EINVAL, "null device pointer in ALP-RD column");
  a->n_vecs = c->n_vecs;
  a->right_bw = c->right_bit_width;
  a->left_bw = c->left_bit_width;
  a->dict_size = c->dictionary_size;
  for (int i = 0; i < 8; ++i) a->dict[i] = c->dictionary[i];
  a->right = static_cast<const typename uint_of<T>::type *>(c->right_packed);
  a->left = c->left_packed;
  a->exceptions_offsets = c->exceptions_offsets;
  a->exceptions = c->exceptions;
  a->positions = c->positions;
  a->counts = c->counts;
  return FLSGPU_OK;
}

template <typename T, bool NT>
struct AlprdDecompress {
  static int launch(const AlprdArgs<T> &a, T *out, hipStream_t stream);
};

// Decode through LDS in S slot parts per lane, exceptions applied in LDS,
// 16-byte non-temporal stores (out must be 16-byte aligned).
template <typename T, int S>
struct AlprdDecompressStaged {
  static int launch(const AlprdArgs<T> &a, T *out, hipStream_t stream);
};

}  // namespace flsgpu_hip
