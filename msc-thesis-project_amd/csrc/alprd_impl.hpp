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
