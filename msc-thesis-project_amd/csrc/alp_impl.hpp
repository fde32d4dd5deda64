// Declarations shared by the ALP kernel instantiation units (alp_inst.hip)
// and the ALP C ABI (alp_capi.hip).
#pragma once

#include "fls_impl.hpp"

namespace flsgpu_hip {

// Exception layouts: the reference's ALPColumn (per-vector counts, positions
// in ascending order, src/flsgpu/structs.cuh:50-65) and ALPExtendedColumn
// (G-ALP: per-lane (count << 10) | offset headers, structs.cuh:67-84, 231-306).
enum AlpLayout : int { ALP_STD = 0, ALP_GALP = 1 };

template <typename T>
struct AlpArgs {
  using UT = typename uint_of<T>::type;
  FforArgs<UT> ffor;
  const uint8_t *factor_indices;
  const uint8_t *fraction_indices;
  const uint64_t *exceptions_offsets;
  const T *exceptions;
  const uint16_t *positions;
  const uint16_t *counts;          // ALP_STD
  const uint16_t *offsets_counts;  // ALP_GALP
};

template <typename T>
struct MultiAlpArgs {
  AlpArgs<T> col[MAX_COLS];
  uint32_t n_cols;
};

template <typename T, int S, bool NT, int LAYOUT, bool PATCH>
struct AlpDecompress {
  static int launch(const AlpArgs<T> &a, T *out, hipStream_t stream);
};
// Decode through LDS, exceptions applied in LDS by the whole wave, 16-byte
// stores (out must be 16-byte aligned). PATCH false skips the exceptions.
template <typename T, int S, int LAYOUT, bool PATCH>
struct AlpDecompressStaged {
  static int launch(const AlpArgs<T> &a, T *out, hipStream_t stream);
};
template <typename T, int LAYOUT, bool PATCH>
struct AlpAny {
  static int launch(const AlpArgs<T> &a, T magic, uint8_t *d_found, hipStream_t stream);
};
template <typename T, int LAYOUT, bool PATCH>
struct AlpMulti {
  static int launch(const MultiAlpArgs<T> &a, T value, uint8_t *d_found, hipStream_t stream);
};

}  // namespace flsgpu_hip
