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

// Checks an ALP descriptor of the given layout and converts it to kernel
// arguments.
template <typename T>
int make_alp_args(const flsgpu_alp_column *c, int layout, AlpArgs<T> *a) {
  using UT = typename uint_of<T>::type;
  if (!c) return set_error(FLSGPU_EINVAL, "null column");
  const flsgpu_bp_column &bp = c->ffor.bp;
  if (bp.n_values != bp.n_vecs * VALUES_PER_VECTOR)
    return set_error(FLSGPU_EINVAL, "n_values (%llu) must equal n_vecs * 1024 (n_vecs = %llu)",
                     (unsigned long long)bp.n_values, (unsigned long long)bp.n_vecs);
  if (bp.n_vecs > 0) {
    if (!bp.packed_array || !bp.bit_widths || !bp.vector_offsets || !c->ffor.bases ||
        !c->factor_indices || !c->fraction_indices || !c->exceptions_offsets)
      return set_error(FLSGPU_EINVAL, "null device pointer in ALP column");
    if (layout == ALP_STD && !c->counts)
      return set_error(FLSGPU_EINVAL, "ALPColumn needs counts");
    if (layout == ALP_GALP && !c->offsets_counts)
      return set_error(FLSGPU_EINVAL, "ALPExtendedColumn needs offsets_counts");
    if (c->n_exceptions > 0 && (!c->exceptions || !c->positions))
      return set_error(FLSGPU_EINVAL, "null exception arrays");
  }
  a->ffor = FforArgs<UT>{static_cast<const UT *>(bp.packed_array), bp.bit_widths,
                         bp.vector_offsets, static_cast<const UT *>(c->ffor.bases), bp.n_vecs};
  a->factor_indices = c->factor_indices;
  a->fraction_indices = c->fraction_indices;
  a->exceptions_offsets = c->exceptions_offsets;
  a->exceptions = static_cast<const T *>(c->exceptions);
  a->positions = c->positions;
  a->counts = c->counts;
  a->offsets_counts = c->offsets_counts;
  return FLSGPU_OK;
}

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
