// Declarations shared by the FastLanes/FFOR kernel instantiation units
// (fls_inst.hip, compiled once per kernel family and variant so the
// per-bit-width specialisations build in parallel) and the C ABI (fls_capi.hip).
#pragma once

#include "common.hpp"

namespace flsgpu_hip {

template <typename UT>
struct FforArgs {
  const UT *packed;
  const uint8_t *bit_widths;
  const uint64_t *vector_offsets;
  const UT *bases;  // nullptr for plain bit-packing
  uint64_t n_vecs;
};

constexpr int MAX_COLS = 10;
template <typename UT>
struct MultiFforArgs {
  FforArgs<UT> col[MAX_COLS];
  uint32_t n_cols;
};

template <typename UT, int S, bool NT>
struct FlsDecompress {
  static int launch(const FforArgs<UT> &a, UT *out, hipStream_t stream);
};
// Decode through LDS with 16-byte stores (out must be 16-byte aligned).
template <typename UT, int S>
struct FlsDecompressStaged {
  static int launch(const FforArgs<UT> &a, UT *out, hipStream_t stream);
};
template <typename UT, int S>
struct FlsAny {
  static int launch(const FforArgs<UT> &a, UT magic, uint8_t *d_found, hipStream_t stream);
};
template <typename UT, int S, int PRED>
struct FlsFilterSum {
  static int launch(const FforArgs<UT> &a, UT lo, UT hi, uint64_t *d_result, hipStream_t stream);
};
template <typename UT>
struct FlsCompute {
  static int launch(const FforArgs<UT> &a, uint32_t reps, uint8_t *d_found, hipStream_t stream);
};
template <typename UT>
struct FlsMulti {
  static int launch(const MultiFforArgs<UT> &a, UT value, uint8_t *d_found, hipStream_t stream);
};

}  // namespace flsgpu_hip
