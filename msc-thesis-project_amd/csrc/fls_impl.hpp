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

// Checks a BP/FFOR descriptor (bases may be null for plain bit-packing) and
// converts it to kernel arguments.
template <typename UT>
int make_ffor_args(const flsgpu_bp_column *bp, const void *bases, FforArgs<UT> *a) {
  if (!bp) return set_error(FLSGPU_EINVAL, "null column");
  if (bp->n_values != bp->n_vecs * VALUES_PER_VECTOR)
    return set_error(FLSGPU_EINVAL, "n_values (%llu) must equal n_vecs * 1024 (n_vecs = %llu)",
                     (unsigned long long)bp->n_values, (unsigned long long)bp->n_vecs);
  if (bp->n_vecs > 0 && (!bp->packed_array || !bp->bit_widths || !bp->vector_offsets))
    return set_error(FLSGPU_EINVAL, "null device pointer in column");
  *a = FforArgs<UT>{static_cast<const UT *>(bp->packed_array), bp->bit_widths, bp->vector_offsets,
                    static_cast<const UT *>(bases), bp->n_vecs};
  return FLSGPU_OK;
}

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
  // The same two launches over partials the caller provides: room for
  // `slots` blocks (the grid is capped at that); allocates nothing.
  static int launch_partials(const FforArgs<UT> &a, UT lo, UT hi, uint64_t *d_result,
                             uint64_t *partials, unsigned slots, hipStream_t stream);
  // The scan launch alone (its block partials in `partials`, *grid_out
  // blocks).
  static int launch_scan(const FforArgs<UT> &a, UT lo, UT hi, uint64_t *partials, unsigned slots,
                         hipStream_t stream, unsigned *grid_out);
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
