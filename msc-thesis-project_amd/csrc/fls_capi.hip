// C-ABI entry points of the FastLanes / FFOR kernels: argument checking,
// variant resolution and dispatch to the per-variant launchers. Mirrors
// bindings::{decompress,query,compute,query_multi}_column for BP/FFOR columns
// (src/generated-bindings/kernel-bindings.cuh:11-175), with status codes in
// place of exceptions.
#include "fls_impl.hpp"

namespace flsgpu_hip {

struct VariantShape {
  int S;
  bool nt;
  bool staged;
};

static int resolve_variant(int variant, int dtype_bits, VariantShape *vs) {
  if (variant == FLSGPU_VARIANT_AUTO)
    variant = flsgpu_get_tuning(dtype_bits == 32 ? FLSGPU_ENC_FFOR_U32 : FLSGPU_ENC_FFOR_U64, 255);
  switch (variant) {
    case FLSGPU_VARIANT_LANE: *vs = {1, false, false}; return FLSGPU_OK;
    case FLSGPU_VARIANT_LANE_NT: *vs = {1, true, false}; return FLSGPU_OK;
    case FLSGPU_VARIANT_HALF_LANE: *vs = {2, false, false}; return FLSGPU_OK;
    case FLSGPU_VARIANT_HALF_LANE_NT: *vs = {2, true, false}; return FLSGPU_OK;
    case FLSGPU_VARIANT_QUARTER_LANE: *vs = {4, false, false}; return FLSGPU_OK;
    case FLSGPU_VARIANT_QUARTER_LANE_NT: *vs = {4, true, false}; return FLSGPU_OK;
    case FLSGPU_VARIANT_HALF_LANE_STAGED: *vs = {dtype_bits == 64 ? 4 : 2, true, true}; return FLSGPU_OK;
    case FLSGPU_VARIANT_QUARTER_LANE_STAGED: *vs = {4, true, true}; return FLSGPU_OK;
    case FLSGPU_VARIANT_EIGHTH_LANE_STAGED: *vs = {8, true, true}; return FLSGPU_OK;
    default: return set_error(FLSGPU_EUNSUPPORTED, "unknown variant %d", variant);
  }
}

template <typename UT>
static int launch_decompress(const FforArgs<UT> &a, UT *out, int variant, hipStream_t stream) {
  if (!out && a.n_vecs) return set_error(FLSGPU_EINVAL, "null output");
  VariantShape vs;
  int rc = resolve_variant(variant, lanes<UT>::W, &vs);
  if (rc) return rc;
  if (vs.staged && (reinterpret_cast<uintptr_t>(out) % 16) == 0) {
    if constexpr (lanes<UT>::W == 32) {
      if (vs.S == 2) return FlsDecompressStaged<UT, 2>::launch(a, out, stream);
    }
    if (vs.S == 8) return FlsDecompressStaged<UT, 8>::launch(a, out, stream);
    return FlsDecompressStaged<UT, 4>::launch(a, out, stream);
  }
  if (vs.S == 8) vs.S = 4;  // unaligned output: the NT variant of the nearest split
  if (vs.S == 1) return vs.nt ? FlsDecompress<UT, 1, true>::launch(a, out, stream)
                              : FlsDecompress<UT, 1, false>::launch(a, out, stream);
  if (vs.S == 2) return vs.nt ? FlsDecompress<UT, 2, true>::launch(a, out, stream)
                              : FlsDecompress<UT, 2, false>::launch(a, out, stream);
  return vs.nt ? FlsDecompress<UT, 4, true>::launch(a, out, stream)
               : FlsDecompress<UT, 4, false>::launch(a, out, stream);
}

template <typename UT>
static int launch_any(const FforArgs<UT> &a, UT magic, uint8_t *d_found, int variant,
                      hipStream_t stream) {
  if (!d_found) return set_error(FLSGPU_EINVAL, "null result pointer");
  VariantShape vs;
  int rc = resolve_variant(variant, lanes<UT>::W, &vs);
  if (rc) return rc;
  return FlsAny<UT, 1>::launch(a, magic, d_found, stream);
}

template <typename UT>
static int launch_filter_sum(const FforArgs<UT> &a, int predicate, uint64_t lo, uint64_t hi,
                             uint64_t *d_result, int variant, hipStream_t stream) {
  if (!d_result) return set_error(FLSGPU_EINVAL, "null result pointer");
  VariantShape vs;
  int rc = resolve_variant(variant, lanes<UT>::W, &vs);
  if (rc) return rc;
  if (predicate == FLSGPU_PRED_RANGE)
    return FlsFilterSum<UT, 1, FLSGPU_PRED_RANGE>::launch(a, UT(lo), UT(hi), d_result, stream);
  if (predicate == FLSGPU_PRED_OFFSET_LOWER_HALF)
    return FlsFilterSum<UT, 1, FLSGPU_PRED_OFFSET_LOWER_HALF>::launch(a, UT(lo), UT(hi), d_result,
                                                                     stream);
  if (predicate == FLSGPU_PRED_ALL)
    return FlsFilterSum<UT, 1, FLSGPU_PRED_ALL>::launch(a, UT(lo), UT(hi), d_result, stream);
  return set_error(FLSGPU_EINVAL, "unknown predicate %d", predicate);
}

template <typename UT>
static int launch_compute(const FforArgs<UT> &a, uint32_t reps, uint8_t *d_found, int variant,
                          hipStream_t stream) {
  if (!d_found) return set_error(FLSGPU_EINVAL, "null result pointer");
  VariantShape vs;
  int rc = resolve_variant(variant, lanes<UT>::W, &vs);
  if (rc) return rc;
  return FlsCompute<UT>::launch(a, reps, d_found, stream);
}

template <typename UT>
static int launch_multi(const flsgpu_ffor_column *cols, uint32_t n_cols, UT value,
                        uint8_t *d_found, hipStream_t stream) {
  if (!cols || !d_found) return set_error(FLSGPU_EINVAL, "null argument");
  if (n_cols < 1 || n_cols > MAX_COLS)
    return set_error(FLSGPU_EINVAL, "n_cols must be in 1..%d, got %u", MAX_COLS, n_cols);
  MultiFforArgs<UT> m{};
  m.n_cols = n_cols;
  for (uint32_t c = 0; c < n_cols; ++c) {
    if (!cols[c].bases && cols[c].bp.n_vecs) return set_error(FLSGPU_EINVAL, "null bases");
    int rc = make_ffor_args<UT>(&cols[c].bp, cols[c].bases, &m.col[c]);
    if (rc) return rc;
    if (m.col[c].n_vecs != m.col[0].n_vecs)
      return set_error(FLSGPU_EINVAL, "all columns must have the same number of vectors");
  }
  return FlsMulti<UT>::launch(m, value, d_found, stream);
}

}  // namespace flsgpu_hip

using namespace flsgpu_hip;

extern "C" {

int flsgpu_decompress_bp_u32(const flsgpu_bp_column *col, uint32_t *out, int variant,
                             flsgpu_stream_t stream) {
  FforArgs<uint32_t> a;
  int rc = make_ffor_args<uint32_t>(col, nullptr, &a);
  return rc ? rc : launch_decompress(a, out, variant, stream);
}
int flsgpu_decompress_bp_u64(const flsgpu_bp_column *col, uint64_t *out, int variant,
                             flsgpu_stream_t stream) {
  FforArgs<uint64_t> a;
  int rc = make_ffor_args<uint64_t>(col, nullptr, &a);
  return rc ? rc : launch_decompress(a, out, variant, stream);
}
int flsgpu_decompress_ffor_u32(const flsgpu_ffor_column *col, uint32_t *out, int variant,
                               flsgpu_stream_t stream) {
  if (!col || (!col->bases && col->bp.n_vecs)) return set_error(FLSGPU_EINVAL, "null column/bases");
  FforArgs<uint32_t> a;
  int rc = make_ffor_args<uint32_t>(&col->bp, col->bases, &a);
  return rc ? rc : launch_decompress(a, out, variant, stream);
}
int flsgpu_decompress_ffor_u64(const flsgpu_ffor_column *col, uint64_t *out, int variant,
                               flsgpu_stream_t stream) {
  if (!col || (!col->bases && col->bp.n_vecs)) return set_error(FLSGPU_EINVAL, "null column/bases");
  FforArgs<uint64_t> a;
  int rc = make_ffor_args<uint64_t>(&col->bp, col->bases, &a);
  return rc ? rc : launch_decompress(a, out, variant, stream);
}

int flsgpu_query_any_bp_u32(const flsgpu_bp_column *col, uint32_t magic, uint8_t *d_found,
                            int variant, flsgpu_stream_t stream) {
  FforArgs<uint32_t> a;
  int rc = make_ffor_args<uint32_t>(col, nullptr, &a);
  return rc ? rc : launch_any(a, magic, d_found, variant, stream);
}
int flsgpu_query_any_bp_u64(const flsgpu_bp_column *col, uint64_t magic, uint8_t *d_found,
                            int variant, flsgpu_stream_t stream) {
  FforArgs<uint64_t> a;
  int rc = make_ffor_args<uint64_t>(col, nullptr, &a);
  return rc ? rc : launch_any(a, magic, d_found, variant, stream);
}
int flsgpu_query_any_ffor_u32(const flsgpu_ffor_column *col, uint32_t magic, uint8_t *d_found,
                              int variant, flsgpu_stream_t stream) {
  if (!col || (!col->bases && col->bp.n_vecs)) return set_error(FLSGPU_EINVAL, "null column/bases");
  FforArgs<uint32_t> a;
  int rc = make_ffor_args<uint32_t>(&col->bp, col->bases, &a);
  return rc ? rc : launch_any(a, magic, d_found, variant, stream);
}
int flsgpu_query_any_ffor_u64(const flsgpu_ffor_column *col, uint64_t magic, uint8_t *d_found,
                              int variant, flsgpu_stream_t stream) {
  if (!col || (!col->bases && col->bp.n_vecs)) return set_error(FLSGPU_EINVAL, "null column/bases");
  FforArgs<uint64_t> a;
  int rc = make_ffor_args<uint64_t>(&col->bp, col->bases, &a);
  return rc ? rc : launch_any(a, magic, d_found, variant, stream);
}

int flsgpu_query_filter_sum_bp_u32(const flsgpu_bp_column *col, int predicate, uint64_t lo,
                                   uint64_t hi, uint64_t *d_result, int variant,
                                   flsgpu_stream_t stream) {
  FforArgs<uint32_t> a;
  int rc = make_ffor_args<uint32_t>(col, nullptr, &a);
  return rc ? rc : launch_filter_sum(a, predicate, lo, hi, d_result, variant, stream);
}
int flsgpu_query_filter_sum_bp_u64(const flsgpu_bp_column *col, int predicate, uint64_t lo,
                                   uint64_t hi, uint64_t *d_result, int variant,
                                   flsgpu_stream_t stream) {
  FforArgs<uint64_t> a;
  int rc = make_ffor_args<uint64_t>(col, nullptr, &a);
  return rc ? rc : launch_filter_sum(a, predicate, lo, hi, d_result, variant, stream);
}
int flsgpu_query_filter_sum_ffor_u32(const flsgpu_ffor_column *col, int predicate, uint64_t lo,
                                     uint64_t hi, uint64_t *d_result, int variant,
                                     flsgpu_stream_t stream) {
  if (!col || (!col->bases && col->bp.n_vecs)) return set_error(FLSGPU_EINVAL, "null column/bases");
  FforArgs<uint32_t> a;
  int rc = make_ffor_args<uint32_t>(&col->bp, col->bases, &a);
  return rc ? rc : launch_filter_sum(a, predicate, lo, hi, d_result, variant, stream);
}
int flsgpu_query_filter_sum_ffor_u64(const flsgpu_ffor_column *col, int predicate, uint64_t lo,
                                     uint64_t hi, uint64_t *d_result, int variant,
                                     flsgpu_stream_t stream) {
  if (!col || (!col->bases && col->bp.n_vecs)) return set_error(FLSGPU_EINVAL, "null column/bases");
  FforArgs<uint64_t> a;
  int rc = make_ffor_args<uint64_t>(&col->bp, col->bases, &a);
  return rc ? rc : launch_filter_sum(a, predicate, lo, hi, d_result, variant, stream);
}

int flsgpu_compute_bp_u32(const flsgpu_bp_column *col, uint32_t n_repetitions,
                          uint8_t *d_result, int variant, flsgpu_stream_t stream) {
  FforArgs<uint32_t> a;
  int rc = make_ffor_args<uint32_t>(col, nullptr, &a);
  return rc ? rc : launch_compute(a, n_repetitions, d_result, variant, stream);
}
int flsgpu_compute_bp_u64(const flsgpu_bp_column *col, uint32_t n_repetitions,
                          uint8_t *d_result, int variant, flsgpu_stream_t stream) {
  FforArgs<uint64_t> a;
  int rc = make_ffor_args<uint64_t>(col, nullptr, &a);
  return rc ? rc : launch_compute(a, n_repetitions, d_result, variant, stream);
}
int flsgpu_compute_ffor_u32(const flsgpu_ffor_column *col, uint32_t n_repetitions,
                            uint8_t *d_result, int variant, flsgpu_stream_t stream) {
  if (!col || (!col->bases && col->bp.n_vecs)) return set_error(FLSGPU_EINVAL, "null column/bases");
  FforArgs<uint32_t> a;
  int rc = make_ffor_args<uint32_t>(&col->bp, col->bases, &a);
  return rc ? rc : launch_compute(a, n_repetitions, d_result, variant, stream);
}
int flsgpu_compute_ffor_u64(const flsgpu_ffor_column *col, uint32_t n_repetitions,
                            uint8_t *d_result, int variant, flsgpu_stream_t stream) {
  if (!col || (!col->bases && col->bp.n_vecs)) return set_error(FLSGPU_EINVAL, "null column/bases");
  FforArgs<uint64_t> a;
  int rc = make_ffor_args<uint64_t>(&col->bp, col->bases, &a);
  return rc ? rc : launch_compute(a, n_repetitions, d_result, variant, stream);
}

int flsgpu_query_multi_column_ffor_u32(const flsgpu_ffor_column *cols, uint32_t n_cols,
                                       uint32_t value, uint8_t *d_found, int variant,
                                       flsgpu_stream_t stream) {
  (void)variant;
  return launch_multi<uint32_t>(cols, n_cols, value, d_found, stream);
}
int flsgpu_query_multi_column_ffor_u64(const flsgpu_ffor_column *cols, uint32_t n_cols,
                                       uint64_t value, uint8_t *d_found, int variant,
                                       flsgpu_stream_t stream) {
  (void)variant;
  return launch_multi<uint64_t>(cols, n_cols, value, d_found, stream);
}

}  // extern "C"
