// C-ABI entry points of the ALP / G-ALP kernels. Mirrors
// bindings::{decompress,query,query_multi}_column for ALPColumn and
// ALPExtendedColumn (src/generated-bindings/kernel-bindings.cuh:11-175) and
// the patcher/column pairing of generate-kernel-bindings.py:355-357.
#include "alp_impl.hpp"

namespace flsgpu_hip {

// Patchers valid for each layout; all patching patchers produce the same
// output and map to the native per-lane patch path.
static int patch_mode(int layout, int patcher, bool *patch) {
  if (layout == ALP_STD) {
    switch (patcher) {
      case FLSGPU_PATCHER_NONE:
      case FLSGPU_PATCHER_DUMMY: *patch = false; return FLSGPU_OK;
      case FLSGPU_PATCHER_STATELESS:
      case FLSGPU_PATCHER_STATEFUL: *patch = true; return FLSGPU_OK;
      default: break;
    }
    return set_error(FLSGPU_EUNSUPPORTED,
                     "patcher %d is not valid for ALPColumn (use none/dummy/stateless/stateful)",
                     patcher);
  }
  switch (patcher) {
    case FLSGPU_PATCHER_NAIVE:
    case FLSGPU_PATCHER_NAIVE_BRANCHLESS:
    case FLSGPU_PATCHER_PREFETCH_POSITION:
    case FLSGPU_PATCHER_PREFETCH_ALL:
    case FLSGPU_PATCHER_PREFETCH_ALL_BRANCHLESS: *patch = true; return FLSGPU_OK;
    default: break;
  }
  return set_error(FLSGPU_EUNSUPPORTED,
                   "patcher %d is not valid for ALPExtendedColumn (use naive, naive-branchless, "
                   "prefetch-position, prefetch-all, prefetch-all-branchless)",
                   patcher);
}

static int alp_shape(int variant, int encoding, int *S, bool *nt, bool *staged) {
  if (variant == FLSGPU_VARIANT_AUTO) variant = flsgpu_get_tuning(encoding, 255);
  *staged = false;
  const bool wide = encoding == FLSGPU_ENC_ALP_F64 || encoding == FLSGPU_ENC_ALP_EXT_F64;
  switch (variant) {
    case FLSGPU_VARIANT_HALF_LANE_STAGED:
      *S = wide ? 4 : 2;
      *nt = true;
      *staged = true;
      return FLSGPU_OK;
    case FLSGPU_VARIANT_QUARTER_LANE_STAGED:
      *S = 4;
      *nt = true;
      *staged = true;
      return FLSGPU_OK;
    case FLSGPU_VARIANT_EIGHTH_LANE_STAGED:
      *S = wide ? 8 : 2;
      *nt = true;
      *staged = true;
      return FLSGPU_OK;
    case FLSGPU_VARIANT_LANE: *S = 1; *nt = false; return FLSGPU_OK;
    case FLSGPU_VARIANT_LANE_NT: *S = 1; *nt = true; return FLSGPU_OK;
    case FLSGPU_VARIANT_HALF_LANE: *S = 2; *nt = false; return FLSGPU_OK;
    case FLSGPU_VARIANT_HALF_LANE_NT: *S = 2; *nt = true; return FLSGPU_OK;
    case FLSGPU_VARIANT_QUARTER_LANE: *S = 4; *nt = false; return FLSGPU_OK;
    case FLSGPU_VARIANT_QUARTER_LANE_NT: *S = 4; *nt = true; return FLSGPU_OK;
    default: return set_error(FLSGPU_EUNSUPPORTED, "unknown variant %d", variant);
  }
}

template <typename T, int S, bool NT>
static int decompress_s(const AlpArgs<T> &a, T *out, int layout, bool patch, hipStream_t st) {
  if (!patch) return AlpDecompress<T, S, NT, ALP_STD, false>::launch(a, out, st);
  if (layout == ALP_GALP) return AlpDecompress<T, S, NT, ALP_GALP, true>::launch(a, out, st);
  return AlpDecompress<T, S, NT, ALP_STD, true>::launch(a, out, st);
}

template <typename T, int S>
static int decompress_staged(const AlpArgs<T> &a, T *out, int layout, bool patch, hipStream_t st) {
  if (!patch) return AlpDecompressStaged<T, S, ALP_STD, false>::launch(a, out, st);
  if (layout == ALP_GALP) return AlpDecompressStaged<T, S, ALP_GALP, true>::launch(a, out, st);
  return AlpDecompressStaged<T, S, ALP_STD, true>::launch(a, out, st);
}

template <typename T>
static int alp_decompress(const flsgpu_alp_column *c, T *out, int layout, int patcher,
                          int variant, hipStream_t st) {
  AlpArgs<T> a;
  int rc = make_alp_args<T>(c, layout, &a);
  if (rc) return rc;
  if (!out && a.ffor.n_vecs) return set_error(FLSGPU_EINVAL, "null output");
  bool patch;
  if ((rc = patch_mode(layout, patcher, &patch))) return rc;
  int S;
  bool nt, staged;
  const int enc = sizeof(T) == 4 ? (layout == ALP_GALP ? FLSGPU_ENC_ALP_EXT_F32 : FLSGPU_ENC_ALP_F32)
                                 : (layout == ALP_GALP ? FLSGPU_ENC_ALP_EXT_F64 : FLSGPU_ENC_ALP_F64);
  if ((rc = alp_shape(variant, enc, &S, &nt, &staged))) return rc;
  if (staged && (reinterpret_cast<uintptr_t>(out) % 16) == 0) {
    if constexpr (sizeof(T) == 4) {
      if (S == 2) return decompress_staged<T, 2>(a, out, layout, patch, st);
    } else {
      if (S == 8) return decompress_staged<T, 8>(a, out, layout, patch, st);
    }
    return decompress_staged<T, 4>(a, out, layout, patch, st);
  }
  if (S == 8) S = 4;  // unaligned output: the NT variant of the nearest split
  if (S == 1) return nt ? decompress_s<T, 1, true>(a, out, layout, patch, st)
                        : decompress_s<T, 1, false>(a, out, layout, patch, st);
  if (S == 2) return nt ? decompress_s<T, 2, true>(a, out, layout, patch, st)
                        : decompress_s<T, 2, false>(a, out, layout, patch, st);
  return nt ? decompress_s<T, 4, true>(a, out, layout, patch, st)
            : decompress_s<T, 4, false>(a, out, layout, patch, st);
}

template <typename T>
static int alp_any(const flsgpu_alp_column *c, T magic, uint8_t *d_found, int layout,
                   int patcher, hipStream_t st) {
  AlpArgs<T> a;
  int rc = make_alp_args<T>(c, layout, &a);
  if (rc) return rc;
  if (!d_found) return set_error(FLSGPU_EINVAL, "null result pointer");
  bool patch;
  if ((rc = patch_mode(layout, patcher, &patch))) return rc;
  if (!patch) return AlpAny<T, ALP_STD, false>::launch(a, magic, d_found, st);
  if (layout == ALP_GALP) return AlpAny<T, ALP_GALP, true>::launch(a, magic, d_found, st);
  return AlpAny<T, ALP_STD, true>::launch(a, magic, d_found, st);
}

template <typename T>
static int alp_multi(const flsgpu_alp_column *cols, uint32_t n_cols, T value, uint8_t *d_found,
                     int patcher, hipStream_t st) {
  if (!cols || !d_found) return set_error(FLSGPU_EINVAL, "null argument");
  if (n_cols < 1 || n_cols > MAX_COLS)
    return set_error(FLSGPU_EINVAL, "n_cols must be in 1..%d, got %u", MAX_COLS, n_cols);
  const int layout = patcher >= FLSGPU_PATCHER_NAIVE ? ALP_GALP : ALP_STD;
  bool patch;
  int rc = patch_mode(layout, patcher, &patch);
  if (rc) return rc;
  MultiAlpArgs<T> m{};
  m.n_cols = n_cols;
  for (uint32_t c = 0; c < n_cols; ++c) {
    if ((rc = make_alp_args<T>(&cols[c], layout, &m.col[c]))) return rc;
    if (m.col[c].ffor.n_vecs != m.col[0].ffor.n_vecs)
      return set_error(FLSGPU_EINVAL, "all columns must have the same number of vectors");
  }
  if (!patch) return AlpMulti<T, ALP_STD, false>::launch(m, value, d_found, st);
  if (layout == ALP_GALP) return AlpMulti<T, ALP_GALP, true>::launch(m, value, d_found, st);
  return AlpMulti<T, ALP_STD, true>::launch(m, value, d_found, st);
}

}  // namespace flsgpu_hip

using namespace flsgpu_hip;

extern "C" {

int flsgpu_decompress_alp_f32(const flsgpu_alp_column *col, float *out, int patcher, int variant,
                              flsgpu_stream_t stream) {
  return alp_decompress<float>(col, out, ALP_STD, patcher, variant, stream);
}
int flsgpu_decompress_alp_f64(const flsgpu_alp_column *col, double *out, int patcher, int variant,
                              flsgpu_stream_t stream) {
  return alp_decompress<double>(col, out, ALP_STD, patcher, variant, stream);
}
int flsgpu_decompress_alp_ext_f32(const flsgpu_alp_column *col, float *out, int patcher,
                                  int variant, flsgpu_stream_t stream) {
  return alp_decompress<float>(col, out, ALP_GALP, patcher, variant, stream);
}
int flsgpu_decompress_alp_ext_f64(const flsgpu_alp_column *col, double *out, int patcher,
                                  int variant, flsgpu_stream_t stream) {
  return alp_decompress<double>(col, out, ALP_GALP, patcher, variant, stream);
}

int flsgpu_query_any_alp_f32(const flsgpu_alp_column *col, float magic, uint8_t *d_found,
                             int patcher, int variant, flsgpu_stream_t stream) {
  (void)variant;
  return alp_any<float>(col, magic, d_found, ALP_STD, patcher, stream);
}
int flsgpu_query_any_alp_f64(const flsgpu_alp_column *col, double magic, uint8_t *d_found,
                             int patcher, int variant, flsgpu_stream_t stream) {
  (void)variant;
  return alp_any<double>(col, magic, d_found, ALP_STD, patcher, stream);
}
int flsgpu_query_any_alp_ext_f32(const flsgpu_alp_column *col, float magic, uint8_t *d_found,
                                 int patcher, int variant, flsgpu_stream_t stream) {
  (void)variant;
  return alp_any<float>(col, magic, d_found, ALP_GALP, patcher, stream);
}
int flsgpu_query_any_alp_ext_f64(const flsgpu_alp_column *col, double magic, uint8_t *d_found,
                                 int patcher, int variant, flsgpu_stream_t stream) {
  (void)variant;
  return alp_any<double>(col, magic, d_found, ALP_GALP, patcher, stream);
}

int flsgpu_query_multi_column_alp_f32(const flsgpu_alp_column *cols, uint32_t n_cols, float value,
                                      uint8_t *d_found, int patcher, int variant,
                                      flsgpu_stream_t stream) {
  (void)variant;
  return alp_multi<float>(cols, n_cols, value, d_found, patcher, stream);
}
int flsgpu_query_multi_column_alp_f64(const flsgpu_alp_column *cols, uint32_t n_cols,
                                      double value, uint8_t *d_found, int patcher, int variant,
                                      flsgpu_stream_t stream) {
  (void)variant;
  return alp_multi<double>(cols, n_cols, value, d_found, patcher, stream);
}

}  // extern "C"
