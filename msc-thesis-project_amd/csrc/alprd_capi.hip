// C-ABI entry points of the ALP-RD decoder (no reference GPU counterpart:
// src/alp/rd.hpp:146-172 is a CPU decoder of unpacked parts).
#include "alprd_impl.hpp"

namespace flsgpu_hip {

template <typename T>
static int alprd_decompress(const flsgpu_alprd_column *c, T *out, int variant, hipStream_t st) {
  constexpr unsigned W = sizeof(T) * 8;
  AlprdArgs<T> a;
  int rc = make_rd_args<T>(c, &a);
  if (rc) return rc;
  if (c->n_vecs && !out) return set_error(FLSGPU_EINVAL, "null output");
  if (variant == FLSGPU_VARIANT_AUTO)
    variant = flsgpu_get_tuning(W == 32 ? FLSGPU_ENC_ALPRD_F32 : FLSGPU_ENC_ALPRD_F64, 255);
  if (variant < FLSGPU_VARIANT_LANE || variant >= FLSGPU_VARIANT_COUNT)
    return set_error(FLSGPU_EUNSUPPORTED, "unknown variant %d", variant);
  // The staged variants (S = 2 for float; 4, or 8 for EIGHTH, for double) decode through
  // LDS when the output is 16-byte aligned; otherwise they and the other
  // non-temporal variants use the one-thread-per-lane kernel with
  // non-temporal stores. Lane splitting without staging is not provided for
  // ALP-RD (each thread loads its lane's uint16 left lanes once).
  const bool staged = variant == FLSGPU_VARIANT_HALF_LANE_STAGED ||
                      variant == FLSGPU_VARIANT_QUARTER_LANE_STAGED ||
                      variant == FLSGPU_VARIANT_EIGHTH_LANE_STAGED;
  if (staged && (reinterpret_cast<uintptr_t>(out) % 16) == 0 && a.n_vecs) {
    if constexpr (W == 64) {
      if (variant == FLSGPU_VARIANT_EIGHTH_LANE_STAGED)
        return AlprdDecompressStaged<T, 8>::launch(a, out, st);
    }
    return AlprdDecompressStaged<T, W == 32 ? 2 : 4>::launch(a, out, st);
  }
  const bool nt = staged || variant == FLSGPU_VARIANT_LANE_NT ||
                  variant == FLSGPU_VARIANT_HALF_LANE_NT ||
                  variant == FLSGPU_VARIANT_QUARTER_LANE_NT;
  return nt ? AlprdDecompress<T, true>::launch(a, out, st) : AlprdDecompress<T, false>::launch(a, out, st);
}

}  // namespace flsgpu_hip

using namespace flsgpu_hip;

extern "C" {
int flsgpu_decompress_alprd_f32(const flsgpu_alprd_column *col, float *out, int variant,
                                flsgpu_stream_t stream) {
  return alprd_decompress<float>(col, out, variant, stream);
}
int flsgpu_decompress_alprd_f64(const flsgpu_alprd_column *col, double *out, int variant,
                                flsgpu_stream_t stream) {
  return alprd_decompress<double>(col, out, variant, stream);
}
}
