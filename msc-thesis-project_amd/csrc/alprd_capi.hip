// C-ABI entry points of the ALP-RD decoder (no reference GPU counterpart:
// src/alp/rd.hpp:146-172 is a CPU decoder of unpacked parts).
#include "alprd_impl.hpp"

namespace flsgpu_hip {

template <typename T>
static int alprd_decompress(const flsgpu_alprd_column *c, T *out, int variant, hipStream_t st) {
  constexpr unsigned W = sizeof(T) * 8;
  if (!c) return set_error(FLSGPU_EINVAL, "null column");
  if (c->n_values != c->n_vecs * VALUES_PER_VECTOR)
    return set_error(FLSGPU_EINVAL, "n_values must equal n_vecs * 1024");
  if (c->right_bit_width < W - 16 || c->right_bit_width > W - 1)
    return set_error(FLSGPU_EINVAL, "right_bit_width %u outside %u..%u", c->right_bit_width,
                     W - 16, W - 1);
  if (c->left_bit_width < 1 || c->left_bit_width > 3)
    return set_error(FLSGPU_EINVAL, "left_bit_width %u outside 1..3", c->left_bit_width);
  if (c->dictionary_size < 1 || c->dictionary_size > 8 ||
      c->dictionary_size > (1u << c->left_bit_width))
    return set_error(FLSGPU_EINVAL, "dictionary_size %u invalid", c->dictionary_size);
  if (c->n_vecs && (!c->right_packed || !c->left_packed || !c->exceptions_offsets || !c->counts || !out))
    return set_error(FLSGPU_EINVAL, "null device pointer");
  AlprdArgs<T> a;
  a.n_vecs = c->n_vecs;
  a.right_bw = c->right_bit_width;
  a.left_bw = c->left_bit_width;
  a.dict_size = c->dictionary_size;
  for (int i = 0; i < 8; ++i) a.dict[i] = c->dictionary[i];
  a.right = static_cast<const typename uint_of<T>::type *>(c->right_packed);
  a.left = c->left_packed;
  a.exceptions_offsets = c->exceptions_offsets;
  a.exceptions = c->exceptions;
  a.positions = c->positions;
  a.counts = c->counts;
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
