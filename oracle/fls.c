/*
 * Oracle (test infrastructure only): FastLanes bit-packing restated as a
 * per-(lane, slot) bit-position formula instead of the reference's generated
 * per-width code.
 *
 *   reference unpack : src/fls/unpack.cpp:17249-17441 (dispatch), bodies such
 *                      as unffor_3bw_32ow_32crw_1uf src/fls/unffor.cpp:2982
 *   reference pack   : src/fls/pack.cpp:29645-29834, ffor.cpp:29771-29966
 *   column loop      : src/engine/data.cuh:287-313
 *
 * Lane l, slot j of a 1024-value vector of W-bit words lives at logical index
 * l + N_LANES*j, and at stream bits [j*bw, (j+1)*bw) of lane l, whose r-th
 * W-bit word is in[l + N_LANES*r] (least significant bit first).
 */
#include <pthread.h>
#include <string.h>

#include "oracle.h"

#define DEFINE_FLS(W, UT)                                                                          \
  static inline UT mask_##W(unsigned bw) {                                                         \
    return bw >= (W) ? (UT)~(UT)0 : (UT)(((UT)1 << bw) - 1);                                       \
  }                                                                                                \
  void orc_fls_unpack_u##W(const UT *in, UT *out, unsigned bw) {                                   \
    const unsigned n_lanes = 1024 / (W);                                                           \
    const UT m = mask_##W(bw);                                                                     \
    for (unsigned l = 0; l < n_lanes; ++l) {                                                       \
      for (unsigned j = 0; j < (W); ++j) {                                                         \
        UT v = 0;                                                                                  \
        if (bw != 0) {                                                                             \
          const unsigned p = j * bw, r = p / (W), s = p % (W);                                     \
          v = (UT)(in[l + n_lanes * r] >> s);                                                      \
          if (s + bw > (W)) v |= (UT)(in[l + n_lanes * (r + 1)] << ((W) - s));                     \
          v &= m;                                                                                  \
        }                                                                                          \
        out[l + n_lanes * j] = v;                                                                  \
      }                                                                                            \
    }                                                                                              \
  }                                                                                                \
  void orc_fls_pack_u##W(const UT *in, UT *out, unsigned bw) {                                     \
    const unsigned n_lanes = 1024 / (W);                                                           \
    const UT m = mask_##W(bw);                                                                     \
    memset(out, 0, sizeof(UT) * n_lanes * bw);                                                     \
    if (bw == 0) return;                                                                           \
    for (unsigned l = 0; l < n_lanes; ++l) {                                                       \
      for (unsigned j = 0; j < (W); ++j) {                                                         \
        const UT v = in[l + n_lanes * j] & m;                                                      \
        const unsigned p = j * bw, r = p / (W), s = p % (W);                                       \
        out[l + n_lanes * r] |= (UT)(v << s);                                                      \
        if (s + bw > (W)) out[l + n_lanes * (r + 1)] |= (UT)(v >> ((W) - s));                      \
      }                                                                                            \
    }                                                                                              \
  }                                                                                                \
  void orc_fls_ffor_u##W(const UT *in, UT *out, unsigned bw, UT base) {                            \
    UT tmp[1024];                                                                                  \
    for (unsigned i = 0; i < 1024; ++i) tmp[i] = (UT)(in[i] - base);                               \
    orc_fls_pack_u##W(tmp, out, bw);                                                               \
  }                                                                                                \
  void orc_fls_unffor_u##W(const UT *in, UT *out, unsigned bw, UT base) {                          \
    orc_fls_unpack_u##W(in, out, bw);                                                              \
    for (unsigned i = 0; i < 1024; ++i) out[i] = (UT)(out[i] + base);                              \
  }

DEFINE_FLS(8, uint8_t)
DEFINE_FLS(16, uint16_t)
DEFINE_FLS(32, uint32_t)
DEFINE_FLS(64, uint64_t)

void orc_ffor_column_decode_u32(size_t n_vecs, const uint32_t *packed, const uint8_t *bit_widths,
                                const uint64_t *vector_offsets, const uint32_t *bases,
                                uint32_t *out) {
  for (size_t v = 0; v < n_vecs; ++v) {
    orc_fls_unffor_u32(packed + vector_offsets[v], out + v * 1024, bit_widths[v],
                       bases ? bases[v] : 0u);
  }
}

void orc_ffor_column_decode_u64(size_t n_vecs, const uint64_t *packed, const uint8_t *bit_widths,
                                const uint64_t *vector_offsets, const uint64_t *bases,
                                uint64_t *out) {
  for (size_t v = 0; v < n_vecs; ++v) {
    orc_fls_unffor_u64(packed + vector_offsets[v], out + v * 1024, bit_widths[v],
                       bases ? bases[v] : 0u);
  }
}

struct mt_u32_job {
  size_t begin, end;
  const uint32_t *packed;
  const uint8_t *bit_widths;
  const uint64_t *vector_offsets;
  const uint32_t *bases;
  uint32_t *out;
};

static void *mt_u32_worker(void *arg) {
  struct mt_u32_job *j = (struct mt_u32_job *)arg;
  for (size_t v = j->begin; v < j->end; ++v) {
    orc_fls_unffor_u32(j->packed + j->vector_offsets[v], j->out + v * 1024, j->bit_widths[v],
                       j->bases ? j->bases[v] : 0u);
  }
  return NULL;
}

void orc_ffor_column_decode_u32_mt(size_t n_vecs, const uint32_t *packed, const uint8_t *bit_widths,
                                   const uint64_t *vector_offsets, const uint32_t *bases,
                                   uint32_t *out, int n_threads) {
  if (n_threads < 1) n_threads = 1;
  if (n_threads > 256) n_threads = 256;
  pthread_t th[256];
  struct mt_u32_job jobs[256];
  for (int t = 0; t < n_threads; ++t) {
    jobs[t].begin = n_vecs * (size_t)t / (size_t)n_threads;
    jobs[t].end = n_vecs * (size_t)(t + 1) / (size_t)n_threads;
    jobs[t].packed = packed;
    jobs[t].bit_widths = bit_widths;
    jobs[t].vector_offsets = vector_offsets;
    jobs[t].bases = bases;
    jobs[t].out = out;
    pthread_create(&th[t], NULL, mt_u32_worker, &jobs[t]);
  }
  for (int t = 0; t < n_threads; ++t) pthread_join(th[t], NULL);
}
