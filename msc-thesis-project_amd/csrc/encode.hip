// Device-side encoding for gfx950: FFOR bit-packing of integer columns and
// ALP encoding of float columns with per-vector (exponent, factor) chosen
// elsewhere (e.g. by the learned selector), in the standard or the G-ALP
// exception layout. The output bytes equal those of the host encoder
// (flsgpu_alp_encode_with_params_*, flsgpu_ffor_pack_*), which follows:
//   encode_simdized        src/alp/encoder.hpp:302-379 (exceptions, fill rule)
//   analyze_ffor           src/alp/encoder.hpp:104-115 (base = min, width of max - min)
//   FastLanes ffor/pack    src/fls/ffor.cpp:29771-29966
//   G-ALP layout           src/flsgpu/structs.cuh:231-306
//
// Three launches on the caller's stream:
//   k_enc_analyze  one group of N_LANES threads per vector (one thread per
//                  FastLanes lane): encodes the vector, finds its exceptions,
//                  fill value, base and bit width; per-block sums of packed
//                  words and exceptions.
//   k_enc_scan     one block: exclusive scan of the block sums, totals.
//   k_enc_pack     recomputes each vector, places it at its scanned offsets,
//                  packs the lane's words and writes the exceptions.
#include <limits>

#include "common.hpp"

namespace flsgpu_hip {
namespace {

template <typename T>
struct Enc;
template <>
struct Enc<float> {
  using UT = uint32_t;
  using ST = int32_t;
  static constexpr UT SIGN_MASK = 0x7FFFFFFFu, EXP_MASK = 0x7F800000u, NEG_ZERO = 0x80000000u;
  static constexpr float MAGIC = 12582912.0f;
  static constexpr float UPPER = 9223372036854774784.0f;
};
template <>
struct Enc<double> {
  using UT = uint64_t;
  using ST = int64_t;
  static constexpr UT SIGN_MASK = 0x7FFFFFFFFFFFFFFFull, EXP_MASK = 0x7FF0000000000000ull,
                      NEG_ZERO = 0x8000000000000000ull;
  static constexpr double MAGIC = 6755399441055744.0;
  static constexpr double UPPER = 9223372036854774784.0;
};

// 10^e on the encoder side (src/alp/constants.hpp EXP_ARR).
static __constant__ float c_enc_exp_f32[11] = {1.0,       10.0,       100.0,        1000.0,
                                               10000.0,   100000.0,   1000000.0,    10000000.0,
                                               100000000.0, 1000000000.0, 10000000000.0};
static __constant__ double c_enc_exp_f64[19] = {1e0,  1e1,  1e2,  1e3,  1e4,  1e5,  1e6,
                                                1e7,  1e8,  1e9,  1e10, 1e11, 1e12, 1e13,
                                                1e14, 1e15, 1e16, 1e17, 1e18};

__device__ __forceinline__ float enc_exp(float, int e) { return c_enc_exp_f32[e]; }
__device__ __forceinline__ double enc_exp(double, int e) { return c_enc_exp_f64[e]; }

// Float -> integer conversion with the host encoder's semantics: an
// out-of-range or NaN input gives the x86-64 "integer indefinite" value
// (INT_MIN). The hardware conversion would saturate instead.
__device__ __forceinline__ int32_t to_int(float t) {
  return (t >= -2147483648.0f && t < 2147483648.0f) ? int32_t(t) : INT32_MIN;
}
__device__ __forceinline__ int64_t to_int(double t) {
  return (t >= -9223372036854775808.0 && t < 9223372036854775808.0) ? int64_t(t) : INT64_MIN;
}

#pragma clang fp contract(off)
// encode_simdized's per-value step: special values (inf, NaN, -0.0) are
// replaced by ENCODING_UPPER_LIMIT, then the unchecked encode_value
// (encoder.hpp:76-84): two products, rounding by the magic number, cast.
// Sets *exc when the value does not round-trip through decode_value.
template <typename T>
__device__ __forceinline__ typename Enc<T>::ST encode_one(T v, int fac, int exp, bool *exc) {
  using E = Enc<T>;
  using UT = typename E::UT;
  using ST = typename E::ST;
  const UT bits = __builtin_bit_cast(UT, v);
  const bool special = (bits & E::SIGN_MASK) >= E::EXP_MASK || bits == E::NEG_ZERO;
  const T x = special ? T(E::UPPER) : v;
  T t = x * enc_exp(T(0), exp) * alp_fraction<T>(unsigned(fac));
  t = t + E::MAGIC - E::MAGIC;
  const ST e = to_int(t);
  const ST s = ST(UT(UT(e) * alp_factor<T>(unsigned(fac))));
  *exc = T(s) * alp_fraction<T>(unsigned(exp)) != x;
  return e;
}

template <typename UT>
__device__ __forceinline__ int bit_count(UT d) {
  if constexpr (sizeof(UT) == 8) return d == 0 ? 0 : 64 - __clzll(d);
  else return d == 0 ? 0 : 32 - __clz(d);
}

// Reductions over the N_LANES threads of one vector (aligned groups inside
// a wavefront; xor offsets below N_LANES stay inside the group).
template <int NL, typename V, class Op>
__device__ __forceinline__ V group_reduce(V v, Op op) {
#pragma unroll
  for (int off = NL / 2; off > 0; off >>= 1) v = op(v, __shfl_xor(v, off, WAVE));
  return v;
}
template <int NL>
__device__ __forceinline__ unsigned group_exclusive_scan(unsigned v, int lane) {
  unsigned incl = v;
#pragma unroll
  for (int off = 1; off < NL; off <<= 1) {
    const unsigned o = __shfl_up(incl, off, NL);
    if (lane >= off) incl += o;
  }
  return incl - v;
}

// Per-vector analysis shared by both passes. For ALP (IS_ALP) the lane's
// exception mask is returned in *m; mn/mx are over the vector's encoded
// integers after exceptions take the fill value.
template <typename T, bool IS_ALP>
struct VecInfo {
  using UT = typename uint_of<T>::type;
  using ST = std::conditional_t<IS_ALP, typename int_of<T>::type, UT>;
  UT m;           // this lane's exception slots
  unsigned n_exc; // exceptions in the vector
  ST fill;
  ST mn, mx;
};

template <typename T, bool IS_ALP>
__device__ __forceinline__ VecInfo<T, IS_ALP> analyze_vector(const T *__restrict vin, int lane,
                                                             int fac, int exp) {
  using UT = typename uint_of<T>::type;
  using L = lanes<UT>;
  using Info = VecInfo<T, IS_ALP>;
  using ST = typename Info::ST;
  Info r{};
  ST mn = std::numeric_limits<ST>::max(), mx = std::numeric_limits<ST>::min();
  UT m = 0;
  for (int j = 0; j < L::W; ++j) {
    const T v = vin[lane + L::N * j];
    if constexpr (IS_ALP) {
      bool exc;
      const ST e = encode_one<T>(v, fac, exp, &exc);
      if (exc) {
        m |= UT(UT(1) << j);
      } else {
        mn = e < mn ? e : mn;
        mx = e > mx ? e : mx;
      }
    } else {
      mn = v < mn ? v : mn;
      mx = v > mx ? v : mx;
    }
  }
  auto vmin = [](ST a, ST b) { return a < b ? a : b; };
  auto vmax = [](ST a, ST b) { return a > b ? a : b; };
  r.mn = group_reduce<L::N>(mn, vmin);
  r.mx = group_reduce<L::N>(mx, vmax);
  r.m = m;
  if constexpr (IS_ALP) {
    r.n_exc = group_reduce<L::N>(unsigned(__builtin_popcountll(uint64_t(m))),
                                 [](unsigned a, unsigned b) { return a + b; });
    r.fill = 0;
    if (r.n_exc) {
      // First non-exception position; the fill is its encoded value, except
      // when it is slot 1023 or absent (encoder.hpp:352-358 finds no index
      // mismatch then and keeps 0).
      const UT nonexc = UT(~m);
      const unsigned mine = nonexc ? unsigned(lane + L::N * __builtin_ctzll(uint64_t(nonexc))) : 1024u;
      const unsigned p =
          group_reduce<L::N>(mine, [](unsigned a, unsigned b) { return a < b ? a : b; });
      if (p < 1023u) {
        bool dummy;
        r.fill = encode_one<T>(vin[p], fac, exp, &dummy);
      }
      r.mn = r.n_exc == 1024u ? r.fill : vmin(r.mn, r.fill);
      r.mx = r.n_exc == 1024u ? r.fill : vmax(r.mx, r.fill);
    }
  }
  return r;
}

// The (exponent, factor) pair a vector is encoded with: the caller's, when
// it indexes the tables (exponent <= 10 for floats, <= 18 for doubles;
// factor <= exponent, and <= 9 for floats, whose factor table has ten
// entries, src/alp/constants.hpp:49-55), else (0, 0). Any pair decodes
// losslessly (values that do not round-trip become exceptions), so an
// out-of-range request costs compression ratio, never correctness or a
// read outside the tables.
template <typename T>
__device__ __forceinline__ void checked_params(unsigned e_in, unsigned f_in, int *exp, int *fac) {
  constexpr unsigned MAX_E = sizeof(T) == 4 ? 10u : 18u;
  constexpr unsigned MAX_F = sizeof(T) == 4 ? 9u : 18u;
  const bool ok = e_in <= MAX_E && f_in <= e_in && f_in <= MAX_F;
  *exp = ok ? int(e_in) : 0;
  *fac = ok ? int(f_in) : 0;
}

struct EncOut {
  void *packed;
  uint8_t *bit_widths;
  uint64_t *vector_offsets;
  void *bases;
  uint8_t *factor_indices;
  uint8_t *fraction_indices;
  uint64_t *exceptions_offsets;
  void *exceptions;
  uint16_t *positions;
  uint16_t *counts;
  uint16_t *offsets_counts;
};

constexpr int ENC_BLOCK = BLOCK;

template <typename T, bool IS_ALP>
__global__ __launch_bounds__(ENC_BLOCK) void k_enc_analyze(const T *__restrict in, uint64_t n_vecs,
                                                           const uint8_t *__restrict exps,
                                                           const uint8_t *__restrict facs,
                                                           EncOut o, uint64_t *block_sums,
                                                           uint64_t *fills) {
  using UT = typename uint_of<T>::type;
  using L = lanes<UT>;
  constexpr int VPB = ENC_BLOCK / L::N;
  __shared__ uint64_t s_words[VPB], s_exc[VPB];
  const int lane = threadIdx.x % L::N, vl = threadIdx.x / L::N;
  const uint64_t v = uint64_t(blockIdx.x) * VPB + vl;
  uint64_t words = 0, n_exc = 0;
  if (v < n_vecs) {
    int fac = 0, exp = 0;
    if constexpr (IS_ALP) checked_params<T>(exps[v], facs[v], &exp, &fac);
    const auto r = analyze_vector<T, IS_ALP>(in + v * VALUES_PER_VECTOR, lane, fac, exp);
    const int bw = bit_count<UT>(UT(UT(r.mx) - UT(r.mn)));
    words = uint64_t(bw) * L::N;
    if constexpr (IS_ALP) n_exc = r.n_exc;
    if (lane == 0) {
      o.bit_widths[v] = uint8_t(bw);
      static_cast<UT *>(o.bases)[v] = UT(r.mn);
      if constexpr (IS_ALP) {
        o.counts[v] = uint16_t(r.n_exc);
        fills[v] = uint64_t(UT(r.fill));
      }
    }
  }
  if (lane == 0) {
    s_words[vl] = words;
    s_exc[vl] = n_exc;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t a = 0, b = 0;
    for (int i = 0; i < VPB; ++i) {
      a += s_words[i];
      b += s_exc[i];
    }
    block_sums[2 * blockIdx.x] = a;
    block_sums[2 * blockIdx.x + 1] = b;
  }
}

// Exclusive scan of the per-block (words, exceptions) pairs in place; the
// totals go to totals[0..1].
__global__ __launch_bounds__(1024) void k_enc_scan(uint64_t *sums, uint64_t n_blocks,
                                                   uint64_t *totals) {
  __shared__ uint64_t s_a[1024], s_b[1024];
  uint64_t carry_a = 0, carry_b = 0;
  for (uint64_t base = 0; base < n_blocks; base += 1024) {
    const uint64_t i = base + threadIdx.x;
    const uint64_t a = i < n_blocks ? sums[2 * i] : 0, b = i < n_blocks ? sums[2 * i + 1] : 0;
    s_a[threadIdx.x] = a;
    s_b[threadIdx.x] = b;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
      const uint64_t xa = threadIdx.x >= unsigned(off) ? s_a[threadIdx.x - off] : 0;
      const uint64_t xb = threadIdx.x >= unsigned(off) ? s_b[threadIdx.x - off] : 0;
      __syncthreads();
      s_a[threadIdx.x] += xa;
      s_b[threadIdx.x] += xb;
      __syncthreads();
    }
    if (i < n_blocks) {
      sums[2 * i] = carry_a + s_a[threadIdx.x] - a;
      sums[2 * i + 1] = carry_b + s_b[threadIdx.x] - b;
    }
    carry_a += s_a[1023];
    carry_b += s_b[1023];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    totals[0] = carry_a;
    totals[1] = carry_b;
    totals[2] = 0;  // set by the pack pass when a G-ALP lane overflows
  }
}

// Appends the low `bw` bits of v to the lane's bit stream, storing every
// completed word at out[NL * r].
template <typename UT, int NL>
struct LanePacker {
  UT *out;
  int bw;
  UT cur = 0;
  int filled = 0;
  int r = 0;
  __device__ __forceinline__ void push(UT v) {
    constexpr int W = int(sizeof(UT) * 8);
    cur |= UT(v << filled);
    filled += bw;
    if (filled >= W) {
      out[NL * r++] = cur;
      filled -= W;
      cur = filled > 0 ? UT(v >> (bw - filled)) : UT(0);
    }
  }
};

template <typename T, bool IS_ALP, bool GALP>
__global__ __launch_bounds__(ENC_BLOCK) void k_enc_pack(const T *__restrict in, uint64_t n_vecs,
                                                        const uint8_t *__restrict exps,
                                                        const uint8_t *__restrict facs,
                                                        EncOut o, const uint64_t *block_offsets,
                                                        const uint64_t *fills, uint64_t *totals) {
  using UT = typename uint_of<T>::type;
  using L = lanes<UT>;
  using ST = std::conditional_t<IS_ALP, typename int_of<T>::type, UT>;
  constexpr int VPB = ENC_BLOCK / L::N;
  __shared__ uint64_t s_words[VPB], s_exc[VPB];
  __shared__ UT s_mask[VPB][L::N];
  const int lane = threadIdx.x % L::N, vl = threadIdx.x / L::N;
  const uint64_t v = uint64_t(blockIdx.x) * VPB + vl;
  const bool live = v < n_vecs;
  int bw = 0;
  unsigned n_exc = 0;
  if (live) {
    bw = o.bit_widths[v];
    if constexpr (IS_ALP) n_exc = o.counts[v];
  }
  if (lane == 0) {
    s_words[vl] = live ? uint64_t(bw) * L::N : 0;
    s_exc[vl] = n_exc;
  }
  __syncthreads();
  uint64_t woff = block_offsets[2 * blockIdx.x], eoff = block_offsets[2 * blockIdx.x + 1];
  for (int i = 0; i < vl; ++i) {
    woff += s_words[i];
    eoff += s_exc[i];
  }
  UT m = 0;
  ST fill = 0;
  int fac = 0, exp = 0;
  const T *vin = in + v * VALUES_PER_VECTOR;
  if (live) {
    if constexpr (IS_ALP) {
      checked_params<T>(exps[v], facs[v], &exp, &fac);
      fill = ST(UT(fills[v]));
      for (int j = 0; j < L::W; ++j) {
        bool exc;
        (void)encode_one<T>(vin[lane + L::N * j], fac, exp, &exc);
        if (exc) m |= UT(UT(1) << j);
      }
    }
  }
  s_mask[vl][lane] = m;
  __syncthreads();
  if (!live) return;
  const UT base = static_cast<const UT *>(o.bases)[v];
  if (lane == 0) {
    o.vector_offsets[v] = woff;
    if constexpr (IS_ALP) {
      o.exceptions_offsets[v] = eoff;
      o.factor_indices[v] = uint8_t(fac);
      o.fraction_indices[v] = uint8_t(exp);
    }
  }
  // G-ALP: the lane's exceptions are contiguous, after those of lower lanes.
  unsigned lane_off = 0;
  if constexpr (IS_ALP && GALP) {
    const unsigned cnt = unsigned(__builtin_popcountll(uint64_t(m)));
    lane_off = group_exclusive_scan<L::N>(cnt, lane);
    if (cnt > 63u) totals[2] = 1;  // the 16-bit header cannot hold it
    o.offsets_counts[v * L::N + lane] = uint16_t((cnt << 10) | (lane_off & 0x3FFu));
  }
  LanePacker<UT, L::N> pk{static_cast<UT *>(o.packed) + woff + lane, bw};
  unsigned k_lane = 0;
  for (int j = 0; j < L::W; ++j) {
    const T x = vin[lane + L::N * j];
    ST e;
    if constexpr (IS_ALP) {
      bool exc;
      e = encode_one<T>(x, fac, exp, &exc);
      if (exc) {
        const unsigned pos = unsigned(lane + L::N * j);
        unsigned rank;
        if constexpr (GALP) {
          rank = lane_off + k_lane++;
        } else {
          // position order: exceptions of lower lanes at slots <= j, of
          // higher lanes (and this lane) at slots < j
          rank = 0;
          const UT le = UT((UT(1) << j) - 1);
          const UT lt = j + 1 >= L::W ? UT(~UT(0)) : UT((UT(1) << (j + 1)) - 1);
          for (int l2 = 0; l2 < L::N; ++l2)
            rank += unsigned(__builtin_popcountll(uint64_t(s_mask[vl][l2] & (l2 < lane ? lt : le))));
        }
        static_cast<T *>(o.exceptions)[eoff + rank] = x;
        o.positions[eoff + rank] = uint16_t(pos);
        e = fill;
      }
    } else {
      e = x;
    }
    if (bw) pk.push(UT(UT(e) - base));
  }
}

template <typename T, bool IS_ALP>
int encode_launch(const T *in, uint64_t n_vecs, const uint8_t *exps, const uint8_t *facs,
                  bool galp, const EncOut &o, uint64_t *d_totals, void *workspace,
                  hipStream_t stream) {
  using UT = typename uint_of<T>::type;
  constexpr int VPB = ENC_BLOCK / lanes<UT>::N;
  if (n_vecs == 0) return set_error(FLSGPU_EINVAL, "empty column");
  if (!in || !d_totals || !workspace || !o.packed || !o.bit_widths || !o.vector_offsets ||
      !o.bases)
    return set_error(FLSGPU_EINVAL, "null pointer");
  if (IS_ALP && (!exps || !facs || !o.factor_indices || !o.fraction_indices ||
                 !o.exceptions_offsets || !o.exceptions || !o.positions || !o.counts ||
                 (galp && !o.offsets_counts)))
    return set_error(FLSGPU_EINVAL, "null pointer in ALP column");
  const uint64_t blocks = (n_vecs + VPB - 1) / VPB;
  if (blocks > 0x7FFFFFFFull) return set_error(FLSGPU_EINVAL, "column too large");
  uint64_t *sums = static_cast<uint64_t *>(workspace);
  uint64_t *fills = sums + 2 * blocks;
  hipLaunchKernelGGL((k_enc_analyze<T, IS_ALP>), dim3(unsigned(blocks)), dim3(ENC_BLOCK), 0,
                     stream, in, n_vecs, exps, facs, o, sums, fills);
  hipLaunchKernelGGL(k_enc_scan, dim3(1), dim3(1024), 0, stream, sums, blocks, d_totals);
  bool launched = false;
  if constexpr (IS_ALP) {
    if (galp) {
      hipLaunchKernelGGL((k_enc_pack<T, true, true>), dim3(unsigned(blocks)), dim3(ENC_BLOCK), 0,
                         stream, in, n_vecs, exps, facs, o, sums, fills, d_totals);
      launched = true;
    }
  }
  if (!launched)
    hipLaunchKernelGGL((k_enc_pack<T, IS_ALP, false>), dim3(unsigned(blocks)), dim3(ENC_BLOCK), 0,
                       stream, in, n_vecs, exps, facs, o, sums, fills, d_totals);
  return check_hip(hipGetLastError(), "encode launch");
}

template <typename T>
int alp_encode_entry(const T *in, uint64_t n_vecs, const uint8_t *exps, const uint8_t *facs,
                     int galp, flsgpu_alp_column *col, uint64_t *d_totals, void *workspace,
                     hipStream_t stream) {
  if (!col) return set_error(FLSGPU_EINVAL, "null column");
  EncOut o{const_cast<void *>(col->ffor.bp.packed_array),
           const_cast<uint8_t *>(col->ffor.bp.bit_widths),
           const_cast<uint64_t *>(col->ffor.bp.vector_offsets),
           const_cast<void *>(col->ffor.bases),
           const_cast<uint8_t *>(col->factor_indices),
           const_cast<uint8_t *>(col->fraction_indices),
           const_cast<uint64_t *>(col->exceptions_offsets),
           const_cast<void *>(col->exceptions),
           const_cast<uint16_t *>(col->positions),
           const_cast<uint16_t *>(col->counts),
           const_cast<uint16_t *>(col->offsets_counts)};
  // G-ALP output keeps the per-vector counts too (the host encoder's column
  // has them; the decoder ignores them in that layout).
  const int rc = encode_launch<T, true>(in, n_vecs, exps, facs, galp != 0, o, d_totals, workspace,
                                        stream);
  if (rc) return rc;
  col->ffor.bp.n_values = n_vecs * VALUES_PER_VECTOR;
  col->ffor.bp.n_vecs = n_vecs;
  return FLSGPU_OK;
}

template <typename UT>
int ffor_encode_entry(const UT *in, uint64_t n_vecs, flsgpu_ffor_column *col, uint64_t *d_totals,
                      void *workspace, hipStream_t stream) {
  if (!col) return set_error(FLSGPU_EINVAL, "null column");
  EncOut o{const_cast<void *>(col->bp.packed_array),
           const_cast<uint8_t *>(col->bp.bit_widths),
           const_cast<uint64_t *>(col->bp.vector_offsets),
           const_cast<void *>(col->bases),
           nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  const int rc = encode_launch<UT, false>(in, n_vecs, nullptr, nullptr, false, o, d_totals,
                                          workspace, stream);
  if (rc) return rc;
  col->bp.n_values = n_vecs * VALUES_PER_VECTOR;
  col->bp.n_vecs = n_vecs;
  return FLSGPU_OK;
}

}  // namespace
}  // namespace flsgpu_hip

using namespace flsgpu_hip;

extern "C" {

size_t flsgpu_encode_workspace_bytes(uint64_t n_vecs) {
  // (words, exceptions) per block of at least 8 vectors, then one fill
  // value per vector
  return size_t(2 * ((n_vecs + 7) / 8 + 1) + n_vecs) * sizeof(uint64_t);
}

int flsgpu_alp_encode_device_f32(const float *in, uint64_t n_vecs, const uint8_t *exps,
                                 const uint8_t *facs, int galp, flsgpu_alp_column *col,
                                 uint64_t *d_totals, void *d_workspace, flsgpu_stream_t stream) {
  return alp_encode_entry<float>(in, n_vecs, exps, facs, galp, col, d_totals, d_workspace, stream);
}
int flsgpu_alp_encode_device_f64(const double *in, uint64_t n_vecs, const uint8_t *exps,
                                 const uint8_t *facs, int galp, flsgpu_alp_column *col,
                                 uint64_t *d_totals, void *d_workspace, flsgpu_stream_t stream) {
  return alp_encode_entry<double>(in, n_vecs, exps, facs, galp, col, d_totals, d_workspace, stream);
}
int flsgpu_ffor_encode_device_u32(const uint32_t *in, uint64_t n_vecs, flsgpu_ffor_column *col,
                                  uint64_t *d_totals, void *d_workspace, flsgpu_stream_t stream) {
  return ffor_encode_entry<uint32_t>(in, n_vecs, col, d_totals, d_workspace, stream);
}
int flsgpu_ffor_encode_device_u64(const uint64_t *in, uint64_t n_vecs, flsgpu_ffor_column *col,
                                  uint64_t *d_totals, void *d_workspace, flsgpu_stream_t stream) {
  return ffor_encode_entry<uint64_t>(in, n_vecs, col, d_totals, d_workspace, stream);
}

}  // extern "C"
