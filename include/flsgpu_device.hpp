/*
 * flsgpu device-side decoder API (HIP, gfx950), header-only.
 *
 * For kernels that fuse decoding into their own work. It keeps the contract
 * of the reference's device decompressors: construct
 * DecompressorT(column, vector_index, lane), then call unpack_next_into(out)
 * repeatedly; each call yields the next UNPACK_N_VALUES values of FastLanes
 * lane `lane` for each of UNPACK_N_VECTORS consecutive vectors, value i of
 * vector v in out[v * UNPACK_N_VALUES + i]. After n calls the thread has
 * produced slots [0, n * UNPACK_N_VALUES) of its lane, and slot j of lane l
 * is element l + N_LANES * j of the 1024-value vector.
 *   BPDecompressor / FFORDecompressor   src/flsgpu/fls.cuh:620-650
 *   ALPDecompressor (+ exception patch) src/flsgpu/alp.cuh:463-485
 *   FillWarpMapping, write_registers_to_global, MagicChecker
 *                                        src/engine/device-utils.cuh:72-143
 *
 * Differences from the reference, all supersets of its behaviour:
 *   - the vectors of one decompressor may have different bit widths and
 *     need not be packed back to back (the reference assumes one width for
 *     all UNPACK_N_VECTORS vectors, fls.cuh:600-606);
 *   - vectors at or past column.n_vecs, and vectors whose bit width exceeds
 *     W, decode as zeros; their data is not read;
 *   - no word past a vector's own bit_width * N_LANES words is read (no
 *     over-read padding, as for the library's kernels);
 *   - exception positions may come in any order within a vector (standard
 *     layout) or within a lane (G-ALP layout); the reference's stateful
 *     patchers assume ascending positions.
 * The FACT/FRAC tables are compiled in; out-of-table indices give 0.
 *
 * Columns are the C-ABI descriptors of flsgpu.h (device pointers) wrapped as
 * typed views with the reference's field names (src/flsgpu/structs.cuh:33-84).
 * Include from a .hip file built with hipcc --offload-arch=gfx950; no library
 * call is needed.
 */
#ifndef FLSGPU_DEVICE_HPP
#define FLSGPU_DEVICE_HPP

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "flsgpu.h"

namespace flsgpu {
namespace device {

// Index types of the reference (src/flsgpu/device-types.cuh:6-9).
typedef uint8_t vbw_t;
typedef uint16_t lane_t;
typedef uint32_t si_t;
typedef uint32_t vi_t;

constexpr int VALUES_PER_VECTOR = 1024;

template <typename T> struct same_width_uint;
template <> struct same_width_uint<uint32_t> { typedef uint32_t type; };
template <> struct same_width_uint<uint64_t> { typedef uint64_t type; };
template <> struct same_width_uint<int32_t> { typedef uint32_t type; };
template <> struct same_width_uint<int64_t> { typedef uint64_t type; };
template <> struct same_width_uint<float> { typedef uint32_t type; };
template <> struct same_width_uint<double> { typedef uint64_t type; };

template <typename T> struct same_width_int;
template <> struct same_width_int<float> { typedef int32_t type; };
template <> struct same_width_int<double> { typedef int64_t type; };

template <typename T> __host__ __device__ constexpr int get_n_lanes() {
  return VALUES_PER_VECTOR / (8 * int(sizeof(T)));
}
template <typename T> __host__ __device__ constexpr int get_values_per_lane() {
  return 8 * int(sizeof(T));
}

// ---- column views ---------------------------------------------------------

template <typename T> struct BPColumn {
  typedef typename same_width_uint<T>::type UINT_T;
  uint64_t n_values = 0;
  uint64_t n_vecs = 0;
  const UINT_T *packed_array = nullptr;
  const uint8_t *bit_widths = nullptr;
  const uint64_t *vector_offsets = nullptr;
  BPColumn() = default;
  __host__ __device__ explicit BPColumn(const flsgpu_bp_column &c)
      : n_values(c.n_values), n_vecs(c.n_vecs),
        packed_array(static_cast<const UINT_T *>(c.packed_array)), bit_widths(c.bit_widths),
        vector_offsets(c.vector_offsets) {}
};

template <typename T> struct FFORColumn {
  typedef typename same_width_uint<T>::type UINT_T;
  BPColumn<T> bp;
  const UINT_T *bases = nullptr;
  FFORColumn() = default;
  __host__ __device__ explicit FFORColumn(const flsgpu_ffor_column &c)
      : bp(c.bp), bases(static_cast<const UINT_T *>(c.bases)) {}
};

// Standard exception layout: counts[v] exceptions at exceptions_offsets[v].
template <typename T> struct ALPColumn {
  typedef typename same_width_uint<T>::type UINT_T;
  FFORColumn<UINT_T> ffor;
  const uint8_t *factor_indices = nullptr;
  const uint8_t *fraction_indices = nullptr;
  uint64_t n_exceptions = 0;
  const uint64_t *exceptions_offsets = nullptr;
  const T *exceptions = nullptr;
  const uint16_t *positions = nullptr;
  const uint16_t *counts = nullptr;
  ALPColumn() = default;
  __host__ __device__ explicit ALPColumn(const flsgpu_alp_column &c)
      : ffor(c.ffor), factor_indices(c.factor_indices), fraction_indices(c.fraction_indices),
        n_exceptions(c.n_exceptions), exceptions_offsets(c.exceptions_offsets),
        exceptions(static_cast<const T *>(c.exceptions)), positions(c.positions),
        counts(c.counts) {}
};

// G-ALP layout: per lane, offsets_counts[v * N_LANES + lane] =
// (count << 10) | offset of the lane's run inside the vector's list.
template <typename T> struct ALPExtendedColumn {
  typedef typename same_width_uint<T>::type UINT_T;
  FFORColumn<UINT_T> ffor;
  const uint8_t *factor_indices = nullptr;
  const uint8_t *fraction_indices = nullptr;
  uint64_t n_exceptions = 0;
  const uint64_t *exceptions_offsets = nullptr;
  const T *exceptions = nullptr;
  const uint16_t *positions = nullptr;
  const uint16_t *offsets_counts = nullptr;
  ALPExtendedColumn() = default;
  __host__ __device__ explicit ALPExtendedColumn(const flsgpu_alp_column &c)
      : ffor(c.ffor), factor_indices(c.factor_indices), fraction_indices(c.fraction_indices),
        n_exceptions(c.n_exceptions), exceptions_offsets(c.exceptions_offsets),
        exceptions(static_cast<const T *>(c.exceptions)), positions(c.positions),
        offsets_counts(c.offsets_counts) {}
};

// ---- bit unpacking -------------------------------------------------------

// Streams one lane of one vector: slot j is bits [j*bw, (j+1)*bw) of the
// lane's words in[0], in[N_LANES], in[2*N_LANES], ... (LSB first). No shift
// by the full word width is ever executed.
template <typename UINT_T> struct LaneStream {
  static constexpr int W = 8 * int(sizeof(UINT_T));
  static constexpr int N_LANES = get_n_lanes<UINT_T>();
  const UINT_T *in = nullptr;
  UINT_T mask = 0;
  uint32_t bw = 0;
  uint32_t offset = 0;

  // Returns false for a bit width above W (the stream then yields zeros).
  __device__ __forceinline__ bool init(const UINT_T *lane_words, uint32_t bit_width) {
    in = lane_words;
    const bool ok = bit_width <= uint32_t(W);
    bw = ok ? bit_width : 0u;
    mask = bw == uint32_t(W) ? ~UINT_T(0) : ((UINT_T(1) << bw) - UINT_T(1));
    offset = 0;
    return ok;
  }

  __device__ __forceinline__ UINT_T next() {
    if (bw == 0) return UINT_T(0);
    UINT_T v = in[0] >> offset;
    const uint32_t end = offset + bw;
    if (end > uint32_t(W)) v |= in[N_LANES] << (uint32_t(W) - offset);
    if (end >= uint32_t(W)) {
      in += N_LANES;
      offset = end - uint32_t(W);
    } else {
      offset = end;
    }
    return v & mask;
  }
};

// The per-vector state of UNPACK_N_VECTORS consecutive vectors of a column.
// live[v]: vector v exists and has a valid bit width; other vectors decode as
// all zeros.
template <typename UINT_T, unsigned UNPACK_N_VECTORS> struct VectorStreams {
  LaneStream<UINT_T> s[UNPACK_N_VECTORS];
  bool live[UNPACK_N_VECTORS];

  template <typename T>
  __device__ __forceinline__ void init(const BPColumn<T> &c, vi_t vector_index, lane_t lane) {
#pragma unroll
    for (unsigned v = 0; v < UNPACK_N_VECTORS; ++v) {
      const uint64_t vec = uint64_t(vector_index) + v;
      if (vec < c.n_vecs) {
        live[v] = s[v].init(c.packed_array + c.vector_offsets[vec] + lane, c.bit_widths[vec]);
      } else {
        s[v].init(c.packed_array, 0);
        live[v] = false;
      }
    }
  }
};

// ---- ALP value and tables ------------------------------------------------

// ALP decode of one value: (T)(INT)((UINT)(u + base) * (UINT)FACT[f]) * FRAC[e]
// (src/flsgpu/alp.cuh:40-45, src/alp/constants.hpp:49-55,76-144). The f32
// factor table carries 10^10 wrapped to int32 at index 10, as the library's.
__device__ __forceinline__ uint32_t alp_fact(float, unsigned i) {
  constexpr int32_t F[11] = {1, 10, 100, 1000, 10000, 100000, 1000000, 10000000, 100000000,
                             1000000000, 1410065408};
  return i < 11 ? uint32_t(F[i]) : 0u;
}
__device__ __forceinline__ uint64_t alp_fact(double, unsigned i) {
  uint64_t f = 1;
  if (i > 18) return 0;
  for (unsigned k = 0; k < i; ++k) f *= 10u;
  return f;
}
__device__ __forceinline__ float alp_frac(float, unsigned i) {
  // double literals narrowed to float, as the reference's table is written
  constexpr float F[11] = {1.0,  0.1,  0.01, 0.001, 0.0001, 0.00001, 0.000001,
                           1e-7, 1e-8, 1e-9, 1e-10};
  return i < 11 ? F[i] : 0.0f;
}
__device__ __forceinline__ double alp_frac(double, unsigned i) {
  constexpr double F[21] = {1.0,   0.1,   0.01,  0.001, 0.0001, 1e-5,  1e-6,
                            1e-7,  1e-8,  1e-9,  1e-10, 1e-11,  1e-12, 1e-13,
                            1e-14, 1e-15, 1e-16, 1e-17, 1e-18,  1e-19, 1e-20};
  return i < 21 ? F[i] : 0.0;
}

template <typename T>
__device__ __forceinline__ T alp_decode_value(typename same_width_uint<T>::type packed,
                                              typename same_width_uint<T>::type base,
                                              typename same_width_uint<T>::type fact, T frac) {
  typedef typename same_width_uint<T>::type UINT_T;
  typedef typename same_width_int<T>::type INT_T;
  const INT_T s = INT_T(UINT_T(UINT_T(packed + base) * fact));
  return T(s) * frac;  // one rounding conversion, one multiply: nothing to contract
}

// ---- decompressors -------------------------------------------------------

// device::BPDecompressor (src/flsgpu/fls.cuh:620-634).
template <typename T, unsigned UNPACK_N_VECTORS, unsigned UNPACK_N_VALUES>
struct BPDecompressor {
  typedef typename same_width_uint<T>::type UINT_T;
  VectorStreams<UINT_T, UNPACK_N_VECTORS> streams;

  __device__ __forceinline__ BPDecompressor(const BPColumn<T> &column, vi_t vector_index,
                                            lane_t lane) {
    streams.init(column, vector_index, lane);
  }

  __device__ __forceinline__ void unpack_next_into(T *__restrict out) {
#pragma unroll
    for (unsigned i = 0; i < UNPACK_N_VALUES; ++i)
#pragma unroll
      for (unsigned v = 0; v < UNPACK_N_VECTORS; ++v)
        out[UNPACK_N_VALUES * v + i] = T(streams.s[v].next());
  }
};

// device::FFORDecompressor (fls.cuh:636-650): value + base mod 2^W.
template <typename T, unsigned UNPACK_N_VECTORS, unsigned UNPACK_N_VALUES>
struct FFORDecompressor {
  typedef typename same_width_uint<T>::type UINT_T;
  VectorStreams<UINT_T, UNPACK_N_VECTORS> streams;
  UINT_T bases[UNPACK_N_VECTORS];

  __device__ __forceinline__ FFORDecompressor(const FFORColumn<T> &column, vi_t vector_index,
                                              lane_t lane) {
    streams.init(column.bp, vector_index, lane);
#pragma unroll
    for (unsigned v = 0; v < UNPACK_N_VECTORS; ++v) {
      const uint64_t vec = uint64_t(vector_index) + v;
      bases[v] = streams.live[v] ? column.bases[vec] : UINT_T(0);
    }
  }

  __device__ __forceinline__ void unpack_next_into(T *__restrict out) {
#pragma unroll
    for (unsigned i = 0; i < UNPACK_N_VALUES; ++i)
#pragma unroll
      for (unsigned v = 0; v < UNPACK_N_VECTORS; ++v)
        out[UNPACK_N_VALUES * v + i] = T(UINT_T(streams.s[v].next() + bases[v]));
  }
};

// One lane's exceptions of one vector: a bit mask of the lane's slots that
// hold an exception, and the run of the exception list to search on a hit.
template <typename T> struct LaneExceptions {
  typedef typename same_width_uint<T>::type UINT_T;
  static constexpr int N_LANES = get_n_lanes<T>();
  uint64_t slots = 0;           // bit j: slot j of this lane is an exception
  const uint16_t *pos = nullptr;
  const T *val = nullptr;
  uint32_t count = 0;

  __device__ __forceinline__ void scan(lane_t lane) {
    slots = 0;
    for (uint32_t k = 0; k < count; ++k) {
      const uint32_t p = pos[k] & 1023u;
      if (p % N_LANES == lane) slots |= uint64_t(1) << (p / N_LANES);
    }
  }

  // Replaces `value` if slot j of the lane is an exception.
  __device__ __forceinline__ T patch(lane_t lane, uint32_t j, T value) const {
    if (!((slots >> j) & 1u)) return value;
    const uint32_t target = uint32_t(lane) + uint32_t(N_LANES) * j;
    for (uint32_t k = 0; k < count; ++k)
      if ((pos[k] & 1023u) == target) return val[k];
    return value;
  }
};

template <typename T> struct ExceptionSource;
template <typename T> struct ExceptionSource<ALPColumn<T>> {
  __device__ __forceinline__ static void locate(const ALPColumn<T> &c, uint64_t vec, lane_t,
                                                LaneExceptions<T> *e) {
    const uint64_t eo = c.exceptions_offsets[vec];
    e->pos = c.positions + eo;
    e->val = c.exceptions + eo;
    e->count = c.counts[vec];
  }
};
template <typename T> struct ExceptionSource<ALPExtendedColumn<T>> {
  __device__ __forceinline__ static void locate(const ALPExtendedColumn<T> &c, uint64_t vec,
                                                lane_t lane, LaneExceptions<T> *e) {
    const uint32_t oc = c.offsets_counts[vec * uint64_t(get_n_lanes<T>()) + lane];
    const uint64_t start = c.exceptions_offsets[vec] + (oc & 0x3FFu);
    e->pos = c.positions + start;
    e->val = c.exceptions + start;
    e->count = oc >> 10;
  }
};

// device::ALPDecompressor (src/flsgpu/alp.cuh:463-485) for ALPColumn<T>
// (standard layout) or ALPExtendedColumn<T> (G-ALP). PATCH = false leaves
// the exceptions unpatched (the reference's Dummy patcher, alp.cuh:53-62).
template <typename T, unsigned UNPACK_N_VECTORS, unsigned UNPACK_N_VALUES, typename ColumnT,
          bool PATCH = true>
struct ALPDecompressor {
  typedef typename same_width_uint<T>::type UINT_T;
  VectorStreams<UINT_T, UNPACK_N_VECTORS> streams;
  UINT_T bases[UNPACK_N_VECTORS];
  UINT_T fact[UNPACK_N_VECTORS];
  T frac[UNPACK_N_VECTORS];
  LaneExceptions<T> exc[PATCH ? UNPACK_N_VECTORS : 1];
  lane_t lane;
  uint32_t slot = 0;

  __device__ __forceinline__ ALPDecompressor(const ColumnT &column, vi_t vector_index,
                                             lane_t a_lane)
      : lane(a_lane) {
    streams.init(column.ffor.bp, vector_index, lane);
#pragma unroll
    for (unsigned v = 0; v < UNPACK_N_VECTORS; ++v) {
      const uint64_t vec = uint64_t(vector_index) + v;
      const bool live = streams.live[v];
      bases[v] = live ? column.ffor.bases[vec] : UINT_T(0);
      fact[v] = live ? alp_fact(T(0), column.factor_indices[vec]) : UINT_T(0);
      frac[v] = live ? alp_frac(T(0), column.fraction_indices[vec]) : T(0);
      if constexpr (PATCH) {
        if (live) {
          ExceptionSource<ColumnT>::locate(column, vec, lane, &exc[v]);
          exc[v].scan(lane);
        }
      }
    }
  }

  __device__ __forceinline__ void unpack_next_into(T *__restrict out) {
#pragma unroll
    for (unsigned i = 0; i < UNPACK_N_VALUES; ++i) {
#pragma unroll
      for (unsigned v = 0; v < UNPACK_N_VECTORS; ++v) {
        T x = alp_decode_value<T>(streams.s[v].next(), bases[v], fact[v], frac[v]);
        if constexpr (PATCH) x = exc[v].patch(lane, slot + i, x);
        x = streams.live[v] ? x : T(0);
        out[UNPACK_N_VALUES * v + i] = x;
      }
    }
    slot += UNPACK_N_VALUES;
  }
};

// ---- kernel-side helpers ---------------------------------------------------

// FillWarpMapping (src/engine/device-utils.cuh:72-95): thread t of a block of
// blockDim.x threads decodes lane t % N_LANES of vectors starting at the
// returned index; a block covers (blockDim.x / N_LANES) * UNPACK_N_VECTORS
// vectors. blockDim.x must be a multiple of N_LANES (64 works for every type).
template <typename T, unsigned UNPACK_N_VECTORS> struct FillWarpMapping {
  static constexpr uint32_t N_LANES = get_n_lanes<T>();
  static constexpr uint32_t N_VALUES_IN_LANE = get_values_per_lane<T>();
  __device__ __forceinline__ lane_t get_lane() const { return lane_t(threadIdx.x % N_LANES); }
  __device__ __forceinline__ vi_t get_vector_index() const {
    const uint32_t per_block = (blockDim.x / N_LANES) * UNPACK_N_VECTORS;
    return vi_t(per_block * blockIdx.x + (threadIdx.x / N_LANES) * UNPACK_N_VECTORS);
  }
};
template <typename T, unsigned UNPACK_N_VECTORS>
using VectorToWarpMapping = FillWarpMapping<T, UNPACK_N_VECTORS>;

// write_registers_to_global (device-utils.cuh:109-120); `out` points at the
// first vector's first element.
template <typename T, unsigned UNPACK_N_VECTORS, unsigned UNPACK_N_VALUES, unsigned N_LANES>
__device__ __forceinline__ void write_registers_to_global(lane_t lane, si_t index_offset,
                                                          const T *__restrict registers,
                                                          T *__restrict out) {
#pragma unroll
  for (unsigned v = 0; v < UNPACK_N_VECTORS; ++v)
#pragma unroll
    for (unsigned w = 0; w < UNPACK_N_VALUES; ++w)
      out[lane + (index_offset + w) * N_LANES + v * VALUES_PER_VECTOR] =
          registers[w + v * UNPACK_N_VALUES];
}

// MagicChecker (device-utils.cuh:122-143): writes true to *out if any checked
// value equals the magic value; never writes false.
template <typename T, unsigned N_VALUES> struct MagicChecker {
  const T magic_value;
  bool no_magic_found = true;
  __device__ __forceinline__ explicit MagicChecker(T magic) : magic_value(magic) {}
  __device__ __forceinline__ void check(const T *__restrict registers) {
#pragma unroll
    for (unsigned i = 0; i < N_VALUES; ++i) no_magic_found &= registers[i] != magic_value;
  }
  __device__ __forceinline__ void write_result(bool *__restrict out) {
    if (!no_magic_found) *out = true;
  }
};

}  // namespace device
}  // namespace flsgpu

#endif  // FLSGPU_DEVICE_HPP
