// FastLanes bit-unpacking / FFOR kernels for gfx950: decompress, "any ==
// magic" query, fused filter + sum, the synthetic compute kernel and the
// multi-column query. Included only by fls_inst.hip.
//
// Reference behaviour: src/engine/kernels.cuh:13-153 (decompress/query/compute
// kernels and host launchers), src/flsgpu/fls.cuh:17-41 (BP/FFOR functors:
// value + base, wrapping mod 2^W), src/engine/device-utils.cuh:109-143
// (output index lane + N_LANES*j + 1024*v, MagicChecker),
// src/generated-bindings/multi-column-device-kernels.cuh:10-443.
#pragma once

#include <algorithm>

#include "fls_impl.hpp"
#include "kernels.hpp"
#include "stage.hpp"

namespace flsgpu_hip {

// ---------------------------------------------------------------------------
// Policies: what happens to each unpacked value.
// ---------------------------------------------------------------------------
template <typename UT, bool NT>
struct DecompressPolicy {
  const UT *in;
  UT *out;
  UT base;
  template <int BW, int J0, int NJ>
  __device__ __forceinline__ void run() {
    unpack_slots<UT, BW, J0, NJ>(in, [&](int j, UT v) {
      store<UT, NT>(out + j * lanes<UT>::N, UT(v + base));
    });
  }
  __device__ __forceinline__ void run_generic(int bw, int j0, int nj) {
    unpack_runtime<UT>(in, bw, j0, nj, [&](int j, UT v) {
      store<UT, NT>(out + j * lanes<UT>::N, UT(v + base));
    });
  }
};

// Writes value + base to the lane's column of the wave's LDS chunk: slot j
// at lds[(j - j0) * N_LANES].
template <typename UT>
struct StagePolicy {
  const UT *in;
  UT *lds;
  UT base;
  template <int BW, int J0, int NJ>
  __device__ __forceinline__ void run() {
    unpack_slots<UT, BW, J0, NJ>(in, [&](int j, UT v) { lds[(j - J0) * lanes<UT>::N] = UT(v + base); });
  }
  __device__ __forceinline__ void run_generic(int bw, int j0, int nj) {
    unpack_runtime<UT>(in, bw, j0, nj, [&](int j, UT v) { lds[(j - j0) * lanes<UT>::N] = UT(v + base); });
  }
};

template <typename UT>
struct AnyPolicy {
  const UT *in;
  UT base;
  UT magic;
  bool found;
  template <int BW, int J0, int NJ>
  __device__ __forceinline__ void run() {
    unpack_slots<UT, BW, J0, NJ>(in, [&](int, UT v) { found |= UT(v + base) == magic; });
  }
  __device__ __forceinline__ void run_generic(int bw, int j0, int nj) {
    unpack_runtime<UT>(in, bw, j0, nj, [&](int, UT v) { found |= UT(v + base) == magic; });
  }
};

// Fused filter + sum. With no wrap-around in (offset + base) the sum of the
// passing values is sum(offsets) + count * base, so the per-value work is an
// unsigned range test and one narrow add on the FOR offset; the base is
// applied once per lane (the ALU cost per value, not HBM, otherwise bounds
// this kernel). Columns whose base + 2^BW - 1 wraps take the per-value path.
template <typename UT, int PRED>
struct FilterSumPolicy {
  const UT *in;
  UT base;
  UT lo, hi;
  uint64_t count, sum;
  template <int BW, int J0, int NJ>
  __device__ __forceinline__ void run() {
    constexpr int W = lanes<UT>::W;
    constexpr UT MAXOFF = BW == W ? UT(~UT(0)) : UT((UT(1) << BW) - 1);
    // per-lane offset sums stay below 2^32 when BW + log2(NJ) <= 32
    using Acc = std::conditional_t<(W == 32 && BW + 6 <= 32), uint32_t, uint64_t>;
    const bool no_wrap = base <= UT(UT(~UT(0)) - MAXOFF);
    if constexpr (PRED == FLSGPU_PRED_OFFSET_LOWER_HALF) {
      if constexpr (BW > 0) {  // no row passes at bit width 0
        if (W == 64 || no_wrap) {  // 64-bit sums are mod 2^64 anyway
          Acc acc = 0;
          uint32_t cnt = 0;
          unpack_slots<UT, BW, J0, NJ>(in, [&](int, UT off) {
            const bool pass = (off >> (BW - 1)) == 0;
            cnt += pass;
            acc += pass ? Acc(off) : Acc(0);
          });
          count += cnt;
          sum += uint64_t(acc) + uint64_t(cnt) * uint64_t(base);
        } else {
          unpack_slots<UT, BW, J0, NJ>(in, [&](int, UT off) {
            const bool pass = (off >> (BW - 1)) == 0;
            count += pass;
            sum += pass ? uint64_t(UT(off + base)) : 0;
          });
        }
      }
    } else if constexpr (PRED == FLSGPU_PRED_ALL) {
      if (W == 64 || no_wrap) {
        Acc acc = 0;
        unpack_slots<UT, BW, J0, NJ>(in, [&](int, UT off) { acc += Acc(off); });
        count += NJ;
        sum += uint64_t(acc) + uint64_t(NJ) * uint64_t(base);
      } else {
        unpack_slots<UT, BW, J0, NJ>(in, [&](int, UT off) { sum += uint64_t(UT(off + base)); });
        count += NJ;
      }
    } else {
      if (no_wrap) {
        // lo <= base + off < hi  <=>  off - olo < span (unsigned)
        const UT from = lo > base ? lo : base;
        const UT olo = UT(from - base);
        const UT span = hi > from ? UT(hi - from) : UT(0);
        Acc acc = 0;
        uint32_t cnt = 0;
        unpack_slots<UT, BW, J0, NJ>(in, [&](int, UT off) {
          const bool pass = UT(off - olo) < span;
          cnt += pass;
          acc += pass ? Acc(off) : Acc(0);
        });
        count += cnt;
        sum += uint64_t(acc) + uint64_t(cnt) * uint64_t(base);
      } else {
        unpack_slots<UT, BW, J0, NJ>(in, [&](int, UT off) {
          const UT v = UT(off + base);
          const bool pass = (v >= lo) & (v < hi);
          count += pass;
          sum += pass ? uint64_t(v) : 0;
        });
      }
    }
  }
  __device__ __forceinline__ void run_generic(int bw, int j0, int nj) {
    const UT half = bw > 0 ? UT(UT(1) << (bw - 1)) : UT(0);
    unpack_runtime<UT>(in, bw, j0, nj, [&](int, UT off) {
      const UT v = UT(off + base);
      bool pass;
      if constexpr (PRED == FLSGPU_PRED_OFFSET_LOWER_HALF) pass = off < half;
      else if constexpr (PRED == FLSGPU_PRED_ALL) pass = true;
      else pass = (v >= lo) & (v < hi);
      count += pass;
      sum += pass ? uint64_t(v) : 0;
    });
  }
};

template <typename UT>
struct ComputePolicy {
  const UT *in;
  UT base;
  uint32_t reps;
  UT runtime_zero;
  bool found;
  template <int BW, int J0, int NJ>
  __device__ __forceinline__ void run() {
    unpack_slots<UT, BW, J0, NJ>(in, [&](int, UT v) {
      UT x = UT(v + base);
      for (uint32_t k = 0; k < reps; ++k) {
        x *= UT(3);
        x <<= 3;
        x += UT(3);
        x >>= 3;
        x &= runtime_zero;
      }
      found |= x == UT(1);
    });
  }
  __device__ __forceinline__ void run_generic(int bw, int j0, int nj) {
    unpack_runtime<UT>(in, bw, j0, nj, [&](int, UT v) {
      UT x = UT(v + base);
      for (uint32_t k = 0; k < reps; ++k) {
        x *= UT(3);
        x <<= 3;
        x += UT(3);
        x >>= 3;
        x &= runtime_zero;
      }
      found |= x == UT(1);
    });
  }
};

// Bit j - J0 of `mask` is set when slot j of this lane equals `value`.
template <typename UT>
struct MatchMaskPolicy {
  const UT *in;
  UT base;
  UT value;
  uint64_t mask;
  template <int BW, int J0, int NJ>
  __device__ __forceinline__ void run() {
    unpack_slots<UT, BW, J0, NJ>(in, [&](int j, UT v) {
      mask |= uint64_t(UT(v + base) == value) << (j - J0);
    });
  }
  __device__ __forceinline__ void run_generic(int bw, int j0, int nj) {
    unpack_runtime<UT>(in, bw, j0, nj, [&](int j, UT v) {
      mask |= uint64_t(UT(v + base) == value) << (j - j0);
    });
  }
};

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------
// Decodes wave-unit `wave` (VPW vectors, slot part wave % S) of a column.
template <typename UT, int S, bool NT>
__device__ __forceinline__ void ffor_decompress_unit(const FforArgs<UT> &a, UT *__restrict out,
                                                     uint64_t wave) {
  using L = lanes<UT>;
  const UnitMap<UT, S> u = map_unit<UT, S>(wave);
  if (u.vec >= a.n_vecs) return;
  DecompressPolicy<UT, NT> pol{a.packed + a.vector_offsets[u.vec] + u.lane,
                               out + u.vec * VALUES_PER_VECTOR + u.lane,
                               a.bases ? a.bases[u.vec] : UT(0)};
  if (!dispatch_bw<UT, S>(pol, a.bit_widths[u.vec], u.part)) {
    // Bit width above W: the vector decodes as zeros.
    for (int j = u.part * (L::W / S); j < (u.part + 1) * (L::W / S); ++j)
      out[u.vec * VALUES_PER_VECTOR + u.lane + j * L::N] = UT(0);
  }
}

template <typename UT, int S, bool NT>
__global__ __launch_bounds__(BLOCK) void k_ffor_decompress(FforArgs<UT> a, UT *__restrict out) {
  ffor_decompress_unit<UT, S, NT>(a, out, (uint64_t(blockIdx.x) * BLOCK + threadIdx.x) / WAVE);
}

// Staged variant: one wave-unit, decoded into the wave's LDS stage `lds`
// (VPW * CHUNK elements, 16-byte aligned), stored 16 bytes per lane.
template <typename UT, int S>
__device__ __forceinline__ void ffor_staged_unit(const FforArgs<UT> &a, UT *__restrict out,
                                                 uint64_t wave, UT *lds) {
  using L = lanes<UT>;
  using Sh = StageShape<UT, S>;
  const UnitMap<UT, S> u = map_unit<UT, S>(wave);
  const int vv = int(threadIdx.x % WAVE) / L::N;
  if (u.vec < a.n_vecs) {
    StagePolicy<UT> pol{a.packed + a.vector_offsets[u.vec] + u.lane, lds + vv * Sh::CHUNK + u.lane,
                        a.bases ? a.bases[u.vec] : UT(0)};
    if (!dispatch_bw<UT, S>(pol, a.bit_widths[u.vec], u.part)) {
      for (int jj = 0; jj < Sh::NJ; ++jj) lds[vv * Sh::CHUNK + u.lane + jj * L::N] = UT(0);
    }
  }
  wave_lds_sync();
  store_staged<UT, S>(lds, out, wave / S, u.part, a.n_vecs);
}

template <typename UT, int S>
__global__ __launch_bounds__(BLOCK) void k_ffor_decompress_staged(FforArgs<UT> a,
                                                                 UT *__restrict out) {
  __shared__ __align__(16) UT s_stage[BLOCK / WAVE][lanes<UT>::VPW * StageShape<UT, S>::CHUNK];
  ffor_staged_unit<UT, S>(a, out, (uint64_t(blockIdx.x) * BLOCK + threadIdx.x) / WAVE,
                          s_stage[threadIdx.x / WAVE]);
}

// Blocks of the scanning kernels (query any, fused filter + sum): enough to
// fill every CU (8 blocks of 4 waves on each of 256 CUs); every wave then
// walks many units, loading the next unit's metadata while it reduces the
// current one. One unit per wave (a grid of n_waves / 4 blocks) made the
// narrow widths latency-bound: a wave lived for two dependent memory round
// trips and a block reduction to decode a few words per lane (u32 width 1
// filter + sum over 2^28 values: 114 us, now 67 us).
constexpr unsigned SCAN_GRID = 2048;

// Calls f(unit, bit_width, vector_offset, base) for this wave's units
// wave0, wave0 + stride, ... below n_waves whose vector exists.
template <typename UT, int S, class F>
__device__ __forceinline__ void for_each_unit_prefetched(const FforArgs<UT> &a, uint64_t n_waves,
                                                         F &&f) {
  const uint64_t stride = uint64_t(gridDim.x) * (BLOCK / WAVE);
  uint64_t wave = (uint64_t(blockIdx.x) * BLOCK + threadIdx.x) / WAVE;
  UnitMap<UT, S> u = map_unit<UT, S>(wave);
  bool live = wave < n_waves && u.vec < a.n_vecs;
  int bw = live ? int(a.bit_widths[u.vec]) : 0;
  uint64_t off = live ? a.vector_offsets[u.vec] : 0;
  UT base = (live && a.bases) ? a.bases[u.vec] : UT(0);
  while (wave < n_waves) {
    const uint64_t wn = wave + stride;
    const UnitMap<UT, S> un = map_unit<UT, S>(wn);
    const bool live_n = wn < n_waves && un.vec < a.n_vecs;
    const int bw_n = live_n ? int(a.bit_widths[un.vec]) : 0;
    const uint64_t off_n = live_n ? a.vector_offsets[un.vec] : 0;
    const UT base_n = (live_n && a.bases) ? a.bases[un.vec] : UT(0);
    if (live) f(u, bw, off, base);
    wave = wn;
    u = un;
    live = live_n;
    bw = bw_n;
    off = off_n;
    base = base_n;
  }
}

template <typename UT, int S>
__global__ __launch_bounds__(BLOCK) void k_ffor_any(FforArgs<UT> a, UT magic, uint64_t n_waves,
                                                   Scratch s) {
  bool found = false;
  for_each_unit_prefetched<UT, S>(a, n_waves, [&](const UnitMap<UT, S> &u, int bw, uint64_t off,
                                                  UT base) {
    AnyPolicy<UT> pol{a.packed + off + u.lane, base, magic, false};
    if (dispatch_bw<UT, S>(pol, bw, u.part)) found |= pol.found;
    else found |= magic == UT(0);
  });
  block_partial2<RedOp::Or>(found, 0, s.partials);
}

template <typename UT, int S, int PRED>
__global__ __launch_bounds__(BLOCK) void k_ffor_filter_sum(FforArgs<UT> a, UT lo, UT hi,
                                                          uint64_t n_waves, Scratch s) {
  uint64_t count = 0, sum = 0;
  for_each_unit_prefetched<UT, S>(a, n_waves, [&](const UnitMap<UT, S> &u, int bw, uint64_t off,
                                                  UT base) {
    FilterSumPolicy<UT, PRED> pol{a.packed + off + u.lane, base, lo, hi, 0, 0};
    dispatch_bw<UT, S>(pol, bw, u.part);
    count += pol.count;
    sum += pol.sum;
  });
  block_partial2<RedOp::Add>(count, sum, s.partials);
}

template <typename UT>
__global__ __launch_bounds__(BLOCK) void k_ffor_compute(FforArgs<UT> a, uint32_t reps,
                                                       UT runtime_zero, uint64_t n_waves,
                                                       Scratch s) {
  bool found = false;
  const uint64_t stride = uint64_t(gridDim.x) * (BLOCK / WAVE);
  for (uint64_t wave = (uint64_t(blockIdx.x) * BLOCK + threadIdx.x) / WAVE; wave < n_waves;
       wave += stride) {
    const UnitMap<UT, 1> u = map_unit<UT, 1>(wave);
    if (u.vec < a.n_vecs) {
      ComputePolicy<UT> pol{a.packed + a.vector_offsets[u.vec] + u.lane,
                            a.bases ? a.bases[u.vec] : UT(0), reps, runtime_zero, false};
      dispatch_bw<UT, 1>(pol, a.bit_widths[u.vec], 0);
      found |= pol.found;
    }
  }
  block_partial2<RedOp::Or>(found, 0, s.partials);
}

// One thread per (vector, lane): the answer is 1 iff some lane has every slot
// equal to `value` in every column.
template <typename UT>
__global__ __launch_bounds__(BLOCK) void k_ffor_multi(MultiFforArgs<UT> a, UT value,
                                                     uint64_t n_waves, Scratch s) {
  using L = lanes<UT>;
  constexpr uint64_t FULL = L::W == 64 ? ~uint64_t(0) : ((uint64_t(1) << L::W) - 1);
  bool found = false;
  const uint64_t stride = uint64_t(gridDim.x) * (BLOCK / WAVE);
  const uint64_t n_vecs = a.col[0].n_vecs;
  for (uint64_t wave = (uint64_t(blockIdx.x) * BLOCK + threadIdx.x) / WAVE; wave < n_waves;
       wave += stride) {
    const UnitMap<UT, 1> u = map_unit<UT, 1>(wave);
    if (u.vec < n_vecs) {
      uint64_t all = FULL;
      for (uint32_t c = 0; c < a.n_cols; ++c) {
        const FforArgs<UT> &k = a.col[c];
        MatchMaskPolicy<UT> pol{k.packed + k.vector_offsets[u.vec] + u.lane,
                                k.bases ? k.bases[u.vec] : UT(0), value, 0};
        if (!dispatch_bw<UT, 1>(pol, k.bit_widths[u.vec], 0))
          pol.mask = value == UT(0) ? FULL : 0;
        all &= pol.mask;
      }
      found |= all == FULL;
    }
  }
  block_partial2<RedOp::Or>(found, 0, s.partials);
}

// ---------------------------------------------------------------------------
// Launchers (explicitly instantiated in fls_inst.hip)
// ---------------------------------------------------------------------------
template <typename UT, int S, bool NT>
int FlsDecompress<UT, S, NT>::launch(const FforArgs<UT> &a, UT *out, hipStream_t stream) {
  if (a.n_vecs == 0) return FLSGPU_OK;
  const uint64_t blocks = (n_waves_for<UT, S>(a.n_vecs) + (BLOCK / WAVE) - 1) / (BLOCK / WAVE);
  if (blocks > 0x7FFFFFFFull) return set_error(FLSGPU_EINVAL, "column too large");
  hipLaunchKernelGGL((k_ffor_decompress<UT, S, NT>), dim3(unsigned(blocks)), dim3(BLOCK), 0,
                     stream, a, out);
  return check_hip(hipGetLastError(), "k_ffor_decompress launch");
}

template <typename UT, int S>
int FlsDecompressStaged<UT, S>::launch(const FforArgs<UT> &a, UT *out, hipStream_t stream) {
  if (a.n_vecs == 0) return FLSGPU_OK;
  const uint64_t blocks = (n_waves_for<UT, S>(a.n_vecs) + (BLOCK / WAVE) - 1) / (BLOCK / WAVE);
  if (blocks > 0x7FFFFFFFull) return set_error(FLSGPU_EINVAL, "column too large");
  hipLaunchKernelGGL((k_ffor_decompress_staged<UT, S>), dim3(unsigned(blocks)), dim3(BLOCK), 0,
                     stream, a, out);
  return check_hip(hipGetLastError(), "k_ffor_decompress_staged launch");
}

template <typename UT, int S>
int FlsAny<UT, S>::launch(const FforArgs<UT> &a, UT magic, uint8_t *d_found, hipStream_t stream) {
  Scratch s;
  int rc = get_scratch(stream, &s);
  if (rc) return rc;
  const uint64_t n_waves = n_waves_for<UT, S>(a.n_vecs);
  const unsigned grid = std::min(reduce_grid(n_waves), SCAN_GRID);
  hipLaunchKernelGGL((k_ffor_any<UT, S>), dim3(grid), dim3(BLOCK), 0, stream, a, magic, n_waves,
                     s);
  if ((rc = check_hip(hipGetLastError(), "k_ffor_any launch"))) return rc;
  return reduce_finish_flag(s, grid, d_found, stream);
}

template <typename UT, int S, int PRED>
int FlsFilterSum<UT, S, PRED>::launch(const FforArgs<UT> &a, UT lo, UT hi, uint64_t *d_result,
                                      hipStream_t stream) {
  Scratch s;
  int rc = get_scratch(stream, &s);
  if (rc) return rc;
  return launch_partials(a, lo, hi, d_result, s.partials, MAX_REDUCE_BLOCKS, stream);
}

template <typename UT, int S, int PRED>
int FlsFilterSum<UT, S, PRED>::launch_partials(const FforArgs<UT> &a, UT lo, UT hi,
                                               uint64_t *d_result, uint64_t *partials,
                                               unsigned slots, hipStream_t stream) {
  const uint64_t n_waves = n_waves_for<UT, S>(a.n_vecs);
  const unsigned grid = std::min({reduce_grid(n_waves), SCAN_GRID, slots});
  if (grid == 0) return set_error(FLSGPU_EINVAL, "no room for block partials");
  const Scratch s{partials};
  hipLaunchKernelGGL((k_ffor_filter_sum<UT, S, PRED>), dim3(grid), dim3(BLOCK), 0, stream, a, lo,
                     hi, n_waves, s);
  int rc = check_hip(hipGetLastError(), "k_ffor_filter_sum launch");
  if (rc) return rc;
  return reduce_finish_pair(s, grid, d_result, stream);
}

template <typename UT>
int FlsCompute<UT>::launch(const FforArgs<UT> &a, uint32_t reps, uint8_t *d_found,
                           hipStream_t stream) {
  Scratch s;
  int rc = get_scratch(stream, &s);
  if (rc) return rc;
  const uint64_t n_waves = n_waves_for<UT, 1>(a.n_vecs);
  const unsigned grid = reduce_grid(n_waves);
  hipLaunchKernelGGL((k_ffor_compute<UT>), dim3(grid), dim3(BLOCK), 0, stream, a, reps, UT(0),
                     n_waves, s);
  if ((rc = check_hip(hipGetLastError(), "k_ffor_compute launch"))) return rc;
  return reduce_finish_flag(s, grid, d_found, stream);
}

template <typename UT>
int FlsMulti<UT>::launch(const MultiFforArgs<UT> &a, UT value, uint8_t *d_found,
                         hipStream_t stream) {
  Scratch s;
  int rc = get_scratch(stream, &s);
  if (rc) return rc;
  const uint64_t n_waves = n_waves_for<UT, 1>(a.col[0].n_vecs);
  const unsigned grid = reduce_grid(n_waves);
  hipLaunchKernelGGL((k_ffor_multi<UT>), dim3(grid), dim3(BLOCK), 0, stream, a, value, n_waves,
                     s);
  if ((rc = check_hip(hipGetLastError(), "k_ffor_multi launch"))) return rc;
  return reduce_finish_flag(s, grid, d_found, stream);
}

}  // namespace flsgpu_hip
