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

// The range-scan kernel's predicates: the public FLSGPU_PRED_* plus value ==
// lo, which serves the "any value == magic" query (its count of matching
// rows is reduced as a flag).
constexpr int PRED_EQUAL = 3;

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


// Fused filter + sum, runtime-width path (wavefronts whose vectors differ in
// bit width): exact per value.
template <typename UT, int PRED>
struct FilterSumPolicy {
  const UT *in;
  UT base;
  UT lo, hi;
  uint64_t count, sum;
  __device__ __forceinline__ void run_generic(int bw, int j0, int nj) {
    const UT half = bw > 0 ? UT(UT(1) << (bw - 1)) : UT(0);
    unpack_runtime<UT>(in, bw, j0, nj, [&](int, UT off) {
      const UT v = UT(off + base);
      bool pass;
      if constexpr (PRED == FLSGPU_PRED_OFFSET_LOWER_HALF) pass = off < half;
      else if constexpr (PRED == FLSGPU_PRED_ALL) pass = true;
      else if constexpr (PRED == PRED_EQUAL) pass = v == lo;
      else pass = (v >= lo) & (v < hi);
      count += pass;
      sum += pass ? uint64_t(v) : 0;
    });
  }
};

// Bit (q mod 32) of word q / 32 of a span's 32-bit word stream is set for
// every slot j in [J0, J0 + NJ) whose bit B sits at span position
// q = j*BW + B - 32*H0 (H0 = the span's first 32-bit lane word).
template <int BW, int J0, int NJ, int B, int H0>
__host__ __device__ constexpr uint32_t slot_bit_mask(int word) {
  uint32_t m = 0;
  for (int j = J0; j < J0 + NJ; ++j) {
    const int q = j * BW + B - 32 * H0;
    if (q >= 32 * word && q < 32 * word + 32) m |= 1u << (q - 32 * word);
  }
  return m;
}

__host__ __device__ constexpr int ilog2(int x) { return x <= 1 ? 0 : 1 + ilog2(x / 2); }

// Calls f(field) for slots [J0, J0 + NJ) of one lane at a width below 32,
// each field read from the lane's stream as 32-bit words (for 64-bit types
// the low half of each word first; a field spans at most two of them).
template <typename UT, int BW, int J0, int NJ, class F>
__device__ __forceinline__ void span_fields32(const UT *w, F &&f) {
  static_assert(BW > 0 && BW < 32, "32-bit fields");
  constexpr int W = lanes<UT>::W;
  using Sp = SlotSpan<UT, BW, J0, NJ>;
  constexpr int H0 = Sp::R0 * (W / 32);
  constexpr int H = Sp::NR * (W / 32);
  auto hw = [&](auto ic) -> uint32_t {
    constexpr int i = decltype(ic)::value;
    if constexpr (i >= H) return 0u;
    else if constexpr (W == 32) return uint32_t(w[i]);
    else return uint32_t(uint64_t(w[i / 2]) >> (32 * (i % 2)));
  };
  static_for<NJ>([&](auto jc) {
    constexpr int q = (J0 + int(decltype(jc)::value)) * BW - 32 * H0;
    constexpr int i = q / 32;
    constexpr int sh = q % 32;
    uint32_t x;
    if constexpr (sh + BW <= 32) {
      x = hw(std::integral_constant<int, i>{}) >> sh;
    } else {
      x = __builtin_amdgcn_alignbit(hw(std::integral_constant<int, i + 1>{}),
                                    hw(std::integral_constant<int, i>{}), uint32_t(sh));
    }
    f(x & ((1u << BW) - 1u));
  });
}

// Count of slots [J0, J0 + NJ) of one lane that pass PRED (OFFSET_LOWER_HALF:
// the FOR offset's top bit is clear; ALL) and the sum of their offsets mod
// 2^64, from the span's words in registers. The work is on the packed words,
// not per value where a bit trick allows:
//  - the passing count is NJ minus a popcount of the top-bit positions;
//  - OFFSET_LOWER_HALF at BW <= 4 (ALL at BW <= 2): the offset sum is a
//    popcount per bit plane b (bit b of every passing field, weighted 2^b);
//    bit b of a field passes when the top bit, BW-1-b positions higher, is
//    clear (one funnel shift of the stream);
//  - otherwise one signed field extract per value (v_bfe_i32, or a funnel
//    shift first for a field across two words): a passing offset is the
//    non-negative sign-extended field, so max(field, 0) is the term.
// Fields below 32 bits are read from the lane's stream as 32-bit words (for
// 64-bit types the low half of each word first), so they span at most two.
template <typename UT, int BW, int J0, int NJ, int PRED>
__device__ __forceinline__ void span_offset_stats(const UT *w, uint32_t &cnt, uint64_t &offsum) {
  constexpr int W = lanes<UT>::W;
  using Sp = SlotSpan<UT, BW, J0, NJ>;
  constexpr int H0 = Sp::R0 * (W / 32);
  constexpr int H = Sp::NR * (W / 32);
  auto hw = [&](auto ic) -> uint32_t {
    constexpr int i = decltype(ic)::value;
    if constexpr (i >= H) return 0u;
    else if constexpr (W == 32) return uint32_t(w[i]);
    else return uint32_t(uint64_t(w[i / 2]) >> (32 * (i % 2)));
  };
  constexpr bool LOWER = PRED == FLSGPU_PRED_OFFSET_LOWER_HALF;
  if constexpr (BW == 0) {
    cnt = LOWER ? 0 : NJ;  // no offset is below 2^-1
    offsum = 0;
    return;
  } else {
    if constexpr (LOWER) {
      uint32_t top = 0;
      static_for<H>([&](auto ic) {
        constexpr uint32_t M = slot_bit_mask<BW, J0, NJ, BW - 1, H0>(decltype(ic)::value);
        if constexpr (M != 0) top += uint32_t(__builtin_popcount(hw(ic) & M));
      });
      cnt = uint32_t(NJ) - top;
    } else {
      cnt = NJ;
    }
    if constexpr (W == 64 && BW >= 32) {
      uint64_t acc = 0;
      emit_slot_words<UT, BW, J0, NJ>(w, [&](int, UT off) {
        if constexpr (LOWER) acc += (off >> (BW - 1)) ? uint64_t(0) : uint64_t(off);
        else acc += uint64_t(off);
      });
      offsum = acc;
    } else if constexpr (LOWER && BW == 1) {
      offsum = 0;  // a passing 1-bit offset is 0
    } else if constexpr ((LOWER && BW <= 4) || (!LOWER && BW <= 2)) {
      uint32_t acc = 0;
      constexpr int NB = LOWER ? BW - 1 : BW;
      static_for<NB>([&](auto bc) {
        constexpr int b = decltype(bc)::value;
        uint32_t c = 0;
        static_for<H>([&](auto ic) {
          constexpr int i = decltype(ic)::value;
          constexpr uint32_t P = slot_bit_mask<BW, J0, NJ, b, H0>(i);
          if constexpr (P != 0) {
            uint32_t x = hw(ic) & P;
            if constexpr (LOWER) {
              const uint32_t t =
                  __builtin_amdgcn_alignbit(hw(std::integral_constant<int, i + 1>{}), hw(ic),
                                            uint32_t(BW - 1 - b));
              x &= ~t;
            }
            c += uint32_t(__builtin_popcount(x));
          }
        });
        acc += c << b;
      });
      offsum = acc;
    } else {
      constexpr bool FITS = LOWER ? (BW - 1 + ilog2(NJ) <= 32) : (BW + ilog2(NJ) <= 32);
      using Acc = std::conditional_t<FITS, uint32_t, uint64_t>;
      Acc acc = 0;
      static_for<NJ>([&](auto jc) {
        constexpr int q = (J0 + int(decltype(jc)::value)) * BW - 32 * H0;
        constexpr int i = q / 32;
        constexpr int s = q % 32;
        uint32_t x;
        if constexpr (s + BW <= 32) {
          x = hw(std::integral_constant<int, i>{}) >> s;
        } else {
          x = __builtin_amdgcn_alignbit(hw(std::integral_constant<int, i + 1>{}),
                                        hw(std::integral_constant<int, i>{}), uint32_t(s));
        }
        if constexpr (LOWER) {
          const int32_t v = BW == 32 ? int32_t(x) : int32_t(x << (32 - BW)) >> (32 - BW);
          acc += Acc(uint32_t(v > 0 ? v : 0));
        } else {
          acc += Acc(BW == 32 ? x : (x & ((1u << BW) - 1)));
        }
      });
      offsum = acc;
    }
  }
}

// Filter + sum of slots [J0, J0 + NJ) of one lane (words in registers):
// adds the count and sum (mod 2^64) of the passing values. For the offset
// predicates the sum of the passing values is the sum of their offsets +
// count * base, so the base is applied once: exact for every 64-bit vector
// (sums are mod 2^64 anyway) and for a 32-bit vector whose base + 2^BW - 1
// does not wrap; for one that does, a passing value base + off that wraps is
// 2^32 less than the term counted, so the count of passing offsets >=
// 2^32 - base (one compare per value) corrects the sum. The range predicate
// compares offsets with the range shifted by the base while base + off
// cannot wrap, and the wrapped value itself otherwise.
template <typename UT, int BW, int J0, int NJ, int PRED>
__device__ __forceinline__ void span_filter_sum(const UT *w, UT base, UT lo, UT hi,
                                                uint64_t &count, uint64_t &sum) {
  constexpr int W = lanes<UT>::W;
  constexpr UT MAXOFF = BW == W ? UT(~UT(0)) : UT((UT(1) << BW) - 1);
  const bool no_wrap = base <= UT(UT(~UT(0)) - MAXOFF);
  if constexpr (PRED == PRED_EQUAL) {
    // base + off == lo (mod 2^W)  <=>  off == lo - base, a count of one
    // offset (none when that is not a BW-bit offset)
    const UT want = UT(lo - base);
    uint32_t cnt = 0;
    if (want <= MAXOFF) {
      if constexpr (BW > 0 && BW < 32)
        span_fields32<UT, BW, J0, NJ>(w, [&](uint32_t off) { cnt += off == uint32_t(want); });
      else
        emit_slot_words<UT, BW, J0, NJ>(w, [&](int, UT off) { cnt += off == want; });
    }
    count += cnt;
    sum += uint64_t(cnt) * uint64_t(lo);
  } else if constexpr (PRED == FLSGPU_PRED_RANGE) {
    using Acc = std::conditional_t<(W == 32 && BW + ilog2(NJ) <= 32), uint32_t, uint64_t>;
    if (no_wrap) {
      // lo <= base + off < hi  <=>  off - olo < span (unsigned)
      const UT from = lo > base ? lo : base;
      const UT olo = UT(from - base);
      const UT span = hi > from ? UT(hi - from) : UT(0);
      Acc acc = 0;
      uint32_t cnt = 0;
      emit_slot_words<UT, BW, J0, NJ>(w, [&](int, UT off) {
        const bool pass = UT(off - olo) < span;
        cnt += pass;
        acc += pass ? Acc(off) : Acc(0);
      });
      count += cnt;
      sum += uint64_t(acc) + uint64_t(cnt) * uint64_t(base);
    } else {
      emit_slot_words<UT, BW, J0, NJ>(w, [&](int, UT off) {
        const UT v = UT(off + base);
        const bool pass = (v >= lo) & (v < hi);
        count += pass;
        sum += pass ? uint64_t(v) : 0;
      });
    }
  } else {
    uint32_t cnt;
    uint64_t offsum;
    span_offset_stats<UT, BW, J0, NJ, PRED>(w, cnt, offsum);
    count += cnt;
    sum += offsum + uint64_t(cnt) * uint64_t(base);
    if constexpr (W == 32 && BW > 0) {  // 64-bit sums are mod 2^64 anyway
      if (!no_wrap) {
        const UT first_wrap = UT(UT(0) - base);
        uint32_t n_wrap = 0;
        emit_slot_words<UT, BW, J0, NJ>(w, [&](int, UT off) {
          bool pass = off >= first_wrap;
          if constexpr (PRED == FLSGPU_PRED_OFFSET_LOWER_HALF) pass &= (off >> (BW - 1)) == 0;
          n_wrap += pass;
        });
        sum -= uint64_t(n_wrap) << 32;
      }
    }
  }
}

template <typename UT, int BW>
__device__ __forceinline__ bool sum_no_wrap(UT base) {
  constexpr int W = lanes<UT>::W;
  constexpr UT MAXOFF = BW == W ? UT(~UT(0)) : UT((UT(1) << BW) - 1);
  return W == 64 || base <= UT(UT(~UT(0)) - MAXOFF);
}

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

// Cap on the grid of the range-scan kernel (fused filter + sum, query any):
// the resident grid is smaller, and the allocation-free entry point's
// static partials hold this many blocks.
constexpr unsigned SCAN_GRID = 2048;

// Fused filter + sum (and, with PRED_EQUAL, query any). Every wave takes
// ranges of `units` consecutive units (units * VPW vectors, at most 64): lane
// i loads vector i's bit width, offset and base with one coalesced load each,
// and each unit's threads take theirs by shuffle. When all of a range's
// vectors have one width, its units go in batches of as many as fit about 32
// registers of packed words (at most 8 units), the next batch's loads issued
// before the current batch is reduced (two buffers), so a wave always has a
// batch of loads in flight. A lane's words beyond that budget (64-bit types
// from width 17) are read in 2 or 4 chunks of slots, one unit at a time.
// Ranges of mixed widths go unit by unit. The host picks `units` so that the
// ranges fill the resident waves about once, and caps the grid at what fits
// on the device at once.
template <typename UT>
constexpr int SCAN_RANGE_UNITS = WAVE / lanes<UT>::VPW;  // 64 vectors

constexpr int SCAN_DEAD = 255;

__host__ __device__ constexpr int pow2_floor(int x) { return x < 2 ? 1 : 2 * pow2_floor(x / 2); }

template <typename UT, int BW>
struct ScanShape {
  static constexpr int W = lanes<UT>::W;
  static constexpr int LANE_DW = BW * W / 32;  // 32-bit words per lane
  static constexpr int C = LANE_DW <= 32 ? 1 : (LANE_DW <= 64 ? 2 : 4);  // chunks per lane
  static constexpr int NJ = W / C;                                       // slots per chunk
  static constexpr int MAXKB = 8;  // each unit's address is held until its loads issue
  static constexpr int FIT = LANE_DW == 0 ? MAXKB : 32 / LANE_DW;
  static constexpr int KB = C > 1 ? 1 : pow2_floor(FIT < MAXKB ? FIT : MAXKB);  // units per batch
};

// A range's metadata, lane i holding vector i's.
template <typename UT>
struct RangeMeta {
  uint64_t off;
  UT base;
  uint32_t nv;  // vectors in the range
  int vsub;     // this thread's vector within a unit
};

// Whether this thread's vector of unit u of the range exists, and its offset
// (safe_off, a vector of the same width, when it does not).
template <typename UT>
__device__ __forceinline__ bool range_unit_off(const RangeMeta<UT> &r, uint32_t u,
                                               uint64_t safe_off, uint64_t &off) {
  const uint32_t src = u * uint32_t(lanes<UT>::VPW) + uint32_t(r.vsub);
  const uint64_t o = __shfl(r.off, int(src % WAVE), WAVE);
  const bool live = src < r.nv;
  off = live ? o : safe_off;
  return live;
}

// The base of this thread's vector of unit u of the range.
template <typename UT>
__device__ __forceinline__ UT range_unit_base(const RangeMeta<UT> &r, uint32_t u) {
  const uint32_t src = u * uint32_t(lanes<UT>::VPW) + uint32_t(r.vsub);
  return __shfl(r.base, int(src % WAVE), WAVE);
}

// Filter + sum of one unit of width BW at `off` (adds nothing unless live),
// its lane read in chunks of slots.
template <typename UT, int PRED, int BW>
__device__ __forceinline__ void unit_filter_sum(const FforArgs<UT> &a, uint64_t off, UT base,
                                                bool live, int lane, UT lo, UT hi,
                                                uint64_t &count, uint64_t &sum) {
  using Sh = ScanShape<UT, BW>;
  static_for<Sh::C>([&](auto cc) {
    constexpr int J0 = int(decltype(cc)::value) * Sh::NJ;
    UT w[SlotSpan<UT, BW, J0, Sh::NJ>::NR1];
    load_slot_words<UT, BW, J0, Sh::NJ>(a.packed + off + lane, w);
    uint64_t c = 0, t = 0;
    span_filter_sum<UT, BW, J0, Sh::NJ, PRED>(w, base, lo, hi, c, t);
    count += live ? c : 0;
    sum += live ? t : 0;
  });
}

// Filter + sum of the n_units units of a range whose vectors all have width
// BW (lane 0 holds the range's first vector).
template <typename UT, int PRED, int BW>
__device__ __forceinline__ void range_uniform(const FforArgs<UT> &a, const RangeMeta<UT> &r,
                                              uint32_t n_units, int lane, UT lo, UT hi,
                                              uint64_t &count, uint64_t &sum) {
  constexpr int W = lanes<UT>::W;
  using Sh = ScanShape<UT, BW>;
  // The equality test works per value on whole words (no bit-plane form):
  // batches of at most 64 slots keep its registers in bounds.
  constexpr int KB = PRED == PRED_EQUAL && Sh::KB > 64 / W ? 64 / W : Sh::KB;
  const uint64_t safe_off = uniform_u64(r.off);
  if constexpr (Sh::C == 1) {
    using Words = UT[KB][SlotSpan<UT, BW, 0, W>::NR1];
    // Loads are unconditional (units past the range read the vector at
    // safe_off and add nothing): a load under the branch that also guards a
    // batch's arithmetic would be merged into it, serialising the batches.
    // A unit's base is fetched when the batch is reduced, so a buffer holds
    // only words.
    auto load = [&](Words &w, uint32_t u0) {
      static_for<KB>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        uint64_t off;
        range_unit_off<UT>(r, u0 + k, safe_off, off);
        load_slot_words<UT, BW, 0, W>(a.packed + off + lane, w[k]);
      });
    };
    auto reduce = [&](const Words &w, uint32_t u0) {
      static_for<KB>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        const uint32_t src = (u0 + k) * uint32_t(lanes<UT>::VPW) + uint32_t(r.vsub);
        uint64_t c = 0, t = 0;
        span_filter_sum<UT, BW, 0, W, PRED>(w[k], range_unit_base<UT>(r, u0 + k), lo, hi, c, t);
        count += src < r.nv ? c : 0;
        sum += src < r.nv ? t : 0;
      });
    };
    // Every load is issued unconditionally, past the range too: after a
    // branch that may or may not issue loads the compiler waits for all of
    // them (the load counter drains in issue order), so a guarded prefetch
    // would make the wait for the current batch also wait for the next one.
    const uint32_t nb = (n_units + KB - 1) / KB;
    Words w0, w1;
    load(w0, 0);
#pragma unroll 1
    for (uint32_t b = 0; b < nb; b += 2) {
      load(w1, (b + 1) * KB);
      reduce(w0, b * KB);
      load(w0, (b + 2) * KB);
      // a batch past the range is loaded but not waited for
      if (b + 1 < nb) reduce(w1, (b + 1) * KB);
    }
  } else {
#pragma unroll 1
    for (uint32_t u = 0; u < n_units; ++u) {
      uint64_t off;
      const bool live = range_unit_off<UT>(r, u, safe_off, off);
      unit_filter_sum<UT, PRED, BW>(a, off, range_unit_base<UT>(r, u), live, lane, lo, hi, count,
                                    sum);
    }
  }
}

#define FLSGPU_FS_RANGE_CASE(B)                                                  \
  case B:                                                                       \
    if constexpr ((B) <= lanes<UT>::W)                                           \
      range_uniform<UT, PRED, (B)>(a, r, n_units, lane, lo, hi, count, sum);     \
    break;
#define FLSGPU_FS_UNIT_CASE(B)                                                   \
  case B:                                                                       \
    if constexpr ((B) <= lanes<UT>::W)                                           \
      unit_filter_sum<UT, PRED, (B)>(a, off, base, live, lane, lo, hi, count, sum); \
    break;

// A vector whose width is above W decodes as zeros (as in the decoder); the
// fused filter + sum leaves it out, the equality query counts its rows when
// the needle is zero.
template <typename UT, int PRED>
__device__ __forceinline__ void over_wide(bool live, UT lo, uint64_t &count) {
  if constexpr (PRED == PRED_EQUAL) count += live && lo == UT(0) ? lanes<UT>::W : 0;
}

template <typename UT, int PRED>
__device__ __forceinline__ void range_dispatch(const FforArgs<UT> &a, const RangeMeta<UT> &r,
                                               uint32_t n_units, int bw, int lane, UT lo, UT hi,
                                               uint64_t &count, uint64_t &sum) {
  if constexpr (lanes<UT>::W == 32) {
    switch (bw) {
      FLSGPU_CASES_0_32(FLSGPU_FS_RANGE_CASE)
      default: over_wide<UT, PRED>(bw != SCAN_DEAD, lo, count); break;
    }
  } else {
    switch (bw) {
      FLSGPU_CASES_0_32(FLSGPU_FS_RANGE_CASE)
      FLSGPU_CASES_33_64(FLSGPU_FS_RANGE_CASE)
      default: over_wide<UT, PRED>(bw != SCAN_DEAD, lo, count); break;
    }
  }
}

template <typename UT, int PRED>
__device__ __forceinline__ void unit_dispatch(const FforArgs<UT> &a, uint64_t off, UT base,
                                              bool live, int bw, int lane, UT lo, UT hi,
                                              uint64_t &count, uint64_t &sum) {
  if constexpr (lanes<UT>::W == 32) {
    switch (bw) {
      FLSGPU_CASES_0_32(FLSGPU_FS_UNIT_CASE)
      default: over_wide<UT, PRED>(live, lo, count); break;
    }
  } else {
    switch (bw) {
      FLSGPU_CASES_0_32(FLSGPU_FS_UNIT_CASE)
      FLSGPU_CASES_33_64(FLSGPU_FS_UNIT_CASE)
      default: over_wide<UT, PRED>(live, lo, count); break;
    }
  }
}
#undef FLSGPU_FS_RANGE_CASE
#undef FLSGPU_FS_UNIT_CASE

// Waves per SIMD the register allocator is held to: the offset-statistics
// kernels fit three without spilling vector registers, the 64-bit equality
// test (a 64-bit compare per value) two; the range kernel is left
// unconstrained.
template <typename UT, int PRED>
constexpr int FS_WAVES =
    PRED == FLSGPU_PRED_RANGE ? 1 : (PRED == PRED_EQUAL && lanes<UT>::W == 64 ? 2 : 3);

template <typename UT, int S, int PRED>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(FS_WAVES<UT, PRED>, 8)))
void k_ffor_filter_sum(FforArgs<UT> a, UT lo, UT hi, uint32_t units, Scratch s) {
  static_assert(S == 1, "the fused filter + sum reads whole lanes");
  using L = lanes<UT>;
  const int t = int(threadIdx.x % WAVE);
  const int lane = t % L::N;
  const uint64_t span = uint64_t(units) * L::VPW;  // vectors per range
  const uint64_t n_ranges = (a.n_vecs + span - 1) / span;
  const uint64_t stride = uint64_t(gridDim.x) * (BLOCK / WAVE);
  uint64_t count = 0, sum = 0;
  for (uint64_t rr = (uint64_t(blockIdx.x) * BLOCK + threadIdx.x) / WAVE; rr < n_ranges;
       rr += stride) {
    const uint64_t v0 = rr * span;
    RangeMeta<UT> r;
    r.nv = uint32_t(a.n_vecs - v0 < span ? a.n_vecs - v0 : span);
    r.vsub = t / L::N;
    int mbw = SCAN_DEAD;
    r.off = 0;
    r.base = UT(0);
    if (uint32_t(t) < r.nv) {
      mbw = a.bit_widths[v0 + t];
      r.off = a.vector_offsets[v0 + t];
      r.base = a.bases ? a.bases[v0 + t] : UT(0);
    }
    const uint32_t n_units = (r.nv + L::VPW - 1) / L::VPW;
    // Lane 0 holds the range's first vector.
    const int b0 = __builtin_amdgcn_readfirstlane(mbw);
    if (__builtin_amdgcn_ballot_w64(mbw != b0 && mbw != SCAN_DEAD) == 0) {
      range_dispatch<UT, PRED>(a, r, n_units, b0, lane, lo, hi, count, sum);
    } else {
#pragma unroll 1
      for (uint32_t u = 0; u < n_units; ++u) {
        const uint32_t src = u * uint32_t(L::VPW) + uint32_t(r.vsub);
        const int bw = __shfl(mbw, int(src % WAVE), WAVE);  // SCAN_DEAD past the range
        const uint64_t first = uniform_u64(__shfl(r.off, int(u * L::VPW), WAVE));
        uint64_t off;
        const bool live = range_unit_off<UT>(r, u, first, off);
        const UT base = range_unit_base<UT>(r, u);
        // The unit's first vector (lane 0's) exists.
        const int b1 = __builtin_amdgcn_readfirstlane(bw);
        if (__builtin_amdgcn_ballot_w64(live && bw != b1) == 0) {
          unit_dispatch<UT, PRED>(a, off, base, live, b1, lane, lo, hi, count, sum);
        } else if (live && bw <= L::W) {
          FilterSumPolicy<UT, PRED> pol{a.packed + off + lane, base, lo, hi, 0, 0};
          pol.run_generic(bw, 0, L::W);
          count += pol.count;
          sum += pol.sum;
        } else {
          over_wide<UT, PRED>(live, lo, count);
        }
      }
    }
  }
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

// "Any value == magic": the range scan counting rows equal to magic, the
// counts reduced as a flag.
template <typename UT, int S>
int FlsAny<UT, S>::launch(const FforArgs<UT> &a, UT magic, uint8_t *d_found, hipStream_t stream) {
  Scratch s;
  int rc = get_scratch(stream, &s);
  if (rc) return rc;
  unsigned grid = 0;
  rc = FlsFilterSum<UT, 1, PRED_EQUAL>::launch_scan(a, magic, magic, s.partials, MAX_REDUCE_BLOCKS,
                                                    stream, &grid);
  if (rc) return rc;
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
int FlsFilterSum<UT, S, PRED>::launch_scan(const FforArgs<UT> &a, UT lo, UT hi,
                                           uint64_t *partials, unsigned slots, hipStream_t stream,
                                           unsigned *grid_out) {
  static const unsigned resident = [] {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_ffor_filter_sum<UT, S, PRED>, BLOCK,
                                                     0) != hipSuccess || per_cu < 1)
      per_cu = 1;
    return unsigned(per_cu) * unsigned(num_cus());
  }();
  // Units per range: the ranges fill the resident waves about once.
  const uint64_t n_units = n_waves_for<UT, 1>(a.n_vecs);
  const uint64_t waves = uint64_t(resident) * (BLOCK / WAVE);
  const uint32_t units = uint32_t(
      std::clamp<uint64_t>((n_units + waves - 1) / waves, 1, SCAN_RANGE_UNITS<UT>));
  const uint64_t n_ranges = (n_units + units - 1) / units;
  const uint64_t blocks = (n_ranges + (BLOCK / WAVE) - 1) / (BLOCK / WAVE);
  const unsigned grid = unsigned(
      std::min<uint64_t>({std::max<uint64_t>(blocks, 1), resident, SCAN_GRID, slots}));
  if (grid == 0) return set_error(FLSGPU_EINVAL, "no room for block partials");
  const Scratch s{partials};
  hipLaunchKernelGGL((k_ffor_filter_sum<UT, S, PRED>), dim3(grid), dim3(BLOCK), 0, stream, a, lo,
                     hi, units, s);
  *grid_out = grid;
  return check_hip(hipGetLastError(), "k_ffor_filter_sum launch");
}

template <typename UT, int S, int PRED>
int FlsFilterSum<UT, S, PRED>::launch_partials(const FforArgs<UT> &a, UT lo, UT hi,
                                               uint64_t *d_result, uint64_t *partials,
                                               unsigned slots, hipStream_t stream) {
  unsigned grid = 0;
  int rc = launch_scan(a, lo, hi, partials, slots, stream, &grid);
  if (rc) return rc;
  return reduce_finish_pair(Scratch{partials}, grid, d_result, stream);
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
