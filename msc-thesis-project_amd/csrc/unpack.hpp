// FastLanes lane unpacking specialised per bit width for gfx950.
//
// Semantics follow the reference CPU unpacker (src/fls/unffor.cpp, e.g.
// unffor_3bw_32ow_32crw_1uf :2982) and the device unpackers of
// src/flsgpu/fls.cuh:43-618: the value in lane l, slot j of a vector packed at
// width BW is bits [j*BW, (j+1)*BW) of the lane's word stream
// in[l + N_LANES*r], least significant bit first.
//
// Design for CDNA4 (not a translation of the reference's unpacker family):
// every bit width gets its own fully unrolled body, so word indices and shift
// amounts are immediates. There is no runtime shift by W (the reference's
// branchless unpackers rely on CUDA clamping `x >> 32` to 0, which AMD
// hardware does not do: v_lshrrev_b32 masks the count to 5 bits). A value
// straddling two words becomes one v_alignbit_b32 (or a 64-bit funnel).
// A lane may be split over S threads (S = 1, 2, 4, 8); each handles a
// contiguous, compile-time slot range [J0, J0 + NJ) and loads only the words
// that range touches.
#pragma once

#include <utility>

#include "common.hpp"

namespace flsgpu_hip {

template <int... Js, class F>
__device__ __forceinline__ void static_for_impl(std::integer_sequence<int, Js...>, F &&f) {
  (f(std::integral_constant<int, Js>{}), ...);
}
template <int N, class F>
__device__ __forceinline__ void static_for(F &&f) {
  static_for_impl(std::make_integer_sequence<int, N>{}, f);
}

struct NoHook {
  __device__ __forceinline__ void operator()() const {}
};

// The words of one lane that slots [J0, J0 + NJ) at bit width BW touch:
// lane words R0 .. R0 + NR - 1 (none at width 0).
template <typename UT, int BW, int J0, int NJ>
struct SlotSpan {
  static constexpr int W = lanes<UT>::W;
  static constexpr int R0 = BW == 0 ? 0 : (J0 * BW) / W;
  static constexpr int NR = BW == 0 ? 0 : ((J0 + NJ) * BW - 1) / W - R0 + 1;
  static constexpr int NR1 = NR > 0 ? NR : 1;  // array extent
};

// Loads the span's words of one lane; `in` points at the lane's word 0 of the
// vector (packed + vector_offset + lane).
template <typename UT, int BW, int J0, int NJ>
__device__ __forceinline__ void load_slot_words(const UT *__restrict in, UT *w) {
  using Sp = SlotSpan<UT, BW, J0, NJ>;
  static_for<Sp::NR>([&](auto rc) {
    constexpr int r = decltype(rc)::value;
    w[r] = in[(Sp::R0 + r) * lanes<UT>::N];
  });
}

// Calls emit(j, value) for slots J0..J0+NJ-1 from the span's words, already
// in registers (w[0] = lane word R0).
template <typename UT, int BW, int J0, int NJ, class Emit>
__device__ __forceinline__ void emit_slot_words(const UT *w, Emit &&emit) {
  constexpr int W = lanes<UT>::W;
  static_assert(BW >= 0 && BW <= W, "bit width out of range");
  if constexpr (BW == 0) {
    static_for<NJ>([&](auto jc) { emit(J0 + int(decltype(jc)::value), UT(0)); });
  } else {
    constexpr int R0 = SlotSpan<UT, BW, J0, NJ>::R0;
    constexpr UT MASK = BW == W ? UT(~UT(0)) : UT((UT(1) << BW) - 1);
    static_for<NJ>([&](auto jc) {
      constexpr int j = J0 + int(decltype(jc)::value);
      constexpr int p = j * BW;
      constexpr int r = p / W - R0;
      constexpr int s = p % W;
      UT v;
      if constexpr (s + BW > W) {
        v = UT(w[r] >> s) | UT(w[r + 1] << (W - s));
      } else {
        v = UT(w[r] >> s);
      }
      if constexpr (BW < W) v &= MASK;
      emit(j, v);
    });
  }
}

// Calls emit(j, value) for slots J0..J0+NJ-1 of one lane. `in` points at the
// lane's word 0 of the vector (packed + vector_offset + lane). `after_loads`
// runs once every packed word load has been issued and before any is used:
// loads it issues are in flight together with the packed words, and waiting
// for the packed words does not wait for them first (the load counter drains
// in issue order).
template <typename UT, int BW, int J0, int NJ, class Emit, class Hook = NoHook>
__device__ __forceinline__ void unpack_slots(const UT *__restrict in, Emit &&emit,
                                             Hook &&after_loads = Hook{}) {
  UT w[SlotSpan<UT, BW, J0, NJ>::NR1];
  load_slot_words<UT, BW, J0, NJ>(in, w);
  after_loads();
  emit_slot_words<UT, BW, J0, NJ>(w, emit);
}

// Dispatch on a wave-uniform bit width and slot part. Policy provides
//   template <int BW, int J0, int NJ> __device__ void run();
template <typename UT, int S, int BW, class Policy>
__device__ __forceinline__ void dispatch_part(Policy &pol, int part) {
  constexpr int W = lanes<UT>::W;
  constexpr int NJ = W / S;
  if constexpr (S == 1) {
    pol.template run<BW, 0, NJ>();
  } else if constexpr (S == 2) {
    if (part == 0) pol.template run<BW, 0, NJ>();
    else pol.template run<BW, NJ, NJ>();
  } else if constexpr (S == 4) {
    switch (part) {
      case 0: pol.template run<BW, 0, NJ>(); break;
      case 1: pol.template run<BW, NJ, NJ>(); break;
      case 2: pol.template run<BW, 2 * NJ, NJ>(); break;
      default: pol.template run<BW, 3 * NJ, NJ>(); break;
    }
  } else {
    static_assert(S == 8, "S must be 1, 2, 4 or 8");
    switch (part) {
      case 0: pol.template run<BW, 0, NJ>(); break;
      case 1: pol.template run<BW, NJ, NJ>(); break;
      case 2: pol.template run<BW, 2 * NJ, NJ>(); break;
      case 3: pol.template run<BW, 3 * NJ, NJ>(); break;
      case 4: pol.template run<BW, 4 * NJ, NJ>(); break;
      case 5: pol.template run<BW, 5 * NJ, NJ>(); break;
      case 6: pol.template run<BW, 6 * NJ, NJ>(); break;
      default: pol.template run<BW, 7 * NJ, NJ>(); break;
    }
  }
}

#define FLSGPU_CASE(B)                                                          \
  case B:                                                                      \
    if constexpr ((B) <= lanes<UT>::W) dispatch_part<UT, S, (B)>(pol, part);    \
    break;
#define FLSGPU_CASES_0_32(X)                                                   \
  X(0) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) \
  X(15) X(16) X(17) X(18) X(19) X(20) X(21) X(22) X(23) X(24) X(25) X(26) X(27)  \
  X(28) X(29) X(30) X(31) X(32)
#define FLSGPU_CASES_33_64(X)                                                  \
  X(33) X(34) X(35) X(36) X(37) X(38) X(39) X(40) X(41) X(42) X(43) X(44) X(45)  \
  X(46) X(47) X(48) X(49) X(50) X(51) X(52) X(53) X(54) X(55) X(56) X(57) X(58)  \
  X(59) X(60) X(61) X(62) X(63) X(64)

// `bw` must be wave-uniform here (an SGPR value).
template <typename UT, int S, class Policy>
__device__ __forceinline__ bool dispatch_uniform(Policy &pol, int bw, int part) {
  if constexpr (lanes<UT>::W == 32) {
    switch (bw) {
      FLSGPU_CASES_0_32(FLSGPU_CASE)
      default: return false;
    }
  } else {
    switch (bw) {
      FLSGPU_CASES_0_32(FLSGPU_CASE)
      FLSGPU_CASES_33_64(FLSGPU_CASE)
      default: return false;
    }
  }
  return true;
}

// Value of slot j (runtime) of a lane packed at runtime width bw (bw <= W),
// loading the one or two words it spans.
template <typename UT>
__device__ __forceinline__ UT extract_runtime(const UT *__restrict in_lane, int bw, int j) {
  constexpr int W = lanes<UT>::W;
  constexpr int NL = lanes<UT>::N;
  if (bw == 0) return UT(0);
  const int p = j * bw;
  const int r = p / W;
  const int s = p % W;
  UT v = UT(in_lane[r * NL] >> s);
  if (s + bw > W) v |= UT(in_lane[(r + 1) * NL] << (W - s));
  if (bw < W) v &= UT((UT(1) << bw) - 1);
  return v;
}

// Runtime-width unpacking of slots [j0, j0 + nj): the path for wavefronts
// whose vectors have different bit widths.
template <typename UT, class Emit>
__device__ __forceinline__ void unpack_runtime(const UT *__restrict in, int bw, int j0, int nj,
                                               Emit &&emit) {
  for (int j = j0; j < j0 + nj; ++j) emit(j, extract_runtime<UT>(in, bw, j));
}

// Dispatch on the lane's bit width. When every vector of the wavefront has
// the same width (always for single-width columns) the specialised body runs
// from a scalar branch; otherwise the policy's runtime-width path runs. The
// two are separate on purpose: a divergent loop over specialised bodies would
// make the register allocator size every body for the loop (measured: 146
// instead of 54 VGPRs for the 64-bit quarter-lane decoder). Returns false for
// lanes whose width exceeds W (nothing was emitted).
template <typename UT, int S, class Policy>
__device__ __forceinline__ bool dispatch_bw(Policy &pol, int bw, int part) {
  constexpr int W = lanes<UT>::W;
  const int b0 = __builtin_amdgcn_readfirstlane(bw);
  if (__builtin_amdgcn_ballot_w64(bw != b0) == 0) return dispatch_uniform<UT, S>(pol, b0, part);
  if (bw > W) return false;
  pol.run_generic(bw, part * (W / S), W / S);
  return true;
}

}  // namespace flsgpu_hip
