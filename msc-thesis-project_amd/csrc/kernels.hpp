// Work mapping and cross-block reduction shared by all decoder kernels.
//
// Mapping (replaces FillWarpMapping / FillWarpThreadblockMapping of
// src/engine/device-utils.cuh:24-99): one FastLanes lane per thread, so a
// 64-wide wavefront covers VPW = 64 / N_LANES vectors (2 for 32-bit, 4 for
// 64-bit types). A lane may be split over S waves ("parts"); the part index
// is wave-uniform, so every branch on it is scalar. Vectors are never skipped:
// the reference's integer-division grid drops trailing vectors
// (device-utils.cuh:35); here the grid is rounded up and the tail masked.
#pragma once

#include "unpack.hpp"

namespace flsgpu_hip {

template <typename UT, int S>
struct UnitMap {
  uint64_t vec;
  int lane;
  int part;
};

template <typename UT, int S>
__device__ __forceinline__ UnitMap<UT, S> map_unit(uint64_t wave_index) {
  using L = lanes<UT>;
  UnitMap<UT, S> u;
  u.part = __builtin_amdgcn_readfirstlane(int(wave_index % S));
  const uint64_t group = wave_index / S;
  const int t = int(threadIdx.x % WAVE);
  u.lane = t % L::N;
  u.vec = group * L::VPW + uint64_t(t / L::N);
  return u;
}

template <typename UT, int S>
__host__ __device__ constexpr uint64_t n_waves_for(uint64_t n_vecs) {
  return ((n_vecs + lanes<UT>::VPW - 1) / lanes<UT>::VPW) * S;
}

// A value the caller knows to be the same in every lane of the wave, moved
// to scalar registers so that loads indexed by it are scalar loads.
__device__ __forceinline__ uint64_t uniform_u64(uint64_t x) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane(uint32_t(x));
  const uint32_t hi = __builtin_amdgcn_readfirstlane(uint32_t(x >> 32));
  return (uint64_t(hi) << 32) | lo;
}

// Stands in for the exception arrays of an empty list (see list_entry).
static __device__ uint64_t k_no_exception[1];  // zero; never written

// Address of entry k of an exception list of `count` entries starting at
// `first`: past the list's end its first entry, and for an empty list a zero
// stub, so a thread can load an entry with no branch and never reads outside
// the arrays. (A branch between a load and its use makes the compiler's wait
// before the use also wait for every load issued since, and the load
// counter drains in issue order.)
template <typename E>
__device__ __forceinline__ const E *list_entry(const E *arr, uint64_t first, unsigned k,
                                               unsigned count) {
  return count == 0 ? reinterpret_cast<const E *>(k_no_exception) : arr + first + (k < count ? k : 0);
}

enum class RedOp { Add, Or, And };

template <RedOp OP>
__device__ __forceinline__ uint64_t red_combine(uint64_t a, uint64_t b) {
  if constexpr (OP == RedOp::Add) return a + b;
  else if constexpr (OP == RedOp::Or) return a | b;
  else return a & b;
}

template <RedOp OP>
__device__ __forceinline__ uint64_t red_identity() {
  return OP == RedOp::And ? ~uint64_t(0) : uint64_t(0);
}

template <RedOp OP>
__device__ __forceinline__ uint64_t wave_reduce(uint64_t v) {
#pragma unroll
  for (int off = WAVE / 2; off > 0; off >>= 1) v = red_combine<OP>(v, __shfl_xor(v, off, WAVE));
  return v;
}

// Cross-block reductions run as two launches: every block of the work
// kernel reduces its waves in LDS and stores one partial pair
// (block_partial2); a one-block finishing kernel (reduce_finish_*, capi.hip)
// combines the partials and writes the result. The kernel boundary orders
// the partials before the finish, so no block waits on a fence or a ticket:
// on gfx950 an agent-scope fence is an L2 write-back plus invalidate of the
// XCD (buffer_wbl2 / buffer_inv), and per-block tickets keep every block
// alive for a round trip to the coherence point; at small bit widths either
// cost more than the decoding.
template <RedOp OP>
__device__ __forceinline__ void block_partial2(uint64_t v0, uint64_t v1,
                                               uint64_t *__restrict partials) {
  __shared__ uint64_t sh[BLOCK / WAVE][2];
  v0 = wave_reduce<OP>(v0);
  v1 = wave_reduce<OP>(v1);
  const int w = threadIdx.x / WAVE;
  if ((threadIdx.x % WAVE) == 0) {
    sh[w][0] = v0;
    sh[w][1] = v1;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t a = red_identity<OP>(), b = red_identity<OP>();
    for (int i = 0; i < BLOCK / WAVE; ++i) {
      a = red_combine<OP>(a, sh[i][0]);
      b = red_combine<OP>(b, sh[i][1]);
    }
    partials[2 * blockIdx.x] = a;
    partials[2 * blockIdx.x + 1] = b;
  }
}

}  // namespace flsgpu_hip
