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

// Grid-wide reduction of two values in two levels, so that no counter sees
// more than REDUCE_GROUP (or gridDim / REDUCE_GROUP) arrivals: every block
// reduces in LDS and writes a partial; the last block of each group of
// REDUCE_GROUP blocks (atomic ticket per group) reduces the group's partials
// and writes a group partial; the last group (top ticket) reduces those and
// calls done(total0, total1) from thread 0. Each last arrival re-arms its
// ticket, so consecutive launches need no reset. Add/Or/And on integers are
// order-independent, so the result is deterministic.
//
// Cross-block visibility without __threadfence(): on MI355X an agent-scope
// fence is an L2 writeback plus an L2 invalidate of the XCD (buffer_wbl2 /
// buffer_inv), which at one fence per block cost more than the decoding
// itself. Partials are instead written with
// agent-scope atomic stores (written through to the device coherence point),
// the writer waits for their completion (vmcnt) before taking its ticket, and
// the reader, which only reads after winning the ticket, uses agent-scope
// atomic loads.
__device__ __forceinline__ void wait_stores_complete() {
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
}

template <RedOp OP, class Done>
__device__ __forceinline__ void grid_reduce2(uint64_t v0, uint64_t v1, Scratch s, Done &&done) {
  static_assert(REDUCE_GROUP == WAVE, "one wave reduces one group");
  __shared__ uint64_t sh[BLOCK / WAVE][2];
  __shared__ int am_last, am_top;
  v0 = wave_reduce<OP>(v0);
  v1 = wave_reduce<OP>(v1);
  const int w = threadIdx.x / WAVE;
  if ((threadIdx.x % WAVE) == 0) {
    sh[w][0] = v0;
    sh[w][1] = v1;
  }
  if (threadIdx.x == 0) am_top = 0;
  __syncthreads();
  const unsigned g = blockIdx.x / REDUCE_GROUP;
  const unsigned g0 = g * REDUCE_GROUP;
  const unsigned gsize = gridDim.x - g0 < unsigned(REDUCE_GROUP) ? gridDim.x - g0 : REDUCE_GROUP;
  if (threadIdx.x == 0) {
    uint64_t a = red_identity<OP>(), b = red_identity<OP>();
    for (int i = 0; i < BLOCK / WAVE; ++i) {
      a = red_combine<OP>(a, sh[i][0]);
      b = red_combine<OP>(b, sh[i][1]);
    }
    __hip_atomic_store(&s.partials[2 * blockIdx.x], a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&s.partials[2 * blockIdx.x + 1], b, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
    wait_stores_complete();
    const unsigned prev = __hip_atomic_fetch_add(&s.group_counters[g], 1u, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT);
    am_last = (prev == gsize - 1);
  }
  __syncthreads();
  if (!am_last) return;
  // Group level: wave 0 holds one partial per lane.
  if (w == 0) {
    uint64_t a = red_identity<OP>(), b = red_identity<OP>();
    if (threadIdx.x < gsize) {
      const unsigned i = g0 + threadIdx.x;
      a = __hip_atomic_load(&s.partials[2 * i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      b = __hip_atomic_load(&s.partials[2 * i + 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    a = wave_reduce<OP>(a);
    b = wave_reduce<OP>(b);
    if (threadIdx.x == 0) {
      __hip_atomic_store(&s.group_counters[g], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&s.group_partials[2 * g], a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&s.group_partials[2 * g + 1], b, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
      wait_stores_complete();
      const unsigned n_groups = (gridDim.x + REDUCE_GROUP - 1) / REDUCE_GROUP;
      const unsigned prev =
          __hip_atomic_fetch_add(s.counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      am_top = (prev == n_groups - 1);
    }
  }
  __syncthreads();
  if (!am_top) return;
  // Top level: every group partial, in a fixed order per thread.
  const unsigned n_groups = (gridDim.x + REDUCE_GROUP - 1) / REDUCE_GROUP;
  uint64_t a = red_identity<OP>(), b = red_identity<OP>();
  for (unsigned i = threadIdx.x; i < n_groups; i += BLOCK) {
    a = red_combine<OP>(
        a, __hip_atomic_load(&s.group_partials[2 * i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    b = red_combine<OP>(b, __hip_atomic_load(&s.group_partials[2 * i + 1], __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT));
  }
  a = wave_reduce<OP>(a);
  b = wave_reduce<OP>(b);
  __syncthreads();
  if ((threadIdx.x % WAVE) == 0) {
    sh[w][0] = a;
    sh[w][1] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t ta = red_identity<OP>(), tb = red_identity<OP>();
    for (int i = 0; i < BLOCK / WAVE; ++i) {
      ta = red_combine<OP>(ta, sh[i][0]);
      tb = red_combine<OP>(tb, sh[i][1]);
    }
    done(ta, tb);
    __hip_atomic_store(s.counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

}  // namespace flsgpu_hip
