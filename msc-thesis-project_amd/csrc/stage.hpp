// LDS staging of a wave's output: the wave decodes its slot range of VPW
// vectors into LDS (one thread per FastLanes lane, 4- or 8-byte writes), then
// writes each vector's range, which is contiguous in the output, with 16-byte
// non-temporal stores. ALP exceptions are applied in LDS in between, by the
// whole wave from coalesced loads of the vector's exception list, so no
// output element is written twice.
#pragma once

#include "common.hpp"

namespace flsgpu_hip {

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// Orders this wave's LDS writes before its later LDS reads by other lanes.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Elements per vector in one part (slot range) of S parts.
template <typename UT, int S>
struct StageShape {
  static constexpr int NJ = lanes<UT>::W / S;
  static constexpr int CHUNK = lanes<UT>::N * NJ;
  static constexpr int PIECES = CHUNK * int(sizeof(UT)) / 16;  // 16-byte pieces
};

// Writes the wave's staged chunks: vector vv of group `group`, part `part`,
// goes to out[vec * 1024 + part * CHUNK ...]. Every lane of the wave calls it.
template <typename T, int S>
__device__ __forceinline__ void store_staged(const T *lds, T *__restrict out, uint64_t group,
                                             int part, uint64_t n_vecs) {
  using UT = typename uint_of<T>::type;
  using Sh = StageShape<UT, S>;
  constexpr int VPW = lanes<UT>::VPW;
  const int t = threadIdx.x % WAVE;
#pragma unroll
  for (int i = t; i < VPW * Sh::PIECES; i += WAVE) {
    const int vv = i / Sh::PIECES, p = i % Sh::PIECES;
    const uint64_t vec = group * VPW + vv;
    if (vec < n_vecs) {
      const u32x4 val = reinterpret_cast<const u32x4 *>(lds + vv * Sh::CHUNK)[p];
      u32x4 *dst = reinterpret_cast<u32x4 *>(out + vec * VALUES_PER_VECTOR + part * Sh::CHUNK) + p;
      __builtin_nontemporal_store(val, dst);
    }
  }
}

}  // namespace flsgpu_hip
