// ALP-RD decode kernels for gfx950 (the reference has only a CPU decoder of
// unpacked parts, src/alp/rd.hpp:146-172, with no caller).
//
//   out[i] = ((UT)dictionary[left_index[i]] << right_bw) | right[i]
//   exceptions: out[pos] = ((UT)exception_left << right_bw) | right[pos]
//
// Right parts are FastLanes words of the float's width packed at right_bw
// (16 compile-time cases, W-16 .. W-1); left dictionary indices are uint16
// FastLanes words packed at left_bw (1..3 bits). The thread that owns lane l
// of the W-bit layout needs, for slot j, the uint16 lane l + N_LANES*(j % Q)
// at slot j / Q (Q = 64 / N_LANES); its Q uint16 lanes hold at most
// 16 * 3 = 48 bits each, so each is loaded once into one 64-bit register. The
// dictionary, pre-shifted by right_bw, sits in LDS (8 entries per block).
#pragma once

#include "alprd_impl.hpp"
#include "kernels.hpp"
#include "stage.hpp"
#include "unpack.hpp"

namespace flsgpu_hip {

// One thread per lane. Exceptions are patched in registers: bit j of `emask`
// marks slot j of this lane as an exception whose left part is the next
// 16 bits of `lq` (the lane's first RD_REG exceptions in slot order), so
// the value is stored once, already patched.
constexpr int RD_REG = 4;

template <typename T, bool NT>
struct RdPolicy {
  using UT = typename uint_of<T>::type;
  static constexpr int NL = lanes<UT>::N;
  static constexpr int Q = 64 / NL;
  const UT *rin;
  T *out;
  const UT *dshift;  // LDS, 8 entries
  uint64_t lstream[Q];
  unsigned lbw;
  uint64_t emask;
  uint64_t lq;
  template <int BW, int J0, int NJ>
  __device__ __forceinline__ void run() {
    const uint64_t lmask = (uint64_t(1) << lbw) - 1;
    unpack_slots<UT, BW, J0, NJ>(rin, [&](int j, UT r) {
      const unsigned idx = unsigned(lstream[j % Q] >> ((j / Q) * lbw)) & unsigned(lmask);
      const bool ex = (emask >> j) & 1u;
      const UT v = ex ? UT(UT(UT(lq & 0xFFFFu) << BW) | r) : UT(dshift[idx] | r);
      lq = ex ? (lq >> 16) : lq;
      store<T, NT>(out + j * NL, __builtin_bit_cast(T, v));
    });
  }
};

// right_bw is one value per column, so the dispatch is a scalar switch.
template <typename UT, class Policy>
__device__ __forceinline__ bool dispatch_right_bw(Policy &pol, int bw) {
  constexpr int W = lanes<UT>::W;
  const int b = __builtin_amdgcn_readfirstlane(bw);
  switch (b - (W - 16)) {
#define FLSGPU_RD_CASE(K) \
  case K: pol.template run<W - 16 + K, 0, W>(); return true;
    FLSGPU_RD_CASE(0) FLSGPU_RD_CASE(1) FLSGPU_RD_CASE(2) FLSGPU_RD_CASE(3)
    FLSGPU_RD_CASE(4) FLSGPU_RD_CASE(5) FLSGPU_RD_CASE(6) FLSGPU_RD_CASE(7)
    FLSGPU_RD_CASE(8) FLSGPU_RD_CASE(9) FLSGPU_RD_CASE(10) FLSGPU_RD_CASE(11)
    FLSGPU_RD_CASE(12) FLSGPU_RD_CASE(13) FLSGPU_RD_CASE(14) FLSGPU_RD_CASE(15)
#undef FLSGPU_RD_CASE
    default: return false;
  }
}

// The block fills the pre-shifted dictionary in LDS before any wave decodes.
template <typename T>
__device__ __forceinline__ void alprd_fill_dictionary(const AlprdArgs<T> &a,
                                                      typename uint_of<T>::type *dshift) {
  using UT = typename uint_of<T>::type;
  if (threadIdx.x < 8)
    dshift[threadIdx.x] = threadIdx.x < a.dict_size ? UT(UT(a.dict[threadIdx.x]) << a.right_bw) : UT(0);
}

// Exception entries per vector that the one-thread-per-lane decoder stages in
// LDS: the vector's N_LANES threads load the first RD_LIST entries of its
// list together (64 / N_LANES each, coalesced) as (left << 16) | position.
// Every thread then reads the staged list for its lane's entries: the first
// RD_REG (in slot order) are patched in registers during the decode; any
// further ones, and entries past RD_LIST, are stored after the decode (the
// same thread's later store to the same address wins). Reading the whole
// list from global memory per thread and patching after the decode made the
// time grow with the exception count (2^26 doubles: 201 us at 4 exceptions
// per vector, 289 us at 18).
constexpr unsigned RD_LIST = 64;

// `list`: this wave's LDS, VPW * RD_LIST entries. Every lane of the wave
// calls this (the LDS hand-over is wave-wide); lanes past the column end
// decode nothing.
template <typename T, bool NT>
__device__ __forceinline__ void alprd_decompress_unit(const AlprdArgs<T> &a, T *__restrict out,
                                                      uint64_t wave,
                                                      const typename uint_of<T>::type *dshift,
                                                      uint32_t *list) {
  using UT = typename uint_of<T>::type;
  using L = lanes<UT>;
  constexpr int Q = 64 / L::N;
  constexpr int PFN = int(RD_LIST) / L::N;
  const UnitMap<UT, 1> u = map_unit<UT, 1>(wave);
  const int vv = int(threadIdx.x % WAVE) / L::N;
  bool live = u.vec < a.n_vecs;
  uint64_t eo = 0;
  unsigned count = 0;
  uint32_t ent[PFN];
#pragma unroll
  for (int i = 0; i < PFN; ++i) ent[i] = 0;
  T *vout = out + u.vec * VALUES_PER_VECTOR;
  const UT *rin = a.right + u.vec * uint64_t(a.right_bw) * L::N + u.lane;
  RdPolicy<T, NT> pol;
  if (live) {
    eo = a.exceptions_offsets[u.vec];
    count = a.counts[u.vec];
#pragma unroll
    for (int i = 0; i < PFN; ++i) {
      const unsigned k = unsigned(u.lane + L::N * i);
      if (k < count)
        ent[i] = (a.positions[eo + k] & 1023u) | (uint32_t(a.exceptions[eo + k]) << 16);
    }
    const uint16_t *lin = a.left + u.vec * uint64_t(a.left_bw) * 64;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      uint64_t s = 0;
      for (unsigned r = 0; r < a.left_bw; ++r)
        s |= uint64_t(lin[u.lane + L::N * q + 64 * r]) << (16 * r);
      pol.lstream[q] = s;
    }
  }
#pragma unroll
  for (int i = 0; i < PFN; ++i) list[vv * RD_LIST + u.lane + L::N * i] = ent[i];
  wave_lds_sync();
  const uint32_t *vlist = list + vv * RD_LIST;
  const unsigned n_list = count < RD_LIST ? count : RD_LIST;
  // this lane's exception slots among the staged entries
  uint64_t own = 0;
  for (unsigned k = 0; k < n_list; ++k) {
    const unsigned pos = vlist[k] & 1023u;
    if (int(pos % L::N) == u.lane) own |= uint64_t(1) << (pos / L::N);
  }
  // the first RD_REG of them go to registers, in slot order
  uint64_t emask = own;
  for (int extra = __builtin_popcountll(own) - RD_REG; extra > 0; --extra)
    emask &= ~(uint64_t(1) << (63 - __builtin_clzll(emask)));
  uint64_t lq = 0;
  if (emask) {
    for (unsigned k = 0; k < n_list; ++k) {
      const uint32_t e = vlist[k];
      const unsigned pos = e & 1023u;
      const unsigned slot = pos / L::N;
      if (int(pos % L::N) != u.lane || !((emask >> slot) & 1u)) continue;
      const int rank = __builtin_popcountll(emask & ((uint64_t(1) << slot) - 1));
      lq = (lq & ~(uint64_t(0xFFFFu) << (16 * rank))) | (uint64_t(e >> 16) << (16 * rank));
    }
  }
  if (live) {
    pol.rin = rin;
    pol.out = vout + u.lane;
    pol.dshift = dshift;
    pol.lbw = a.left_bw;
    pol.emask = emask;
    pol.lq = lq;
    live = dispatch_right_bw<UT>(pol, int(a.right_bw));  // false only for widths rejected on the host
  }
  if (!live) return;
  // entries the registers did not take: stored over the decoded values
  if (own != emask) {
    for (unsigned k = 0; k < n_list; ++k) {
      const uint32_t e = vlist[k];
      const unsigned pos = e & 1023u;
      if (int(pos % L::N) != u.lane || ((emask >> (pos / L::N)) & 1u)) continue;
      const UT right = extract_runtime<UT>(rin, int(a.right_bw), int(pos / L::N));
      store<T, NT>(vout + pos, __builtin_bit_cast(T, UT(UT(UT(e >> 16) << a.right_bw) | right)));
    }
  }
  constexpr unsigned RD_PF = 8;
  for (unsigned k0 = RD_LIST; k0 < count; k0 += RD_PF) {
    unsigned pos[RD_PF];
    uint16_t left[RD_PF];
#pragma unroll
    for (unsigned i = 0; i < RD_PF; ++i) {
      pos[i] = 0xFFFFu;  // no entry
      if (k0 + i < count) {
        pos[i] = a.positions[eo + k0 + i] & 1023u;
        left[i] = a.exceptions[eo + k0 + i];
      }
    }
#pragma unroll
    for (unsigned i = 0; i < RD_PF; ++i) {
      if (pos[i] > 1023u || int(pos[i] % L::N) != u.lane) continue;
      const UT right = extract_runtime<UT>(rin, int(a.right_bw), int(pos[i] / L::N));
      const UT v = UT(UT(left[i]) << a.right_bw) | right;
      store<T, NT>(vout + pos[i], __builtin_bit_cast(T, v));
    }
  }
}

template <typename T, bool NT>
__global__ __launch_bounds__(BLOCK) void k_alprd_decompress(AlprdArgs<T> a, T *__restrict out) {
  using UT = typename uint_of<T>::type;
  __shared__ UT dshift[8];
  __shared__ uint32_t s_list[BLOCK / WAVE][lanes<UT>::VPW * RD_LIST];
  alprd_fill_dictionary<T>(a, dshift);
  __syncthreads();
  alprd_decompress_unit<T, NT>(a, out, (uint64_t(blockIdx.x) * BLOCK + threadIdx.x) / WAVE, dshift,
                               s_list[threadIdx.x / WAVE]);
}

template <typename T>
struct RdStagePolicy {
  using UT = typename uint_of<T>::type;
  static constexpr int NL = lanes<UT>::N;
  static constexpr int Q = 64 / NL;
  const UT *rin;
  T *lds;  // this lane's slot J0 in the wave's stage
  const UT *dshift;
  uint64_t lstream[Q];
  unsigned lbw;
  template <int BW, int J0, int NJ>
  __device__ __forceinline__ void run() {
    const uint64_t lmask = (uint64_t(1) << lbw) - 1;
    unpack_slots<UT, BW, J0, NJ>(rin, [&](int j, UT r) {
      const unsigned idx = unsigned(lstream[j % Q] >> ((j / Q) * lbw)) & unsigned(lmask);
      lds[(j - J0) * NL] = __builtin_bit_cast(T, UT(dshift[idx] | r));
    });
  }
};

// Right width (W-16 .. W-1, wave-uniform) and slot part.
template <typename UT, int S, class Policy>
__device__ __forceinline__ bool dispatch_right_bw_part(Policy &pol, int bw, int part) {
  constexpr int W = lanes<UT>::W;
  const int b = __builtin_amdgcn_readfirstlane(bw);
  switch (b - (W - 16)) {
#define FLSGPU_RD_CASE(K) \
  case K: dispatch_part<UT, S, W - 16 + K>(pol, part); return true;
    FLSGPU_RD_CASE(0) FLSGPU_RD_CASE(1) FLSGPU_RD_CASE(2) FLSGPU_RD_CASE(3)
    FLSGPU_RD_CASE(4) FLSGPU_RD_CASE(5) FLSGPU_RD_CASE(6) FLSGPU_RD_CASE(7)
    FLSGPU_RD_CASE(8) FLSGPU_RD_CASE(9) FLSGPU_RD_CASE(10) FLSGPU_RD_CASE(11)
    FLSGPU_RD_CASE(12) FLSGPU_RD_CASE(13) FLSGPU_RD_CASE(14) FLSGPU_RD_CASE(15)
#undef FLSGPU_RD_CASE
    default: return false;
  }
}

// Staged ALP-RD decode of one wave-unit: each thread decodes its lane's
// slot part into LDS; the wave then walks each vector's exception list with
// coalesced loads (issued before the decode) and rebuilds the values at the
// positions in its slot range from the staged right part and the stored left
// part; then 16-byte stores.
template <typename T, int S>
__global__ __launch_bounds__(BLOCK) void k_alprd_decompress_staged(AlprdArgs<T> a,
                                                                   T *__restrict out) {
  using UT = typename uint_of<T>::type;
  using L = lanes<UT>;
  using Sh = StageShape<UT, S>;
  constexpr int Q = 64 / L::N;
  __shared__ UT dshift[8];
  __shared__ __align__(16) T s_stage[BLOCK / WAVE][L::VPW * Sh::CHUNK];
  alprd_fill_dictionary<T>(a, dshift);
  __syncthreads();
  const uint64_t wave = (uint64_t(blockIdx.x) * BLOCK + threadIdx.x) / WAVE;
  const UnitMap<UT, S> u = map_unit<UT, S>(wave);
  T *lds = s_stage[threadIdx.x / WAVE];
  const int t = int(threadIdx.x % WAVE), vv = t / L::N;
  const int j0 = u.part * Sh::NJ;
  const uint64_t group = wave / S;
  // the first 64 exceptions of each vector, loaded before the decode
  uint64_t eo[L::VPW];
  unsigned count[L::VPW], pf_pos[L::VPW];
  uint16_t pf_left[L::VPW];
#pragma unroll
  for (int v2 = 0; v2 < L::VPW; ++v2) {
    const uint64_t vec = group * L::VPW + v2;
    count[v2] = 0;
    eo[v2] = 0;
    pf_pos[v2] = 0xFFFFu;  // never inside a slot range
    if (vec < a.n_vecs) {
      eo[v2] = a.exceptions_offsets[vec];
      count[v2] = a.counts[vec];
      if (unsigned(t) < count[v2]) {
        pf_pos[v2] = a.positions[eo[v2] + t] & 1023u;
        pf_left[v2] = a.exceptions[eo[v2] + t];
      }
    }
  }
  if (u.vec < a.n_vecs) {
    RdStagePolicy<T> pol;
    pol.rin = a.right + u.vec * uint64_t(a.right_bw) * L::N + u.lane;
    pol.lds = lds + vv * Sh::CHUNK + u.lane;
    pol.dshift = dshift;
    pol.lbw = a.left_bw;
    const uint16_t *lin = a.left + u.vec * uint64_t(a.left_bw) * 64;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      uint64_t s = 0;
      for (unsigned r = 0; r < a.left_bw; ++r)
        s |= uint64_t(lin[u.lane + L::N * q + 64 * r]) << (16 * r);
      pol.lstream[q] = s;
    }
    dispatch_right_bw_part<UT, S>(pol, int(a.right_bw), u.part);  // widths checked on the host
  }
  wave_lds_sync();
  const unsigned lo = unsigned(L::N * j0), hi = unsigned(L::N * (j0 + Sh::NJ));
  const UT rmask = UT((UT(1) << a.right_bw) - 1);
#pragma unroll
  for (int v2 = 0; v2 < L::VPW; ++v2) {
    if (pf_pos[v2] >= lo && pf_pos[v2] < hi) {
      T &x = lds[v2 * Sh::CHUNK + (pf_pos[v2] - lo)];
      x = __builtin_bit_cast(T, UT((UT(pf_left[v2]) << a.right_bw) |
                                   (__builtin_bit_cast(UT, x) & rmask)));
    }
    for (unsigned k = unsigned(t + WAVE); k < count[v2]; k += WAVE) {
      const unsigned pos = a.positions[eo[v2] + k] & 1023u;
      if (pos >= lo && pos < hi) {
        T &x = lds[v2 * Sh::CHUNK + (pos - lo)];
        x = __builtin_bit_cast(T, UT((UT(a.exceptions[eo[v2] + k]) << a.right_bw) |
                                     (__builtin_bit_cast(UT, x) & rmask)));
      }
    }
  }
  wave_lds_sync();
  store_staged<T, S>(lds, out, group, u.part, a.n_vecs);
}

template <typename T, int S>
int AlprdDecompressStaged<T, S>::launch(const AlprdArgs<T> &a, T *out, hipStream_t stream) {
  using UT = typename uint_of<T>::type;
  if (a.n_vecs == 0) return FLSGPU_OK;
  const uint64_t blocks = (n_waves_for<UT, S>(a.n_vecs) + (BLOCK / WAVE) - 1) / (BLOCK / WAVE);
  if (blocks > 0x7FFFFFFFull) return set_error(FLSGPU_EINVAL, "column too large");
  hipLaunchKernelGGL((k_alprd_decompress_staged<T, S>), dim3(unsigned(blocks)), dim3(BLOCK), 0,
                     stream, a, out);
  return check_hip(hipGetLastError(), "k_alprd_decompress_staged launch");
}

template <typename T, bool NT>
int AlprdDecompress<T, NT>::launch(const AlprdArgs<T> &a, T *out, hipStream_t stream) {
  using UT = typename uint_of<T>::type;
  if (a.n_vecs == 0) return FLSGPU_OK;
  const uint64_t blocks = (n_waves_for<UT, 1>(a.n_vecs) + (BLOCK / WAVE) - 1) / (BLOCK / WAVE);
  if (blocks > 0x7FFFFFFFull) return set_error(FLSGPU_EINVAL, "column too large");
  hipLaunchKernelGGL((k_alprd_decompress<T, NT>), dim3(unsigned(blocks)), dim3(BLOCK), 0, stream,
                     a, out);
  return check_hip(hipGetLastError(), "k_alprd_decompress launch");
}

}  // namespace flsgpu_hip
