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
// Exception entries per vector staged in LDS by the one-thread-per-lane
// decoder (RdListPolicy).
constexpr unsigned RD_LIST = 64;

// The Q uint16 left-index lanes a thread of the W-bit layout needs, each
// later held in one 64-bit register (left_bw <= 3 words of 16 bits). The
// loads are issued with no branch: words past left_bw re-read word 0, and
// left_compose masks them off.
template <typename UT>
struct LeftWords {
  static constexpr int Q = 64 / lanes<UT>::N;
  uint16_t w[Q][3];
  __device__ __forceinline__ void load(const uint16_t *lin, int lane, unsigned lbw) {
    constexpr int NL = lanes<UT>::N;
#pragma unroll
    for (int q = 0; q < Q; ++q)
#pragma unroll
      for (unsigned r = 0; r < 3; ++r) w[q][r] = lin[lane + NL * q + 64 * (r < lbw ? r : 0)];
  }
  __device__ __forceinline__ void compose(unsigned lbw, uint64_t *lstream) const {
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      uint64_t x = 0;
#pragma unroll
      for (unsigned r = 0; r < 3; ++r) x |= r < lbw ? uint64_t(w[q][r]) << (16 * r) : 0;
      lstream[q] = x;
    }
  }
};

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
  template <int BW, class Hook = NoHook>
  __device__ __forceinline__ void decode(Hook &&after_loads = Hook{}) {
    const uint64_t lmask = (uint64_t(1) << lbw) - 1;
    unpack_slots<UT, BW, 0, lanes<UT>::W>(
        rin,
        [&](int j, UT r) {
          const unsigned idx = unsigned(lstream[j % Q] >> ((j / Q) * lbw)) & unsigned(lmask);
          const bool ex = (emask >> j) & 1u;
          const UT v = ex ? UT(UT(UT(lq & 0xFFFFu) << BW) | r) : UT(dshift[idx] | r);
          lq = ex ? (lq >> 16) : lq;
          store<T, NT>(out + j * NL, __builtin_bit_cast(T, v));
        },
        after_loads);
  }
  template <int BW, int J0, int NJ>
  __device__ __forceinline__ void run() {
    decode<BW>();
  }
};

// One thread per lane with the vector's exception list handed over in LDS:
// the vector's N_LANES threads load its first RD_LIST entries (64 / N_LANES
// each) once the right-part words are in flight (prepare, the after-loads
// hook), write them to the wave's LDS list as (left << 16) | position, and
// each thread collects its lane's entries: the first RD_REG (in slot order)
// into the register queue, the rest are stored after the decode.
template <typename T, bool NT>
struct RdListPolicy : RdPolicy<T, NT> {
  using UT = typename uint_of<T>::type;
  static constexpr int NL = lanes<UT>::N;
  static constexpr int PFN = int(RD_LIST) / NL;
  const uint16_t *positions;
  const uint16_t *exceptions;
  uint64_t eo;
  unsigned count, n_list;
  int lane;
  uint32_t *list;         // this wave's LDS list
  const uint32_t *vlist;  // this vector's part of it
  uint64_t own;
  LeftWords<UT> left;
  __device__ __forceinline__ void prepare() {
    left.compose(this->lbw, this->lstream);
    uint16_t p[PFN], l[PFN];
#pragma unroll
    for (int i = 0; i < PFN; ++i) {
      const unsigned k = unsigned(lane + NL * i);
      p[i] = *list_entry(positions, eo, k, count);
      l[i] = *list_entry(exceptions, eo, k, count);
    }
#pragma unroll
    for (int i = 0; i < PFN; ++i)
      list[(vlist - list) + lane + NL * i] = (p[i] & 1023u) | (uint32_t(l[i]) << 16);
    wave_lds_sync();
    own = 0;
    for (unsigned k = 0; k < n_list; ++k) {
      const unsigned pos = vlist[k] & 1023u;
      if (int(pos % NL) == lane) own |= uint64_t(1) << (pos / NL);
    }
    uint64_t em = own;
    for (int extra = __builtin_popcountll(own) - RD_REG; extra > 0; --extra)
      em &= ~(uint64_t(1) << (63 - __builtin_clzll(em)));
    uint64_t lq = 0;
    if (em) {
      for (unsigned k = 0; k < n_list; ++k) {
        const uint32_t e = vlist[k];
        const unsigned pos = e & 1023u;
        const unsigned slot = pos / NL;
        if (int(pos % NL) != lane || !((em >> slot) & 1u)) continue;
        const int rank = __builtin_popcountll(em & ((uint64_t(1) << slot) - 1));
        lq = (lq & ~(uint64_t(0xFFFFu) << (16 * rank))) | (uint64_t(e >> 16) << (16 * rank));
      }
    }
    this->emask = em;
    this->lq = lq;
  }
  template <int BW, int J0, int NJ>
  __device__ __forceinline__ void run() {
    this->template decode<BW>([&] { prepare(); });
  }
};

// The block fills the pre-shifted dictionary in LDS before any wave decodes.
template <typename T>
__device__ __forceinline__ void alprd_fill_dictionary(const AlprdArgs<T> &a,
                                                      typename uint_of<T>::type *dshift) {
  using UT = typename uint_of<T>::type;
  if (threadIdx.x < 8)
    dshift[threadIdx.x] = threadIdx.x < a.dict_size ? UT(UT(a.dict[threadIdx.x]) << a.right_bw) : UT(0);
}

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

// Reading each thread's whole list from global memory and patching after
// the decode made the time grow with the exception count (2^26 doubles: 201
// us at 4 exceptions per vector, 289 us at 18); the LDS hand-over and the
// register queue keep it flat.
// Where lanes past a column's last vector write (see alprd_decompress_unit):
// at most VPW - 1 vectors' lanes per launch. Never read.
static __device__ uint64_t k_rd_sink[1024];

// `list`: this wave's LDS, VPW * RD_LIST entries. The right-part words'
// addresses need no metadata, so they are the first loads the decode issues;
// the list is set up in its after-loads hook, whose LDS hand-over is
// wave-wide. Every lane of the wave takes the same path through it: a lane
// past the column's last vector decodes vector 0 with no exceptions into a
// sink that nothing reads.
template <typename T, bool NT>
__device__ __forceinline__ void alprd_decompress_unit(const AlprdArgs<T> &a, T *__restrict out,
                                                      uint64_t wave,
                                                      const typename uint_of<T>::type *dshift,
                                                      uint32_t *list) {
  using UT = typename uint_of<T>::type;
  using L = lanes<UT>;
  const UnitMap<UT, 1> u = map_unit<UT, 1>(wave);
  const int vv = int(threadIdx.x % WAVE) / L::N;
  // a wave past the column's end leaves at once (all its lanes together);
  // only the last wave can hold lanes on both sides of the end
  if (uniform_u64(wave) * L::VPW >= a.n_vecs) return;
  const bool live = u.vec < a.n_vecs;
  const uint64_t vs = live ? u.vec : 0;
  const uint64_t eo = a.exceptions_offsets[vs];
  const unsigned count = live ? unsigned(a.counts[vs]) : 0u;
  T *vout = live ? out + u.vec * VALUES_PER_VECTOR : reinterpret_cast<T *>(k_rd_sink);
  const UT *rin = a.right + vs * uint64_t(a.right_bw) * L::N + u.lane;
  RdListPolicy<T, NT> pol;
  pol.rin = rin;
  pol.out = vout + u.lane;
  pol.dshift = dshift;
  pol.lbw = a.left_bw;
  pol.emask = 0;
  pol.lq = 0;
  pol.positions = a.positions;
  pol.exceptions = a.exceptions;
  pol.eo = eo;
  pol.count = count;
  pol.n_list = count < RD_LIST ? count : RD_LIST;
  pol.lane = u.lane;
  pol.list = list;
  pol.vlist = list + vv * RD_LIST;
  pol.own = 0;
  pol.left.load(a.left + vs * uint64_t(a.left_bw) * 64, u.lane, a.left_bw);
  if (!dispatch_right_bw<UT>(pol, int(a.right_bw))) return;  // widths are checked on the host
  const uint64_t own = pol.own;
  const unsigned n_list = pol.n_list;
  const uint32_t *vlist = pol.vlist;
  uint64_t emask = pol.emask;
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

// `hook` runs once the right-part words are in flight (see unpack_slots).
template <typename T, class Hook>
struct RdStagePolicy {
  using UT = typename uint_of<T>::type;
  static constexpr int NL = lanes<UT>::N;
  static constexpr int Q = 64 / NL;
  const UT *rin;
  T *lds;  // this lane's slot J0 in the wave's stage
  const UT *dshift;
  uint64_t lstream[Q];
  unsigned lbw;
  LeftWords<UT> left;
  Hook hook;
  template <int BW, int J0, int NJ>
  __device__ __forceinline__ void run() {
    const uint64_t lmask = (uint64_t(1) << lbw) - 1;
    unpack_slots<UT, BW, J0, NJ>(
        rin,
        [&](int j, UT r) {
          const unsigned idx = unsigned(lstream[j % Q] >> ((j / Q) * lbw)) & unsigned(lmask);
          lds[(j - J0) * NL] = __builtin_bit_cast(T, UT(dshift[idx] | r));
        },
        [&] {
          left.compose(lbw, lstream);
          hook();
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
// `dshift`: the pre-shifted dictionary in LDS; `lds`: the wave's stage,
// VPW * CHUNK elements, 16-byte aligned.
template <typename T, int S>
__device__ __forceinline__ void alprd_staged_unit(const AlprdArgs<T> &a, T *__restrict out,
                                                  uint64_t wave,
                                                  const typename uint_of<T>::type *dshift,
                                                  T *lds) {
  using UT = typename uint_of<T>::type;
  using L = lanes<UT>;
  using Sh = StageShape<UT, S>;
  const UnitMap<UT, S> u = map_unit<UT, S>(wave);
  const int t = int(threadIdx.x % WAVE), vv = t / L::N;
  const int j0 = u.part * Sh::NJ;
  const uint64_t group = wave / S;
  // The headers of the group's vectors first (wave-uniform), then each
  // vector's first 64 exceptions (entry t to thread t) once the right-part
  // words are in flight, so one round trip covers both.
  const uint64_t g = uniform_u64(group);
  uint64_t eo[L::VPW];
  unsigned count[L::VPW], pf_pos[L::VPW];
  uint16_t pf_left[L::VPW];
#pragma unroll
  for (int v2 = 0; v2 < L::VPW; ++v2) {
    const uint64_t vec = g * L::VPW + v2;
    const uint64_t vs = vec < a.n_vecs ? vec : 0;
    eo[v2] = a.exceptions_offsets[vs];
    count[v2] = a.counts[vs];
  }
#pragma unroll
  for (int v2 = 0; v2 < L::VPW; ++v2)
    count[v2] = __builtin_amdgcn_readfirstlane(g * L::VPW + v2 < a.n_vecs ? count[v2] : 0u);
  bool entries = false;
  auto hook = [&] {
#pragma unroll
    for (int v2 = 0; v2 < L::VPW; ++v2) {
      pf_pos[v2] = *list_entry(a.positions, eo[v2], unsigned(t), count[v2]);
      pf_left[v2] = *list_entry(a.exceptions, eo[v2], unsigned(t), count[v2]);
    }
    entries = true;
  };
  // a lane past the column's last vector decodes vector 0 into its unused
  // LDS row (store_staged skips it), so there is no branch between the loads
  const uint64_t vs = u.vec < a.n_vecs ? u.vec : 0;
  RdStagePolicy<T, decltype(hook) &> pol{a.right + vs * uint64_t(a.right_bw) * L::N + u.lane,
                                         lds + vv * Sh::CHUNK + u.lane, dshift, {}, a.left_bw,
                                         {}, hook};
  pol.left.load(a.left + vs * uint64_t(a.left_bw) * 64, u.lane, a.left_bw);
  dispatch_right_bw_part<UT, S>(pol, int(a.right_bw), u.part);  // widths checked on the host
  if (!entries) hook();
#pragma unroll
  for (int v2 = 0; v2 < L::VPW; ++v2)
    if (unsigned(t) >= count[v2]) pf_pos[v2] = 0xFFFFu;  // never inside a slot range
    else pf_pos[v2] &= 1023u;
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
__global__ __launch_bounds__(BLOCK) void k_alprd_decompress_staged(AlprdArgs<T> a,
                                                                   T *__restrict out) {
  using UT = typename uint_of<T>::type;
  __shared__ UT dshift[8];
  __shared__ __align__(16) T s_stage[BLOCK / WAVE][lanes<UT>::VPW * StageShape<UT, S>::CHUNK];
  alprd_fill_dictionary<T>(a, dshift);
  __syncthreads();
  alprd_staged_unit<T, S>(a, out, (uint64_t(blockIdx.x) * BLOCK + threadIdx.x) / WAVE, dshift,
                          s_stage[threadIdx.x / WAVE]);
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
