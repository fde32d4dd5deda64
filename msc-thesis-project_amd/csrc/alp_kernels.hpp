// ALP / G-ALP decode kernels for gfx950. Included only by alp_inst.hip.
//
// Value formula, in this exact order (src/flsgpu/alp.cuh:40-45, CPU
// src/alp/falp.cpp, e.g. :95-130):
//   out = (T)(INT)((UINT)(unpacked + base) * (UINT)FACT[factor_idx]) * FRAC[fraction_idx]
// i.e. wrapping unsigned add and multiply, reinterpretation as signed,
// round-to-nearest-even integer->float conversion, one IEEE multiply. No FMA
// contraction (built with -ffp-contract=off), no division by 10^e.
//
// Exception patching (src/flsgpu/alp.cuh:48-461, CPU decoder.hpp:137-145):
// every exception is written by the thread that owns its (lane, slot). That
// thread loads its first few exceptions before decoding (G-ALP: through its
// lane's header; standard layout: through the wave's LDS index of the
// vector's list) and stores them in place of the decoded values; further
// ones it writes after it stored the decoded value at the same address, so
// the final value is the exception by single-thread program order. The
// LDS-staged kernels patch the staged values.
#pragma once

#include <algorithm>

#include "alp_impl.hpp"
#include "kernels.hpp"
#include "stage.hpp"

namespace flsgpu_hip {

template <typename T>
__device__ __forceinline__ T alp_value(typename uint_of<T>::type v, typename uint_of<T>::type base,
                                       typename uint_of<T>::type fact, T frac) {
  using UT = typename uint_of<T>::type;
  using ST = typename int_of<T>::type;
  const ST s = ST(UT(UT(v + base) * fact));
  return T(s) * frac;
}

// The per-vector constants of the value formula. The FACT / FRAC table
// entries are read in the decode's after-loads hook (see unpack_slots): their
// indices arrive with the vector's other metadata, and reading the tables
// before the packed words are requested would add a round trip ahead of them.
template <typename T>
struct AlpConsts {
  using UT = typename uint_of<T>::type;
  UT base;
  unsigned fi, ei;
  UT fact;
  T frac;
  __device__ __forceinline__ void resolve() {
    fact = alp_factor<T>(fi);
    frac = alp_fraction<T>(ei);
  }
  __device__ __forceinline__ T operator()(UT v) const { return alp_value<T>(v, base, fact, frac); }
};

template <typename T>
__device__ __forceinline__ AlpConsts<T> alp_consts(const AlpArgs<T> &a, uint64_t vec) {
  return AlpConsts<T>{a.ffor.bases[vec], a.factor_indices[vec], a.fraction_indices[vec], 0, T(0)};
}

template <typename T, bool NT>
struct AlpDecompressPolicy {
  using UT = typename uint_of<T>::type;
  const UT *in;
  T *out;
  AlpConsts<T> c;
  template <int BW, int J0, int NJ>
  __device__ __forceinline__ void run() {
    unpack_slots<UT, BW, J0, NJ>(
        in, [&](int j, UT v) { store<T, NT>(out + j * lanes<UT>::N, c(v)); }, [&] { c.resolve(); });
  }
  __device__ __forceinline__ void run_generic(int bw, int j0, int nj) {
    c.resolve();
    unpack_runtime<UT>(in, bw, j0, nj,
                       [&](int j, UT v) { store<T, NT>(out + j * lanes<UT>::N, c(v)); });
  }
};

// Decode with exceptions patched in registers: bit j of `emask` marks slot j
// as an exception whose value is the head of the queue q (filled in slot
// order), so each value is stored once. With non-temporal stores an exception
// stored after the decode rewrites part of a line already on its way to HBM
// (2^26 doubles, 10 exceptions per vector, standard layout: 146 -> 133 us).
// RK entries per thread: a few exceptions per lane are typical (2 decimals,
// 1 % exceptions: 0.6 per f64 lane; f32 columns of 1 decimal: 4.5).
template <typename T>
constexpr int alp_reg_k() { return sizeof(T) == 8 ? 4 : 8; }

template <typename T, bool NT>
struct AlpRegPatchPolicy {
  using UT = typename uint_of<T>::type;
  static constexpr int RK = alp_reg_k<T>();
  const UT *in;
  T *out;
  AlpConsts<T> c;
  uint64_t emask;
  T q[RK];
  __device__ __forceinline__ T take(int j, T y) {
    const bool ex = (emask >> j) & 1u;
    const T r = ex ? q[0] : y;
#pragma unroll
    for (int i = 0; i + 1 < RK; ++i) q[i] = ex ? q[i + 1] : q[i];
    return r;
  }
  template <int BW, int J0, int NJ>
  __device__ __forceinline__ void run() {
    unpack_slots<UT, BW, J0, NJ>(
        in, [&](int j, UT v) { store<T, NT>(out + j * lanes<UT>::N, take(j, c(v))); },
        [&] { c.resolve(); });
  }
  __device__ __forceinline__ void run_generic(int bw, int j0, int nj) {
    c.resolve();
    unpack_runtime<UT>(in, bw, j0, nj,
                       [&](int j, UT v) { store<T, NT>(out + j * lanes<UT>::N, take(j, c(v))); });
  }
};

// G-ALP register patching. The lane's first RK list entries are loaded
// before the decode (load_entries, no branch on the lane's count, so the
// wait placed before their first use does not also wait for the packed
// words issued after them) and sorted into the queue once the packed words
// are in flight (prepare, in the after-loads hook): entries in this thread's
// slot range go to the queue by slot rank (a repeated slot keeps the later
// entry). `queued` marks the entries queued.
template <typename T, bool NT>
struct AlpGalpRegPolicy : AlpRegPatchPolicy<T, NT> {
  using UT = typename uint_of<T>::type;
  using Base = AlpRegPatchPolicy<T, NT>;
  static constexpr int RK = Base::RK;
  uint64_t first;
  unsigned count;
  int lane, j0, nj;
  uint64_t queued;
  unsigned raw[RK];
  T val[RK];
  __device__ __forceinline__ void load_entries(const AlpArgs<T> &a) {
#pragma unroll
    for (int k = 0; k < RK; ++k) {
      raw[k] = *list_entry(a.positions, first, unsigned(k), count);
      val[k] = *list_entry(a.exceptions, first, unsigned(k), count);
    }
  }
  __device__ __forceinline__ void prepare() {
    constexpr int NL = lanes<UT>::N;
    unsigned slot[RK];
#pragma unroll
    for (int k = 0; k < RK; ++k) {
      const unsigned sl = raw[k] / unsigned(NL);
      slot[k] = unsigned(k) < count && int(raw[k] % unsigned(NL)) == lane && sl >= unsigned(j0) &&
                        sl < unsigned(j0 + nj)
                    ? sl
                    : 0xFFFFu;  // not queued
    }
    this->emask = 0;
    queued = 0;
#pragma unroll
    for (int k = 0; k < RK; ++k) {
      this->q[k] = T(0);
      if (slot[k] != 0xFFFFu) {
        this->emask |= uint64_t(1) << slot[k];
        queued |= uint64_t(1) << k;
      }
    }
#pragma unroll
    for (int k = 0; k < RK; ++k) {
      if (slot[k] != 0xFFFFu) {
        const int r = __builtin_popcountll(this->emask & ((uint64_t(1) << slot[k]) - 1));
#pragma unroll
        for (int i = 0; i < RK; ++i) this->q[i] = r == i ? val[k] : this->q[i];
      }
    }
  }
  template <int BW, int J0, int NJ>
  __device__ __forceinline__ void run() {
    unpack_slots<UT, BW, J0, NJ>(
        this->in,
        [&](int j, UT v) { store<T, NT>(this->out + j * lanes<UT>::N, this->take(j, this->c(v))); },
        [&] {
          this->c.resolve();
          prepare();
        });
  }
  __device__ __forceinline__ void run_generic(int bw, int j0_, int nj_) {
    prepare();
    Base::run_generic(bw, j0_, nj_);
  }
};

// Bit (j - J0) set when slot j decodes (before patching) to `target`.
template <typename T>
struct AlpMatchPolicy {
  using UT = typename uint_of<T>::type;
  const UT *in;
  AlpConsts<T> c;
  T target;
  uint64_t mask;
  template <int BW, int J0, int NJ>
  __device__ __forceinline__ void run() {
    unpack_slots<UT, BW, J0, NJ>(
        in, [&](int j, UT v) { mask |= uint64_t(c(v) == target) << (j - J0); },
        [&] { c.resolve(); });
  }
  __device__ __forceinline__ void run_generic(int bw, int j0, int nj) {
    c.resolve();
    unpack_runtime<UT>(in, bw, j0, nj,
                       [&](int j, UT v) { mask |= uint64_t(c(v) == target) << (j - j0); });
  }
};

// ---------------------------------------------------------------------------
// Exception lookup.
//
// G-ALP layout: the owner thread reads its lane's header (count << 10 |
// offset) and its exceptions, PF at a time with every load issued before
// the first is used (the lane's exceptions are contiguous).
//
// Standard layout: a vector's exceptions are one list sorted by position, so
// a thread cannot find its own without scanning. The wave instead indexes
// its vectors' lists cooperatively in LDS (coalesced loads of the
// positions): a 1024-bit map of exception positions per vector and a slot
// mask per lane. The owner then walks its mask; an exception's index in the
// list is its rank in the position map (positions are sorted).
// ---------------------------------------------------------------------------
constexpr int ALP_PF = 8;

template <typename UT>
__device__ __forceinline__ void lds_or(UT *p, UT v) {
  if constexpr (sizeof(UT) == 8)
    atomicOr(reinterpret_cast<unsigned long long *>(p), static_cast<unsigned long long>(v));
  else
    atomicOr(reinterpret_cast<unsigned *>(p), static_cast<unsigned>(v));
}

// `skip`: bit k set for list entries k < 64 the caller has applied already.
template <typename T, class F>
__device__ __forceinline__ void galp_own_exceptions(const AlpArgs<T> &a, uint64_t vec, int lane,
                                                    int j0, int nj, F &&f, uint64_t skip = 0) {
  constexpr int NL = lanes<typename uint_of<T>::type>::N;
  const uint64_t eo = a.exceptions_offsets[vec];
  const unsigned oc = a.offsets_counts[vec * NL + lane];
  const unsigned count = oc >> 10;
  const uint64_t first = eo + (oc & 0x3FFu);
  for (unsigned k0 = 0; k0 < count; k0 += ALP_PF) {
    unsigned pos[ALP_PF];
    T val[ALP_PF];
#pragma unroll
    for (int i = 0; i < ALP_PF; ++i) {
      if (k0 + i < count) {
        pos[i] = a.positions[first + k0 + i];
        val[i] = a.exceptions[first + k0 + i];
      }
    }
#pragma unroll
    for (int i = 0; i < ALP_PF; ++i) {
      if (k0 + i < count && !(k0 + i < 64 && ((skip >> (k0 + i)) & 1u))) {
        const int slot = int(pos[i]) / NL;
        if (slot >= j0 && slot < j0 + nj) f(pos[i], val[i]);
      }
    }
  }
}

// The first PF * 64 entries of the exception lists of the wave's VPW vectors
// (entry t + 64 p to thread t of the wave). The per-vector headers are loaded
// first (wave-uniform), then every list entry, and nothing loaded is used
// before the last load is issued: the compiler puts a wait before the first
// use of each loaded value and the load counter drains in issue order, so
// loading and using vector by vector cost two memory round trips per vector
// (2^26 doubles, 10 exceptions per vector, staged: 155 us against 108 us
// without exceptions).
template <typename T, int LAYOUT, int PF>
struct ListPrefetch {
  using UT = typename uint_of<T>::type;
  static constexpr int NL = lanes<UT>::N;
  static constexpr int VPW = lanes<UT>::VPW;
  uint64_t eo[VPW];
  unsigned count[VPW];
  unsigned pos[VPW][PF];  // raw positions (see entry_pos)
  T val[VPW][PF];

  __device__ __forceinline__ void load_headers(const AlpArgs<T> &a, uint64_t group) {
    const uint64_t g = uniform_u64(group);
    unsigned raw[VPW];
#pragma unroll
    for (int v2 = 0; v2 < VPW; ++v2) {
      const uint64_t vec = g * VPW + v2;
      const uint64_t vs = vec < a.ffor.n_vecs ? vec : 0;
      eo[v2] = a.exceptions_offsets[vs];
      if constexpr (LAYOUT == ALP_GALP) raw[v2] = a.offsets_counts[vs * NL + (NL - 1)];
      else raw[v2] = a.counts[vs];
    }
#pragma unroll
    for (int v2 = 0; v2 < VPW; ++v2) {
      // G-ALP lanes are concatenated in order: the last lane's offset + count
      const unsigned c = LAYOUT == ALP_GALP ? (raw[v2] & 0x3FFu) + (raw[v2] >> 10) : raw[v2];
      count[v2] = __builtin_amdgcn_readfirstlane(g * VPW + v2 < a.ffor.n_vecs ? c : 0u);
    }
  }
  // Entry t + 64 p of each list (see entry_pos), loaded with no branch (see
  // list_entry).
  __device__ __forceinline__ void load_entries(const AlpArgs<T> &a) {
    const unsigned t = threadIdx.x % WAVE;
#pragma unroll
    for (int v2 = 0; v2 < VPW; ++v2) {
#pragma unroll
      for (int p = 0; p < PF; ++p) {
        const unsigned k = t + unsigned(p * WAVE);
        pos[v2][p] = *list_entry(a.positions, eo[v2], k, count[v2]);
        val[v2][p] = *list_entry(a.exceptions, eo[v2], k, count[v2]);
      }
    }
  }
  // Raw position of entry t + 64 p of list v2, or 0xFFFF past its end.
  __device__ __forceinline__ unsigned entry_pos(int v2, int p) const {
    return (threadIdx.x % WAVE) + unsigned(p * WAVE) < count[v2] ? pos[v2][p] : 0xFFFFu;
  }
};

// Per-wave LDS index of the standard-layout exception lists of the VPW
// vectors of wave group `group`: a map of exception positions, a slot mask
// per lane and the lists' first 64 values, so that an owner thread finds an
// exception's value in LDS (its list index is its rank in the position map,
// positions being sorted) and reads global memory only past entry 64.
template <typename T>
struct StdIndex {
  using UT = typename uint_of<T>::type;
  static constexpr int NL = lanes<UT>::N;
  static constexpr int VPW = lanes<UT>::VPW;
  uint64_t *bits;  // [VPW][16]: bit p set when position p holds an exception
  UT *lmask;       // [VPW][NL]: bit j set when slot j of the lane holds one
  T *vals;         // [VPW][WAVE]: entries 0..63 of each list

  // All 64 threads of the wave must call this (no early exit before it).
  __device__ __forceinline__ void build(const AlpArgs<T> &a, uint64_t group) {
    const int t = threadIdx.x % WAVE;
    ListPrefetch<T, ALP_STD, 1> lp;
    lp.load_headers(a, group);
    lp.load_entries(a);
    for (int i = t; i < VPW * 16; i += WAVE) bits[i] = 0;
    for (int i = t; i < VPW * NL; i += WAVE) lmask[i] = 0;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int vv = 0; vv < VPW; ++vv) {
      if (lp.entry_pos(vv, 0) != 0xFFFFu) {
        const unsigned pos = lp.pos[vv][0] & 1023u;
        lds_or<uint64_t>(&bits[vv * 16 + (pos >> 6)], uint64_t(1) << (pos & 63));
        lds_or<UT>(&lmask[vv * NL + pos % NL], UT(UT(1) << (pos / NL)));
        vals[vv * WAVE + t] = lp.val[vv][0];
      }
      for (unsigned k = unsigned(t + WAVE); k < lp.count[vv]; k += WAVE) {
        const unsigned pos = a.positions[lp.eo[vv] + k] & 1023u;
        lds_or<uint64_t>(&bits[vv * 16 + (pos >> 6)], uint64_t(1) << (pos & 63));
        lds_or<UT>(&lmask[vv * NL + pos % NL], UT(UT(1) << (pos / NL)));
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    __builtin_amdgcn_wave_barrier();
  }

  // Slot mask of (vector vv of the group, lane) in slots [j0, j0 + nj).
  __device__ __forceinline__ UT slots(int vv, int lane, int j0, int nj) const {
    constexpr int W = lanes<UT>::W;
    UT m = lmask[vv * NL + lane];
    if (nj < W) m &= UT(((UT(1) << nj) - 1) << j0);
    return m;
  }
  // Index in vector vv's list of the exception at position p.
  __device__ __forceinline__ unsigned rank(int vv, unsigned p) const {
    const uint64_t *b = bits + vv * 16;
    unsigned r = unsigned(__builtin_popcountll(b[p >> 6] & ((1ull << (p & 63)) - 1)));
    for (unsigned w = 0; w < (p >> 6); ++w) r += unsigned(__builtin_popcountll(b[w]));
    return r;
  }
  // Value of entry r of vector vv's list (vector `vec` of the column).
  __device__ __forceinline__ T value(const AlpArgs<T> &a, uint64_t vec, int vv, unsigned r) const {
    return r < unsigned(WAVE) ? vals[vv * WAVE + r] : a.exceptions[a.exceptions_offsets[vec] + r];
  }
  // Exceptions of (vector vv of the group, lane) in slots [j0, j0 + nj).
  template <class F>
  __device__ __forceinline__ void own(const AlpArgs<T> &a, uint64_t vec, int vv, int lane, int j0,
                                      int nj, F &&f) const {
    own_slots(a, vec, vv, lane, slots(vv, lane, j0, nj), f);
  }
  // The exceptions at the slots set in m of (vector vv of the group, lane).
  template <class F>
  __device__ __forceinline__ void own_slots(const AlpArgs<T> &a, uint64_t vec, int vv, int lane,
                                            UT m, F &&f) const {
    while (m) {
      const int j = int(__builtin_ctzll(uint64_t(m)));
      m &= UT(m - 1);
      const unsigned p = unsigned(lane + NL * j);
      f(p, value(a, vec, vv, rank(vv, p)));
    }
  }
};

// LDS for the standard-layout index of the calling wave (unused otherwise).
template <typename T, int LAYOUT, bool PATCH>
__device__ __forceinline__ StdIndex<T> wave_std_index() {
  using UT = typename uint_of<T>::type;
  using L = lanes<UT>;
  if constexpr (PATCH && LAYOUT == ALP_STD) {
    __shared__ uint64_t s_bits[BLOCK / WAVE][L::VPW * 16];
    __shared__ UT s_lmask[BLOCK / WAVE][L::VPW * L::N];
    __shared__ T s_vals[BLOCK / WAVE][L::VPW * WAVE];
    const int w = threadIdx.x / WAVE;
    return StdIndex<T>{s_bits[w], s_lmask[w], s_vals[w]};
  } else {
    return StdIndex<T>{nullptr, nullptr, nullptr};
  }
}

// Per-thread scan of the exception list: the owner thread reads its lane's
// exceptions one at a time (G-ALP) or scans the vector's list (standard
// layout). Used by the batched decoder, which keeps this simpler patch path.
template <typename T, int LAYOUT, class F>
__device__ __forceinline__ void scan_own_exceptions(const AlpArgs<T> &a, uint64_t vec, int lane,
                                                    int j0, int nj, F &&f) {
  constexpr int NL = lanes<typename uint_of<T>::type>::N;
  const uint64_t eo = a.exceptions_offsets[vec];
  if constexpr (LAYOUT == ALP_GALP) {
    const unsigned oc = a.offsets_counts[vec * NL + lane];
    const unsigned count = oc >> 10;
    const uint64_t first = eo + (oc & 0x3FFu);
    for (unsigned k = 0; k < count; ++k) {
      const unsigned pos = a.positions[first + k];
      const int slot = int(pos) / NL;
      if (slot >= j0 && slot < j0 + nj) f(pos, a.exceptions[first + k]);
    }
  } else {
    const unsigned count = a.counts[vec];
    for (unsigned k = 0; k < count; ++k) {
      const unsigned pos = a.positions[eo + k];
      const int slot = int(pos) / NL;
      if (int(pos % NL) == lane && slot >= j0 && slot < j0 + nj) f(pos, a.exceptions[eo + k]);
    }
  }
}

// One wave-unit of the batched decoder: decode, then the per-thread scan.
template <typename T, bool NT, int LAYOUT>
__device__ __forceinline__ void alp_decompress_unit_scan(const AlpArgs<T> &a, T *__restrict out,
                                                         uint64_t wave) {
  using UT = typename uint_of<T>::type;
  using L = lanes<UT>;
  const UnitMap<UT, 1> u = map_unit<UT, 1>(wave);
  if (u.vec >= a.ffor.n_vecs) return;
  T *vout = out + u.vec * VALUES_PER_VECTOR;
  AlpDecompressPolicy<T, NT> pol{a.ffor.packed + a.ffor.vector_offsets[u.vec] + u.lane,
                                 vout + u.lane, alp_consts(a, u.vec)};
  if (!dispatch_bw<UT, 1>(pol, a.ffor.bit_widths[u.vec], 0)) {
    for (int j = 0; j < L::W; ++j) vout[u.lane + j * L::N] = T(0);
  }
  scan_own_exceptions<T, LAYOUT>(a, u.vec, u.lane, 0, L::W,
                                 [&](unsigned pos, T e) { store<T, NT>(vout + pos, e); });
}

template <typename T, int S, bool NT, int LAYOUT, bool PATCH>
__device__ __forceinline__ void alp_decompress_unit(const AlpArgs<T> &a, T *__restrict out,
                                                    uint64_t wave) {
  using UT = typename uint_of<T>::type;
  using L = lanes<UT>;
  const UnitMap<UT, S> u = map_unit<UT, S>(wave);
  const int j0 = u.part * (L::W / S);
  constexpr int NJ = L::W / S;
  if constexpr (PATCH && LAYOUT == ALP_STD) {
    StdIndex<T> idx = wave_std_index<T, LAYOUT, PATCH>();
    const bool live = u.vec < a.ffor.n_vecs;
    const uint64_t vs = live ? u.vec : 0;
    // the decode's metadata is loaded before the index's lists, so that one
    // wait covers both
    const uint64_t off = a.ffor.vector_offsets[vs];
    const int bw = a.ffor.bit_widths[vs];
    const UT base = a.ffor.bases[vs];
    const unsigned fi = a.factor_indices[vs], ei = a.fraction_indices[vs];
    idx.build(a, wave / S);
    if (!live) return;
    T *vout = out + u.vec * VALUES_PER_VECTOR;
    const int vv = int((threadIdx.x % WAVE) / L::N);
    AlpRegPatchPolicy<T, NT> pol{a.ffor.packed + off + u.lane, vout + u.lane,
                                 AlpConsts<T>{base, fi, ei, 0, T(0)}, 0, {}};
    // the lane's first RK exception slots of the range go to the queue (list
    // index = rank in the position map), the rest are stored after the decode
    const UT lm = idx.slots(vv, u.lane, j0, NJ);
    UT rest = lm;
#pragma unroll
    for (int i = 0; i < pol.RK; ++i) {
      pol.q[i] = T(0);
      if (rest) {
        const int j = int(__builtin_ctzll(uint64_t(rest)));
        rest &= UT(rest - 1);
        pol.emask |= uint64_t(1) << j;
        pol.q[i] = idx.value(a, u.vec, vv, idx.rank(vv, unsigned(u.lane + L::N * j)));
      }
    }
    if (!dispatch_bw<UT, S>(pol, bw, u.part)) {
      for (int j = j0; j < j0 + NJ; ++j) vout[u.lane + j * L::N] = T(0);
      rest = lm;  // nothing was patched
    }
    idx.own_slots(a, u.vec, vv, u.lane, rest, [&](unsigned pos, T e) { store<T, NT>(vout + pos, e); });
  } else if constexpr (PATCH) {
    if (u.vec >= a.ffor.n_vecs) return;
    T *vout = out + u.vec * VALUES_PER_VECTOR;
    const int bw = a.ffor.bit_widths[u.vec];
    const uint64_t eo = a.exceptions_offsets[u.vec];
    const unsigned oc = a.offsets_counts[u.vec * L::N + u.lane];
    AlpGalpRegPolicy<T, NT> pol;
    pol.in = a.ffor.packed + a.ffor.vector_offsets[u.vec] + u.lane;
    pol.out = vout + u.lane;
    pol.c = alp_consts(a, u.vec);
    pol.first = eo + (oc & 0x3FFu);
    pol.count = oc >> 10;
    pol.lane = u.lane;
    pol.j0 = j0;
    pol.nj = NJ;
    pol.queued = 0;
    pol.load_entries(a);
    constexpr int RK = AlpGalpRegPolicy<T, NT>::RK;
    const unsigned count = pol.count;
    uint64_t queued = 0;
    const bool decoded = dispatch_bw<UT, S>(pol, bw, u.part);
    if (decoded) {
      queued = pol.queued;
    } else {
      for (int j = j0; j < j0 + NJ; ++j) vout[u.lane + j * L::N] = T(0);
    }
    // every entry of a list of at most RK that lies in this thread's range
    // is queued, so only longer lists (or an undecodable width) need the pass
    if (count > unsigned(RK) || !decoded)
      galp_own_exceptions<T>(a, u.vec, u.lane, j0, NJ,
                             [&](unsigned pos, T e) { store<T, NT>(vout + pos, e); }, queued);
  } else {
    if (u.vec >= a.ffor.n_vecs) return;
    T *vout = out + u.vec * VALUES_PER_VECTOR;
    AlpDecompressPolicy<T, NT> pol{a.ffor.packed + a.ffor.vector_offsets[u.vec] + u.lane,
                                   vout + u.lane, alp_consts(a, u.vec)};
    if (!dispatch_bw<UT, S>(pol, a.ffor.bit_widths[u.vec], u.part)) {
      for (int j = j0; j < j0 + NJ; ++j) vout[u.lane + j * L::N] = T(0);
    }
  }
}

// `hook` runs once the packed words are in flight (see unpack_slots).
template <typename T, class Hook = NoHook>
struct AlpStagePolicy {
  using UT = typename uint_of<T>::type;
  const UT *in;
  T *lds;
  AlpConsts<T> c;
  Hook hook;
  template <int BW, int J0, int NJ>
  __device__ __forceinline__ void run() {
    unpack_slots<UT, BW, J0, NJ>(
        in, [&](int j, UT v) { lds[(j - J0) * lanes<UT>::N] = c(v); },
        [&] {
          c.resolve();
          hook();
        });
  }
  __device__ __forceinline__ void run_generic(int bw, int j0, int nj) {
    c.resolve();
    hook();
    unpack_runtime<UT>(in, bw, j0, nj, [&](int j, UT v) { lds[(j - j0) * lanes<UT>::N] = c(v); });
  }
};

// Staged ALP decode of one wave-unit: decode into LDS, then the wave walks
// each vector's exception list with coalesced loads and overwrites the
// staged values whose positions fall in its slot range (positions are unique
// in a vector, so the writes never collide), then 16-byte stores.
// `lds`: the wave's stage, VPW * CHUNK elements, 16-byte aligned.
template <typename T, int S, int LAYOUT, bool PATCH>
__device__ __forceinline__ void alp_staged_unit(const AlpArgs<T> &a, T *__restrict out,
                                                uint64_t wave, T *lds) {
  using UT = typename uint_of<T>::type;
  using L = lanes<UT>;
  using Sh = StageShape<UT, S>;
  const UnitMap<UT, S> u = map_unit<UT, S>(wave);
  const int t = int(threadIdx.x % WAVE), vv = t / L::N;
  const int j0 = u.part * Sh::NJ;
  const uint64_t group = wave / S;
  // The first PF * 64 exceptions of each vector are loaded while the packed
  // words are in flight (the decode's hook), so that one memory round trip
  // covers both; longer lists finish in a loop after the decode. Thread t
  // loads entry t of every vector of the group, so a thread whose own vector
  // does not exist, or does not decode, still loads its entries.
  constexpr int PF = PATCH ? (sizeof(T) == 4 ? 3 : 1) : 1;
  ListPrefetch<T, LAYOUT, PF> lp;
  if constexpr (PATCH) lp.load_headers(a, group);
  bool entries = !PATCH;
  auto hook = [&] {
    if constexpr (PATCH) lp.load_entries(a);
    entries = true;
  };
  // A lane past the column's last vector decodes vector 0 into its unused
  // LDS row, so the code has no branch between the loads (store_staged skips
  // that row).
  const uint64_t vs = u.vec < a.ffor.n_vecs ? u.vec : 0;
  AlpStagePolicy<T, decltype(hook) &> pol{a.ffor.packed + a.ffor.vector_offsets[vs] + u.lane,
                                          lds + vv * Sh::CHUNK + u.lane, alp_consts(a, vs), hook};
  if (!dispatch_bw<UT, S>(pol, a.ffor.bit_widths[vs], u.part)) {
    for (int jj = 0; jj < Sh::NJ; ++jj) lds[vv * Sh::CHUNK + u.lane + jj * L::N] = T(0);
  }
  if (!entries) hook();
  if constexpr (PATCH) {
    wave_lds_sync();
    const unsigned lo = unsigned(L::N * j0), hi = unsigned(L::N * (j0 + Sh::NJ));
#pragma unroll
    for (int v2 = 0; v2 < L::VPW; ++v2) {
#pragma unroll
      for (int p = 0; p < PF; ++p) {
        const unsigned pos = lp.pos[v2][p] & 1023u;
        if (lp.entry_pos(v2, p) != 0xFFFFu && pos >= lo && pos < hi)
          lds[v2 * Sh::CHUNK + (pos - lo)] = lp.val[v2][p];
      }
      for (unsigned k = unsigned(t + PF * WAVE); k < lp.count[v2]; k += WAVE) {
        const unsigned pos = a.positions[lp.eo[v2] + k] & 1023u;
        if (pos >= lo && pos < hi) lds[v2 * Sh::CHUNK + (pos - lo)] = a.exceptions[lp.eo[v2] + k];
      }
    }
  }
  wave_lds_sync();
  store_staged<T, S>(lds, out, wave / S, u.part, a.ffor.n_vecs);
}

template <typename T, int S, int LAYOUT, bool PATCH>
__global__ __launch_bounds__(BLOCK) void k_alp_decompress_staged(AlpArgs<T> a, T *__restrict out) {
  using UT = typename uint_of<T>::type;
  __shared__ __align__(16) T s_stage[BLOCK / WAVE][lanes<UT>::VPW * StageShape<UT, S>::CHUNK];
  alp_staged_unit<T, S, LAYOUT, PATCH>(a, out, (uint64_t(blockIdx.x) * BLOCK + threadIdx.x) / WAVE,
                                       s_stage[threadIdx.x / WAVE]);
}

template <typename T, int S, bool NT, int LAYOUT, bool PATCH>
__global__ __launch_bounds__(BLOCK) void k_alp_decompress(AlpArgs<T> a, T *__restrict out) {
  alp_decompress_unit<T, S, NT, LAYOUT, PATCH>(a, out,
                                              (uint64_t(blockIdx.x) * BLOCK + threadIdx.x) / WAVE);
}

// Slot-match mask of one (vector, lane) after patching; exceptions replace
// the decoded value at their slot. `idx` is the wave's standard-layout
// index (built for this vector's group) when LAYOUT == ALP_STD.
template <typename T, int LAYOUT, bool PATCH>
__device__ __forceinline__ uint64_t alp_lane_match(const AlpArgs<T> &a, uint64_t vec, int lane,
                                                   T target, const StdIndex<T> *idx) {
  using UT = typename uint_of<T>::type;
  using L = lanes<UT>;
  AlpMatchPolicy<T> pol{a.ffor.packed + a.ffor.vector_offsets[vec] + lane, alp_consts(a, vec),
                        target, 0};
  if (!dispatch_bw<UT, 1>(pol, a.ffor.bit_widths[vec], 0)) {
    constexpr uint64_t FULL = L::W == 64 ? ~uint64_t(0) : ((uint64_t(1) << L::W) - 1);
    pol.mask = T(0) == target ? FULL : 0;
  }
  uint64_t m = pol.mask;
  if constexpr (PATCH) {
    auto patch = [&](unsigned pos, T e) {
      const uint64_t bit = uint64_t(1) << (pos / L::N);
      m = (e == target) ? (m | bit) : (m & ~bit);
    };
    if constexpr (LAYOUT == ALP_GALP) galp_own_exceptions<T>(a, vec, lane, 0, L::W, patch);
    else idx->own(a, vec, int((threadIdx.x % WAVE) / L::N), lane, 0, L::W, patch);
  }
  return m;
}

// Smallest v in [0, n] with p(v), for p false then true over [0, n] and p(n)
// true: a gallop from the estimate e (steps 1, 2, 4, ... away from it), then
// a binary search of the bracket. At most about 2 log2(n) evaluations.
template <class P>
__device__ __forceinline__ int64_t first_true(int64_t e, int64_t n, P &&p) {
  int64_t f, t;  // p(f) false (f = -1: none), p(t) true
  if (p(e)) {
    t = e;
    f = -1;
    for (int64_t d = 1; t > 0; d *= 2) {
      const int64_t x = t - d < 0 ? 0 : t - d;
      if (!p(x)) {
        f = x;
        break;
      }
      t = x;
    }
  } else {
    f = e;
    t = n;
    for (int64_t d = 1; f + d < n; d *= 2) {
      if (p(f + d)) {
        t = f + d;
        break;
      }
      f += d;
    }
  }
  while (t - f > 1) {
    const int64_t m = f + (t - f) / 2;
    if (p(m)) t = m;
    else f = m;
  }
  return t;
}

// Largest v in [0, n] with p(v), for p true then false over [0, n] and p(0)
// true (the mirror of first_true).
template <class P>
__device__ __forceinline__ int64_t last_true(int64_t e, int64_t n, P &&p) {
  int64_t t, f;  // p(t) true, p(f) false (f = n + 1: none)
  if (p(e)) {
    t = e;
    f = n + 1;
    for (int64_t d = 1; t < n; d *= 2) {
      const int64_t x = t + d > n ? n : t + d;
      if (!p(x)) {
        f = x;
        break;
      }
      t = x;
    }
  } else {
    f = e;
    t = 0;
    for (int64_t d = 1; f - d > 0; d *= 2) {
      if (p(f - d)) {
        t = f - d;
        break;
      }
      f -= d;
    }
  }
  while (f - t > 1) {
    const int64_t m = t + (f - t) / 2;
    if (p(m)) t = m;
    else f = m;
  }
  return t;
}

// How the query matches the slots of one vector (before patching).
enum AnyHow : int {
  ANY_NONE = 0,      // no slot decodes to the needle
  ANY_INTERVAL = 1,  // exactly the offsets v with v - lo <= span (unsigned)
  ANY_DECODE = 2,    // decode and compare every slot
  ANY_ALL = 3,       // every slot (a width above W decodes to zeros; the needle is zero)
};

// The FOR offsets v in [0, 2^bw) of a vector whose value decodes exactly to
// `target`. When (base + v) * FACT cannot leave the signed range over the
// vector's offsets, the decode's integer is the exact product, increasing in
// v, and the RNE conversion and the multiply by the positive fraction keep
// the order, so those offsets are one interval. Its ends are found with the
// decode formula itself, so they agree with it bit for bit: an estimate from
// the target, then a gallop and a binary search. Otherwise ANY_DECODE.
template <typename T>
__device__ __forceinline__ int alp_target_interval(const AlpConsts<T> &c, T target, int bw,
                                                   typename uint_of<T>::type &lo,
                                                   typename uint_of<T>::type &span) {
  using UT = typename uint_of<T>::type;
  using ST = typename int_of<T>::type;
  constexpr int W = lanes<UT>::W;
  if (bw > W) return T(0) == target ? ANY_ALL : ANY_NONE;
  if (bw >= W - 1) return ANY_DECODE;
  const UT v1 = UT((UT(1) << bw) - 1);  // the largest offset
  constexpr ST ST_MAX = ST(UT(~UT(0)) >> 1);
  const ST y0 = ST(c.base);
  const ST f = ST(c.fact);
  ST p0, p1;
  if (f <= 0 || y0 > ST(ST_MAX - ST(v1)) || __builtin_mul_overflow(y0, f, &p0) ||
      __builtin_mul_overflow(ST(y0 + ST(v1)), f, &p1))
    return ANY_DECODE;
  if (!(target >= c(UT(0)) && target <= c(v1))) return ANY_NONE;  // NaN too
  const double q = double(target) / double(c.frac) / double(f) - double(y0);
  const int64_t e = q <= 0.0 ? 0 : (q >= double(v1) ? int64_t(v1) : int64_t(q));
  const int64_t first = first_true(e, int64_t(v1), [&](int64_t v) { return c(UT(v)) >= target; });
  const int64_t last = last_true(e, int64_t(v1), [&](int64_t v) { return c(UT(v)) <= target; });
  if (first > last) return ANY_NONE;
  lo = UT(first);
  span = UT(last - first);
  return ANY_INTERVAL;
}

// "Any value == magic" over a patched column. The patched column holds the
// exception values at their positions and the decoded values elsewhere, so
// the answer is: some exception value equals magic, or some decoded value at
// a position that is not an exception does.
//
// Work mapping (as for the fused filter + sum of the FFOR kernels): every
// wave takes ranges of up to 64 consecutive vectors. Lane i loads vector i's
// metadata (coalesced) and works out once which of its offsets decode to the
// needle (alp_target_interval): for the common case of a vector whose
// values cannot reach the needle there is nothing to compare, and otherwise
// one interval test per slot, not a decode. The range's exception values,
// one contiguous run when the vectors' lists follow each other, are scanned
// by the whole wave. The units of a range (VPW vectors each, one thread per
// lane) then load their packed words in batches of as many as fit about 32
// registers, the next batch's loads issued before the current one is
// tested, and take their vector's answer by shuffle. A lane with a decoded
// match reads its vector's exception positions to drop matches at exception
// slots, which is rare. A position listed twice counts each of its
// exception values.
template <typename T>
struct AnyLane {
  using UT = typename uint_of<T>::type;
  uint64_t off, eo;
  UT base, lo, span, fact;
  T frac;
  int how;
  unsigned cnt;  // the vector's exception count
};

template <typename UT, int BW>
struct AnyShape {
  static constexpr int W = lanes<UT>::W;
  static constexpr int LANE_DW = BW * W / 32;  // 32-bit words per lane
  static constexpr int C = LANE_DW <= 32 ? 1 : (LANE_DW <= 64 ? 2 : 4);  // chunks per lane
  static constexpr int NJ = W / C;                                       // slots per chunk
  // units per batch: their words within about 32 registers and their slots
  // within 64 (each slot may need a decode and compare; 128 made the 64-bit
  // kernel spill registers to scratch)
  static constexpr int MAXKB = 64 / W;
  static constexpr int FIT = LANE_DW == 0 ? MAXKB : 32 / LANE_DW;
  static constexpr int KB = C > 1 ? 1 : (FIT >= MAXKB ? MAXKB : (FIT >= 2 ? 2 : 1));
};

constexpr int ANY_DEAD_WIDTH = 255;  // a lane past the range

// This thread's lane of the vector at `off`.
template <typename T>
__device__ __forceinline__ const typename uint_of<T>::type *generic_in(const AlpArgs<T> &a,
                                                                      uint64_t off, int lane) {
  return a.ffor.packed + off + lane;
}

// Lane of the range holding this thread's vector of unit u.
template <typename UT>
__device__ __forceinline__ uint32_t any_src(uint32_t u, int vsub) {
  return u * uint32_t(lanes<UT>::VPW) + uint32_t(vsub);
}

// Match mask of slots [J0, J0 + NJ) (bit j - J0) from the lane's words, for
// the vector whose answer is held by lane `sl` (shuffled in here; every
// thread of the wave takes part).
template <typename T, int BW, int J0, int NJ>
__device__ __forceinline__ uint64_t any_mask(const typename uint_of<T>::type *w,
                                             const AnyLane<T> &l, int sl, bool live, T magic) {
  using UT = typename uint_of<T>::type;
  const int how = live ? __shfl(l.how, sl, WAVE) : int(ANY_NONE);
  const UT lo = __shfl(l.lo, sl, WAVE), span = __shfl(l.span, sl, WAVE);
  AlpConsts<T> c{__shfl(l.base, sl, WAVE), 0, 0, __shfl(l.fact, sl, WAVE), __shfl(l.frac, sl, WAVE)};
  uint64_t m = 0;
  if (how == ANY_INTERVAL) {
    emit_slot_words<UT, BW, J0, NJ>(w, [&](int j, UT v) { m |= uint64_t(UT(v - lo) <= span) << (j - J0); });
  } else if (how == ANY_DECODE) {
    emit_slot_words<UT, BW, J0, NJ>(w, [&](int j, UT v) { m |= uint64_t(c(v) == magic) << (j - J0); });
  } else if (how == ANY_ALL) {
    m = NJ == 64 ? ~uint64_t(0) : ((uint64_t(1) << NJ) - 1);
  }
  return m;
}

// Drops from m (slots [j0, j0 + nj) of the thread's lane) the slots that are
// exceptions of its vector.
template <typename T, int LAYOUT>
__device__ __forceinline__ uint64_t any_drop_exceptions(const AlpArgs<T> &a, uint64_t vec,
                                                        int lane, int j0, int nj, uint64_t m,
                                                        uint64_t eo, unsigned cnt) {
  constexpr int NL = lanes<typename uint_of<T>::type>::N;
  uint64_t first = eo;
  unsigned count = cnt;
  if constexpr (LAYOUT == ALP_GALP) {  // the lane's own list
    const unsigned oc = a.offsets_counts[vec * NL + lane];
    first = eo + (oc & 0x3FFu);
    count = oc >> 10;
  }
  for (unsigned k = 0; k < count; ++k) {
    const unsigned pos = unsigned(a.positions[first + k]) & 1023u;
    const int slot = int(pos / unsigned(NL));
    if (int(pos % unsigned(NL)) == lane && slot >= j0 && slot < j0 + nj)
      m &= ~(uint64_t(1) << (slot - j0));
  }
  return m;
}

// Tests one unit's (already loaded) chunk and folds the answer into found.
template <typename T, int LAYOUT, bool PATCH, int BW, int J0, int NJ>
__device__ __forceinline__ void any_test(const AlpArgs<T> &a, const typename uint_of<T>::type *w,
                                         const AnyLane<T> &l, uint64_t v0, uint32_t nv, uint32_t u,
                                         int vsub, int lane, T magic, bool &found) {
  using UT = typename uint_of<T>::type;
  const uint32_t src = any_src<UT>(u, vsub);
  const int sl = int(src % WAVE);
  const bool live = src < nv;
  uint64_t m = any_mask<T, BW, J0, NJ>(w, l, sl, live, magic);
  if constexpr (PATCH) {
    if (__builtin_amdgcn_ballot_w64(m != 0)) {  // rare: every thread takes part in the shuffles
      const uint64_t eo = __shfl(l.eo, sl, WAVE);
      const unsigned cnt = __shfl(l.cnt, sl, WAVE);
      if (m) m = any_drop_exceptions<T, LAYOUT>(a, v0 + src, lane, J0, NJ, m, eo, cnt);
    }
  }
  found |= m != 0;
}

// The range's units when all of its vectors have width BW (lane 0 holds the
// range's first vector).
template <typename T, int LAYOUT, bool PATCH, int BW>
__device__ __forceinline__ void any_range_uniform(const AlpArgs<T> &a, const AnyLane<T> &l,
                                                  uint64_t v0, uint32_t nv, uint32_t n_units,
                                                  int vsub, int lane, T magic, bool &found) {
  using UT = typename uint_of<T>::type;
  constexpr int W = lanes<UT>::W;
  using Sh = AnyShape<UT, BW>;
  constexpr int KB = Sh::KB;
  const uint64_t safe_off = uniform_u64(l.off);
  // A unit past the range reads the vector at safe_off (width BW) and tests
  // nothing; loads are never under a branch (see the fused filter + sum).
  auto unit_off = [&](uint32_t u) {
    const uint32_t src = any_src<UT>(u, vsub);
    const uint64_t o = __shfl(l.off, int(src % WAVE), WAVE);
    return src < nv ? o : safe_off;
  };
  if constexpr (Sh::C == 1) {
    using Words = UT[KB][SlotSpan<UT, BW, 0, W>::NR1];
    auto load = [&](Words &w, uint32_t u0) {
      static_for<KB>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        load_slot_words<UT, BW, 0, W>(generic_in(a, unit_off(u0 + k), lane), w[k]);
      });
    };
    // The batch's masks first; the rare exception-slot check after them, so
    // that its registers do not overlap the words'.
    auto test = [&](const Words &w, uint32_t u0) {
      uint64_t m[KB];
      bool any = false;
      static_for<KB>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        const uint32_t src = any_src<UT>(u0 + k, vsub);
        m[k] = any_mask<T, BW, 0, W>(w[k], l, int(src % WAVE), src < nv, magic);
        any |= m[k] != 0;
      });
      if constexpr (PATCH) {
        if (__builtin_amdgcn_ballot_w64(any)) {
          static_for<KB>([&](auto kc) {
            constexpr int k = decltype(kc)::value;
            const uint32_t src = any_src<UT>(u0 + k, vsub);
            const uint64_t eo = __shfl(l.eo, int(src % WAVE), WAVE);
            const unsigned cnt = __shfl(l.cnt, int(src % WAVE), WAVE);
            if (m[k]) m[k] = any_drop_exceptions<T, LAYOUT>(a, v0 + src, lane, 0, W, m[k], eo, cnt);
          });
        }
      }
      static_for<KB>([&](auto kc) { found |= m[decltype(kc)::value] != 0; });
    };
    const uint32_t nb = (n_units + KB - 1) / KB;
    Words w0, w1;
    load(w0, 0);
#pragma unroll 1
    for (uint32_t b = 0; b < nb; b += 2) {
      load(w1, (b + 1) * KB);
      test(w0, b * KB);
      load(w0, (b + 2) * KB);
      // a batch past the range is loaded but not waited for
      if (b + 1 < nb) test(w1, (b + 1) * KB);
    }
  } else {
#pragma unroll 1
    for (uint32_t u = 0; u < n_units; ++u) {
      const uint64_t off = unit_off(u);
      static_for<Sh::C>([&](auto cc) {
        constexpr int J0 = int(decltype(cc)::value) * Sh::NJ;
        UT w[SlotSpan<UT, BW, J0, Sh::NJ>::NR1];
        load_slot_words<UT, BW, J0, Sh::NJ>(generic_in(a, off, lane), w);
        any_test<T, LAYOUT, PATCH, BW, J0, Sh::NJ>(a, w, l, v0, nv, u, vsub, lane, magic, found);
      });
    }
  }
}

#define FLSGPU_ANY_CASE(B)                                                                  \
  case B:                                                                                  \
    if constexpr ((B) <= lanes<UT>::W)                                                      \
      any_range_uniform<T, LAYOUT, PATCH, (B)>(a, l, v0, nv, n_units, vsub, lane, magic, found); \
    break;

template <typename T, int LAYOUT, bool PATCH>
__device__ __forceinline__ void any_range_dispatch(const AlpArgs<T> &a, const AnyLane<T> &l,
                                                   uint64_t v0, uint32_t nv, uint32_t n_units,
                                                   int bw, int vsub, int lane, T magic,
                                                   bool &found) {
  using UT = typename uint_of<T>::type;
  if constexpr (lanes<UT>::W == 32) {
    switch (bw) {
      FLSGPU_CASES_0_32(FLSGPU_ANY_CASE)
      default: break;
    }
  } else {
    switch (bw) {
      FLSGPU_CASES_0_32(FLSGPU_ANY_CASE)
      FLSGPU_CASES_33_64(FLSGPU_ANY_CASE)
      default: break;
    }
  }
}
#undef FLSGPU_ANY_CASE

// One unit of width BW at `off` (a range of mixed widths), in chunks of slots.
template <typename T, int LAYOUT, bool PATCH, int BW>
__device__ __forceinline__ void any_unit(const AlpArgs<T> &a, const AnyLane<T> &l, uint64_t v0,
                                         uint32_t nv, uint32_t u, uint64_t off, int vsub,
                                         int lane, T magic, bool &found) {
  using UT = typename uint_of<T>::type;
  using Sh = AnyShape<UT, BW>;
  static_for<Sh::C>([&](auto cc) {
    constexpr int J0 = int(decltype(cc)::value) * Sh::NJ;
    UT w[SlotSpan<UT, BW, J0, Sh::NJ>::NR1];
    load_slot_words<UT, BW, J0, Sh::NJ>(generic_in(a, off, lane), w);
    any_test<T, LAYOUT, PATCH, BW, J0, Sh::NJ>(a, w, l, v0, nv, u, vsub, lane, magic, found);
  });
}

#define FLSGPU_ANY_UNIT_CASE(B)                                                              \
  case B:                                                                                   \
    if constexpr ((B) <= lanes<UT>::W)                                                       \
      any_unit<T, LAYOUT, PATCH, (B)>(a, l, v0, nv, u, off, vsub, lane, magic, found);       \
    break;

template <typename T, int LAYOUT, bool PATCH>
__device__ __forceinline__ void any_unit_dispatch(const AlpArgs<T> &a, const AnyLane<T> &l,
                                                  uint64_t v0, uint32_t nv, uint32_t u,
                                                  uint64_t off, int bw, int vsub, int lane,
                                                  T magic, bool &found) {
  using UT = typename uint_of<T>::type;
  if constexpr (lanes<UT>::W == 32) {
    switch (bw) {
      FLSGPU_CASES_0_32(FLSGPU_ANY_UNIT_CASE)
      default: {  // a width above W: zeros
        const uint32_t src = any_src<UT>(u, vsub);
        found |= src < nv && __shfl(l.how, int(src % WAVE), WAVE) == ANY_ALL;
        break;
      }
    }
  } else {
    switch (bw) {
      FLSGPU_CASES_0_32(FLSGPU_ANY_UNIT_CASE)
      FLSGPU_CASES_33_64(FLSGPU_ANY_UNIT_CASE)
      default: {
        const uint32_t src = any_src<UT>(u, vsub);
        found |= src < nv && __shfl(l.how, int(src % WAVE), WAVE) == ANY_ALL;
        break;
      }
    }
  }
}
#undef FLSGPU_ANY_UNIT_CASE

template <typename T, int LAYOUT, bool PATCH>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(3, 8)))
void k_alp_any(AlpArgs<T> a, T magic, uint32_t units, Scratch s) {
  using UT = typename uint_of<T>::type;
  using L = lanes<UT>;
  const int t = int(threadIdx.x % WAVE);
  const int lane = t % L::N, vsub = t / L::N;
  const uint64_t span = uint64_t(units) * L::VPW;  // vectors per range
  const uint64_t n_vecs = a.ffor.n_vecs;
  const uint64_t n_ranges = (n_vecs + span - 1) / span;
  const uint64_t stride = uint64_t(gridDim.x) * (BLOCK / WAVE);
  bool found = false;
  for (uint64_t rr = (uint64_t(blockIdx.x) * BLOCK + threadIdx.x) / WAVE; rr < n_ranges;
       rr += stride) {
    const uint64_t v0 = rr * span;
    const uint32_t nv = uint32_t(n_vecs - v0 < span ? n_vecs - v0 : span);
    const uint32_t n_units = (nv + L::VPW - 1) / L::VPW;
    // lane t: vector v0 + t
    AnyLane<T> l{};
    int bw = ANY_DEAD_WIDTH;
    if (uint32_t(t) < nv) {
      const uint64_t v = v0 + uint64_t(t);
      bw = a.ffor.bit_widths[v];
      l.off = a.ffor.vector_offsets[v];
      l.base = a.ffor.bases[v];
      AlpConsts<T> c{l.base, a.factor_indices[v], a.fraction_indices[v], 0, T(0)};
      if constexpr (PATCH) {
        l.eo = a.exceptions_offsets[v];
        if constexpr (LAYOUT == ALP_GALP) {  // the lanes' lists follow each other
          const unsigned oc = a.offsets_counts[v * L::N + (L::N - 1)];
          l.cnt = (oc & 0x3FFu) + (oc >> 10);
        } else {
          l.cnt = a.counts[v];
        }
      }
      c.resolve();
      l.fact = c.fact;
      l.frac = c.frac;
      l.how = alp_target_interval<T>(c, magic, bw, l.lo, l.span);
    }
    if constexpr (PATCH) {
      // the range's exception values: one run when the lists follow each other
      const uint64_t eo_next = __shfl(l.eo, (t + 1) % WAVE, WAVE);
      const bool gap = uint32_t(t + 1) < nv && l.eo + l.cnt != eo_next;
      if (__builtin_amdgcn_ballot_w64(gap) == 0) {
        const uint64_t e0 = __shfl(l.eo, 0, WAVE);
        const int last = int(nv) - 1;
        const uint64_t e1 = __shfl(l.eo, last, WAVE) + __shfl(l.cnt, last, WAVE);
        for (uint64_t k0 = e0; k0 < e1; k0 += uint64_t(WAVE) * ALP_PF) {
          T e[ALP_PF];
#pragma unroll
          for (int i = 0; i < ALP_PF; ++i) {
            const uint64_t k = k0 + uint64_t(t + i * WAVE);
            e[i] = a.exceptions[k < e1 ? k : e0];
          }
#pragma unroll
          for (int i = 0; i < ALP_PF; ++i)
            found |= k0 + uint64_t(t + i * WAVE) < e1 && e[i] == magic;
        }
      } else {
        for (uint32_t v = 0; v < nv; ++v) {
          const uint64_t eo = __shfl(l.eo, int(v), WAVE);
          const unsigned cnt = __shfl(l.cnt, int(v), WAVE);
          for (unsigned k = unsigned(t); k < cnt; k += WAVE) found |= a.exceptions[eo + k] == magic;
        }
      }
    }
    // Lane 0 holds the range's first vector.
    const int b0 = __builtin_amdgcn_readfirstlane(bw);
    if (__builtin_amdgcn_ballot_w64(bw != b0 && bw != ANY_DEAD_WIDTH) == 0) {
      any_range_dispatch<T, LAYOUT, PATCH>(a, l, v0, nv, n_units, b0, vsub, lane, magic, found);
    } else {
#pragma unroll 1
      for (uint32_t u = 0; u < n_units; ++u) {
        const uint32_t src = any_src<UT>(u, vsub);
        const int sl = int(src % WAVE);
        const bool live = src < nv;
        const int ubw = __shfl(bw, sl, WAVE);
        const uint64_t first = uniform_u64(__shfl(l.off, int(u * L::VPW), WAVE));
        const uint64_t o = __shfl(l.off, sl, WAVE);
        const uint64_t off = live ? o : first;
        // The unit's first vector (lane 0's) exists.
        const int b1 = __builtin_amdgcn_readfirstlane(ubw);
        if (__builtin_amdgcn_ballot_w64(live && ubw != b1) == 0) {
          any_unit_dispatch<T, LAYOUT, PATCH>(a, l, v0, nv, u, off, b1, vsub, lane, magic, found);
        } else {
          // divergent widths: the runtime-width path, every slot of the lane
          const int how = live ? __shfl(l.how, sl, WAVE) : int(ANY_NONE);
          const UT lo = __shfl(l.lo, sl, WAVE), sp = __shfl(l.span, sl, WAVE);
          AlpConsts<T> c{__shfl(l.base, sl, WAVE), 0, 0, __shfl(l.fact, sl, WAVE),
                         __shfl(l.frac, sl, WAVE)};
          const uint64_t eo = __shfl(l.eo, sl, WAVE);
          const unsigned cnt = __shfl(l.cnt, sl, WAVE);
          uint64_t m = 0;
          if (how == ANY_INTERVAL || how == ANY_DECODE) {
            unpack_runtime<UT>(a.ffor.packed + off + lane, ubw, 0, L::W, [&](int j, UT v) {
              const bool hit = how == ANY_INTERVAL ? UT(v - lo) <= sp : c(v) == magic;
              m |= uint64_t(hit) << j;
            });
          } else if (how == ANY_ALL) {
            m = L::W == 64 ? ~uint64_t(0) : ((uint64_t(1) << L::W) - 1);
          }
          if constexpr (PATCH) {
            if (m) m = any_drop_exceptions<T, LAYOUT>(a, v0 + src, lane, 0, L::W, m, eo, cnt);
          }
          found |= m != 0;
        }
      }
    }
  }
  block_partial2<RedOp::Or>(found, 0, s.partials);
}

template <typename T, int LAYOUT, bool PATCH>
__global__ __launch_bounds__(BLOCK) void k_alp_multi(MultiAlpArgs<T> a, T value, uint64_t n_waves,
                                                    Scratch s) {
  using UT = typename uint_of<T>::type;
  using L = lanes<UT>;
  constexpr uint64_t FULL = L::W == 64 ? ~uint64_t(0) : ((uint64_t(1) << L::W) - 1);
  bool found = false;
  const uint64_t stride = uint64_t(gridDim.x) * (BLOCK / WAVE);
  StdIndex<T> idx = wave_std_index<T, LAYOUT, PATCH>();
  for (uint64_t wave = (uint64_t(blockIdx.x) * BLOCK + threadIdx.x) / WAVE; wave < n_waves;
       wave += stride) {
    const UnitMap<UT, 1> u = map_unit<UT, 1>(wave);
    uint64_t all = FULL;
    for (uint32_t c = 0; c < a.n_cols; ++c) {
      // every thread of the wave takes part in building the index
      if constexpr (PATCH && LAYOUT == ALP_STD) idx.build(a.col[c], wave);
      if (u.vec < a.col[0].ffor.n_vecs)
        all &= alp_lane_match<T, LAYOUT, PATCH>(a.col[c], u.vec, u.lane, value, &idx);
    }
    if (u.vec < a.col[0].ffor.n_vecs) found |= all == FULL;
  }
  block_partial2<RedOp::Or>(found, 0, s.partials);
}

template <typename T, int S, bool NT, int LAYOUT, bool PATCH>
int AlpDecompress<T, S, NT, LAYOUT, PATCH>::launch(const AlpArgs<T> &a, T *out,
                                                   hipStream_t stream) {
  using UT = typename uint_of<T>::type;
  if (a.ffor.n_vecs == 0) return FLSGPU_OK;
  const uint64_t blocks = (n_waves_for<UT, S>(a.ffor.n_vecs) + (BLOCK / WAVE) - 1) / (BLOCK / WAVE);
  if (blocks > 0x7FFFFFFFull) return set_error(FLSGPU_EINVAL, "column too large");
  hipLaunchKernelGGL((k_alp_decompress<T, S, NT, LAYOUT, PATCH>), dim3(unsigned(blocks)),
                     dim3(BLOCK), 0, stream, a, out);
  return check_hip(hipGetLastError(), "k_alp_decompress launch");
}

template <typename T, int S, int LAYOUT, bool PATCH>
int AlpDecompressStaged<T, S, LAYOUT, PATCH>::launch(const AlpArgs<T> &a, T *out,
                                                     hipStream_t stream) {
  using UT = typename uint_of<T>::type;
  if (a.ffor.n_vecs == 0) return FLSGPU_OK;
  const uint64_t blocks = (n_waves_for<UT, S>(a.ffor.n_vecs) + (BLOCK / WAVE) - 1) / (BLOCK / WAVE);
  if (blocks > 0x7FFFFFFFull) return set_error(FLSGPU_EINVAL, "column too large");
  hipLaunchKernelGGL((k_alp_decompress_staged<T, S, LAYOUT, PATCH>), dim3(unsigned(blocks)),
                     dim3(BLOCK), 0, stream, a, out);
  return check_hip(hipGetLastError(), "k_alp_decompress_staged launch");
}

template <typename T, int LAYOUT, bool PATCH>
int AlpAny<T, LAYOUT, PATCH>::launch(const AlpArgs<T> &a, T magic, uint8_t *d_found,
                                     hipStream_t stream) {
  using UT = typename uint_of<T>::type;
  using L = lanes<UT>;
  Scratch s;
  int rc = get_scratch(stream, &s);
  if (rc) return rc;
  static const unsigned resident = [] {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_alp_any<T, LAYOUT, PATCH>, BLOCK,
                                                     0) != hipSuccess || per_cu < 1)
      per_cu = 1;
    return unsigned(per_cu) * unsigned(num_cus());
  }();
  // Units per range: the ranges fill the resident waves about once.
  const uint64_t n_units = n_waves_for<UT, 1>(a.ffor.n_vecs);
  const uint64_t waves = uint64_t(resident) * (BLOCK / WAVE);
  const uint32_t units =
      uint32_t(std::clamp<uint64_t>((n_units + waves - 1) / waves, 1, WAVE / L::VPW));
  const uint64_t n_ranges = (n_units + units - 1) / units;
  const uint64_t blocks = (n_ranges + (BLOCK / WAVE) - 1) / (BLOCK / WAVE);
  const unsigned grid =
      unsigned(std::min<uint64_t>({std::max<uint64_t>(blocks, 1), resident, 2048}));
  hipLaunchKernelGGL((k_alp_any<T, LAYOUT, PATCH>), dim3(grid), dim3(BLOCK), 0, stream, a, magic,
                     units, s);
  if ((rc = check_hip(hipGetLastError(), "k_alp_any launch"))) return rc;
  return reduce_finish_flag(s, grid, d_found, stream);
}

template <typename T, int LAYOUT, bool PATCH>
int AlpMulti<T, LAYOUT, PATCH>::launch(const MultiAlpArgs<T> &a, T value, uint8_t *d_found,
                                       hipStream_t stream) {
  using UT = typename uint_of<T>::type;
  Scratch s;
  int rc = get_scratch(stream, &s);
  if (rc) return rc;
  const uint64_t n_waves = n_waves_for<UT, 1>(a.col[0].ffor.n_vecs);
  const unsigned grid = reduce_grid(n_waves);
  hipLaunchKernelGGL((k_alp_multi<T, LAYOUT, PATCH>), dim3(grid), dim3(BLOCK), 0, stream, a,
                     value, n_waves, s);
  if ((rc = check_hip(hipGetLastError(), "k_alp_multi launch"))) return rc;
  return reduce_finish_flag(s, grid, d_found, stream);
}

}  // namespace flsgpu_hip
