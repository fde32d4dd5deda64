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

// "Any value == magic" over a patched column. The patched column holds the
// exception values at their positions and the decoded values elsewhere, so
// the answer is: some exception value equals magic, or some decoded value
// at a position that is not an exception does. The first part reads the
// exception values only (the vector's threads together, coalesced); the
// positions are read only by a lane with a decoded match, to drop matches at
// exception positions, which is rare. (Building the standard layout's LDS
// index for every vector, as the full patch does, made the query slower than
// the decode: 2^26 doubles, 91 us.) A position listed twice counts each of
// its exception values.
// Threads per lane of the query: two for doubles (the 64-slot decode and
// compare otherwise needs ~184 VGPRs), one for floats.
template <typename T>
constexpr int alp_any_split() { return sizeof(T) == 8 ? 2 : 1; }

template <typename T, int LAYOUT, bool PATCH>
__global__ __launch_bounds__(BLOCK) void k_alp_any(AlpArgs<T> a, T magic, uint64_t n_waves,
                                                  Scratch s) {
  using UT = typename uint_of<T>::type;
  using L = lanes<UT>;
  constexpr int S = alp_any_split<T>();
  constexpr int NJ = L::W / S;
  constexpr uint64_t FULL = NJ == 64 ? ~uint64_t(0) : ((uint64_t(1) << NJ) - 1);
  bool found = false;
  const uint64_t stride = uint64_t(gridDim.x) * (BLOCK / WAVE);
  for (uint64_t wave = (uint64_t(blockIdx.x) * BLOCK + threadIdx.x) / WAVE; wave < n_waves;
       wave += stride) {
    const UnitMap<UT, S> u = map_unit<UT, S>(wave);
    if (u.vec >= a.ffor.n_vecs) continue;
    const int j0 = u.part * NJ;
    AlpMatchPolicy<T> pol{a.ffor.packed + a.ffor.vector_offsets[u.vec] + u.lane,
                          alp_consts(a, u.vec), magic, 0};
    if (!dispatch_bw<UT, S>(pol, a.ffor.bit_widths[u.vec], u.part))
      pol.mask = T(0) == magic ? FULL : 0;
    uint64_t m = pol.mask;  // bit (j - j0): slot j decodes to magic
    if constexpr (PATCH) {
      const uint64_t eo = a.exceptions_offsets[u.vec];
      unsigned count, first, k0, step;
      if constexpr (LAYOUT == ALP_GALP) {
        const unsigned oc = a.offsets_counts[u.vec * L::N + u.lane];
        count = oc >> 10;
        first = oc & 0x3FFu;
        k0 = 0;  // the lane's own list
        step = 1;
      } else {
        count = a.counts[u.vec];
        first = 0;
        k0 = unsigned(u.lane);  // the vector's list, shared by its lanes
        step = unsigned(L::N);
      }
      if (u.part == 0) {  // the exception values, ALP_PF loads in flight
        for (unsigned k = k0; k < count; k += ALP_PF * step) {
          T e[ALP_PF];
#pragma unroll
          for (int i = 0; i < ALP_PF; ++i)
            e[i] = k + i * step < count ? a.exceptions[eo + first + k + i * step] : T(0);
#pragma unroll
          for (int i = 0; i < ALP_PF; ++i) found |= k + i * step < count && e[i] == magic;
        }
      }
      if (m) {
        for (unsigned k = 0; k < count; ++k) {
          const unsigned pos = unsigned(a.positions[eo + first + k]) & 1023u;
          const int slot = int(pos / unsigned(L::N));
          if (int(pos % unsigned(L::N)) == u.lane && slot >= j0 && slot < j0 + NJ)
            m &= ~(uint64_t(1) << (slot - j0));
        }
      }
    }
    found |= m != 0;
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
  Scratch s;
  int rc = get_scratch(stream, &s);
  if (rc) return rc;
  const uint64_t n_waves = n_waves_for<UT, alp_any_split<T>()>(a.ffor.n_vecs);
  // a CU-filling grid whose waves walk several units each (2^26 doubles:
  // 56 us at 2048 blocks, 61 us with one unit per wave)
  const unsigned grid = std::min(reduce_grid(n_waves), 2048u);
  hipLaunchKernelGGL((k_alp_any<T, LAYOUT, PATCH>), dim3(grid), dim3(BLOCK), 0, stream, a, magic,
                     n_waves, s);
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
