// Host-side ALP / G-ALP / ALP-RD / FFOR encoder (multi-threaded C++), the
// producer of the columns the HIP decoders read. Behaviour follows the
// reference encoder so that its columns are byte-identical:
//   sampler            src/alp/sampler.hpp:11-48
//   encode_value       src/alp/encoder.hpp:76-84
//   decode_value       src/alp/decoder.hpp:121-124
//   top-k search       src/alp/encoder.hpp:134-230 (incl. the ALP_RD switch :208-211)
//   per-vector choice  src/alp/encoder.hpp:236-300
//   encode_simdized    src/alp/encoder.hpp:302-379
//   analyze_ffor       src/alp/encoder.hpp:104-115
//   column assembly    src/alp/alp-bindings.cu:83-206 (sizes :19-43)
//   G-ALP layout       src/flsgpu/structs.cuh:231-306
//   ALP-RD             src/alp/rd.hpp:12-180 (dictionary ties: ascending left value)
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <thread>
#include <type_traits>
#include <vector>

#include <sched.h>

#include "flsgpu.h"
#include "flsgpu_host.h"

namespace {

int set_err(int code, const char *msg);

template <typename T>
struct Alp;

template <>
struct Alp<float> {
  using UT = uint32_t;
  using ST = int32_t;
  static constexpr int MAX_EXP = 10;
  // Factor indices 0..9 only: the reference's float factor table has ten
  // entries (src/alp/constants.hpp:55), so its search reading index 10 at
  // exp_ref = factor_idx = 10 (encoder.hpp:153-154) is out of bounds. That
  // one pair is skipped here.
  static constexpr int MAX_FAC = 9;
  static constexpr uint64_t EXC_BITS = 32;
  static constexpr uint64_t RD_LIMIT = 22 * 32;
  static constexpr float MAGIC = 12582912.0f;
  static constexpr UT SIGN_MASK = 0x7FFFFFFFu, EXP_MASK = 0x7F800000u, NEG_ZERO = 0x80000000u;
  static constexpr float FRAC[11] = {1.0,      0.1,       0.01,       0.001,       0.0001,      0.00001,
                                     0.000001, 0.0000001, 0.00000001, 0.000000001, 0.0000000001};
  static constexpr float EXP10[11] = {1.0,       10.0,       100.0,        1000.0,
                                      10000.0,   100000.0,   1000000.0,    10000000.0,
                                      100000000.0, 1000000000.0, 10000000000.0};
  static constexpr int32_t FACT[10] = {1, 10, 100, 1000, 10000, 100000, 1000000, 10000000,
                                       100000000, 1000000000};
  // encode_value's "impossible" path returns ENCODING_UPPER_LIMIT converted
  // to int32_t, an out-of-range conversion that folds to INT32_MAX.
  static constexpr ST IMPOSSIBLE = std::numeric_limits<int32_t>::max();
};

template <>
struct Alp<double> {
  using UT = uint64_t;
  using ST = int64_t;
  static constexpr int MAX_EXP = 18;
  static constexpr int MAX_FAC = 18;
  static constexpr uint64_t EXC_BITS = 64;
  static constexpr uint64_t RD_LIMIT = 48 * 32;
  static constexpr double MAGIC = 6755399441055744.0;
  static constexpr UT SIGN_MASK = 0x7FFFFFFFFFFFFFFFull, EXP_MASK = 0x7FF0000000000000ull,
                      NEG_ZERO = 0x8000000000000000ull;
  static constexpr double FRAC[21] = {1.0,     1e-1,  1e-2,  1e-3,  1e-4,  1e-5,  1e-6,
                                      1e-7,    1e-8,  1e-9,  1e-10, 1e-11, 1e-12, 1e-13,
                                      1e-14,   1e-15, 1e-16, 1e-17, 1e-18, 1e-19, 1e-20};
  static constexpr double EXP10[24] = {1e0,  1e1,  1e2,  1e3,  1e4,  1e5,  1e6,  1e7,
                                       1e8,  1e9,  1e10, 1e11, 1e12, 1e13, 1e14, 1e15,
                                       1e16, 1e17, 1e18, 1e19, 1e20, 1e21, 1e22, 1e23};
  static constexpr int64_t FACT[19] = {1LL,
                                       10LL,
                                       100LL,
                                       1000LL,
                                       10000LL,
                                       100000LL,
                                       1000000LL,
                                       10000000LL,
                                       100000000LL,
                                       1000000000LL,
                                       10000000000LL,
                                       100000000000LL,
                                       1000000000000LL,
                                       10000000000000LL,
                                       100000000000000LL,
                                       1000000000000000LL,
                                       10000000000000000LL,
                                       100000000000000000LL,
                                       1000000000000000000LL};
  static constexpr ST IMPOSSIBLE = 9223372036854774784LL;
};

constexpr double ENC_UPPER = 9223372036854774784.0;
constexpr double ENC_LOWER = -9223372036854774784.0;
constexpr int POS_BITS = 16;
constexpr size_t ROWGROUP = 102400;
constexpr size_t JUMP = 12;  // (ROWGROUP / 8) / 1024
constexpr int MAX_K = 5;

template <typename T>
bool impossible(T n) {
  return !std::isfinite(n) || std::isnan(n) || double(n) > ENC_UPPER || double(n) < ENC_LOWER ||
         (n == T(0) && std::signbit(n));
}

template <typename T>
typename Alp<T>::ST encode_value(T v, int fac, int exp, bool safe) {
  using A = Alp<T>;
  T t = v * A::EXP10[exp] * A::FRAC[fac];
  if (safe && impossible(t)) return A::IMPOSSIBLE;
  t = t + A::MAGIC - A::MAGIC;
  return static_cast<typename A::ST>(t);
}

template <typename T>
T decode_value(typename Alp<T>::ST e, int fac, int exp) {
  using A = Alp<T>;
  using UT = typename A::UT;
  const auto s = static_cast<typename A::ST>(UT(e) * UT(A::FACT[fac]));
  return T(s) * A::FRAC[exp];
}

template <typename UT>
unsigned bit_count(UT x) {
  unsigned n = 0;
  while (x) {
    ++n;
    x >>= 1;
  }
  return n;
}

template <typename T>
size_t first_level_sample(const T *data, size_t n, T *out) {
  const size_t portion = std::min(ROWGROUP, n);
  const size_t avail = size_t(std::ceil(double(portion) / 1024.0));
  size_t k = 0, at = 0;
  for (size_t vi = 0; vi < avail; ++vi) {
    const size_t cur = std::min<size_t>(n - at, 1024);
    if (vi % JUMP != 0) {
      at += cur;
      continue;
    }
    const size_t step = size_t(std::max(1, int32_t(std::ceil(double(cur) / 32.0))));
    if (cur < 32 && k != 0) {
      at += cur;
      continue;
    }
    for (size_t i = 0; i < cur; i += step) out[k++] = data[at + i];
    at += cur;
  }
  return k;
}

// One (exp, fac) candidate over n samples: (estimated bits, valid).
template <typename T>
uint64_t estimate(const T *smp, size_t n, int e, int f, unsigned *n_exc, unsigned *n_ok) {
  using A = Alp<T>;
  using ST = typename A::ST;
  using UT = typename A::UT;
  ST mx = std::numeric_limits<ST>::min(), mn = std::numeric_limits<ST>::max();
  unsigned exc = 0, ok = 0;
  for (size_t i = 0; i < n; ++i) {
    const ST en = encode_value<T>(smp[i], f, e, true);
    if (decode_value<T>(en, f, e) == smp[i]) {
      ++ok;
      mx = std::max(mx, en);
      mn = std::min(mn, en);
    } else {
      ++exc;
    }
  }
  *n_exc = exc;
  *n_ok = ok;
  return bit_count<UT>(UT(UT(mx) - UT(mn)));
}

// Exhaustive search on one sample vector (the inner loops of
// find_top_k_combinations). Returns the estimated size; writes the choice.
template <typename T>
uint64_t search_exhaustive(const T *smp, size_t n, int *out_e, int *out_f) {
  using A = Alp<T>;
  const uint64_t ss = n;
  uint64_t best = ss * (A::EXC_BITS + POS_BITS) + ss * A::EXC_BITS;
  int fe = 0, ff = 0;
  for (int e = A::MAX_EXP; e >= 0; --e) {
    for (int f = std::min(e, A::MAX_FAC); f >= 0; --f) {
      unsigned exc, ok;
      const uint64_t bits = estimate<T>(smp, n, e, f, &exc, &ok);
      if (ok < 2) continue;
      const uint64_t est = ss * bits + uint64_t(exc) * (A::EXC_BITS + POS_BITS);
      if (est < best || (est == best && fe < e) || (est == best && fe == e && ff < f)) {
        best = est;
        fe = e;
        ff = f;
      }
    }
  }
  *out_e = fe;
  *out_f = ff;
  return best;
}

// Returns k (0 => ALP_RD) and the ordered top-k (exp, fac) pairs.
template <typename T>
int top_k(const T *sample, size_t n, int *exps, int *facs) {
  using A = Alp<T>;
  const uint64_t n_vec = uint64_t(std::ceil(T(n) / T(32)));
  const size_t ss = std::min<size_t>(n, 32);
  const uint64_t worst = ss * (A::EXC_BITS + POS_BITS) + ss * A::EXC_BITS;
  uint64_t best = worst;
  int counts[A::MAX_EXP + 1][A::MAX_EXP + 1] = {};
  for (uint64_t s = 0; s < n_vec; ++s) {
    int e, f;
    const uint64_t est = search_exhaustive<T>(sample + s * ss, ss, &e, &f);
    best = std::min(best, est);
    counts[e][f]++;
  }
  if (best >= A::RD_LIMIT) return 0;
  struct C {
    int count, e, f;
  };
  std::vector<C> cs;
  for (int e = 0; e <= A::MAX_EXP; ++e)
    for (int f = 0; f <= A::MAX_EXP; ++f)
      if (counts[e][f]) cs.push_back({counts[e][f], e, f});
  std::sort(cs.begin(), cs.end(), [](const C &a, const C &b) {
    if (a.count != b.count) return a.count > b.count;
    if (a.e != b.e) return a.e > b.e;
    return a.f > b.f;
  });
  const int k = std::min<int>(MAX_K, int(cs.size()));
  for (int i = 0; i < k; ++i) {
    exps[i] = cs[i].e;
    facs[i] = cs[i].f;
  }
  return k;
}

template <typename T>
void best_combination(const int *exps, const int *facs, int k, const T *vec, int *oe, int *of) {
  using A = Alp<T>;
  T smp[32];
  for (int i = 0; i < 32; ++i) smp[i] = vec[i * 32];
  uint64_t best = 0;
  int be = exps[0], bf = facs[0], worse = 0;
  for (int c = 0; c < k; ++c) {
    unsigned exc, ok;
    const uint64_t bits = estimate<T>(smp, 32, exps[c], facs[c], &exc, &ok);
    const uint64_t est = 32 * bits + uint64_t(exc) * (A::EXC_BITS + POS_BITS);
    if (c == 0) {
      best = est;
      continue;
    }
    if (est >= best) {
      if (++worse == 2) break;
      continue;
    }
    best = est;
    be = exps[c];
    bf = facs[c];
    worse = 0;
  }
  *oe = be;
  *of = bf;
}

// encode_simdized: returns the exception count.
template <typename T>
unsigned encode_vector(const T *vec, int fac, int exp, typename Alp<T>::ST *enc, T *exc,
                       uint16_t *pos) {
  using A = Alp<T>;
  using UT = typename A::UT;
  using ST = typename A::ST;
  uint16_t idx[1024];
  unsigned n_exc = 0;
  for (unsigned i = 0; i < 1024; ++i) {
    UT bits;
    std::memcpy(&bits, &vec[i], sizeof(bits));
    const bool special = (bits & A::SIGN_MASK) >= A::EXP_MASK || bits == A::NEG_ZERO;
    const T x = special ? T(ENC_UPPER) : vec[i];
    enc[i] = encode_value<T>(x, fac, exp, false);
    const bool is_exc = decode_value<T>(enc[i], fac, exp) != x;
    idx[n_exc] = uint16_t(i);
    n_exc += is_exc;
  }
  // Fill value: encoded value of the first slot i whose index entry differs
  // from i (encoder.hpp:352-358); the entry at n_exc was last written with
  // 1023 unless slot 1023 is itself an exception.
  ST fill = 0;
  for (unsigned i = 0; i < 1024; ++i) {
    const unsigned entry = i < n_exc ? idx[i] : (i == n_exc ? 1023u : i);
    if (entry != i) {
      fill = enc[i];
      break;
    }
  }
  for (unsigned k = 0; k < n_exc; ++k) {
    enc[idx[k]] = fill;
    exc[k] = vec[idx[k]];
    pos[k] = idx[k];
  }
  return n_exc;
}

template <typename UT>
void fls_pack(const UT *in, UT *out, unsigned bw) {
  constexpr unsigned W = sizeof(UT) * 8, NL = 1024 / W;
  std::memset(out, 0, sizeof(UT) * NL * bw);
  if (bw == 0) return;
  const UT mask = bw >= W ? UT(~UT(0)) : UT((UT(1) << bw) - 1);
  for (unsigned j = 0; j < W; ++j) {
    const unsigned p = j * bw, r = p / W, s = p % W;
    for (unsigned l = 0; l < NL; ++l) {
      const UT v = in[l + NL * j] & mask;
      out[l + NL * r] |= UT(v << s);
      if (s + bw > W) out[l + NL * (r + 1)] |= UT(v >> (W - s));
    }
  }
}

// Threads for n_threads <= 0: the CPUs this process may run on, capped by
// OMP_NUM_THREADS when that is set (the usual way a host limits a job).
int default_threads() {
  int n = int(std::max(1u, std::thread::hardware_concurrency()));
  cpu_set_t set;
  if (sched_getaffinity(0, sizeof(set), &set) == 0) n = std::max(1, CPU_COUNT(&set));
  if (const char *e = std::getenv("OMP_NUM_THREADS")) {
    const int m = std::atoi(e);
    if (m > 0) n = std::min(n, m);
  }
  return n;
}

template <class F>
void parallel_for(size_t n, int n_threads, F &&f) {
  if (n_threads <= 0) n_threads = default_threads();
  n_threads = int(std::min<size_t>(size_t(n_threads), std::max<size_t>(n, 1)));
  if (n_threads <= 1) {
    f(size_t(0), n);
    return;
  }
  std::vector<std::thread> th;
  for (int t = 0; t < n_threads; ++t) {
    const size_t b = n * size_t(t) / size_t(n_threads), e = n * size_t(t + 1) / size_t(n_threads);
    th.emplace_back([&, b, e] { f(b, e); });
  }
  for (auto &x : th) x.join();
}

// Encodes each vector with the given (exp, fac): per-vector encode_simdized,
// analyze_ffor and FFOR packing, then concatenation in vector order.
template <typename T>
int encode_with_params(const T *in, uint64_t n_vecs, const uint8_t *exps, const uint8_t *facs,
                       flsgpu_alp_host_column *col, int n_threads) {
  using A = Alp<T>;
  using UT = typename A::UT;
  using ST = typename A::ST;
  constexpr unsigned NL = 1024 / (sizeof(T) * 8);
  std::vector<uint16_t> cnt(n_vecs);
  std::vector<uint8_t> bws(n_vecs);
  std::vector<UT> bases(n_vecs);
  std::vector<UT> packed_tmp(n_vecs * 1024);
  std::vector<T> exc_tmp(n_vecs * 1024);
  std::vector<uint16_t> pos_tmp(n_vecs * 1024);
  parallel_for(n_vecs, n_threads, [&](size_t b, size_t e) {
    std::vector<ST> enc(1024);
    for (size_t v = b; v < e; ++v) {
      cnt[v] = uint16_t(encode_vector<T>(in + v * 1024, facs[v], exps[v], enc.data(),
                                         exc_tmp.data() + v * 1024, pos_tmp.data() + v * 1024));
      ST mn = std::numeric_limits<ST>::max(), mx = std::numeric_limits<ST>::min();
      for (int i = 0; i < 1024; ++i) {
        mn = std::min(mn, enc[i]);
        mx = std::max(mx, enc[i]);
      }
      bws[v] = uint8_t(bit_count<UT>(UT(UT(mx) - UT(mn))));
      bases[v] = UT(mn);
      std::vector<UT> d(1024);
      for (int i = 0; i < 1024; ++i) d[i] = UT(UT(enc[i]) - UT(mn));
      fls_pack<UT>(d.data(), packed_tmp.data() + v * 1024, bws[v]);
    }
  });
  UT *packed = static_cast<UT *>(col->packed);
  T *exc = static_cast<T *>(col->exceptions);
  UT *out_bases = static_cast<UT *>(col->bases);
  uint64_t w = 0, x = 0, vec_bytes = 0;
  for (size_t v = 0; v < n_vecs; ++v) {
    const uint64_t nw = uint64_t(bws[v]) * NL;
    std::memcpy(packed + w, packed_tmp.data() + v * 1024, sizeof(UT) * nw);
    std::memcpy(exc + x, exc_tmp.data() + v * 1024, sizeof(T) * cnt[v]);
    std::memcpy(col->positions + x, pos_tmp.data() + v * 1024, sizeof(uint16_t) * cnt[v]);
    col->bit_widths[v] = bws[v];
    col->vector_offsets[v] = w;
    out_bases[v] = bases[v];
    col->factor_indices[v] = facs[v];
    col->fraction_indices[v] = exps[v];
    col->exceptions_offsets[v] = x;
    col->counts[v] = cnt[v];
    w += nw;
    x += cnt[v];
    vec_bytes += nw * sizeof(T) + uint64_t(cnt[v]) * (sizeof(T) + 2);
  }
  col->n_values = n_vecs * 1024;
  col->n_vecs = n_vecs;
  col->n_packed_words = w;
  col->n_exceptions = x;
  col->compressed_bytes_alp = vec_bytes + n_vecs * (3 + sizeof(T));
  col->compressed_bytes_galp = vec_bytes + n_vecs * (3 + sizeof(T) + 2 * NL);
  return FLSGPU_OK;
}

template <typename T>
int alp_encode(const T *in, uint64_t n_values, flsgpu_alp_host_column *col, int n_threads) {
  if (!in || !col) return set_err(FLSGPU_EINVAL, "null argument");
  if (n_values % 1024) return set_err(FLSGPU_EINVAL, "n_values must be a multiple of 1024");
  const uint64_t n_vecs = n_values / 1024;
  std::vector<T> sample(std::min<size_t>(n_values, ROWGROUP) + 64);
  const size_t ns = first_level_sample<T>(in, n_values, sample.data());
  int ke[MAX_K], kf[MAX_K];
  const int k = top_k<T>(sample.data(), ns, ke, kf);
  if (k == 0) return FLSGPU_SCHEME_ALPRD;
  std::vector<uint8_t> exps(n_vecs), facs(n_vecs);
  parallel_for(n_vecs, n_threads, [&](size_t b, size_t e) {
    for (size_t v = b; v < e; ++v) {
      int be = ke[0], bf = kf[0];
      if (k > 1) best_combination<T>(ke, kf, k, in + v * 1024, &be, &bf);
      exps[v] = uint8_t(be);
      facs[v] = uint8_t(bf);
    }
  });
  return encode_with_params<T>(in, n_vecs, exps.data(), facs.data(), col, n_threads);
}

template <typename T>
int galp_convert(uint64_t n_vecs, const T *exc, const uint16_t *pos, const uint16_t *counts,
                 const uint64_t *offsets, T *out_exc, uint16_t *out_pos, uint16_t *out_oc) {
  constexpr unsigned NL = 1024 / (sizeof(T) * 8);
  for (uint64_t v = 0; v < n_vecs; ++v) {
    const uint64_t o = offsets[v];
    const unsigned c = counts[v];
    unsigned lane_count[NL] = {};
    for (unsigned k = 0; k < c; ++k) lane_count[pos[o + k] % NL]++;
    unsigned start[NL];
    unsigned acc = 0;
    for (unsigned l = 0; l < NL; ++l) {
      if (lane_count[l] > 63)
        return set_err(FLSGPU_EINVAL,
                       "a lane holds more than 63 exceptions: the 16-bit G-ALP header "
                       "(count << 10 | offset) cannot represent it");
      start[l] = acc;
      out_oc[v * NL + l] = uint16_t((lane_count[l] << 10) | acc);
      acc += lane_count[l];
    }
    for (unsigned k = 0; k < c; ++k) {  // stable partition by lane
      const unsigned l = pos[o + k] % NL;
      out_exc[o + start[l]] = exc[o + k];
      out_pos[o + start[l]] = pos[o + k];
      ++start[l];
    }
  }
  return FLSGPU_OK;
}

// ---- ALP-RD ----------------------------------------------------------------
template <typename T>
struct RdChoice {
  unsigned right_bw, left_bw, dict_size;
  uint16_t dict[8];
};

template <typename T>
RdChoice<T> rd_dictionary(const T *sample, size_t n) {
  using UT = typename Alp<T>::UT;
  constexpr unsigned W = sizeof(T) * 8;
  RdChoice<T> best{};
  double best_est = std::numeric_limits<double>::max();
  std::vector<uint32_t> lefts(n);
  auto build = [&](unsigned rbw, RdChoice<T> *out) {
    for (size_t i = 0; i < n; ++i) {
      UT b;
      std::memcpy(&b, &sample[i], sizeof(b));
      lefts[i] = uint32_t(b >> rbw);
    }
    std::sort(lefts.begin(), lefts.end());
    std::vector<std::pair<int, uint32_t>> runs;  // (count, value)
    for (size_t i = 0; i < n;) {
      size_t j = i;
      while (j < n && lefts[j] == lefts[i]) ++j;
      runs.push_back({int(j - i), lefts[i]});
      i = j;
    }
    std::stable_sort(runs.begin(), runs.end(),
                     [](const std::pair<int, uint32_t> &a, const std::pair<int, uint32_t> &b) {
                       return a.first > b.first;
                     });
    uint32_t exc = 0;
    for (size_t i = 8; i < runs.size(); ++i) exc += uint32_t(runs[i].first);
    const unsigned ds = unsigned(std::min<size_t>(8, runs.size()));
    const unsigned lbw = std::max(1u, unsigned(std::ceil(std::log2(double(ds)))));
    if (out) {
      out->right_bw = rbw;
      out->left_bw = lbw;
      out->dict_size = ds;
      for (unsigned i = 0; i < ds; ++i) out->dict[i] = uint16_t(runs[i].second);
    }
    return double(rbw + lbw) + double(uint32_t(exc * 32u)) / double(n);
  };
  unsigned best_rbw = W - 1;
  for (unsigned i = 1; i <= 16; ++i) {
    const double est = build(W - i, nullptr);
    if (est < best_est) {
      best_est = est;
      best_rbw = W - i;
    }
  }
  build(best_rbw, &best);
  return best;
}

template <typename T>
int alprd_encode(const T *in, uint64_t n_values, flsgpu_alprd_host_column *col, int n_threads) {
  using UT = typename Alp<T>::UT;
  constexpr unsigned W = sizeof(T) * 8, NL = 1024 / W;
  if (!in || !col) return set_err(FLSGPU_EINVAL, "null argument");
  if (n_values % 1024) return set_err(FLSGPU_EINVAL, "n_values must be a multiple of 1024");
  const uint64_t n_vecs = n_values / 1024;
  std::vector<T> sample(std::min<size_t>(n_values, ROWGROUP) + 64);
  const size_t ns = first_level_sample<T>(in, n_values, sample.data());
  if (ns == 0) return set_err(FLSGPU_EINVAL, "empty column");
  const RdChoice<T> ch = rd_dictionary<T>(sample.data(), ns);
  col->right_bit_width = ch.right_bw;
  col->left_bit_width = ch.left_bw;
  col->dictionary_size = ch.dict_size;
  for (unsigned i = 0; i < 8; ++i) col->dictionary[i] = i < ch.dict_size ? ch.dict[i] : 0;
  std::vector<uint16_t> cnt(n_vecs);
  std::vector<uint16_t> exc_tmp(n_vecs * 1024), pos_tmp(n_vecs * 1024);
  UT *right = static_cast<UT *>(col->right_packed);
  const UT rmask = UT((UT(1) << ch.right_bw) - 1);
  parallel_for(n_vecs, n_threads, [&](size_t b, size_t e) {
    std::vector<UT> r(1024);
    std::vector<uint16_t> l(1024);
    for (size_t v = b; v < e; ++v) {
      unsigned c = 0;
      for (unsigned i = 0; i < 1024; ++i) {
        UT bits;
        std::memcpy(&bits, &in[v * 1024 + i], sizeof(bits));
        r[i] = bits & rmask;
        const uint16_t lp = uint16_t(bits >> ch.right_bw);
        unsigned d = 0;
        while (d < ch.dict_size && ch.dict[d] != lp) ++d;
        if (d < ch.dict_size) {
          l[i] = uint16_t(d);
        } else {
          l[i] = 0;
          exc_tmp[v * 1024 + c] = lp;
          pos_tmp[v * 1024 + c] = uint16_t(i);
          ++c;
        }
      }
      cnt[v] = uint16_t(c);
      fls_pack<UT>(r.data(), right + v * size_t(ch.right_bw) * NL, ch.right_bw);
      fls_pack<uint16_t>(l.data(), col->left_packed + v * size_t(ch.left_bw) * 64, ch.left_bw);
    }
  });
  uint64_t x = 0;
  for (uint64_t v = 0; v < n_vecs; ++v) {
    std::memcpy(col->exceptions + x, exc_tmp.data() + v * 1024, 2 * cnt[v]);
    std::memcpy(col->positions + x, pos_tmp.data() + v * 1024, 2 * cnt[v]);
    col->exceptions_offsets[v] = x;
    col->counts[v] = cnt[v];
    x += cnt[v];
  }
  col->n_values = n_values;
  col->n_vecs = n_vecs;
  col->n_exceptions = x;
  col->compressed_bytes = n_vecs * (size_t(ch.right_bw) * NL * sizeof(UT) + size_t(ch.left_bw) * 128 +
                                    2 /* count */) +
                          x * 4 + 8 * 2 /* dictionary */;
  return FLSGPU_OK;
}

thread_local const char *g_msg = "";
int set_err(int code, const char *msg) {
  g_msg = msg;
  return code;
}

}  // namespace

extern "C" {

const char *flsgpu_encoder_last_error(void) { return g_msg; }

int flsgpu_alp_encode_f32(const float *in, uint64_t n_values, flsgpu_alp_host_column *col,
                          int n_threads) {
  return alp_encode<float>(in, n_values, col, n_threads);
}
int flsgpu_alp_encode_f64(const double *in, uint64_t n_values, flsgpu_alp_host_column *col,
                          int n_threads) {
  return alp_encode<double>(in, n_values, col, n_threads);
}
int flsgpu_alp_encode_with_params_f32(const float *in, uint64_t n_values, const uint8_t *exps,
                                      const uint8_t *facs, flsgpu_alp_host_column *col,
                                      int n_threads) {
  if (!in || !col || !exps || !facs || n_values % 1024)
    return set_err(FLSGPU_EINVAL, "bad argument");
  for (uint64_t v = 0; v < n_values / 1024; ++v)
    if (exps[v] > Alp<float>::MAX_EXP || facs[v] > exps[v] || facs[v] > Alp<float>::MAX_FAC)
      return set_err(FLSGPU_EINVAL, "exponent/factor out of range");
  return encode_with_params<float>(in, n_values / 1024, exps, facs, col, n_threads);
}
int flsgpu_alp_encode_with_params_f64(const double *in, uint64_t n_values, const uint8_t *exps,
                                      const uint8_t *facs, flsgpu_alp_host_column *col,
                                      int n_threads) {
  if (!in || !col || !exps || !facs || n_values % 1024)
    return set_err(FLSGPU_EINVAL, "bad argument");
  for (uint64_t v = 0; v < n_values / 1024; ++v)
    if (exps[v] > Alp<double>::MAX_EXP || facs[v] > exps[v])
      return set_err(FLSGPU_EINVAL, "exponent/factor out of range");
  return encode_with_params<double>(in, n_values / 1024, exps, facs, col, n_threads);
}

int flsgpu_alp_search_f32(const float *samples, uint64_t n_vecs, uint8_t *exps, uint8_t *facs,
                          uint64_t *est_bits, int n_threads) {
  if (!samples || !exps || !facs) return set_err(FLSGPU_EINVAL, "null argument");
  parallel_for(n_vecs, n_threads, [&](size_t b, size_t e) {
    for (size_t v = b; v < e; ++v) {
      int ee, ff;
      const uint64_t est = search_exhaustive<float>(samples + v * 32, 32, &ee, &ff);
      exps[v] = uint8_t(ee);
      facs[v] = uint8_t(ff);
      if (est_bits) est_bits[v] = est;
    }
  });
  return FLSGPU_OK;
}
int flsgpu_alp_search_f64(const double *samples, uint64_t n_vecs, uint8_t *exps, uint8_t *facs,
                          uint64_t *est_bits, int n_threads) {
  if (!samples || !exps || !facs) return set_err(FLSGPU_EINVAL, "null argument");
  parallel_for(n_vecs, n_threads, [&](size_t b, size_t e) {
    for (size_t v = b; v < e; ++v) {
      int ee, ff;
      const uint64_t est = search_exhaustive<double>(samples + v * 32, 32, &ee, &ff);
      exps[v] = uint8_t(ee);
      facs[v] = uint8_t(ff);
      if (est_bits) est_bits[v] = est;
    }
  });
  return FLSGPU_OK;
}

int flsgpu_galp_convert_f32(uint64_t n_vecs, const float *exceptions, const uint16_t *positions,
                            const uint16_t *counts, const uint64_t *exceptions_offsets,
                            float *out_exceptions, uint16_t *out_positions,
                            uint16_t *out_offsets_counts) {
  return galp_convert<float>(n_vecs, exceptions, positions, counts, exceptions_offsets,
                             out_exceptions, out_positions, out_offsets_counts);
}
int flsgpu_galp_convert_f64(uint64_t n_vecs, const double *exceptions, const uint16_t *positions,
                            const uint16_t *counts, const uint64_t *exceptions_offsets,
                            double *out_exceptions, uint16_t *out_positions,
                            uint16_t *out_offsets_counts) {
  return galp_convert<double>(n_vecs, exceptions, positions, counts, exceptions_offsets,
                              out_exceptions, out_positions, out_offsets_counts);
}

int flsgpu_alprd_encode_f32(const float *in, uint64_t n_values, flsgpu_alprd_host_column *col,
                            int n_threads) {
  return alprd_encode<float>(in, n_values, col, n_threads);
}
int flsgpu_alprd_encode_f64(const double *in, uint64_t n_values, flsgpu_alprd_host_column *col,
                            int n_threads) {
  return alprd_encode<double>(in, n_values, col, n_threads);
}

int flsgpu_ffor_pack_u32(const uint32_t *values, uint64_t n_vecs, const uint8_t *bit_widths,
                         const uint32_t *bases, uint32_t *packed, uint64_t *vector_offsets) {
  uint64_t w = 0;
  for (uint64_t v = 0; v < n_vecs; ++v) {
    if (bit_widths[v] > 32) return set_err(FLSGPU_EINVAL, "bit width > 32");
    uint32_t d[1024];
    for (int i = 0; i < 1024; ++i) d[i] = values[v * 1024 + i] - (bases ? bases[v] : 0u);
    fls_pack<uint32_t>(d, packed + w, bit_widths[v]);
    vector_offsets[v] = w;
    w += uint64_t(bit_widths[v]) * 32;
  }
  return FLSGPU_OK;
}
int flsgpu_ffor_pack_u64(const uint64_t *values, uint64_t n_vecs, const uint8_t *bit_widths,
                         const uint64_t *bases, uint64_t *packed, uint64_t *vector_offsets) {
  uint64_t w = 0;
  for (uint64_t v = 0; v < n_vecs; ++v) {
    if (bit_widths[v] > 64) return set_err(FLSGPU_EINVAL, "bit width > 64");
    uint64_t d[1024];
    for (int i = 0; i < 1024; ++i) d[i] = values[v * 1024 + i] - (bases ? bases[v] : 0ull);
    fls_pack<uint64_t>(d, packed + w, bit_widths[v]);
    vector_offsets[v] = w;
    w += uint64_t(bit_widths[v]) * 16;
  }
  return FLSGPU_OK;
}

}  // extern "C"
