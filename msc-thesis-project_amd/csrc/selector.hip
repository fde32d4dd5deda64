// Learned ALP parameter proposal: per-vector features and a small MLP
// (32 -> 64 -> 64 -> 19 + 19 + 2) with hand-written forward, backward and
// Adam kernels. No reference counterpart: the reference chooses (exponent,
// factor) by sampling the first rowgroup and searching every pair
// (src/alp/encoder.hpp:134-300); that search stays the fallback and labels
// the training data. Decoding is unaffected: whatever the proposal, the
// encoder stores exceptions for values that do not round-trip.
//
// The network is tiny (8 872 parameters) and each kernel is one thread per
// sample (or per parameter for the weight gradients), with the parameters
// staged in LDS; the work is latency-bound, so MFMA tiling would not pay.
#include <cmath>
#include <limits>

#include "common.hpp"
#include "flsgpu_selector.h"

namespace flsgpu_hip {

constexpr int NIN = FLSGPU_SEL_INPUTS, NH = FLSGPU_SEL_HIDDEN, NOUT = FLSGPU_SEL_OUTPUTS;
// 10^e for the encoder side (src/alp/constants.hpp EXP_ARR), used by the
// selector's round-trip features.
static __constant__ float c_exp10_f32[11] = {1.0,       10.0,       100.0,        1000.0,
                                             10000.0,   100000.0,   1000000.0,    10000000.0,
                                             100000000.0, 1000000000.0, 10000000000.0};
static __constant__ double c_exp10_f64[19] = {1e0,  1e1,  1e2,  1e3,  1e4,  1e5,  1e6,
                                              1e7,  1e8,  1e9,  1e10, 1e11, 1e12, 1e13,
                                              1e14, 1e15, 1e16, 1e17, 1e18};

constexpr int NE = 19, NF = 19, NS = 2;
constexpr int OW1 = 0, OB1 = OW1 + NIN * NH, OW2 = OB1 + NH, OB2 = OW2 + NH * NH, OW3 = OB2 + NH,
              OB3 = OW3 + NH * NOUT, NPARAM = OB3 + NOUT;
static_assert(NPARAM == FLSGPU_SEL_PARAMS, "parameter layout");

// ALP round trip of x at (exponent k, factor 0), the SAFE encode of
// src/alp/encoder.hpp:76-84 followed by decoder.hpp:121-124. Products and
// sums stay separately rounded, as on the encoder (no FMA contraction).
#pragma clang fp contract(off)
template <typename T>
__device__ __forceinline__ bool roundtrips(T x, int k);
template <>
__device__ __forceinline__ bool roundtrips<double>(double x, int k) {
  double t = x * c_exp10_f64[k];
  if (!isfinite(t) || t > 9223372036854774784.0 || t < -9223372036854774784.0 ||
      (t == 0.0 && signbit(t)))
    return false;
  t = t + 6755399441055744.0 - 6755399441055744.0;
  const int64_t e = int64_t(t);
  return double(e) * c_frac_f64[k] == x;
}
template <>
__device__ __forceinline__ bool roundtrips<float>(float x, int k) {
  if (k > 10) return false;
  float t = x * c_exp10_f32[k];
  if (!isfinite(t) || double(t) > 9223372036854774784.0 || double(t) < -9223372036854774784.0 ||
      (t == 0.0f && signbit(t)))
    return false;
  t = t + 12582912.0f - 12582912.0f;
  if (!(t > -2147483648.0f && t < 2147483648.0f)) return false;
  const int32_t e = int32_t(t);
  return float(e) * c_frac_f32[k] == x;
}

// 32 features from the 32 samples at slots 0, 32, ..., 992 of a vector (the
// samples of src/alp/encoder.hpp:248-249).
template <typename T>
__global__ __launch_bounds__(BLOCK) void k_features(const T *__restrict col, uint64_t n_vecs,
                                                   uint32_t vstride, uint32_t sstride,
                                                   float *__restrict feats) {
  const uint64_t v = uint64_t(blockIdx.x) * BLOCK + threadIdx.x;
  if (v >= n_vecs) return;
  T s[32];
  for (int i = 0; i < 32; ++i) s[i] = col[v * vstride + uint64_t(sstride) * i];
  float *f = feats + v * NIN;
  int first_full = 19;
  for (int k = 0; k < 19; ++k) {
    int ok = 0;
    for (int i = 0; i < 32; ++i) ok += roundtrips<T>(s[i], k);
    f[k] = float(ok) / 32.0f;
    if (ok == 32 && first_full == 19) first_full = k;
  }
  double mx = 0.0, mn = 1e308, sl = 0.0, lo = 1e308, hi = -1e308, sum = 0.0, sq = 0.0;
  int zeros = 0, neg = 0, special = 0, nfin = 0, same = 0, distinct = 0;
  for (int i = 0; i < 32; ++i) {
    const double x = double(s[i]);
    if (!isfinite(x) || (x == 0.0 && signbit(x))) {
      ++special;
      continue;
    }
    ++nfin;
    const double a = fabs(x);
    zeros += a == 0.0;
    neg += x < 0.0;
    if (a > 0.0) {
      mx = fmax(mx, a);
      mn = fmin(mn, a);
      sl += log10(a);
    }
    lo = fmin(lo, x);
    hi = fmax(hi, x);
    sum += x;
    sq += x * x;
    same += s[i] == s[0];
    bool seen = false;
    for (int j = 0; j < i; ++j) seen |= s[j] == s[i];
    distinct += !seen;
  }
  const int nz = nfin - zeros;
  f[19] = mx > 0.0 ? float(log10(mx) / 10.0) : -1.0f;
  f[20] = nz > 0 ? float(log10(mn) / 10.0) : -1.0f;
  f[21] = nz > 0 ? float(sl / nz / 10.0) : -1.0f;
  f[22] = float(zeros) / 32.0f;
  f[23] = float(neg) / 32.0f;
  f[24] = float(special) / 32.0f;
  f[25] = nfin > 1 ? float(fmin(log2(hi - lo + 1.0), 1024.0) / 64.0) : 0.0f;
  f[26] = float(same) / 32.0f;
  f[27] = float(distinct) / 32.0f;
  int k90 = 19;
  for (int k = 18; k >= 0; --k)
    if (f[k] >= 0.9f) k90 = k;
  f[28] = float(k90) / 19.0f;
  f[29] = float(first_full) / 19.0f;
  f[30] = 1.0f;
  const double var = nfin > 0 ? fmax(sq / nfin - (sum / nfin) * (sum / nfin), 0.0) : 0.0;
  f[31] = var > 0.0 ? float(fmin(log10(sqrt(var) + 1e-300), 160.0) / 10.0) : -1.0f;
}

__device__ __forceinline__ void stage_params(const float *__restrict p, float *sp) {
  for (int i = threadIdx.x; i < NPARAM; i += BLOCK) sp[i] = p[i];
  __syncthreads();
}

// Forward pass; H1/H2 (post-ReLU) and Y (logits) are stored for backward.
__global__ __launch_bounds__(BLOCK) void k_forward(const float *__restrict p,
                                                  const float *__restrict x, uint32_t n,
                                                  float *__restrict h1o, float *__restrict h2o,
                                                  float *__restrict yo) {
  __shared__ float sp[NPARAM];
  stage_params(p, sp);
  const uint32_t r = blockIdx.x * BLOCK + threadIdx.x;
  if (r >= n) return;
  float in[NIN], h1[NH], h2[NH];
  for (int i = 0; i < NIN; ++i) in[i] = x[uint64_t(r) * NIN + i];
  for (int j = 0; j < NH; ++j) {
    float a = sp[OB1 + j];
    for (int i = 0; i < NIN; ++i) a += in[i] * sp[OW1 + i * NH + j];
    h1[j] = fmaxf(a, 0.0f);
  }
  for (int j = 0; j < NH; ++j) {
    float a = sp[OB2 + j];
    for (int i = 0; i < NH; ++i) a += h1[i] * sp[OW2 + i * NH + j];
    h2[j] = fmaxf(a, 0.0f);
  }
  for (int j = 0; j < NH; ++j) {
    if (h1o) h1o[uint64_t(r) * NH + j] = h1[j];
    if (h2o) h2o[uint64_t(r) * NH + j] = h2[j];
  }
  for (int k = 0; k < NOUT; ++k) {
    float a = sp[OB3 + k];
    for (int i = 0; i < NH; ++i) a += h2[i] * sp[OW3 + i * NOUT + k];
    yo[uint64_t(r) * NOUT + k] = a;
  }
}

// Softmax cross-entropy over the three heads; dY = (softmax - onehot) / n.
__global__ __launch_bounds__(BLOCK) void k_output_grad(const float *__restrict y,
                                                      const int32_t *__restrict labels, uint32_t n,
                                                      float *__restrict dy,
                                                      float *__restrict loss_per_row) {
  const uint32_t r = blockIdx.x * BLOCK + threadIdx.x;
  if (r >= n) return;
  const int heads[3][2] = {{0, NE}, {NE, NF}, {NE + NF, NS}};
  float loss = 0.0f;
  for (int h = 0; h < 3; ++h) {
    const int o = heads[h][0], m = heads[h][1];
    const float *yy = y + uint64_t(r) * NOUT + o;
    float mx = yy[0];
    for (int k = 1; k < m; ++k) mx = fmaxf(mx, yy[k]);
    float z = 0.0f;
    for (int k = 0; k < m; ++k) z += expf(yy[k] - mx);
    const int lab = labels[uint64_t(r) * 3 + h];
    for (int k = 0; k < m; ++k) {
      const float pk = expf(yy[k] - mx) / z;
      dy[uint64_t(r) * NOUT + o + k] = (pk - (k == lab ? 1.0f : 0.0f)) / float(n);
    }
    loss += -(yy[lab] - mx - logf(z));
  }
  loss_per_row[r] = loss;
}

// dH2 = (dY W3^T) * [H2 > 0], dH1 = (dH2 W2^T) * [H1 > 0]
__global__ __launch_bounds__(BLOCK) void k_hidden_grad(const float *__restrict p,
                                                      const float *__restrict h1,
                                                      const float *__restrict h2,
                                                      const float *__restrict dy, uint32_t n,
                                                      float *__restrict dh1,
                                                      float *__restrict dh2) {
  __shared__ float sp[NPARAM];
  stage_params(p, sp);
  const uint32_t r = blockIdx.x * BLOCK + threadIdx.x;
  if (r >= n) return;
  float g[NOUT], d2[NH];
  for (int k = 0; k < NOUT; ++k) g[k] = dy[uint64_t(r) * NOUT + k];
  for (int i = 0; i < NH; ++i) {
    float a = 0.0f;
    for (int k = 0; k < NOUT; ++k) a += g[k] * sp[OW3 + i * NOUT + k];
    d2[i] = h2[uint64_t(r) * NH + i] > 0.0f ? a : 0.0f;
    dh2[uint64_t(r) * NH + i] = d2[i];
  }
  for (int j = 0; j < NH; ++j) {
    float a = 0.0f;
    for (int i = 0; i < NH; ++i) a += d2[i] * sp[OW2 + j * NH + i];
    dh1[uint64_t(r) * NH + j] = h1[uint64_t(r) * NH + j] > 0.0f ? a : 0.0f;
  }
}

// One thread per parameter: sums over the batch in a fixed order.
__global__ __launch_bounds__(BLOCK) void k_weight_grad(const float *__restrict x,
                                                      const float *__restrict h1,
                                                      const float *__restrict h2,
                                                      const float *__restrict dh1,
                                                      const float *__restrict dh2,
                                                      const float *__restrict dy, uint32_t n,
                                                      float *__restrict grad) {
  const int q = blockIdx.x * BLOCK + threadIdx.x;
  if (q >= NPARAM) return;
  const float *a;
  const float *b;
  int sa, sb, ia, ib;
  if (q < OB1) {
    ia = (q - OW1) / NH; ib = (q - OW1) % NH; a = x; sa = NIN; b = dh1; sb = NH;
  } else if (q < OW2) {
    ia = -1; ib = q - OB1; a = nullptr; sa = 0; b = dh1; sb = NH;
  } else if (q < OB2) {
    ia = (q - OW2) / NH; ib = (q - OW2) % NH; a = h1; sa = NH; b = dh2; sb = NH;
  } else if (q < OW3) {
    ia = -1; ib = q - OB2; a = nullptr; sa = 0; b = dh2; sb = NH;
  } else if (q < OB3) {
    ia = (q - OW3) / NOUT; ib = (q - OW3) % NOUT; a = h2; sa = NH; b = dy; sb = NOUT;
  } else {
    ia = -1; ib = q - OB3; a = nullptr; sa = 0; b = dy; sb = NOUT;
  }
  float acc = 0.0f;
  if (a) {
    for (uint32_t r = 0; r < n; ++r) acc += a[uint64_t(r) * sa + ia] * b[uint64_t(r) * sb + ib];
  } else {
    for (uint32_t r = 0; r < n; ++r) acc += b[uint64_t(r) * sb + ib];
  }
  grad[q] = acc;
}

__global__ __launch_bounds__(BLOCK) void k_adam(float *__restrict p, const float *__restrict g,
                                               float *__restrict m, float *__restrict v,
                                               float lr, float b1, float b2, float eps,
                                               float c1, float c2) {
  const int q = blockIdx.x * BLOCK + threadIdx.x;
  if (q >= NPARAM) return;
  const float gq = g[q];
  const float mq = b1 * m[q] + (1.0f - b1) * gq;
  const float vq = b2 * v[q] + (1.0f - b2) * gq * gq;
  m[q] = mq;
  v[q] = vq;
  p[q] -= lr * (mq / c1) / (sqrtf(vq / c2) + eps);
}

// Logits of one row from LDS-staged parameters.
__device__ __forceinline__ void mlp_logits(const float *sp, const float *__restrict xr,
                                           float (&y)[NOUT]) {
  float in[NIN], h1[NH], h2[NH];
  for (int i = 0; i < NIN; ++i) in[i] = xr[i];
  for (int j = 0; j < NH; ++j) {
    float a = sp[OB1 + j];
    for (int i = 0; i < NIN; ++i) a += in[i] * sp[OW1 + i * NH + j];
    h1[j] = fmaxf(a, 0.0f);
  }
  for (int j = 0; j < NH; ++j) {
    float a = sp[OB2 + j];
    for (int i = 0; i < NH; ++i) a += h1[i] * sp[OW2 + i * NH + j];
    h2[j] = fmaxf(a, 0.0f);
  }
  for (int k = 0; k < NOUT; ++k) {
    float a = sp[OB3 + k];
    for (int i = 0; i < NH; ++i) a += h2[i] * sp[OW3 + i * NOUT + k];
    y[k] = a;
  }
}

// Forward + argmax: exponent, then the best factor <= exponent, then scheme.
__global__ __launch_bounds__(BLOCK) void k_predict(const float *__restrict p,
                                                  const float *__restrict x, uint32_t n,
                                                  uint8_t *__restrict exps,
                                                  uint8_t *__restrict facs,
                                                  uint8_t *__restrict scheme) {
  __shared__ float sp[NPARAM];
  stage_params(p, sp);
  const uint32_t r = blockIdx.x * BLOCK + threadIdx.x;
  if (r >= n) return;
  float y[NOUT];
  mlp_logits(sp, x + uint64_t(r) * NIN, y);
  int be = 0;
  for (int k = 1; k < NE; ++k)
    if (y[k] > y[be]) be = k;
  int bf = 0;
  for (int k = 1; k <= be; ++k)
    if (y[NE + k] > y[NE + bf]) bf = k;
  exps[r] = uint8_t(be);
  facs[r] = uint8_t(bf);
  if (scheme) scheme[r] = y[NE + NF + 1] > y[NE + NF] ? 1 : 0;
}

// ---- verification: the search's size estimate (src/alp/encoder.hpp:153-181)
template <typename T>
struct SearchTraits;
template <>
struct SearchTraits<double> {
  using ST = int64_t;
  using UT = uint64_t;
  static constexpr int MAXE = 18;
  static constexpr int MAXF = 18;
  static constexpr uint64_t EXCSZ = 64;
  static constexpr uint64_t RD_THRESHOLD = 48 * 32;
  // encode_value<SAFE> (encoder.hpp:76-84): two separate products, the
  // impossible-to-encode check, rounding by the magic number.
  __device__ static ST encode(double x, int f, int e) {
    double t = x * c_exp10_f64[e] * c_frac_f64[f];
    if (!isfinite(t) || t > 9223372036854774784.0 || t < -9223372036854774784.0 ||
        (t == 0.0 && signbit(t)))
      return ST(9223372036854774784ULL);
    t = t + 6755399441055744.0 - 6755399441055744.0;
    return ST(t);
  }
  // decode_value (decoder.hpp:121-124): wrapping integer product, then the
  // int -> double conversion and one product.
  __device__ static double decode(ST v, int f, int e) {
    const ST s = ST(UT(v) * UT(c_fact_f64[f]));
    return double(s) * c_frac_f64[e];
  }
  __device__ static int bits(UT d) { return d == 0 ? 0 : 64 - __clzll(d); }
};
template <>
struct SearchTraits<float> {
  using ST = int32_t;
  using UT = uint32_t;
  static constexpr int MAXE = 10;
  // Factor indices 0..9: the float factor table has ten entries
  // (src/alp/constants.hpp:55); the pair (10, 10) is never proposed.
  static constexpr int MAXF = 9;
  static constexpr uint64_t EXCSZ = 32;
  static constexpr uint64_t RD_THRESHOLD = 22 * 32;
  __device__ static ST encode(float x, int f, int e) {
    float t = x * c_exp10_f32[e] * c_frac_f32[f];
    if (!isfinite(t) || double(t) > 9223372036854774784.0 ||
        double(t) < -9223372036854774784.0 || (t == 0.0f && signbit(t)))
      return INT32_MAX;
    t = t + 12582912.0f - 12582912.0f;
    // Out of int32 range the x86-64 conversion yields INT32_MIN; the value
    // only feeds a round-trip test that fails either way.
    return (t >= -2147483648.0f && t < 2147483648.0f) ? ST(t) : INT32_MIN;
  }
  __device__ static float decode(ST v, int f, int e) {
    const ST s = ST(UT(v) * UT(c_fact_f32[f]));
    return float(s) * c_frac_f32[e];
  }
  __device__ static int bits(UT d) { return d == 0 ? 0 : 32 - __clz(d); }
};

template <typename T>
__device__ __forceinline__ uint64_t estimate(const T (&s)[32], int e, int f) {
  using Tr = SearchTraits<T>;
  using ST = typename Tr::ST;
  using UT = typename Tr::UT;
  int exc = 0;
  ST mx = std::numeric_limits<ST>::min(), mn = std::numeric_limits<ST>::max();
#pragma unroll
  for (int i = 0; i < 32; ++i) {
    const ST v = Tr::encode(s[i], f, e);
    if (Tr::decode(v, f, e) == s[i]) {
      mx = v > mx ? v : mx;
      mn = v < mn ? v : mn;
    } else {
      ++exc;
    }
  }
  if (32 - exc < 2) return ~uint64_t(0);
  return 32ull * uint64_t(Tr::bits(UT(UT(mx) - UT(mn)))) + uint64_t(exc) * (Tr::EXCSZ + 16);
}

// One thread per vector: logits (kept in LDS for runtime indexing), the
// n_cand best (exponent, factor) pairs by y[e] + y[19 + f] (ties to the pair
// met first, exponent then factor descending), each verified on the samples.
template <typename T>
__global__ __launch_bounds__(BLOCK) void k_select(const float *__restrict p,
                                                 const float *__restrict x,
                                                 const T *__restrict col, uint64_t n_vecs,
                                                 uint32_t vstride, uint32_t sstride, int n_cand,
                                                 uint8_t *__restrict exps, uint8_t *__restrict facs,
                                                 uint8_t *__restrict scheme,
                                                 uint64_t *__restrict est_out) {
  using Tr = SearchTraits<T>;
  __shared__ float sp[NPARAM];
  __shared__ float sy[NOUT][BLOCK];
  stage_params(p, sp);
  const uint64_t r = uint64_t(blockIdx.x) * BLOCK + threadIdx.x;
  if (r >= n_vecs) return;
  {
    float y[NOUT];
    mlp_logits(sp, x + r * NIN, y);
    for (int k = 0; k < NOUT; ++k) sy[k][threadIdx.x] = y[k];
  }
  T s[32];
#pragma unroll
  for (int i = 0; i < 32; ++i) s[i] = col[r * vstride + uint64_t(sstride) * i];
  constexpr uint64_t worst = 32 * (Tr::EXCSZ + 16) + 32 * Tr::EXCSZ;
  int fe = 0, ff = 0;
  uint64_t fest = worst;
  uint64_t taken[3] = {0, 0, 0};
  for (int c = 0; c < n_cand; ++c) {
    int bi = -1, be = 0, bf = 0;
    float bs = 0.0f;
    int i = 0;
    for (int e = Tr::MAXE; e >= 0; --e) {
      const float ye = sy[e][threadIdx.x];
      for (int f = e < Tr::MAXF ? e : Tr::MAXF; f >= 0; --f, ++i) {
        if ((taken[i >> 6] >> (i & 63)) & 1) continue;
        const float sc = ye + sy[NE + f][threadIdx.x];
        if (bi < 0 || sc > bs) {
          bi = i;
          bs = sc;
          be = e;
          bf = f;
        }
      }
    }
    taken[bi >> 6] |= uint64_t(1) << (bi & 63);
    const uint64_t est = estimate<T>(s, be, bf);
    if (est == ~uint64_t(0)) continue;
    if (est < fest || (est == fest && fe < be) || (est == fest && fe == be && ff < bf)) {
      fest = est;
      fe = be;
      ff = bf;
    }
  }
  exps[r] = uint8_t(fe);
  facs[r] = uint8_t(ff);
  if (scheme) scheme[r] = fest >= Tr::RD_THRESHOLD ? 1 : 0;
  if (est_out) est_out[r] = fest;
}

static unsigned blocks_for(uint64_t n) { return unsigned((n + BLOCK - 1) / BLOCK); }

}  // namespace flsgpu_hip

using namespace flsgpu_hip;

template <typename T>
static int select_launch(const float *params, const float *features, const T *values,
                         uint64_t n_vecs, uint32_t vector_stride, uint32_t sample_stride,
                         uint32_t n_candidates, uint8_t *exps, uint8_t *facs, uint8_t *scheme,
                         uint64_t *est_bits, flsgpu_stream_t stream) {
  if (n_vecs == 0) return FLSGPU_OK;
  if (!params || !features || !values || !exps || !facs)
    return set_error(FLSGPU_EINVAL, "null argument");
  if (sample_stride == 0 || vector_stride < 31 * sample_stride + 1)
    return set_error(FLSGPU_EINVAL, "strides must place 32 distinct samples inside a vector");
  constexpr uint32_t pairs = (SearchTraits<T>::MAXE + 1) * (SearchTraits<T>::MAXE + 2) / 2 -
                             (SearchTraits<T>::MAXE - SearchTraits<T>::MAXF);
  if (n_candidates == 0) return set_error(FLSGPU_EINVAL, "n_candidates must be at least 1");
  if (n_candidates > pairs) n_candidates = pairs;
  hipLaunchKernelGGL((k_select<T>), dim3(blocks_for(n_vecs)), dim3(BLOCK), 0, stream, params,
                     features, values, n_vecs, vector_stride, sample_stride, int(n_candidates),
                     exps, facs, scheme, est_bits);
  return check_hip(hipGetLastError(), "k_select launch");
}

extern "C" {

int flsgpu_sel_features_f64(const double *values, uint64_t n_vecs, uint32_t vector_stride,
                            uint32_t sample_stride, float *features, flsgpu_stream_t stream) {
  if (n_vecs == 0) return FLSGPU_OK;
  if (!values || !features) return set_error(FLSGPU_EINVAL, "null argument");
  if (sample_stride == 0 || vector_stride < 31 * sample_stride + 1)
    return set_error(FLSGPU_EINVAL, "strides must place 32 distinct samples inside a vector");
  hipLaunchKernelGGL((k_features<double>), dim3(blocks_for(n_vecs)), dim3(BLOCK), 0, stream,
                     values, n_vecs, vector_stride, sample_stride, features);
  return check_hip(hipGetLastError(), "k_features launch");
}
int flsgpu_sel_features_f32(const float *values, uint64_t n_vecs, uint32_t vector_stride,
                            uint32_t sample_stride, float *features, flsgpu_stream_t stream) {
  if (n_vecs == 0) return FLSGPU_OK;
  if (!values || !features) return set_error(FLSGPU_EINVAL, "null argument");
  if (sample_stride == 0 || vector_stride < 31 * sample_stride + 1)
    return set_error(FLSGPU_EINVAL, "strides must place 32 distinct samples inside a vector");
  hipLaunchKernelGGL((k_features<float>), dim3(blocks_for(n_vecs)), dim3(BLOCK), 0, stream,
                     values, n_vecs, vector_stride, sample_stride, features);
  return check_hip(hipGetLastError(), "k_features launch");
}

int flsgpu_sel_forward(const float *params, const float *x, uint32_t n, float *h1, float *h2,
                       float *logits, flsgpu_stream_t stream) {
  if (n == 0) return FLSGPU_OK;
  if (!params || !x || !logits) return set_error(FLSGPU_EINVAL, "null argument");
  hipLaunchKernelGGL(k_forward, dim3(blocks_for(n)), dim3(BLOCK), 0, stream, params, x, n, h1, h2,
                     logits);
  return check_hip(hipGetLastError(), "k_forward launch");
}

int flsgpu_sel_backward(const float *params, const float *x, const float *h1, const float *h2,
                        const float *logits, const int32_t *labels, uint32_t n, float *dlogits,
                        float *dh1, float *dh2, float *grad, float *loss_per_row,
                        flsgpu_stream_t stream) {
  if (n == 0) return FLSGPU_OK;
  if (!params || !x || !h1 || !h2 || !logits || !labels || !dlogits || !dh1 || !dh2 || !grad ||
      !loss_per_row)
    return set_error(FLSGPU_EINVAL, "null argument");
  hipLaunchKernelGGL(k_output_grad, dim3(blocks_for(n)), dim3(BLOCK), 0, stream, logits, labels, n,
                     dlogits, loss_per_row);
  hipLaunchKernelGGL(k_hidden_grad, dim3(blocks_for(n)), dim3(BLOCK), 0, stream, params, h1, h2,
                     dlogits, n, dh1, dh2);
  hipLaunchKernelGGL(k_weight_grad, dim3(blocks_for(NPARAM)), dim3(BLOCK), 0, stream, x, h1, h2,
                     dh1, dh2, dlogits, n, grad);
  return check_hip(hipGetLastError(), "selector backward launch");
}

int flsgpu_sel_adam(float *params, const float *grad, float *m, float *v, float lr, float beta1,
                    float beta2, float eps, uint32_t step, flsgpu_stream_t stream) {
  if (!params || !grad || !m || !v || step == 0) return set_error(FLSGPU_EINVAL, "bad argument");
  const float c1 = 1.0f - powf(beta1, float(step));
  const float c2 = 1.0f - powf(beta2, float(step));
  hipLaunchKernelGGL(k_adam, dim3(blocks_for(NPARAM)), dim3(BLOCK), 0, stream, params, grad, m, v,
                     lr, beta1, beta2, eps, c1, c2);
  return check_hip(hipGetLastError(), "k_adam launch");
}

int flsgpu_sel_predict(const float *params, const float *x, uint32_t n, uint8_t *exps,
                       uint8_t *facs, uint8_t *scheme, flsgpu_stream_t stream) {
  if (n == 0) return FLSGPU_OK;
  if (!params || !x || !exps || !facs) return set_error(FLSGPU_EINVAL, "null argument");
  hipLaunchKernelGGL(k_predict, dim3(blocks_for(n)), dim3(BLOCK), 0, stream, params, x, n, exps,
                     facs, scheme);
  return check_hip(hipGetLastError(), "k_predict launch");
}

int flsgpu_sel_select_f32(const float *params, const float *features, const float *values,
                          uint64_t n_vecs, uint32_t vector_stride, uint32_t sample_stride,
                          uint32_t n_candidates, uint8_t *exps, uint8_t *facs, uint8_t *scheme,
                          uint64_t *est_bits, flsgpu_stream_t stream) {
  return select_launch<float>(params, features, values, n_vecs, vector_stride, sample_stride,
                              n_candidates, exps, facs, scheme, est_bits, stream);
}
int flsgpu_sel_select_f64(const float *params, const float *features, const double *values,
                          uint64_t n_vecs, uint32_t vector_stride, uint32_t sample_stride,
                          uint32_t n_candidates, uint8_t *exps, uint8_t *facs, uint8_t *scheme,
                          uint64_t *est_bits, flsgpu_stream_t stream) {
  return select_launch<double>(params, features, values, n_vecs, vector_stride, sample_stride,
                               n_candidates, exps, facs, scheme, est_bits, stream);
}

}  // extern "C"
