// Brute-force checks, on the host, of the arithmetic the range-scan kernels
// use in place of a per-value decode:
//  - the fused filter + sum's per-lane statistics (csrc/fls_kernels.hpp:
//    popcounts of the top-bit positions, bit planes, signed field extracts,
//    the wrap-around correction, the range and equality predicates), for
//    every bit width of both element types and every chunk shape, against
//    a per-value restatement over the same words;
//  - the ALP query's interval of matching offsets (csrc/alp_kernels.hpp:
//    alp_target_interval), against the decode of every offset.
// Prints one line per part and exits nonzero on any mismatch.
#include <random>

#include "../../msc-thesis-project_amd/csrc/alp_kernels.hpp"
#include "../../msc-thesis-project_amd/csrc/fls_kernels.hpp"

namespace flsgpu_hip {
int set_error(int code, const char *fmt, ...) {
  std::fprintf(stderr, "error %d: %s\n", code, fmt);
  return code;
}
int check_hip(hipError_t e, const char *) { return e == hipSuccess ? 0 : 1; }
int get_scratch(hipStream_t, Scratch *) { return 1; }
unsigned reduce_grid(uint64_t) { return 1; }
int num_cus() { return 256; }
int reduce_finish_flag(const Scratch &, unsigned, uint8_t *, hipStream_t) { return 1; }
int reduce_finish_pair(const Scratch &, unsigned, uint64_t *, hipStream_t) { return 1; }
}  // namespace flsgpu_hip
using namespace flsgpu_hip;

static std::mt19937_64 rng(2026);
static long bad = 0, checks = 0;

// Count and sum (mod 2^64) of the passing values of one chunk, per value.
template <typename UT, int BW, int J0, int NJ, int PRED>
static void brute(const UT *w, UT base, UT lo, UT hi, uint64_t &count, uint64_t &sum) {
  emit_slot_words<UT, BW, J0, NJ>(w, [&](int, UT off) {
    const UT v = UT(off + base);
    bool pass;
    if constexpr (PRED == FLSGPU_PRED_OFFSET_LOWER_HALF) pass = BW > 0 && (off >> (BW > 0 ? BW - 1 : 0)) == 0;
    else if constexpr (PRED == FLSGPU_PRED_ALL) pass = true;
    else if constexpr (PRED == PRED_EQUAL) pass = v == lo;
    else pass = v >= lo && v < hi;
    count += pass;
    sum += pass ? uint64_t(v) : 0;
  });
}

template <typename UT, int BW, int C, int PRED>
static void check_chunks() {
  constexpr int W = lanes<UT>::W, NJ = W / C;
  static_for<C>([&](auto cc) {
    constexpr int J0 = int(decltype(cc)::value) * NJ;
    using Sp = SlotSpan<UT, BW, J0, NJ>;
    for (int it = 0; it < 24; ++it) {
      UT w[Sp::NR1];
      for (int r = 0; r < Sp::NR1; ++r) w[r] = UT(rng());
      const UT maxoff = BW == W ? UT(~UT(0)) : UT((UT(1) << BW) - 1);
      // bases that never wrap, that may wrap, and the extremes
      const UT bases[4] = {UT(rng() % 1000), UT(rng()), UT(~UT(0) - maxoff), UT(~UT(0))};
      for (UT base : bases) {
        UT lo = UT(rng()), hi = UT(rng());
        if (it % 3 == 0) {  // a needle / range taken from the values
          uint64_t k = 0;
          const uint64_t pick = rng() % NJ;
          emit_slot_words<UT, BW, J0, NJ>(w, [&](int, UT off) {
            if (k++ == pick) lo = UT(off + base);
          });
          hi = UT(lo + UT(rng() % 64));
        }
        if (lo > hi) std::swap(lo, hi);
        uint64_t c0 = 0, s0 = 0, c1 = 0, s1 = 0;
        brute<UT, BW, J0, NJ, PRED>(w, base, lo, hi, c0, s0);
        span_filter_sum<UT, BW, J0, NJ, PRED>(w, base, lo, hi, c1, s1);
        // the equality query uses only the count
        const bool ok = c0 == c1 && (PRED == PRED_EQUAL || s0 == s1);
        ++checks;
        if (!ok && bad++ < 10)
          std::printf("filter-sum mismatch: W=%d BW=%d J0=%d NJ=%d pred=%d base=%llu: %llu/%llu vs %llu/%llu\n",
                      W, BW, J0, NJ, PRED, (unsigned long long)base, (unsigned long long)c0,
                      (unsigned long long)s0, (unsigned long long)c1, (unsigned long long)s1);
      }
    }
  });
}

template <typename UT, int BW>
static void check_width() {
  constexpr int W = lanes<UT>::W;
  using Sh = ScanShape<UT, BW>;
  for (int pred : {0, 1, 2, 3}) (void)pred;
  check_chunks<UT, BW, Sh::C, FLSGPU_PRED_OFFSET_LOWER_HALF>();
  check_chunks<UT, BW, Sh::C, FLSGPU_PRED_ALL>();
  check_chunks<UT, BW, Sh::C, FLSGPU_PRED_RANGE>();
  check_chunks<UT, BW, Sh::C, PRED_EQUAL>();
  if constexpr (Sh::C == 1 && W == 64) check_chunks<UT, BW, 2, FLSGPU_PRED_OFFSET_LOWER_HALF>();
}

template <typename T>
static void check_interval(int iters) {
  using UT = typename uint_of<T>::type;
  using ST = typename int_of<T>::type;
  constexpr int W = sizeof(T) * 8;
  long n[4] = {0, 0, 0, 0};
  for (int it = 0; it < iters; ++it) {
    const int bw = int(rng() % 17);
    const UT v1 = UT((UT(1) << bw) - 1);
    const int kind = int(rng() % 4);
    const ST y0 = kind == 0   ? ST(int64_t(rng() % 41) - 20)
                  : kind == 1 ? ST(int64_t(rng() % 2000001) - 1000000)
                  : kind == 2 ? ST(ST(UT(rng()) >> 1) - ST(rng() % 1000))
                              : ST(UT(rng()));
    AlpConsts<T> c{UT(y0), unsigned(rng() % (W == 32 ? 10 : 19)), unsigned(rng() % (W == 32 ? 11 : 21)),
                   0, T(0)};
    c.resolve();
    T target;
    switch (rng() % 6) {
      case 0: case 1: case 2: target = c(UT(rng() % (uint64_t(v1) + 1))); break;
      case 3: target = c(UT(0)); break;
      case 4: target = c(v1); break;
      default: target = T(double(int64_t(rng() % 2000001) - 1000000) / 7.0); break;
    }
    UT lo = 0, span = 0;
    const int how = alp_target_interval<T>(c, target, bw, lo, span);
    ++n[how];
    if (how == ANY_DECODE) continue;
    ++checks;
    for (uint64_t v = 0; v <= uint64_t(v1); ++v) {
      const bool want = c(UT(v)) == target;
      const bool got = how == ANY_INTERVAL && UT(UT(v) - lo) <= span;
      if (want != got) {
        if (bad++ < 10)
          std::printf("interval mismatch: W=%d bw=%d base=%lld v=%llu how=%d\n", W, bw,
                      (long long)y0, (unsigned long long)v, how);
        break;
      }
    }
  }
  std::printf("interval f%d: %d vectors (none %ld, interval %ld, decode %ld)\n", W, iters, n[0], n[1], n[2]);
}

int main() {
  static_for<33>([&](auto bc) { check_width<uint32_t, int(decltype(bc)::value)>(); });
  static_for<65>([&](auto bc) { check_width<uint64_t, int(decltype(bc)::value)>(); });
  std::printf("filter-sum statistics: %ld checks\n", checks);
  check_interval<float>(4000);
  check_interval<double>(4000);
  std::printf("mismatches: %ld of %ld checks\n", bad, checks);
  return bad ? 1 : 0;
}
