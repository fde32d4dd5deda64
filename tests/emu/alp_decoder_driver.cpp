// Runs the ALP decoders (csrc/alp_kernels.hpp: one to four threads per lane,
// plain and non-temporal stores, and the LDS-staged kernels, each with
// exception patching) on the host emulation over one column, with exact-size
// buffers, once per ballot mode, and counts values that differ from the
// expected ones.
// Built once per element type and layout (-DEMU_T=float|double,
// -DEMU_GALP=0|1) so that the builds run in parallel.
//   alp_decoder_driver <column>
//   column: uint64 width, n_vecs, packed words, exceptions, galp, 0; packed,
//           bit_widths, vector_offsets, bases, factor and fraction indices,
//           exceptions offsets, exceptions, positions, counts (standard) or
//           offsets_counts (G-ALP); expected values
#include "../../msc-thesis-project_amd/csrc/alp_kernels.hpp"

namespace flsgpu_hip {
int set_error(int code, const char *fmt, ...) {
  std::fprintf(stderr, "error %d: %s\n", code, fmt);
  return code;
}
int check_hip(hipError_t e, const char *) { return e == hipSuccess ? 0 : 1; }
int get_scratch(Scratch *) { return 1; }
unsigned reduce_grid(uint64_t) { return 1; }
int num_cus() { return 256; }
int reduce_finish_flag(const Scratch &, unsigned, uint8_t *, hipStream_t) { return 1; }
int reduce_finish_pair(const Scratch &, unsigned, uint64_t *, hipStream_t) { return 1; }
}  // namespace flsgpu_hip
using namespace flsgpu_hip;

template <class X>
static X *get(FILE *f, size_t n) {
  X *p = static_cast<X *>(std::malloc(sizeof(X) * (n ? n : 1)));
  if (n && std::fread(p, sizeof(X), n, f) != n) std::abort();
  return p;
}

template <typename T, int LAYOUT>
static int run(FILE *f, uint64_t nv, uint64_t words, uint64_t ne) {
  using UT = typename uint_of<T>::type;
  constexpr int NL = lanes<UT>::N;
  AlpArgs<T> a{};
  a.ffor.n_vecs = nv;
  a.ffor.packed = get<UT>(f, words);
  a.ffor.bit_widths = get<uint8_t>(f, nv);
  a.ffor.vector_offsets = get<uint64_t>(f, nv);
  a.ffor.bases = get<UT>(f, nv);
  a.factor_indices = get<uint8_t>(f, nv);
  a.fraction_indices = get<uint8_t>(f, nv);
  a.exceptions_offsets = get<uint64_t>(f, nv);
  a.exceptions = get<T>(f, ne);
  a.positions = get<uint16_t>(f, ne);
  if (LAYOUT == ALP_GALP) a.offsets_counts = get<uint16_t>(f, nv * NL);
  else a.counts = get<uint16_t>(f, nv);
  const T *x = get<T>(f, nv * 1024);
  T *out = static_cast<T *>(std::malloc(sizeof(T) * nv * 1024));
  int bad = 0;
  auto check = [&](const char *name) {
    size_t d = 0;
    for (size_t i = 0; i < nv * 1024; ++i) d += std::memcmp(&out[i], &x[i], sizeof(T)) != 0;
    std::printf("%s ballot=%d differences=%zu\n", name, emu_ballot_mode, d);
    bad += d != 0;
  };
  constexpr int SA = sizeof(T) == 8 ? 4 : 2, SB = sizeof(T) == 8 ? 8 : 4;
  for (int mode = 0; mode < 2; ++mode) {
    emu_ballot_mode = mode;
#define RUN(name, ...)                                \
  std::memset(out, 0xAB, sizeof(T) * nv * 1024);      \
  if (__VA_ARGS__::launch(a, out, nullptr)) return 9; \
  check(name);
    RUN("lane", AlpDecompress<T, 1, false, LAYOUT, true>)
    RUN("lane-nt", AlpDecompress<T, 1, true, LAYOUT, true>)
    RUN("half", AlpDecompress<T, 2, false, LAYOUT, true>)
    RUN("half-nt", AlpDecompress<T, 2, true, LAYOUT, true>)
    RUN("quarter", AlpDecompress<T, 4, false, LAYOUT, true>)
    RUN("quarter-nt", AlpDecompress<T, 4, true, LAYOUT, true>)
    RUN("staged-a", AlpDecompressStaged<T, SA, LAYOUT, true>)
    RUN("staged-b", AlpDecompressStaged<T, SB, LAYOUT, true>)
#undef RUN
  }
  return bad;
}

int main(int argc, char **argv) {
  if (argc != 2) return 1;
  FILE *f = std::fopen(argv[1], "rb");
  uint64_t h[6];
  if (!f || std::fread(h, 8, 6, f) != 6) return 1;
  if (h[0] != sizeof(EMU_T) * 8 || h[4] != EMU_GALP) {
    std::fprintf(stderr, "column of width %llu, galp %llu: not the element type and layout this driver decodes\n",
                 (unsigned long long)h[0], (unsigned long long)h[4]);
    return 1;
  }
  return run<EMU_T, EMU_GALP ? ALP_GALP : ALP_STD>(f, h[1], h[2], h[3]);
}
