// Runs the device encoder's three kernels (csrc/encode.hip) on the host
// emulation, with every output buffer sized as flsgpu/encode_device.py sizes
// it, and writes the arrays it produced.
//   encoder_driver <in> <out>
//   in:  uint32 width, n_vecs, galp, kind (0 ALP, 1 FFOR); values;
//        for ALP: exponents[n_vecs], factors[n_vecs] (uint8)
//   out: totals[3] (uint64), packed[words], bit_widths, vector_offsets,
//        bases; for ALP also factor and fraction indices, exceptions
//        offsets, exceptions, positions, counts, offsets_counts (G-ALP)
#include "../../msc-thesis-project_amd/csrc/encode.hip"

namespace flsgpu_hip {
int set_error(int code, const char *fmt, ...) {
  std::fprintf(stderr, "error %d: %s\n", code, fmt);
  return code;
}
int check_hip(hipError_t e, const char *) { return e == hipSuccess ? 0 : 1; }
}  // namespace flsgpu_hip

using namespace flsgpu_hip;

template <class X>
static X *alloc(size_t n) {
  return static_cast<X *>(std::malloc(sizeof(X) * (n ? n : 1)));
}
static void put(FILE *f, const void *p, size_t bytes) { std::fwrite(p, 1, bytes, f); }
static size_t ws_bytes(uint64_t nv) { return size_t(2 * ((nv + 7) / 8 + 1) + nv) * sizeof(uint64_t); }

template <typename T>
static int run_alp(FILE *in, FILE *out, uint64_t nv, bool galp) {
  using UT = typename uint_of<T>::type;
  constexpr int NL = lanes<UT>::N;
  const size_t n = nv * 1024;
  T *x = alloc<T>(n);
  uint8_t *e = alloc<uint8_t>(nv), *f = alloc<uint8_t>(nv);
  if (std::fread(x, sizeof(T), n, in) != n || std::fread(e, 1, nv, in) != nv ||
      std::fread(f, 1, nv, in) != nv)
    return 2;
  UT *packed = alloc<UT>(n), *bases = alloc<UT>(nv);
  uint8_t *bw = alloc<uint8_t>(nv), *fi = alloc<uint8_t>(nv), *fr = alloc<uint8_t>(nv);
  uint64_t *vo = alloc<uint64_t>(nv), *eo = alloc<uint64_t>(nv), *totals = alloc<uint64_t>(3);
  T *exc = alloc<T>(n);
  uint16_t *pos = alloc<uint16_t>(n), *cnt = alloc<uint16_t>(nv);
  uint16_t *oc = galp ? alloc<uint16_t>(nv * NL) : nullptr;
  std::memset(totals, 0, 24);
  void *ws = std::malloc(ws_bytes(nv));
  EncOut o{packed, bw, vo, bases, fi, fr, eo, exc, pos, cnt, oc};
  emu_group = NL;
  if (encode_launch<T, true>(x, nv, e, f, galp, o, totals, ws, nullptr)) return 3;
  const uint64_t words = totals[0], ne = totals[1];
  put(out, totals, 24);
  put(out, packed, sizeof(UT) * words);
  put(out, bw, nv);
  put(out, vo, 8 * nv);
  put(out, bases, sizeof(UT) * nv);
  put(out, fi, nv);
  put(out, fr, nv);
  put(out, eo, 8 * nv);
  put(out, exc, sizeof(T) * ne);
  put(out, pos, 2 * ne);
  put(out, cnt, 2 * nv);
  if (galp) put(out, oc, 2 * nv * NL);
  return 0;
}

template <typename UT>
static int run_ffor(FILE *in, FILE *out, uint64_t nv) {
  const size_t n = nv * 1024;
  UT *x = alloc<UT>(n);
  if (std::fread(x, sizeof(UT), n, in) != n) return 2;
  UT *packed = alloc<UT>(n), *bases = alloc<UT>(nv);
  uint8_t *bw = alloc<uint8_t>(nv);
  uint64_t *vo = alloc<uint64_t>(nv), *totals = alloc<uint64_t>(3);
  std::memset(totals, 0, 24);
  void *ws = std::malloc(ws_bytes(nv));
  EncOut o{packed, bw, vo, bases, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  emu_group = lanes<UT>::N;
  if (encode_launch<UT, false>(x, nv, nullptr, nullptr, false, o, totals, ws, nullptr)) return 3;
  put(out, totals, 24);
  put(out, packed, sizeof(UT) * totals[0]);
  put(out, bw, nv);
  put(out, vo, 8 * nv);
  put(out, bases, sizeof(UT) * nv);
  return 0;
}

int main(int argc, char **argv) {
  if (argc != 3) return 1;
  FILE *in = std::fopen(argv[1], "rb"), *out = std::fopen(argv[2], "wb");
  uint32_t h[4];
  if (!in || !out || std::fread(h, 4, 4, in) != 4) return 1;
  const uint64_t nv = h[1];
  int rc;
  if (h[3] == 1)
    rc = h[0] == 64 ? run_ffor<uint64_t>(in, out, nv) : run_ffor<uint32_t>(in, out, nv);
  else
    rc = h[0] == 64 ? run_alp<double>(in, out, nv, h[2] != 0) : run_alp<float>(in, out, nv, h[2] != 0);
  std::fclose(out);
  return rc;
}
