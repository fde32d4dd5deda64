// Runs the batched decoder (csrc/batch.hip: plan, task list, persistent
// kernel) on the host emulation: reads the columns of a batch from a file
// into exact-size buffers, creates the plan, runs it twice (a plan is
// reusable) and compares every output with the expected values.
//   batch_driver <batch>    EMU_BALLOT=0|1 selects the unpack path
//   batch: uint32 count; per column int32 encoding, uint64 n_vecs, int32
//   patcher, the descriptor's arrays (uint64 byte count + bytes each; ALP-RD
//   first its widths, dictionary size and dictionary), the expected values
#include "../../msc-thesis-project_amd/csrc/batch.hip"

namespace flsgpu_hip {
int set_error(int code, const char *fmt, ...) {
  std::fprintf(stderr, "error %d: %s\n", code, fmt);
  return code;
}
int check_hip(hipError_t e, const char *what) {
  if (e) std::fprintf(stderr, "hip error in %s\n", what);
  return e == hipSuccess ? 0 : 1;
}
int get_scratch(Scratch *) { return 1; }
unsigned reduce_grid(uint64_t) { return 1; }
int num_cus() { return 2; }
int reduce_finish_flag(const Scratch &, unsigned, uint8_t *, hipStream_t) { return 1; }
int reduce_finish_pair(const Scratch &, unsigned, uint64_t *, hipStream_t) { return 1; }
}  // namespace flsgpu_hip

static FILE *g_in;
static void *arr(size_t *nbytes_out = nullptr) {
  uint64_t n;
  if (std::fread(&n, 8, 1, g_in) != 1) std::abort();
  void *p = std::malloc(n ? n : 1);
  if (n && std::fread(p, 1, n, g_in) != n) std::abort();
  if (nbytes_out) *nbytes_out = n;
  return p;
}
template <class X>
static X get() {
  X x;
  if (std::fread(&x, sizeof(X), 1, g_in) != 1) std::abort();
  return x;
}

int main(int argc, char **argv) {
  if (const char *m = std::getenv("EMU_BALLOT")) emu_ballot_mode = std::atoi(m);
  g_in = std::fopen(argv[1], "rb");
  const uint32_t n = get<uint32_t>();
  std::vector<flsgpu_batch_item> items(n);
  std::vector<void *> expected(n);
  std::vector<size_t> exp_bytes(n);
  for (uint32_t i = 0; i < n; ++i) {
    const int enc = get<int32_t>();
    const uint64_t nv = get<uint64_t>(), nvals = nv * 1024;
    items[i].encoding = enc;
    items[i].patcher = get<int32_t>();
    if (enc <= 3) {
      auto *f = new flsgpu_ffor_column{};
      f->bp.n_values = nvals;
      f->bp.n_vecs = nv;
      f->bp.packed_array = arr();
      f->bp.bit_widths = static_cast<const uint8_t *>(arr());
      f->bp.vector_offsets = static_cast<const uint64_t *>(arr());
      f->bases = (enc == 1 || enc == 3) ? arr() : nullptr;
      items[i].column = (enc == 1 || enc == 3) ? static_cast<void *>(f) : static_cast<void *>(&f->bp);
    } else if (enc <= 7) {
      auto *a = new flsgpu_alp_column{};
      a->ffor.bp.n_values = nvals;
      a->ffor.bp.n_vecs = nv;
      a->ffor.bp.packed_array = arr();
      a->ffor.bp.bit_widths = static_cast<const uint8_t *>(arr());
      a->ffor.bp.vector_offsets = static_cast<const uint64_t *>(arr());
      a->ffor.bases = arr();
      a->factor_indices = static_cast<const uint8_t *>(arr());
      a->fraction_indices = static_cast<const uint8_t *>(arr());
      a->exceptions_offsets = static_cast<const uint64_t *>(arr());
      size_t eb;
      a->exceptions = arr(&eb);
      a->n_exceptions = eb / (enc == 4 || enc == 6 ? 4 : 8);
      a->positions = static_cast<const uint16_t *>(arr());
      if (enc == 6 || enc == 7) a->offsets_counts = static_cast<const uint16_t *>(arr());
      else a->counts = static_cast<const uint16_t *>(arr());
      items[i].column = a;
    } else {
      auto *r = new flsgpu_alprd_column{};
      r->n_values = nvals;
      r->n_vecs = nv;
      r->right_bit_width = get<uint32_t>();
      r->left_bit_width = get<uint32_t>();
      r->dictionary_size = get<uint32_t>();
      for (int k = 0; k < 8; ++k) r->dictionary[k] = get<uint16_t>();
      r->right_packed = arr();
      r->left_packed = static_cast<const uint16_t *>(arr());
      r->exceptions_offsets = static_cast<const uint64_t *>(arr());
      r->exceptions = static_cast<const uint16_t *>(arr());
      r->positions = static_cast<const uint16_t *>(arr());
      r->counts = static_cast<const uint16_t *>(arr());
      items[i].column = r;
    }
    expected[i] = arr(&exp_bytes[i]);
    items[i].out = std::malloc(exp_bytes[i]);
    std::memset(items[i].out, 0xAB, exp_bytes[i]);
  }
  flsgpu_batch_plan *plan = nullptr;
  int rc = flsgpu_batch_plan_create(items.data(), n, 0, &plan);
  std::printf("plan rc %d tasks %llu\n", rc, (unsigned long long)flsgpu_batch_plan_tasks(plan));
  if (rc) return 2;
  for (int run = 0; run < 2; ++run) {
    rc = flsgpu_batch_run(plan, nullptr);
    size_t bad = 0;
    for (uint32_t i = 0; i < n; ++i) {
      if (std::memcmp(items[i].out, expected[i], exp_bytes[i]) != 0) {
        ++bad;
        std::printf("item %u (encoding %d): output differs\n", i, int(items[i].encoding));
      }
      std::memset(items[i].out, 0xAB, exp_bytes[i]);
    }
    std::printf("run %d rc %d: %zu of %u items differ\n", run, rc, bad, n);
  }
  return 0;
}
