// The fused filter + sum over a 32-bit FFOR column with its block partials in
// a static device array, for callers that cannot allocate device memory
// (code under hipGraph capture, tight serving loops, harnesses that forbid
// hipMalloc). flsgpu_query_filter_sum_ffor_u32 keeps its partials in a
// per-stream area the library allocates on first use; this variant runs
// the same two launches over `scratch` and allocates nothing.
//
// `scratch` holds NOALLOC_SLOTS partial pairs (32 KiB). In host code its
// name is only a stand-in for the symbol, never a device address: every
// device the library runs on holds its own copy, whose address
// hipGetSymbolAddress returns once per device at initialisation (kept in
// g_scratch_addr) and which the kernels take as an argument. The array
// keeps the last call's partials; each call's blocks write every partial
// the finishing kernel reads, so nothing needs clearing. Calls that overlap
// on one device (on any streams) share the array and race.
#include "fls_impl.hpp"

namespace flsgpu_hip {

constexpr unsigned NOALLOC_SLOTS = 2048;  // the scan kernels' grid cap
static __device__ __align__(256) uint64_t scratch[2 * NOALLOC_SLOTS];
static uint64_t *g_scratch_addr[MAX_DEVICES];

int setup_static_scratch(int dev) {
  if (dev < 0 || dev >= MAX_DEVICES) return set_error(FLSGPU_EINVAL, "device %d out of range", dev);
  if (g_scratch_addr[dev]) return FLSGPU_OK;
  void *p = nullptr;
  const int rc = check_hip(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)),
                           "hipGetSymbolAddress(scratch)");
  if (rc) return rc;
  if (!p) return set_error(FLSGPU_EHIP, "hipGetSymbolAddress(scratch) returned null");
  g_scratch_addr[dev] = static_cast<uint64_t *>(p);
  return FLSGPU_OK;
}

}  // namespace flsgpu_hip

using namespace flsgpu_hip;

extern "C" {

int flsgpu_query_filter_sum_ffor_u32_noalloc(const flsgpu_ffor_column *col, int predicate,
                                             uint64_t lo, uint64_t hi, uint64_t *d_result,
                                             flsgpu_stream_t stream) {
  if (!col || (!col->bases && col->bp.n_vecs)) return set_error(FLSGPU_EINVAL, "null column/bases");
  if (!d_result) return set_error(FLSGPU_EINVAL, "null result pointer");
  FforArgs<uint32_t> a;
  int rc = make_ffor_args<uint32_t>(&col->bp, col->bases, &a);
  if (rc) return rc;
  int dev;
  if ((rc = ready_device(&dev))) return rc;
  uint64_t *partials = g_scratch_addr[dev];
  if (!partials) return set_error(FLSGPU_EHIP, "static scratch of device %d not set up", dev);
  // One partial pair per block: launch_partials caps the grid at the
  // array's slots and its waves stride over the rest of the column, so
  // every column size fits.
  const unsigned grid = NOALLOC_SLOTS;
  switch (predicate) {
    case FLSGPU_PRED_RANGE:
      return FlsFilterSum<uint32_t, 1, FLSGPU_PRED_RANGE>::launch_partials(
          a, uint32_t(lo), uint32_t(hi), d_result, partials, grid, stream);
    case FLSGPU_PRED_OFFSET_LOWER_HALF:
      return FlsFilterSum<uint32_t, 1, FLSGPU_PRED_OFFSET_LOWER_HALF>::launch_partials(
          a, uint32_t(lo), uint32_t(hi), d_result, partials, grid, stream);
    case FLSGPU_PRED_ALL:
      return FlsFilterSum<uint32_t, 1, FLSGPU_PRED_ALL>::launch_partials(
          a, uint32_t(lo), uint32_t(hi), d_result, partials, grid, stream);
    default:
      return set_error(FLSGPU_EINVAL, "unknown predicate %d", predicate);
  }
}

}  // extern "C"
