// Library state behind the C ABI: error reporting, per-device scratch for
// grid reductions, device memory helpers and the autotuner table.
//
// Replaces the reference's CUDA_SAFE_CALL -> exit(EXIT_FAILURE)
// (src/flsgpu/host-utils.cuh:12-32) with status codes, and GPUArray /
// copy_to_device (host-utils.cuh:42-111, src/flsgpu/structs.cuh:103-229) with
// flsgpu_upload / flsgpu_free.
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "common.hpp"

namespace flsgpu_hip {

static thread_local char g_err[512] = "";

int set_error(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

int check_hip(hipError_t e, const char *what) {
  if (e == hipSuccess) return FLSGPU_OK;
  // Reset the runtime's sticky last error so that a later hipGetLastError()
  // after a kernel launch does not report this call's failure again.
  (void)hipGetLastError();
  return set_error(FLSGPU_EHIP, "%s: %s", what, hipGetErrorString(e));
}

// Reduction areas per device, one per stream that ran a reduction. A stream
// handle is only a key here: the area of a destroyed stream stays until
// flsgpu_release_stream, and a new stream that reuses the handle reuses the
// area (in its own stream order).
constexpr size_t MAX_STREAM_SCRATCH = 256;
constexpr size_t SCRATCH_BYTES = sizeof(uint64_t) * 2 * MAX_REDUCE_BLOCKS;
struct StreamScratch {
  hipStream_t stream;
  uint64_t *partials;
};
static std::mutex g_mu;
static std::vector<StreamScratch> g_scratch[MAX_DEVICES];
static bool g_ready[MAX_DEVICES];
static int g_cus[MAX_DEVICES];

static int current_device(int *dev) {
  int rc = check_hip(hipGetDevice(dev), "hipGetDevice");
  if (rc) return rc;
  if (*dev < 0 || *dev >= MAX_DEVICES) return set_error(FLSGPU_EINVAL, "device %d out of range", *dev);
  return FLSGPU_OK;
}

static int init_device_locked(int dev) {
  if (g_ready[dev]) return FLSGPU_OK;
  hipDeviceProp_t prop;
  int rc = check_hip(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
  if (rc) return rc;
  g_cus[dev] = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
  if ((rc = setup_static_scratch(dev))) return rc;
  g_ready[dev] = true;
  return FLSGPU_OK;
}

int ready_device(int *dev) {
  int rc = current_device(dev);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(g_mu);
  return init_device_locked(*dev);
}

int get_scratch(hipStream_t stream, Scratch *s) {
  int dev;
  int rc = current_device(&dev);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(g_mu);
  if ((rc = init_device_locked(dev))) return rc;
  for (const StreamScratch &e : g_scratch[dev]) {
    if (e.stream == stream) {
      s->partials = e.partials;
      return FLSGPU_OK;
    }
  }
  if (g_scratch[dev].size() >= MAX_STREAM_SCRATCH)
    return set_error(FLSGPU_ENOMEM,
                     "reductions on more than %zu streams of device %d: release finished "
                     "streams with flsgpu_release_stream",
                     MAX_STREAM_SCRATCH, dev);
  void *p = nullptr;
  if ((rc = check_hip(hipMalloc(&p, SCRATCH_BYTES), "hipMalloc(scratch)"))) return rc;
  g_scratch[dev].push_back({stream, static_cast<uint64_t *>(p)});
  s->partials = static_cast<uint64_t *>(p);
  return FLSGPU_OK;
}

int num_cus() {
  int dev;
  if (current_device(&dev)) return 256;
  std::lock_guard<std::mutex> lk(g_mu);
  if (init_device_locked(dev)) return 256;
  return g_cus[dev];
}

unsigned reduce_grid(uint64_t n_waves) {
  const uint64_t blocks = (n_waves + (BLOCK / WAVE) - 1) / (BLOCK / WAVE);
  return unsigned(blocks < 1 ? 1 : (blocks < MAX_REDUCE_BLOCKS ? blocks : MAX_REDUCE_BLOCKS));
}

// One block: combines the partials in a fixed order per thread, then across
// threads (deterministic for any grid).
constexpr int FINISH_BLOCK = 1024;
template <bool ADD>
__global__ __launch_bounds__(FINISH_BLOCK) void k_reduce_finish(const uint64_t *__restrict partials,
                                                                unsigned n, void *out) {
  __shared__ uint64_t sh[FINISH_BLOCK / WAVE][2];
  uint64_t a = 0, b = 0;
  for (unsigned i = threadIdx.x; i < n; i += FINISH_BLOCK) {
    if constexpr (ADD) {
      a += partials[2 * i];
      b += partials[2 * i + 1];
    } else {
      a |= partials[2 * i];
    }
  }
  for (int off = WAVE / 2; off > 0; off >>= 1) {
    const uint64_t xa = __shfl_xor(a, off, WAVE), xb = __shfl_xor(b, off, WAVE);
    a = ADD ? a + xa : (a | xa);
    b = ADD ? b + xb : b;
  }
  const int w = threadIdx.x / WAVE;
  if ((threadIdx.x % WAVE) == 0) {
    sh[w][0] = a;
    sh[w][1] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t ta = 0, tb = 0;
    for (int i = 0; i < FINISH_BLOCK / WAVE; ++i) {
      ta = ADD ? ta + sh[i][0] : (ta | sh[i][0]);
      tb += sh[i][1];
    }
    if constexpr (ADD) {
      static_cast<uint64_t *>(out)[0] = ta;
      static_cast<uint64_t *>(out)[1] = tb;
    } else {
      *static_cast<uint8_t *>(out) = ta ? 1 : 0;
    }
  }
}

int reduce_finish_flag(const Scratch &s, unsigned n_blocks, uint8_t *d_flag, hipStream_t stream) {
  hipLaunchKernelGGL(k_reduce_finish<false>, dim3(1), dim3(FINISH_BLOCK), 0, stream, s.partials,
                     n_blocks, static_cast<void *>(d_flag));
  return check_hip(hipGetLastError(), "k_reduce_finish launch");
}

int reduce_finish_pair(const Scratch &s, unsigned n_blocks, uint64_t *d_pair, hipStream_t stream) {
  hipLaunchKernelGGL(k_reduce_finish<true>, dim3(1), dim3(FINISH_BLOCK), 0, stream, s.partials,
                     n_blocks, static_cast<void *>(d_pair));
  return check_hip(hipGetLastError(), "k_reduce_finish launch");
}

// Autotuner table: [encoding][bit width 0..64, 65 = default for the encoding].
constexpr int N_ENCODINGS = 10;
static int g_tuning[N_ENCODINGS][66];
static std::once_flag g_tuning_once;
static void tuning_defaults() {
  // Measured defaults on MI355X (FFOR sweep at 2^28 values): two threads per
  // lane for 32-bit types, four for 64-bit types, streaming stores.
  for (int e = 0; e < N_ENCODINGS; ++e) {
    const bool wide = e == FLSGPU_ENC_BP_U64 || e == FLSGPU_ENC_FFOR_U64 || e == FLSGPU_ENC_ALP_F64 ||
                      e == FLSGPU_ENC_ALP_EXT_F64 || e == FLSGPU_ENC_ALPRD_F64;
    for (int b = 0; b < 66; ++b)
      g_tuning[e][b] = wide ? FLSGPU_VARIANT_QUARTER_LANE_NT : FLSGPU_VARIANT_HALF_LANE_NT;
  }
}

}  // namespace flsgpu_hip

using namespace flsgpu_hip;

extern "C" {

const char *flsgpu_last_error(void) { return g_err; }

int flsgpu_init(int device) {
  int rc = check_hip(hipSetDevice(device), "hipSetDevice");
  if (rc) return rc;
  if (device < 0 || device >= MAX_DEVICES) return set_error(FLSGPU_EINVAL, "bad device");
  std::lock_guard<std::mutex> lk(g_mu);
  return init_device_locked(device);
}

int flsgpu_set_tuning(uint32_t encoding, uint32_t bit_width, int variant) {
  std::call_once(g_tuning_once, tuning_defaults);
  if (encoding >= uint32_t(N_ENCODINGS)) return set_error(FLSGPU_EINVAL, "unknown encoding %u", encoding);
  if (variant <= FLSGPU_VARIANT_AUTO || variant >= FLSGPU_VARIANT_COUNT)
    return set_error(FLSGPU_EINVAL, "variant %d out of range", variant);
  const uint32_t w = (encoding == FLSGPU_ENC_BP_U32 || encoding == FLSGPU_ENC_FFOR_U32 ||
                      encoding == FLSGPU_ENC_ALP_F32 || encoding == FLSGPU_ENC_ALP_EXT_F32 ||
                      encoding == FLSGPU_ENC_ALPRD_F32) ? 32 : 64;
  if (bit_width == 255) {
    g_tuning[encoding][65] = variant;
  } else {
    if (bit_width > w) return set_error(FLSGPU_EINVAL, "bit width %u > %u", bit_width, w);
    g_tuning[encoding][bit_width] = variant;
  }
  return FLSGPU_OK;
}

int flsgpu_get_tuning(uint32_t encoding, uint32_t bit_width) {
  std::call_once(g_tuning_once, tuning_defaults);
  if (encoding >= uint32_t(N_ENCODINGS)) return FLSGPU_VARIANT_LANE;
  if (bit_width == 255 || bit_width > 64) return g_tuning[encoding][65];
  return g_tuning[encoding][bit_width];
}

int flsgpu_upload(void **d_ptr, const void *h_ptr, size_t bytes, size_t pad_bytes,
                  flsgpu_stream_t stream) {
  if (!d_ptr || (!h_ptr && bytes)) return set_error(FLSGPU_EINVAL, "null pointer");
  void *p = nullptr;
  const size_t total = bytes + pad_bytes;
  int rc = check_hip(hipMalloc(&p, total ? total : 1), "hipMalloc");
  if (rc) return rc;
  if (bytes && (rc = check_hip(hipMemcpyAsync(p, h_ptr, bytes, hipMemcpyHostToDevice, stream),
                               "hipMemcpyAsync H2D"))) {
    (void)hipFree(p);
    return rc;
  }
  if (pad_bytes && (rc = check_hip(hipMemsetAsync(static_cast<char *>(p) + bytes, 0, pad_bytes, stream),
                                   "hipMemsetAsync"))) {
    (void)hipFree(p);
    return rc;
  }
  if ((rc = check_hip(hipStreamSynchronize(stream), "hipStreamSynchronize"))) {
    (void)hipFree(p);
    return rc;
  }
  *d_ptr = p;
  return FLSGPU_OK;
}

int flsgpu_download(void *h_ptr, const void *d_ptr, size_t bytes, flsgpu_stream_t stream) {
  if (!bytes) return FLSGPU_OK;
  if (!h_ptr || !d_ptr) return set_error(FLSGPU_EINVAL, "null pointer");
  int rc = check_hip(hipMemcpyAsync(h_ptr, d_ptr, bytes, hipMemcpyDeviceToHost, stream),
                     "hipMemcpyAsync D2H");
  if (rc) return rc;
  return check_hip(hipStreamSynchronize(stream), "hipStreamSynchronize");
}

int flsgpu_release_stream(flsgpu_stream_t stream) {
  int dev;
  int rc = current_device(&dev);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(g_mu);
  auto &v = g_scratch[dev];
  for (size_t i = 0; i < v.size(); ++i) {
    if (v[i].stream != stream) continue;
    // the stream's last reduction may still read the area
    if ((rc = check_hip(stream ? hipStreamSynchronize(stream) : hipDeviceSynchronize(),
                        "synchronize before release")))
      return rc;
    rc = check_hip(hipFree(v[i].partials), "hipFree(scratch)");
    v.erase(v.begin() + i);
    return rc;
  }
  return FLSGPU_OK;
}

int flsgpu_synchronize(flsgpu_stream_t stream) {
  if (!stream) return check_hip(hipDeviceSynchronize(), "hipDeviceSynchronize");
  return check_hip(hipStreamSynchronize(stream), "hipStreamSynchronize");
}

int flsgpu_free(void *d_ptr) {
  if (!d_ptr) return FLSGPU_OK;
  return check_hip(hipFree(d_ptr), "hipFree");
}

}  // extern "C"
