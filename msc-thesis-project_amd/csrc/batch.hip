// Batched, mixed-encoding decode in one persistent launch.
//
// No reference counterpart: the reference launches one kernel per column and
// synchronises after each (src/engine/kernels.cuh:94-112,
// src/engine/multi-column-host-kernels.cuh:31-115). Here a plan holds a device
// table of column descriptors (FFOR/BP, ALP, G-ALP, ALP-RD at any width) and
// a task list; a persistent grid pulls tasks (TASK_WAVES wave-units of one
// column each) from a device-side atomic queue, so small and large columns of
// different encodings share the chip. The queue re-arms itself when the last
// block exits, so the single launch can be captured in a hipGraph and
// replayed with no memset node.
#include <cstdlib>
#include <cstring>
#include <new>
#include <vector>

#include "alp_kernels.hpp"
#include "alprd_kernels.hpp"
#include "fls_kernels.hpp"

namespace flsgpu_hip {

// Wave-units per task: each of the block's 4 waves runs one. Smaller tasks
// balance the tail of a mixed batch better than the queue round trip costs
// (C4, 96 columns: 16 per task 541 us, 8 522 us, 4 513 us).
constexpr int TASK_WAVES = 4;
#ifndef BATCH_NT
#define BATCH_NT true
#endif

struct BatchColumn {
  uint32_t enc;
  uint32_t patch;
  void *out;
  uint64_t first_task;
  union {
    FforArgs<uint32_t> f32;
    FforArgs<uint64_t> f64;
    AlpArgs<float> a32;
    AlpArgs<double> a64;
    AlprdArgs<float> r32;
    AlprdArgs<double> r64;
  } u;
};

// Per-encoding task bodies are inlined into the kernel. As separate
// (non-inlined) functions they follow the call convention, whose
// callee-saved VGPRs sit in blocks of 8 across the whole register file: a
// 64-bit body that needs ~130 caller-saved registers then touches v0..v247
// and the kernel ran one wave per SIMD. Inlined, with a three-wave target,
// the kernel takes 168 VGPRs (a few bytes of spill on the widest paths).
template <typename UT, bool NT>
__device__ __forceinline__ void batch_ffor(const BatchColumn *c, uint64_t w0, int wib) {
  const FforArgs<UT> &a = *reinterpret_cast<const FforArgs<UT> *>(&c->u);
  for (int k = 0; k < TASK_WAVES / 4; ++k)
    ffor_decompress_unit<UT, 1, NT>(a, static_cast<UT *>(c->out), w0 + wib + 4 * k);
}

// G-ALP columns take the per-thread decoder with register patching (no LDS,
// no barriers); standard-layout ones keep the per-thread scan of the list.
template <typename T, int LAYOUT, bool NT>
__device__ __forceinline__ void batch_alp(const BatchColumn *c, uint64_t w0, int wib) {
  const AlpArgs<T> &a = *reinterpret_cast<const AlpArgs<T> *>(&c->u);
  for (int k = 0; k < TASK_WAVES / 4; ++k) {
    if constexpr (LAYOUT == ALP_GALP)
      alp_decompress_unit<T, 1, NT, ALP_GALP, true>(a, static_cast<T *>(c->out), w0 + wib + 4 * k);
    else
      alp_decompress_unit_scan<T, NT, LAYOUT>(a, static_cast<T *>(c->out), w0 + wib + 4 * k);
  }
}

template <typename T, bool NT>
__device__ __forceinline__ void batch_alprd(const BatchColumn *c, uint64_t w0, int wib,
                                         const typename uint_of<T>::type *dict, uint32_t *list) {
  const AlprdArgs<T> &a = *reinterpret_cast<const AlprdArgs<T> *>(&c->u);
  for (int k = 0; k < TASK_WAVES / 4; ++k)
    alprd_decompress_unit<T, NT>(a, static_cast<T *>(c->out), w0 + wib + 4 * k, dict, list);
}

template <bool NT>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(3))) void k_batch(
    const BatchColumn *__restrict cols, const uint32_t *__restrict task_col, uint64_t n_tasks,
    unsigned long long *queue) {
  __shared__ unsigned long long s_task;
  __shared__ uint64_t dict_lds[8];
  static_assert(lanes<uint32_t>::VPW <= lanes<uint64_t>::VPW, "rd_list sized for the widest VPW");
  __shared__ uint32_t rd_list[BLOCK / WAVE][lanes<uint64_t>::VPW * RD_LIST];
  const int wib = int(threadIdx.x / WAVE);
  for (;;) {
    __syncthreads();  // every wave is done with the previous task (and its LDS)
    if (threadIdx.x == 0) s_task = atomicAdd(&queue[0], 1ull);
    __syncthreads();
    const uint64_t t = s_task;
    if (t >= n_tasks) break;
    const BatchColumn *c = cols + task_col[t];
    const uint64_t w0 = (t - c->first_task) * TASK_WAVES;
    switch (c->enc) {
      case FLSGPU_ENC_BP_U32:
      case FLSGPU_ENC_FFOR_U32: batch_ffor<uint32_t, NT>(c, w0, wib); break;
      case FLSGPU_ENC_BP_U64:
      case FLSGPU_ENC_FFOR_U64: batch_ffor<uint64_t, NT>(c, w0, wib); break;
      case FLSGPU_ENC_ALP_F32: batch_alp<float, ALP_STD, NT>(c, w0, wib); break;
      case FLSGPU_ENC_ALP_EXT_F32: batch_alp<float, ALP_GALP, NT>(c, w0, wib); break;
      case FLSGPU_ENC_ALP_F64: batch_alp<double, ALP_STD, NT>(c, w0, wib); break;
      case FLSGPU_ENC_ALP_EXT_F64: batch_alp<double, ALP_GALP, NT>(c, w0, wib); break;
      case FLSGPU_ENC_ALPRD_F32: {
        uint32_t *d = reinterpret_cast<uint32_t *>(dict_lds);
        alprd_fill_dictionary<float>(c->u.r32, d);
        __syncthreads();
        batch_alprd<float, NT>(c, w0, wib, d, rd_list[wib]);
        break;
      }
      case FLSGPU_ENC_ALPRD_F64:
        alprd_fill_dictionary<double>(c->u.r64, dict_lds);
        __syncthreads();
        batch_alprd<double, NT>(c, w0, wib, dict_lds, rd_list[wib]);
        break;
      default:
        break;
    }
  }
  // Re-arm the queue: the last block to leave resets both counters.
  if (threadIdx.x == 0) {
    __threadfence();
    if (atomicAdd(&queue[1], 1ull) == gridDim.x - 1) {
      __hip_atomic_store(&queue[0], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&queue[1], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

}  // namespace flsgpu_hip

using namespace flsgpu_hip;

struct flsgpu_batch_plan {
  BatchColumn *d_cols = nullptr;
  uint32_t *d_task_col = nullptr;
  unsigned long long *d_queue = nullptr;
  uint64_t n_tasks = 0;
  unsigned grid = 1;
  bool nt = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t graph_stream = nullptr;
};

namespace {

uint64_t waves_of(uint32_t enc, uint64_t n_vecs) {
  switch (enc) {
    case FLSGPU_ENC_BP_U32:
    case FLSGPU_ENC_FFOR_U32:
    case FLSGPU_ENC_ALP_F32:
    case FLSGPU_ENC_ALP_EXT_F32:
    case FLSGPU_ENC_ALPRD_F32: return n_waves_for<uint32_t, 1>(n_vecs);
    default: return n_waves_for<uint64_t, 1>(n_vecs);
  }
}

void destroy(flsgpu_batch_plan *p) {
  if (!p) return;
  if (p->exec) (void)hipGraphExecDestroy(p->exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  if (p->graph_stream) (void)hipStreamDestroy(p->graph_stream);
  (void)hipFree(p->d_cols);
  (void)hipFree(p->d_task_col);
  (void)hipFree(p->d_queue);
  delete p;
}

}  // namespace

extern "C" {

int flsgpu_batch_plan_create(const flsgpu_batch_item *items, uint32_t n_items, int variant,
                             flsgpu_batch_plan **plan_out) {
  if (!plan_out || (!items && n_items)) return set_error(FLSGPU_EINVAL, "null argument");
  *plan_out = nullptr;
  std::vector<BatchColumn> cols(n_items);
  std::vector<uint32_t> task_col;
  for (uint32_t i = 0; i < n_items; ++i) {
    const flsgpu_batch_item &it = items[i];
    BatchColumn &c = cols[i];
    std::memset(&c, 0, sizeof(c));
    c.enc = uint32_t(it.encoding);
    c.out = it.out;
    c.first_task = task_col.size();
    if (!it.column) return set_error(FLSGPU_EINVAL, "batch item %u: null column", i);
    int rc = FLSGPU_OK;
    uint64_t n_vecs = 0;
    switch (it.encoding) {
      case FLSGPU_ENC_BP_U32:
        rc = make_ffor_args<uint32_t>(static_cast<const flsgpu_bp_column *>(it.column), nullptr, &c.u.f32);
        n_vecs = c.u.f32.n_vecs;
        break;
      case FLSGPU_ENC_BP_U64:
        rc = make_ffor_args<uint64_t>(static_cast<const flsgpu_bp_column *>(it.column), nullptr, &c.u.f64);
        n_vecs = c.u.f64.n_vecs;
        break;
      case FLSGPU_ENC_FFOR_U32: {
        auto *f = static_cast<const flsgpu_ffor_column *>(it.column);
        rc = f->bp.n_vecs && !f->bases ? set_error(FLSGPU_EINVAL, "null bases")
                                       : make_ffor_args<uint32_t>(&f->bp, f->bases, &c.u.f32);
        n_vecs = c.u.f32.n_vecs;
        break;
      }
      case FLSGPU_ENC_FFOR_U64: {
        auto *f = static_cast<const flsgpu_ffor_column *>(it.column);
        rc = f->bp.n_vecs && !f->bases ? set_error(FLSGPU_EINVAL, "null bases")
                                       : make_ffor_args<uint64_t>(&f->bp, f->bases, &c.u.f64);
        n_vecs = c.u.f64.n_vecs;
        break;
      }
      case FLSGPU_ENC_ALP_F32:
      case FLSGPU_ENC_ALP_EXT_F32:
        rc = make_alp_args<float>(static_cast<const flsgpu_alp_column *>(it.column),
                                  it.encoding == FLSGPU_ENC_ALP_EXT_F32 ? ALP_GALP : ALP_STD,
                                  &c.u.a32);
        n_vecs = c.u.a32.ffor.n_vecs;
        break;
      case FLSGPU_ENC_ALP_F64:
      case FLSGPU_ENC_ALP_EXT_F64:
        rc = make_alp_args<double>(static_cast<const flsgpu_alp_column *>(it.column),
                                   it.encoding == FLSGPU_ENC_ALP_EXT_F64 ? ALP_GALP : ALP_STD,
                                   &c.u.a64);
        n_vecs = c.u.a64.ffor.n_vecs;
        break;
      case FLSGPU_ENC_ALPRD_F32:
        rc = make_rd_args<float>(static_cast<const flsgpu_alprd_column *>(it.column), &c.u.r32);
        n_vecs = c.u.r32.n_vecs;
        break;
      case FLSGPU_ENC_ALPRD_F64:
        rc = make_rd_args<double>(static_cast<const flsgpu_alprd_column *>(it.column), &c.u.r64);
        n_vecs = c.u.r64.n_vecs;
        break;
      default:
        return set_error(FLSGPU_EINVAL, "batch item %u: unknown encoding %d", i, it.encoding);
    }
    if (rc) return rc;
    if (n_vecs && !it.out) return set_error(FLSGPU_EINVAL, "batch item %u: null output", i);
    const bool is_alp = it.encoding >= FLSGPU_ENC_ALP_F32 && it.encoding <= FLSGPU_ENC_ALP_EXT_F64;
    if (is_alp && (it.patcher == FLSGPU_PATCHER_NONE || it.patcher == FLSGPU_PATCHER_DUMMY))
      return set_error(FLSGPU_EUNSUPPORTED,
                       "batch item %u: the batch decoder always patches exceptions", i);
    c.patch = is_alp;
    const uint64_t tasks = (waves_of(c.enc, n_vecs) + TASK_WAVES - 1) / TASK_WAVES;
    for (uint64_t k = 0; k < tasks; ++k) task_col.push_back(i);
  }
  auto *p = new (std::nothrow) flsgpu_batch_plan();
  if (!p) return set_error(FLSGPU_ENOMEM, "out of host memory");
  (void)variant;  // one store flavour is compiled into the batch kernel (BATCH_NT)
  p->nt = BATCH_NT;
  p->n_tasks = task_col.size();
  int rc;
  if ((rc = check_hip(hipMalloc(&p->d_cols, sizeof(BatchColumn) * (cols.size() + 1)), "hipMalloc")) ||
      (rc = check_hip(hipMalloc(&p->d_task_col, sizeof(uint32_t) * (task_col.size() + 1)), "hipMalloc")) ||
      (rc = check_hip(hipMalloc(&p->d_queue, 2 * sizeof(unsigned long long)), "hipMalloc")) ||
      (rc = check_hip(hipMemcpy(p->d_cols, cols.data(), sizeof(BatchColumn) * cols.size(),
                                hipMemcpyHostToDevice), "hipMemcpy")) ||
      (rc = check_hip(hipMemcpy(p->d_task_col, task_col.data(), sizeof(uint32_t) * task_col.size(),
                                hipMemcpyHostToDevice), "hipMemcpy")) ||
      (rc = check_hip(hipMemset(p->d_queue, 0, 2 * sizeof(unsigned long long)), "hipMemset"))) {
    destroy(p);
    return rc;
  }
  const uint64_t max_blocks = uint64_t(num_cus()) * 8;
  p->grid = unsigned(p->n_tasks < max_blocks ? (p->n_tasks ? p->n_tasks : 1) : max_blocks);
  *plan_out = p;
  return FLSGPU_OK;
}

uint64_t flsgpu_batch_plan_tasks(const flsgpu_batch_plan *plan) { return plan ? plan->n_tasks : 0; }

int flsgpu_batch_run(flsgpu_batch_plan *plan, flsgpu_stream_t stream) {
  if (!plan) return set_error(FLSGPU_EINVAL, "null plan");
  if (plan->n_tasks == 0) return FLSGPU_OK;
  hipLaunchKernelGGL((k_batch<BATCH_NT>), dim3(plan->grid), dim3(BLOCK), 0, stream, plan->d_cols,
                     plan->d_task_col, plan->n_tasks, plan->d_queue);
  return check_hip(hipGetLastError(), "k_batch launch");
}

int flsgpu_batch_graph_launch(flsgpu_batch_plan *plan, flsgpu_stream_t stream) {
  if (!plan) return set_error(FLSGPU_EINVAL, "null plan");
  if (plan->n_tasks == 0) return FLSGPU_OK;
  if (!plan->exec) {
    // Capture on a private non-blocking stream: the caller's stream may be
    // the null stream, which cannot be captured. The instantiated graph is
    // then launched on the caller's stream.
    int rc;
    if (!plan->graph_stream &&
        (rc = check_hip(hipStreamCreateWithFlags(&plan->graph_stream, hipStreamNonBlocking),
                        "hipStreamCreateWithFlags")))
      return rc;
    hipStream_t cs = plan->graph_stream;
    if ((rc = check_hip(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal),
                        "hipStreamBeginCapture")))
      return rc;
    const int run_rc = flsgpu_batch_run(plan, cs);
    hipGraph_t g = nullptr;
    rc = check_hip(hipStreamEndCapture(cs, &g), "hipStreamEndCapture");
    if (run_rc) {
      if (g) (void)hipGraphDestroy(g);
      return run_rc;
    }
    if (rc) return rc;
    if ((rc = check_hip(hipGraphInstantiate(&plan->exec, g, nullptr, nullptr, 0),
                        "hipGraphInstantiate"))) {
      (void)hipGraphDestroy(g);
      return rc;
    }
    plan->graph = g;
  }
  return check_hip(hipGraphLaunch(plan->exec, stream), "hipGraphLaunch");
}

int flsgpu_batch_plan_destroy(flsgpu_batch_plan *plan) {
  destroy(plan);
  return FLSGPU_OK;
}

}  // extern "C"
