// Batched, mixed-encoding decode in one persistent launch.
//
// No reference counterpart: the reference launches one kernel per column and
// synchronises after each (src/engine/kernels.cuh:94-112,
// src/engine/multi-column-host-kernels.cuh:31-115). Here a plan holds a device
// table of column descriptors (FFOR/BP, ALP, G-ALP, ALP-RD at any width) and
// a task list (TASK_UNITS wave-units of one column per task, column after
// column). Every wave of a persistent grid walks the task list on its own
// (task t goes to wave t mod the grid's waves), so consecutive tasks, which
// belong to one column, run on different CUs and each wave sees a mix of
// columns and encodings. No block barrier, atomic or queue is involved: a
// wave's loop is bounded by the task count and nothing a wave does waits for
// another wave, so one launch decodes the batch and can be captured in a
// hipGraph as it is.
//
// Each encoding runs an LDS-staged decode (decode into the wave's LDS stage,
// exceptions patched there, 16-byte non-temporal stores) at one lane split
// per element width (batch_split).
#include <cstdlib>
#include <cstring>
#include <new>
#include <vector>

#include "alp_kernels.hpp"
#include "alprd_kernels.hpp"
#include "fls_kernels.hpp"

namespace flsgpu_hip {

// Wave-units per task. One unit per task balances the waves best: on 96
// columns of 2^12..2^24 values (FFOR u32, G-ALP f64, ALP-RD f64) the launch
// took 378 us at 1 unit per task, 388 us at 2, 395-398 us at 4 and 422 us
// at 8; the task list costs 4 bytes per unit.
constexpr int TASK_UNITS = 1;

// Lane split of the batch's wave-units, one per element width: every
// encoding of that width shares one unpack body (see batch_unit). Each unit
// loads its vectors' metadata and exception-list heads, so finer splits pay
// that cost more often: on 96 columns of 2^12..2^24 values (FFOR u32, G-ALP
// f64, ALP-RD f64) 32-bit S = 1 with 64-bit S = 4 took 370-371 us, 32-bit
// S = 2 376-377 us, 32-bit S = 4 396-398 us, 64-bit S = 8 454-457 us (at
// five waves per SIMD 444-447 us), 64-bit S = 2 (16 KB stages, two waves
// per SIMD) 387-392 us.
template <typename UT>
constexpr int batch_split() {
  return sizeof(UT) == 8 ? 4 : 1;
}

// LDS stage per wave: VPW * CHUNK elements of either shape (64-bit at S = 4:
// 4 * 16 * 16 * 8 bytes; 32-bit at S = 1: 2 * 32 * 32 * 4 bytes).
constexpr int BATCH_STAGE_WORDS = 1024;  // 64-bit words
static_assert(lanes<uint64_t>::VPW * StageShape<uint64_t, batch_split<uint64_t>()>::CHUNK * 8 <=
                  BATCH_STAGE_WORDS * 8 &&
              lanes<uint32_t>::VPW * StageShape<uint32_t, batch_split<uint32_t>()>::CHUNK * 4 <=
                  BATCH_STAGE_WORDS * 8,
              "batch stage");

enum BatchKind : uint32_t { BK_FFOR = 0, BK_ALP_STD = 1, BK_ALP_GALP = 2, BK_RD = 3 };

template <typename UT>
struct float_of;
template <>
struct float_of<uint32_t> {
  using type = float;
};
template <>
struct float_of<uint64_t> {
  using type = double;
};

struct BatchColumn {
  uint32_t enc;   // flsgpu_encoding
  uint32_t kind;  // BatchKind
  void *out;
  uint64_t first_task;
  uint64_t n_units;
  union {
    FforArgs<uint32_t> f32;
    FforArgs<uint64_t> f64;
    AlpArgs<float> a32;
    AlpArgs<double> a64;
    AlprdArgs<float> r32;
    AlprdArgs<double> r64;
  } u;
};

// Makes the compiler wait for x at this point (an empty asm that reads it):
// the wait for a loaded value is then placed before the loads that follow,
// not where the value is first used, which may come after them (the load
// counter drains in issue order, so waiting there would also wait for the
// later loads).
template <typename X>
__device__ __forceinline__ void pin(const X &x) {
#if defined(__HIP_DEVICE_COMPILE__)  // the device pass; a host build (tests/emu) has no VGPRs
  asm volatile("" ::"v"(x));
#else
  (void)x;
#endif
}

// The batch's unpack into the stage: the FOR base added and, for ALP
// columns (`alp`, wave-uniform), the integer-to-float formula applied before
// the cell is written, so the stage is not read and written again for it.
template <typename UT>
struct BatchStagePolicy {
  using T = typename float_of<UT>::type;
  using ST = typename int_of<T>::type;
  const UT *in;
  UT *lds;
  UT base;
  bool alp;
  UT fact;
  T frac;
  __device__ __forceinline__ UT value(UT v) const {
    const UT x = UT(v + base);
    return alp ? __builtin_bit_cast(UT, T(ST(UT(x * fact))) * frac) : x;
  }
  template <int BW, int J0, int NJ>
  __device__ __forceinline__ void run() {
    unpack_slots<UT, BW, J0, NJ>(in, [&](int j, UT v) { lds[(j - J0) * lanes<UT>::N] = value(v); });
  }
  __device__ __forceinline__ void run_generic(int bw, int j0, int nj) {
    unpack_runtime<UT>(in, bw, j0, nj, [&](int j, UT v) { lds[(j - j0) * lanes<UT>::N] = value(v); });
  }
};

// One wave-unit (vector group w / S, lane part w % S) of a column of element
// width UT, any encoding. The encodings share one unpack body: the packed
// integers plus the FOR base are decoded into the wave's LDS stage (ALP: with
// the integer-to-float formula applied on the way), ALP-RD threads then join
// the dictionary's left part to their own cells' right parts, the wave
// patches the exceptions it loaded before the decode, and the stage is stored
// with 16-byte non-temporal stores. `dshift`: an ALP-RD column's pre-shifted
// dictionary in LDS.
template <typename UT>
__device__ __forceinline__ void batch_unit(const BatchColumn &c, uint32_t kind, uint64_t w,
                                           UT *stage, const UT *dshift) {
  using T = typename float_of<UT>::type;
  using ST = typename int_of<T>::type;
  using L = lanes<UT>;
  constexpr int S = batch_split<UT>();
  using Sh = StageShape<UT, S>;
  constexpr int VPW = L::VPW;
  static_assert(offsetof(AlpArgs<T>, ffor) == 0, "ALP arguments start with the FFOR ones");
  const FforArgs<UT> &f = *reinterpret_cast<const FforArgs<UT> *>(&c.u);
  const AlpArgs<T> &al = *reinterpret_cast<const AlpArgs<T> *>(&c.u);
  const AlprdArgs<T> &rd = *reinterpret_cast<const AlprdArgs<T> *>(&c.u);
  const UnitMap<UT, S> u = map_unit<UT, S>(w);
  const int t = int(threadIdx.x % WAVE), vv = t / L::N;
  const int j0 = u.part * Sh::NJ;
  const uint64_t g = uniform_u64(w / S);
  const bool is_rd = kind == BK_RD, is_alp = kind == BK_ALP_STD || kind == BK_ALP_GALP;
  const uint64_t n_vecs = is_rd ? rd.n_vecs : f.n_vecs;
  const uint64_t vs = u.vec < n_vecs ? u.vec : 0;  // past the end: vector 0, never stored

  // 1. metadata and exception list headers
  const UT *in;
  int bw;
  UT base = 0;
  unsigned fi = 0, ei = 0;
  uint64_t eo[VPW];
  unsigned raw[VPW];
  LeftWords<UT> left;
#pragma unroll
  for (int v2 = 0; v2 < VPW; ++v2) {
    eo[v2] = 0;
    raw[v2] = 0;
  }
  if (is_rd) {
    in = rd.right + vs * uint64_t(rd.right_bw) * L::N + u.lane;
    bw = int(rd.right_bw);
    left.load(rd.left + vs * uint64_t(rd.left_bw) * 64, u.lane, rd.left_bw);
#pragma unroll
    for (int v2 = 0; v2 < VPW; ++v2) {
      const uint64_t vec = g * VPW + v2, s2 = vec < n_vecs ? vec : 0;
      eo[v2] = rd.exceptions_offsets[s2];
      raw[v2] = rd.counts[s2];
    }
  } else {
    in = f.packed + f.vector_offsets[vs] + u.lane;
    bw = f.bit_widths[vs];
    base = f.bases ? f.bases[vs] : UT(0);
    if (is_alp) {
      fi = al.factor_indices[vs];
      ei = al.fraction_indices[vs];
#pragma unroll
      for (int v2 = 0; v2 < VPW; ++v2) {
        const uint64_t vec = g * VPW + v2, s2 = vec < n_vecs ? vec : 0;
        eo[v2] = al.exceptions_offsets[s2];
        raw[v2] = kind == BK_ALP_GALP ? unsigned(al.offsets_counts[s2 * L::N + (L::N - 1)])
                                      : unsigned(al.counts[s2]);
      }
    }
  }
  unsigned count[VPW];
#pragma unroll
  for (int v2 = 0; v2 < VPW; ++v2) {
    // G-ALP lanes are concatenated in order: the last lane's offset + count
    const unsigned cnt = kind == BK_ALP_GALP ? (raw[v2] & 0x3FFu) + (raw[v2] >> 10) : raw[v2];
    count[v2] = __builtin_amdgcn_readfirstlane(g * VPW + v2 < n_vecs ? cnt : 0u);
  }
  pin(in);
  pin(bw);
  pin(base);

  // 2. the lists' first 64 entries (entry t to thread t), then the packed
  // words: one round trip covers both
  unsigned ppos[VPW];
  UT pval[VPW];
  if (is_rd) {
#pragma unroll
    for (int v2 = 0; v2 < VPW; ++v2) {
      ppos[v2] = *list_entry(rd.positions, eo[v2], unsigned(t), count[v2]);
      pval[v2] = UT(*list_entry(rd.exceptions, eo[v2], unsigned(t), count[v2]));
    }
  } else {
    const UT *exc = reinterpret_cast<const UT *>(al.exceptions);
#pragma unroll
    for (int v2 = 0; v2 < VPW; ++v2) {
      ppos[v2] = *list_entry(al.positions, eo[v2], unsigned(t), count[v2]);
      pval[v2] = *list_entry(exc, eo[v2], unsigned(t), count[v2]);
    }
  }
  UT *row = stage + vv * Sh::CHUNK + u.lane;
  BatchStagePolicy<UT> pol{in, row, base, is_alp, is_alp ? alp_factor<T>(fi) : UT(0),
                           is_alp ? alp_fraction<T>(ei) : T(0)};
  if (!dispatch_bw<UT, S>(pol, bw, u.part)) {
    for (int jj = 0; jj < Sh::NJ; ++jj) row[jj * L::N] = UT(0);
  }

  // 3. ALP-RD: the thread's own cells, the dictionary's left part joined
  if (is_rd) {
    constexpr int Q = 64 / L::N;
    uint64_t lstream[Q];
    left.compose(rd.left_bw, lstream);
    const unsigned lmask = (1u << rd.left_bw) - 1u;
#pragma unroll
    for (int jj = 0; jj < Sh::NJ; ++jj) {
      const int j = j0 + jj;
      uint64_t ls = lstream[0];
#pragma unroll
      for (int q = 1; q < Q; ++q) ls = (j % Q) == q ? lstream[q] : ls;
      const unsigned idx = unsigned(ls >> ((j / Q) * rd.left_bw)) & lmask;
      row[jj * L::N] = UT(dshift[idx] | row[jj * L::N]);
    }
  }

  // 4. exceptions: each position is written once (positions are unique in a
  // vector), by the thread that loaded it
  if (is_alp || is_rd) {
    wave_lds_sync();
    const unsigned lo = unsigned(L::N * j0), hi = unsigned(L::N * (j0 + Sh::NJ));
    const UT rmask = is_rd ? UT((UT(1) << rd.right_bw) - 1) : UT(0);
    const unsigned rbw = is_rd ? rd.right_bw : 0u;
    const uint16_t *positions = is_rd ? rd.positions : al.positions;
#pragma unroll
    for (int v2 = 0; v2 < VPW; ++v2) {
      const unsigned pos = ppos[v2] & 1023u;
      if (unsigned(t) < count[v2] && pos >= lo && pos < hi) {
        UT &x = stage[v2 * Sh::CHUNK + (pos - lo)];
        x = is_rd ? UT(UT(pval[v2] << rbw) | (x & rmask)) : pval[v2];
      }
      for (unsigned k = unsigned(t + WAVE); k < count[v2]; k += WAVE) {
        const unsigned p2 = positions[eo[v2] + k] & 1023u;
        if (p2 >= lo && p2 < hi) {
          UT &x = stage[v2 * Sh::CHUNK + (p2 - lo)];
          x = is_rd ? UT(UT(UT(rd.exceptions[eo[v2] + k]) << rbw) | (x & rmask))
                    : reinterpret_cast<const UT *>(al.exceptions)[eo[v2] + k];
        }
      }
    }
  }
  wave_lds_sync();
  store_staged<UT, S>(stage, static_cast<UT *>(c.out), g, u.part, n_vecs);
}

template <typename UT>
__device__ __forceinline__ void batch_task(const BatchColumn &c, uint64_t u0, uint64_t u1,
                                           UT *stage, UT *dict) {
  using T = typename float_of<UT>::type;
  const uint32_t kind = c.kind;
  if (kind == BK_RD) {  // this column's dictionary, pre-shifted, in the wave's LDS
    const AlprdArgs<T> &rd = *reinterpret_cast<const AlprdArgs<T> *>(&c.u);
    const unsigned lane = threadIdx.x % WAVE;
    wave_lds_sync();  // the previous task's dictionary reads are done
    if (lane < 8) dict[lane] = lane < rd.dict_size ? UT(UT(rd.dict[lane]) << rd.right_bw) : UT(0);
  }
  for (uint64_t w = u0; w < u1; ++w) {
    wave_lds_sync();  // the previous unit's stage reads (and the dictionary writes) are done
    batch_unit<UT>(c, kind, w, stage, dict);
  }
}

__global__ __launch_bounds__(BLOCK) void k_batch(const BatchColumn *__restrict cols,
                                                 const uint32_t *__restrict task_col,
                                                 uint64_t n_tasks) {
  __shared__ __align__(16) uint64_t s_stage[BLOCK / WAVE][BATCH_STAGE_WORDS];
  __shared__ uint64_t s_dict[BLOCK / WAVE][8];
  const int wib = int(threadIdx.x / WAVE);
  const uint64_t stride = uint64_t(gridDim.x) * (BLOCK / WAVE);
  for (uint64_t t = uniform_u64(uint64_t(blockIdx.x) * (BLOCK / WAVE) + wib); t < n_tasks;
       t += stride) {
    const BatchColumn &c = cols[task_col[t]];
    const uint64_t u0 = (t - c.first_task) * TASK_UNITS;
    const uint64_t u1 = u0 + TASK_UNITS < c.n_units ? u0 + TASK_UNITS : c.n_units;
    switch (c.enc) {
      case FLSGPU_ENC_BP_U32:
      case FLSGPU_ENC_FFOR_U32:
      case FLSGPU_ENC_ALP_F32:
      case FLSGPU_ENC_ALP_EXT_F32:
      case FLSGPU_ENC_ALPRD_F32:
        batch_task<uint32_t>(c, u0, u1, reinterpret_cast<uint32_t *>(s_stage[wib]),
                             reinterpret_cast<uint32_t *>(s_dict[wib]));
        break;
      default:
        batch_task<uint64_t>(c, u0, u1, s_stage[wib], s_dict[wib]);
        break;
    }
  }
}

}  // namespace flsgpu_hip

using namespace flsgpu_hip;

struct flsgpu_batch_plan {
  BatchColumn *d_cols = nullptr;
  uint32_t *d_task_col = nullptr;
  uint64_t n_tasks = 0;
  unsigned grid = 1;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t graph_stream = nullptr;
};

namespace {

// Wave-units of a column in the batch (its vectors' groups times the lane
// split).
uint64_t units_of(uint32_t enc, uint64_t n_vecs) {
  switch (enc) {
    case FLSGPU_ENC_BP_U32:
    case FLSGPU_ENC_FFOR_U32:
    case FLSGPU_ENC_ALP_F32:
    case FLSGPU_ENC_ALP_EXT_F32:
    case FLSGPU_ENC_ALPRD_F32: return n_waves_for<uint32_t, batch_split<uint32_t>()>(n_vecs);
    default: return n_waves_for<uint64_t, batch_split<uint64_t>()>(n_vecs);
  }
}

uint32_t kind_of(uint32_t enc) {
  switch (enc) {
    case FLSGPU_ENC_ALP_F32:
    case FLSGPU_ENC_ALP_F64: return BK_ALP_STD;
    case FLSGPU_ENC_ALP_EXT_F32:
    case FLSGPU_ENC_ALP_EXT_F64: return BK_ALP_GALP;
    case FLSGPU_ENC_ALPRD_F32:
    case FLSGPU_ENC_ALPRD_F64: return BK_RD;
    default: return BK_FFOR;
  }
}

void destroy(flsgpu_batch_plan *p) {
  if (!p) return;
  if (p->exec) (void)hipGraphExecDestroy(p->exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  if (p->graph_stream) (void)hipStreamDestroy(p->graph_stream);
  (void)hipFree(p->d_cols);
  (void)hipFree(p->d_task_col);
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
    c.kind = kind_of(c.enc);
    c.n_units = units_of(c.enc, n_vecs);
    const uint64_t tasks = (c.n_units + TASK_UNITS - 1) / TASK_UNITS;
    for (uint64_t k = 0; k < tasks; ++k) task_col.push_back(i);
  }
  auto *p = new (std::nothrow) flsgpu_batch_plan();
  if (!p) return set_error(FLSGPU_ENOMEM, "out of host memory");
  // every encoding runs its one batch shape (batch_split); `variant` selects
  // nothing here
  (void)variant;
  p->n_tasks = task_col.size();
  int rc;
  if ((rc = check_hip(hipMalloc(&p->d_cols, sizeof(BatchColumn) * (cols.size() + 1)), "hipMalloc")) ||
      (rc = check_hip(hipMalloc(&p->d_task_col, sizeof(uint32_t) * (task_col.size() + 1)), "hipMalloc")) ||
      (rc = check_hip(hipMemcpy(p->d_cols, cols.data(), sizeof(BatchColumn) * cols.size(),
                                hipMemcpyHostToDevice), "hipMemcpy")) ||
      (rc = check_hip(hipMemcpy(p->d_task_col, task_col.data(), sizeof(uint32_t) * task_col.size(),
                                hipMemcpyHostToDevice), "hipMemcpy"))) {
    destroy(p);
    return rc;
  }
  // as many blocks as fit on the device at once (every wave then walks
  // its share of the task list), fewer for a small batch
  int per_cu = 0;
  if ((rc = check_hip(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_batch, BLOCK, 0),
                      "hipOccupancyMaxActiveBlocksPerMultiprocessor"))) {
    destroy(p);
    return rc;
  }
  const uint64_t max_blocks = uint64_t(num_cus()) * uint64_t(per_cu > 0 ? per_cu : 1);
  const uint64_t want = (p->n_tasks + (BLOCK / WAVE) - 1) / (BLOCK / WAVE);
  p->grid = unsigned(want < max_blocks ? (want ? want : 1) : max_blocks);
  *plan_out = p;
  return FLSGPU_OK;
}

uint64_t flsgpu_batch_plan_tasks(const flsgpu_batch_plan *plan) { return plan ? plan->n_tasks : 0; }

int flsgpu_batch_run(flsgpu_batch_plan *plan, flsgpu_stream_t stream) {
  if (!plan) return set_error(FLSGPU_EINVAL, "null plan");
  if (plan->n_tasks == 0) return FLSGPU_OK;
  hipLaunchKernelGGL(k_batch, dim3(plan->grid), dim3(BLOCK), 0, stream, plan->d_cols,
                     plan->d_task_col, plan->n_tasks);
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
