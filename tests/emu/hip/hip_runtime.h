// Host-side emulation of the HIP features the library's encoder and ALP
// decoder kernels use, so that their code runs on the CPU (under
// AddressSanitizer) in tests/test_kernel_emulation.py. Not a HIP runtime:
// one OS thread per GPU thread, blocks run one after another,
// __syncthreads is a block barrier, the wave barrier a 64-thread barrier,
// shuffles exchange through a slot array inside groups of `emu_group`
// threads (every shuffle in these kernels stays inside one vector's
// N_LANES-thread group), readfirstlane is the identity and the ballot
// answers with a fixed mode: 0 sends every lane down the uniform-width path
// with its own width, 1 down the generic mixed-width path.
#pragma once
#include <barrier>
#include <climits>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <thread>
#include <type_traits>
#include <vector>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __noinline__
#define __launch_bounds__(...)
#define __constant__
#define __shared__ static
#define __align__(x) __attribute__((aligned(x)))

struct dim3 {
  unsigned x, y, z;
  dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
inline thread_local dim3 threadIdx, blockIdx;
inline dim3 gridDim, blockDim;
typedef struct ihipStream_t *hipStream_t;
typedef int hipError_t;
enum { hipSuccess = 0 };
inline hipError_t hipGetLastError() { return hipSuccess; }
inline const char *hipGetErrorString(hipError_t) { return "emulation"; }

inline int __clz(int x) { return x == 0 ? 32 : __builtin_clz(unsigned(x)); }
inline int __clzll(long long x) { return x == 0 ? 64 : __builtin_clzll((unsigned long long)x); }
#define __builtin_nontemporal_store(v, p) (*(p) = (v))
#define __builtin_amdgcn_readfirstlane(x) (x)
// v_alignbit_b32: the low 32 bits of the 64-bit {hi, lo} shifted right by s (mod 32)
inline uint32_t emu_alignbit(uint32_t hi, uint32_t lo, uint32_t s) {
  return uint32_t(((uint64_t(hi) << 32) | lo) >> (s & 31u));
}
#define __builtin_amdgcn_alignbit(hi, lo, s) emu_alignbit(hi, lo, s)
#define __builtin_amdgcn_fence(...)
inline int emu_ballot_mode = 0;
#define __builtin_amdgcn_ballot_w64(p) ((void)(p), uint64_t(emu_ballot_mode))
inline unsigned long long atomicOr(unsigned long long *p, unsigned long long v) {
  return __atomic_fetch_or(p, v, __ATOMIC_SEQ_CST);
}
inline unsigned atomicOr(unsigned *p, unsigned v) { return __atomic_fetch_or(p, v, __ATOMIC_SEQ_CST); }

struct EmuBlock {
  std::unique_ptr<std::barrier<>> block;
  std::vector<std::unique_ptr<std::barrier<>>> waves;
  std::vector<std::unique_ptr<std::barrier<>>> groups;
  uint64_t slot[1024];
  int group = 64;
};
inline EmuBlock g_emu;
inline int emu_group = 64;

inline void __syncthreads() { g_emu.block->arrive_and_wait(); }
inline void emu_wave_barrier() { g_emu.waves[threadIdx.x / 64]->arrive_and_wait(); }
#define __builtin_amdgcn_wave_barrier() emu_wave_barrier()

template <typename T>
inline T emu_exchange(T v, unsigned src) {
  static_assert(sizeof(T) <= 8, "64-bit shuffles at most");
  const unsigned t = threadIdx.x;
  uint64_t b = 0;
  std::memcpy(&b, &v, sizeof(T));
  g_emu.slot[t] = b;
  auto &bar = *g_emu.groups[t / unsigned(g_emu.group)];
  bar.arrive_and_wait();
  const uint64_t r = g_emu.slot[src];
  bar.arrive_and_wait();
  T out;
  std::memcpy(&out, &r, sizeof(T));
  return out;
}
template <typename T>
inline T __shfl_xor(T v, int mask, int width = 64) {
  (void)width;
  if (unsigned(mask) >= unsigned(g_emu.group)) {
    std::fprintf(stderr, "emulation: xor mask %d leaves the %d-thread group\n", mask, g_emu.group);
    std::abort();
  }
  return emu_exchange(v, threadIdx.x ^ unsigned(mask));
}
template <typename T>
inline T __shfl(T v, int src, int width = 64) {
  const unsigned t = threadIdx.x, s = t - t % unsigned(width) + unsigned(src) % unsigned(width);
  if (s / unsigned(g_emu.group) != t / unsigned(g_emu.group)) {
    std::fprintf(stderr, "emulation: shfl source %u leaves the %d-thread group\n", s, g_emu.group);
    std::abort();
  }
  return emu_exchange(v, s);
}
template <typename T>
inline T __shfl_up(T v, unsigned delta, int width = 64) {
  if (width != g_emu.group) {
    std::fprintf(stderr, "emulation: shfl_up width %d, group %d\n", width, g_emu.group);
    std::abort();
  }
  const unsigned t = threadIdx.x, lane = t % unsigned(width);
  return emu_exchange(v, lane >= delta ? t - delta : t);
}

template <class K, class... A>
inline void emu_launch(K kernel, dim3 grid, dim3 block, A... args) {
  gridDim = grid;
  blockDim = block;
  for (unsigned b = 0; b < grid.x; ++b) {
    g_emu.block = std::make_unique<std::barrier<>>(block.x);
    g_emu.group = emu_group < int(block.x) ? emu_group : int(block.x);
    g_emu.groups.clear();
    g_emu.waves.clear();
    for (unsigned g = 0; g < block.x / unsigned(g_emu.group); ++g)
      g_emu.groups.push_back(std::make_unique<std::barrier<>>(g_emu.group));
    for (unsigned w = 0; w < (block.x + 63) / 64; ++w)
      g_emu.waves.push_back(std::make_unique<std::barrier<>>(block.x - 64 * w < 64 ? block.x - 64 * w : 64));
    std::vector<std::thread> th;
    th.reserve(block.x);
    for (unsigned t = 0; t < block.x; ++t)
      th.emplace_back([=]() {
        threadIdx = dim3(t);
        blockIdx = dim3(b);
        kernel(args...);
      });
    for (auto &x : th) x.join();
  }
}
#define hipLaunchKernelGGL(k, g, b, sh, st, ...) emu_launch(k, dim3(g), dim3(b), __VA_ARGS__)
template <class K>
inline hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor(int *n, K, int, size_t) {
  *n = 2;
  return hipSuccess;
}

// The batch kernel's plan: device memory is host memory here, graph capture
// is unavailable.
enum hipMemcpyKind { hipMemcpyHostToDevice = 1, hipMemcpyDeviceToHost = 2 };
template <class T>
inline hipError_t hipMalloc(T **p, size_t n) {
  *p = static_cast<T *>(std::malloc(n ? n : 1));
  return *p ? hipSuccess : 1;
}
inline hipError_t hipFree(void *p) {
  std::free(p);
  return hipSuccess;
}
inline hipError_t hipMemcpy(void *d, const void *s, size_t n, hipMemcpyKind) {
  std::memcpy(d, s, n);
  return hipSuccess;
}
inline hipError_t hipMemset(void *d, int v, size_t n) {
  std::memset(d, v, n);
  return hipSuccess;
}
inline unsigned long long atomicAdd(unsigned long long *p, unsigned long long v) {
  return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST);
}
inline void __threadfence() { __atomic_thread_fence(__ATOMIC_SEQ_CST); }
#define __HIP_MEMORY_SCOPE_AGENT 0
#define __hip_atomic_store(p, v, order, scope) __atomic_store_n((p), (v), __ATOMIC_SEQ_CST)
#define amdgpu_waves_per_eu(...)
typedef struct ihipGraph_t *hipGraph_t;
typedef struct ihipGraphExec_t *hipGraphExec_t;
enum { hipStreamNonBlocking = 1, hipStreamCaptureModeThreadLocal = 1 };
inline hipError_t hipStreamCreateWithFlags(hipStream_t *, unsigned) { return 1; }
inline hipError_t hipStreamDestroy(hipStream_t) { return 1; }
inline hipError_t hipStreamBeginCapture(hipStream_t, int) { return 1; }
inline hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t *) { return 1; }
inline hipError_t hipGraphInstantiate(hipGraphExec_t *, hipGraph_t, void *, char *, size_t) { return 1; }
inline hipError_t hipGraphLaunch(hipGraphExec_t, hipStream_t) { return 1; }
inline hipError_t hipGraphDestroy(hipGraph_t) { return 1; }
inline hipError_t hipGraphExecDestroy(hipGraphExec_t) { return 1; }
