// Shared device/host definitions for the gfx950 decoder kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <type_traits>

#include "flsgpu.h"

namespace flsgpu_hip {

constexpr int VALUES_PER_VECTOR = 1024;  // consts::VALUES_PER_VECTOR (src/flsgpu/consts.cuh:10)
constexpr int WAVE = 64;                 // CDNA wavefront
constexpr int BLOCK = 256;               // four waves per workgroup

template <typename UT>
struct lanes {
  static constexpr int W = int(sizeof(UT) * 8);
  static constexpr int N = 1024 / W;        // FastLanes lanes per vector
  static constexpr int VPW = WAVE / N;      // vectors per wavefront (one thread per lane)
};

template <typename T>
struct uint_of {
  using type = std::conditional_t<sizeof(T) == 8, uint64_t,
               std::conditional_t<sizeof(T) == 4, uint32_t, uint16_t>>;
};
template <typename T>
struct int_of {
  using type = std::conditional_t<sizeof(T) == 8, int64_t, int32_t>;
};

// ALP tables, one copy per compilation unit (no relocatable device code).
// src/alp/constants.hpp:49-55 (float), :76-144 (double)
// Index 10 is not in the reference's float table (src/alp/constants.hpp:55,
// ten entries) and the encoder never emits it; it holds 10^10 mod 2^32 so
// that a corrupt factor index of 10 reads defined memory.
static __constant__ int32_t c_fact_f32[11] = {1,       10,       100,       1000,
                                              10000,   100000,   1000000,   10000000,
                                              100000000, 1000000000, 1410065408};
static __constant__ float c_frac_f32[11] = {1.0,      0.1,       0.01,       0.001,       0.0001,     0.00001,
                                     0.000001, 0.0000001, 0.00000001, 0.000000001, 0.0000000001};
static __constant__ int64_t c_fact_f64[19] = {1LL,
                                       10LL,
                                       100LL,
                                       1000LL,
                                       10000LL,
                                       100000LL,
                                       1000000LL,
                                       10000000LL,
                                       100000000LL,
                                       1000000000LL,
                                       10000000000LL,
                                       100000000000LL,
                                       1000000000000LL,
                                       10000000000000LL,
                                       100000000000000LL,
                                       1000000000000000LL,
                                       10000000000000000LL,
                                       100000000000000000LL,
                                       1000000000000000000LL};
static __constant__ double c_frac_f64[21] = {1.0,
                                      0.1,
                                      0.01,
                                      0.001,
                                      0.0001,
                                      0.00001,
                                      0.000001,
                                      0.0000001,
                                      0.00000001,
                                      0.000000001,
                                      0.0000000001,
                                      0.00000000001,
                                      0.000000000001,
                                      0.0000000000001,
                                      0.00000000000001,
                                      0.000000000000001,
                                      0.0000000000000001,
                                      0.00000000000000001,
                                      0.000000000000000001,
                                      0.0000000000000000001,
                                      0.00000000000000000001};

template <typename T>
__device__ __forceinline__ typename uint_of<T>::type alp_factor(unsigned idx);
template <>
__device__ __forceinline__ uint32_t alp_factor<float>(unsigned idx) {
  return uint32_t(c_fact_f32[idx]);
}
template <>
__device__ __forceinline__ uint64_t alp_factor<double>(unsigned idx) {
  return uint64_t(c_fact_f64[idx]);
}
template <typename T>
__device__ __forceinline__ T alp_fraction(unsigned idx);
template <>
__device__ __forceinline__ float alp_fraction<float>(unsigned idx) {
  return c_frac_f32[idx];
}
template <>
__device__ __forceinline__ double alp_fraction<double>(unsigned idx) {
  return c_frac_f64[idx];
}

template <typename T, bool NT>
__device__ __forceinline__ void store(T *p, T v) {
  if constexpr (NT) {
    __builtin_nontemporal_store(v, p);
  } else {
    *p = v;
  }
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, WAVE);
  return v;
}

// Scratch owned by the library for cross-block reductions: one partial pair
// per block of the work kernel, one area per (device, stream), so that
// reductions in flight on different streams never share partials.
constexpr int MAX_REDUCE_BLOCKS = 65536;
struct Scratch {
  uint64_t *partials;  // [MAX_REDUCE_BLOCKS * 2]
};

constexpr int MAX_DEVICES = 64;

// Status helpers (capi.hip)
int set_error(int code, const char *fmt, ...);
int check_hip(hipError_t e, const char *what);
// The reduction area of `stream` on the current device (allocated on the
// stream's first reduction, kept until flsgpu_release_stream or exit).
int get_scratch(hipStream_t stream, Scratch *s);
int num_cus();
// Initialises the current device on first use (properties, the static
// scratch address below) and returns its index.
int ready_device(int *dev);
// Looks up this device's copy of the static scratch array (noalloc.hip);
// called once per device, with that device current, by the initialisation.
int setup_static_scratch(int dev);
// Grid for grid-stride reduction kernels: one wave per unit of work, at
// most MAX_REDUCE_BLOCKS blocks.
unsigned reduce_grid(uint64_t n_waves);
// Finishing launches (one block) over the partials of n_blocks blocks:
// *d_flag = (OR of the first values) != 0; d_pair[0..1] = sums of the pairs.
int reduce_finish_flag(const Scratch &s, unsigned n_blocks, uint8_t *d_flag, hipStream_t stream);
int reduce_finish_pair(const Scratch &s, unsigned n_blocks, uint64_t *d_pair, hipStream_t stream);

}  // namespace flsgpu_hip
