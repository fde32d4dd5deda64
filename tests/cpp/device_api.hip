// Test kernels for the device-side decoder API (include/flsgpu_device.hpp):
// the reference's decompress_column and query_column kernel shapes
// (src/engine/kernels.cuh:13-54) written against the header, behind small
// extern "C" launchers that the GPU tests call through ctypes.
#include "flsgpu_device.hpp"

using namespace flsgpu::device;

namespace {

constexpr unsigned BLOCK = 64;

template <typename T, unsigned NV, unsigned NVAL, typename DecompressorT, typename ColumnT>
__global__ __launch_bounds__(BLOCK) void k_decompress(ColumnT column, uint64_t n_vecs, T *out) {
  const FillWarpMapping<T, NV> mapping;
  const lane_t lane = mapping.get_lane();
  const vi_t vector_index = mapping.get_vector_index();
  if (vector_index >= n_vecs) return;
  T registers[NV * NVAL];
  out += uint64_t(vector_index) * VALUES_PER_VECTOR;
  DecompressorT it(column, vector_index, lane);
  const bool all_live = uint64_t(vector_index) + NV <= n_vecs;
  for (si_t i = 0; i < mapping.N_VALUES_IN_LANE; i += NVAL) {
    it.unpack_next_into(registers);
    if (all_live) {
      write_registers_to_global<T, NV, NVAL, FillWarpMapping<T, NV>::N_LANES>(lane, i, registers, out);
    } else {
      for (unsigned v = 0; v < NV; ++v)
        if (uint64_t(vector_index) + v < n_vecs)
          for (unsigned w = 0; w < NVAL; ++w)
            out[lane + (i + w) * FillWarpMapping<T, NV>::N_LANES + v * VALUES_PER_VECTOR] = registers[w + v * NVAL];
    }
  }
}

template <typename T, unsigned NV, unsigned NVAL, typename DecompressorT, typename ColumnT>
__global__ __launch_bounds__(BLOCK) void k_query(ColumnT column, uint64_t n_vecs, T magic,
                                                 bool *out) {
  const FillWarpMapping<T, NV> mapping;
  const lane_t lane = mapping.get_lane();
  const vi_t vector_index = mapping.get_vector_index();
  if (vector_index >= n_vecs) return;
  T registers[NV * NVAL];
  MagicChecker<T, NV * NVAL> checker(magic);
  DecompressorT it(column, vector_index, lane);
  for (si_t i = 0; i < mapping.N_VALUES_IN_LANE; i += NVAL) {
    it.unpack_next_into(registers);
    // vectors past the end decode as zeros; keep them out of the check
    for (unsigned v = 0; v < NV; ++v)
      if (uint64_t(vector_index) + v >= n_vecs)
        for (unsigned w = 0; w < NVAL; ++w) registers[w + v * NVAL] = magic + T(1);
    checker.check(registers);
  }
  checker.write_result(out);
}

template <typename T, unsigned NV>
unsigned grid_for(uint64_t n_vecs) {
  const uint64_t per_block = (BLOCK / get_n_lanes<T>()) * NV;
  return unsigned((n_vecs + per_block - 1) / per_block);
}

template <typename T, unsigned NV, unsigned NVAL, typename DecompressorT, typename ColumnT>
int launch_decompress(const ColumnT &c, uint64_t n_vecs, T *out, hipStream_t s) {
  if (n_vecs == 0) return 0;
  hipLaunchKernelGGL((k_decompress<T, NV, NVAL, DecompressorT, ColumnT>),
                     dim3(grid_for<T, NV>(n_vecs)), dim3(BLOCK), 0, s, c, n_vecs, out);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

template <typename T, unsigned NV, unsigned NVAL, typename DecompressorT, typename ColumnT>
int launch_query(const ColumnT &c, uint64_t n_vecs, T magic, bool *out, hipStream_t s) {
  if (n_vecs == 0) return 0;
  hipLaunchKernelGGL((k_query<T, NV, NVAL, DecompressorT, ColumnT>),
                     dim3(grid_for<T, NV>(n_vecs)), dim3(BLOCK), 0, s, c, n_vecs, magic, out);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

// shape: 0 = (1 vector, 1 value), 1 = (1, 4), 2 = (4, 1), 3 = (4, 2)
template <typename T, template <typename, unsigned, unsigned> class Dec, typename ColumnT>
int decompress_shape(int shape, const ColumnT &c, uint64_t n_vecs, T *out, hipStream_t s) {
  switch (shape) {
    case 0: return launch_decompress<T, 1, 1, Dec<T, 1, 1>>(c, n_vecs, out, s);
    case 1: return launch_decompress<T, 1, 4, Dec<T, 1, 4>>(c, n_vecs, out, s);
    case 2: return launch_decompress<T, 4, 1, Dec<T, 4, 1>>(c, n_vecs, out, s);
    case 3: return launch_decompress<T, 4, 2, Dec<T, 4, 2>>(c, n_vecs, out, s);
    default: return -1;
  }
}

template <typename T, unsigned NV, unsigned NVAL>
using AlpStd = ALPDecompressor<T, NV, NVAL, ALPColumn<T>, true>;
template <typename T, unsigned NV, unsigned NVAL>
using AlpExt = ALPDecompressor<T, NV, NVAL, ALPExtendedColumn<T>, true>;
template <typename T, unsigned NV, unsigned NVAL>
using AlpDummy = ALPDecompressor<T, NV, NVAL, ALPColumn<T>, false>;

template <typename T>
int alp_decompress(const flsgpu_alp_column *c, T *out, int layout, int shape, hipStream_t s) {
  const uint64_t n = c->ffor.bp.n_vecs;
  if (layout == 0) return decompress_shape<T, AlpStd>(shape, ALPColumn<T>(*c), n, out, s);
  if (layout == 1) return decompress_shape<T, AlpExt>(shape, ALPExtendedColumn<T>(*c), n, out, s);
  return decompress_shape<T, AlpDummy>(shape, ALPColumn<T>(*c), n, out, s);
}

}  // namespace

extern "C" {

// bases == NULL in the descriptor: BP decode, else FFOR.
int dapi_decompress_u32(const flsgpu_ffor_column *c, uint32_t *out, int shape, hipStream_t s) {
  if (!c->bases)
    return decompress_shape<uint32_t, BPDecompressor>(shape, BPColumn<uint32_t>(c->bp),
                                                      c->bp.n_vecs, out, s);
  return decompress_shape<uint32_t, FFORDecompressor>(shape, FFORColumn<uint32_t>(*c),
                                                      c->bp.n_vecs, out, s);
}
int dapi_decompress_u64(const flsgpu_ffor_column *c, uint64_t *out, int shape, hipStream_t s) {
  if (!c->bases)
    return decompress_shape<uint64_t, BPDecompressor>(shape, BPColumn<uint64_t>(c->bp),
                                                      c->bp.n_vecs, out, s);
  return decompress_shape<uint64_t, FFORDecompressor>(shape, FFORColumn<uint64_t>(*c),
                                                      c->bp.n_vecs, out, s);
}

// layout: 0 standard (patched), 1 G-ALP (patched), 2 standard unpatched
int dapi_decompress_alp_f32(const flsgpu_alp_column *c, float *out, int layout, int shape,
                            hipStream_t s) {
  return alp_decompress<float>(c, out, layout, shape, s);
}
int dapi_decompress_alp_f64(const flsgpu_alp_column *c, double *out, int layout, int shape,
                            hipStream_t s) {
  return alp_decompress<double>(c, out, layout, shape, s);
}

// query_column shape (1 vector, 1 value) over an FFOR u32 column; *found
// must be cleared by the caller.
int dapi_query_any_ffor_u32(const flsgpu_ffor_column *c, uint32_t magic, bool *found,
                            hipStream_t s) {
  return launch_query<uint32_t, 1, 1, FFORDecompressor<uint32_t, 1, 1>>(
      FFORColumn<uint32_t>(*c), c->bp.n_vecs, magic, found, s);
}
int dapi_query_any_alp_f64(const flsgpu_alp_column *c, double magic, bool *found,
                           hipStream_t s) {
  return launch_query<double, 4, 1, AlpExt<double, 4, 1>>(ALPExtendedColumn<double>(*c),
                                                          c->ffor.bp.n_vecs, magic, found, s);
}

}  // extern "C"
