// Command-line parity driver with the reference's interface (src/test.cu:
// 13 arguments, same order and spellings), written against the C++ host API
// (include/flsgpu.hpp) so it exercises that drop-in layer end to end:
//
//   flsgpu_test <u32|u64|f32|f64> <decompress|query|query-multi-column>
//               <unpack_n_vecs> <unpack_n_vals> <unpacker> <patcher>
//               <start_vbw> <end_vbw> <start_ec> <end_ec> <n_vecs>
//               <n_samples> <print_debug 0|1|2>
//
// Exit status = number of failed runs (0 = all bit-exact), as
// verification::process_results (src/engine/verification.cuh:80-105);
// print_debug 2 exits 0 regardless. Columns are synthetic, as in
// src/engine/data.cuh, but drawn from a fixed-seed mt19937_64 (the reference
// seeds from std::random_device), and every expected output comes from the
// CPU oracle (oracle/liboracle.so). A bad unpacker/patcher combination
// throws std::invalid_argument like the reference's generated bindings.
//
// Test infrastructure: links the oracle as the checker.
#include <cinttypes>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "flsgpu.hpp"

#include "oracle.h"  // the CPU checker (oracle/)

namespace {

std::mt19937_64 g_rng(20240611);

template <typename T>
T uniform(T lo, T hi) {
  std::uniform_int_distribution<T> d(lo, hi);
  return d(g_rng);
}

struct Params {
  enums::DataType data_type;
  enums::Kernel kernel;
  uint32_t n_vec, n_val;
  enums::Unpacker unpacker;
  enums::Patcher patcher;
  vbw_t vbw_min, vbw_max;
  uint16_t ec_min, ec_max;
  size_t n_values;
  uint32_t n_samples;
  enums::Print print;
};

template <typename T>
struct Result {
  bool success;
  std::vector<std::tuple<size_t, T, T>> diffs;
};

template <typename T>
Result<T> compare(const T *want, const T *got, size_t n) {
  using U = typename utils::same_width_uint<T>::type;
  Result<T> r{true, {}};
  for (size_t i = 0; i < n && r.diffs.size() <= 5; ++i) {
    U a, b;
    std::memcpy(&a, &want[i], sizeof(U));
    std::memcpy(&b, &got[i], sizeof(U));
    if (a != b) {
      r.success = false;
      r.diffs.emplace_back(i, want[i], got[i]);
    }
  }
  return r;
}

template <typename T>
int process(const std::vector<Result<T>> &results, bool debug) {
  using U = typename utils::same_width_uint<T>::type;
  int failed = 0;
  for (size_t i = 0; i < results.size(); ++i) {
    if (results[i].success) continue;
    ++failed;
    if (!debug) continue;
    std::fprintf(stderr, "\n Run %zu failed.\n", i);
    for (const auto &[idx, a, b] : results[i].diffs) {
      U ua, ub;
      std::memcpy(&ua, &a, sizeof(U));
      std::memcpy(&ub, &b, sizeof(U));
      std::fprintf(stderr, "[%zu] correct: %" PRIu64 ", found: %" PRIu64 "\n", idx, uint64_t(ua),
                   uint64_t(ub));
    }
  }
  if (debug) {
    if (failed == 0) std::fprintf(stderr, "Compression successful.\n");
    else std::fprintf(stderr, "\n[%d/%zu] Runs failed.\n", failed, results.size());
  }
  return failed;
}

// Random packed words at random per-vector widths (data.cuh
// generate_random_bp_column / generate_random_ffor_column).
template <typename T>
flsgpu::host::FFORColumn<T> random_ffor_column(size_t n_values, vbw_t vmin, vbw_t vmax, T bmin,
                                               T bmax) {
  const size_t nv = utils::get_n_vecs_from_size(n_values);
  auto *bw = new vbw_t[nv];
  auto *off = new size_t[nv];
  size_t words = 0;
  for (size_t v = 0; v < nv; ++v) {
    bw[v] = vbw_t(uniform<unsigned>(vmin, vmax));
    off[v] = words;
    words += size_t(bw[v]) * utils::get_n_lanes<T>();
  }
  auto *packed = new T[std::max<size_t>(words, 1)];
  for (size_t i = 0; i < words; ++i) packed[i] = T(g_rng());
  auto *bases = new T[nv];
  for (size_t v = 0; v < nv; ++v) bases[v] = uniform<T>(bmin, bmax);
  return flsgpu::host::FFORColumn<T>{flsgpu::host::BPColumn<T>{n_values, words, packed, bw, off},
                                     bases};
}

// A column that holds the magic value with probability 1/2 (data.cuh
// generate_binary_ffor_column always leaves it out).
template <typename T>
std::pair<bool, flsgpu::host::FFORColumn<T>> binary_ffor_column(size_t n_values, vbw_t vbw) {
  const size_t nv = utils::get_n_vecs_from_size(n_values);
  const bool present = uniform<int>(0, 1) == 1;
  std::vector<T> data(n_values, T(0));
  auto *bases = new T[nv];
  std::fill(bases, bases + nv, T(0));
  if (vbw == 0) {
    std::fill(bases, bases + nv, T(present));
  } else if (present) {
    data[uniform<size_t>(0, n_values - 1)] = consts::as<T>::MAGIC_NUMBER;
  }
  auto *bw = new vbw_t[nv];
  std::fill(bw, bw + nv, vbw);
  const size_t words = nv * size_t(vbw) * utils::get_n_lanes<T>();
  auto *packed = new T[std::max<size_t>(words, 1)];
  auto *off = new size_t[nv];
  auto *zero = new T[nv]();
  int rc;
  if constexpr (sizeof(T) == 4)
    rc = flsgpu_ffor_pack_u32(data.data(), nv, bw, zero, packed, reinterpret_cast<uint64_t *>(off));
  else
    rc = flsgpu_ffor_pack_u64(data.data(), nv, bw, zero, packed, reinterpret_cast<uint64_t *>(off));
  delete[] zero;
  if (rc) throw std::runtime_error("ffor pack failed");
  return {present, flsgpu::host::FFORColumn<T>{
                       flsgpu::host::BPColumn<T>{n_values, words, packed, bw, off}, bases}};
}

template <typename T>
T *cpu_decode(const flsgpu::host::FFORColumn<T> &c) {
  T *out = new T[c.get_n_values()];
  if constexpr (sizeof(T) == 4)
    orc_ffor_column_decode_u32(c.get_n_vecs(), c.bp.packed_array, c.bp.bit_widths,
                               reinterpret_cast<const uint64_t *>(c.bp.vector_offsets), c.bases, out);
  else
    orc_ffor_column_decode_u64(c.get_n_vecs(), c.bp.packed_array, c.bp.bit_widths,
                               reinterpret_cast<const uint64_t *>(c.bp.vector_offsets), c.bases, out);
  return out;
}

template <typename T>
T *cpu_decode(const flsgpu::host::ALPColumn<T> &c) {
  T *out = new T[c.get_n_values()];
  const auto *eo = reinterpret_cast<const uint64_t *>(c.exceptions_offsets);
  const auto *vo = reinterpret_cast<const uint64_t *>(c.ffor.bp.vector_offsets);
  if constexpr (sizeof(T) == 4)
    orc_alp_column_decode_f32(c.get_n_vecs(), c.ffor.bp.packed_array, c.ffor.bp.bit_widths, vo,
                              c.ffor.bases, c.factor_indices, c.fraction_indices, eo, c.exceptions,
                              c.positions, c.counts, nullptr, out);
  else
    orc_alp_column_decode_f64(c.get_n_vecs(), c.ffor.bp.packed_array, c.ffor.bp.bit_widths, vo,
                              c.ffor.bases, c.factor_indices, c.fraction_indices, eo, c.exceptions,
                              c.positions, c.counts, nullptr, out);
  return out;
}

// data.cuh generate_alp_column: random FFOR words (bases 2..20), factor and
// fraction indices in the lower half of their tables, no exceptions yet.
template <typename T>
flsgpu::host::ALPColumn<T> random_alp_column(size_t n_values, vbw_t vmin, vbw_t vmax) {
  using U = typename utils::same_width_uint<T>::type;
  const size_t nv = utils::get_n_vecs_from_size(n_values);
  flsgpu::host::ALPColumn<T> c{};
  c.ffor = random_ffor_column<U>(n_values, vmin, vmax, U(2), U(20));
  c.factor_indices = new uint8_t[nv];
  c.fraction_indices = new uint8_t[nv];
  for (size_t v = 0; v < nv; ++v) {
    c.factor_indices[v] = uint8_t(uniform<unsigned>(0, consts::as<T>::FACT_ARR_COUNT / 2));
    c.fraction_indices[v] = uint8_t(uniform<unsigned>(0, consts::as<T>::FRAC_ARR_COUNT / 2));
  }
  c.counts = new uint16_t[nv]();
  c.exceptions_offsets = new size_t[nv]();
  c.exceptions = new T[1];
  c.positions = new uint16_t[1];
  c.n_exceptions = 0;
  return c;
}

// data.cuh modify_alp_exception_count: exactly ec exceptions per vector at
// distinct sorted random positions, random bit patterns as values.
template <typename T>
void set_exception_count(flsgpu::host::ALPColumn<T> &c, uint16_t ec) {
  using U = typename utils::same_width_uint<T>::type;
  const size_t nv = c.get_n_vecs();
  delete[] c.exceptions;
  delete[] c.positions;
  c.n_exceptions = nv * ec;
  c.exceptions = new T[std::max<size_t>(c.n_exceptions, 1)];
  c.positions = new uint16_t[std::max<size_t>(c.n_exceptions, 1)];
  std::vector<uint16_t> idx(consts::VALUES_PER_VECTOR);
  for (size_t v = 0; v < nv; ++v) {
    c.counts[v] = ec;
    c.exceptions_offsets[v] = v * ec;
    for (int i = 0; i < consts::VALUES_PER_VECTOR; ++i) idx[i] = uint16_t(i);
    std::shuffle(idx.begin(), idx.end(), g_rng);
    std::sort(idx.begin(), idx.begin() + ec);
    for (uint16_t k = 0; k < ec; ++k) {
      c.positions[v * ec + k] = idx[k];
      const U bits = U(g_rng());
      std::memcpy(&c.exceptions[v * ec + k], &bits, sizeof(T));
    }
  }
}

template <typename T, typename HostColT>
Result<T> run_kernel(const HostColT &column, const T *want, const Params &p, bool expect,
                     T magic) {
  using DevT = typename HostColT::DeviceColumnT;
  if (p.kernel == enums::Kernel::Decompress) {
    auto d = column.copy_to_device();
    T *got = bindings::decompress_column<T, DevT>(d, p.n_vec, p.n_val, p.unpacker, p.patcher,
                                                  p.n_samples);
    flsgpu::host::free_column(d);
    auto r = compare(want, got, column.get_n_values());
    delete[] got;
    return r;
  }
  bool answer;
  if (p.kernel == enums::Kernel::Query) {
    auto d = column.copy_to_device();
    answer = bindings::query_column<T, DevT>(d, p.n_vec, p.n_val, p.unpacker, p.patcher, magic,
                                             p.n_samples);
    flsgpu::host::free_column(d);
  } else {
    answer = bindings::query_multi_column<T, HostColT>(column, p.n_vec, p.n_val, p.unpacker,
                                                       p.patcher, magic, p.n_samples);
  }
  const T a = expect ? T(1) : T(0);
  const T b = answer ? T(1) : T(0);
  return compare(&a, &b, 1);
}

template <typename T>
std::vector<Result<T>> execute_ffor(const Params &p) {
  std::vector<Result<T>> results;
  for (unsigned vbw = p.vbw_min; vbw <= p.vbw_max; ++vbw) {
    const bool multi = p.kernel == enums::Kernel::QueryMultiColumn;
    bool expect = false;
    T magic = consts::as<T>::MAGIC_NUMBER;
    flsgpu::host::FFORColumn<T> column;
    if (p.kernel == enums::Kernel::Query) {
      auto [present, c] = binary_ffor_column<T>(p.n_values, vbw_t(vbw));
      expect = present;
      column = c;
    } else {
      column = random_ffor_column<T>(p.n_values, multi ? p.vbw_min : vbw_t(vbw),
                                     multi ? p.vbw_max : vbw_t(vbw), T(0), T(100));
    }
    if (multi) magic = std::numeric_limits<T>::max();  // never a whole lane
    T *want = p.kernel == enums::Kernel::Decompress ? cpu_decode(column) : nullptr;
    results.push_back(run_kernel<T>(column, want, p, expect, magic));
    delete[] want;
    flsgpu::host::free_column(column);
    if (multi) break;
  }
  return results;
}

template <typename T>
std::vector<Result<T>> execute_alp(const Params &p) {
  using P = enums::Patcher;
  std::vector<Result<T>> results;
  auto column = random_alp_column<T>(p.n_values, p.vbw_min, p.vbw_max);
  for (unsigned ec = p.ec_min; ec <= p.ec_max; ++ec) {
    set_exception_count(column, uint16_t(ec));
    T *want = cpu_decode(column);
    bool expect = false;
    T magic = consts::as<T>::MAGIC_NUMBER;
    if (p.kernel == enums::Kernel::Query) {
      // data.cuh get_value_to_query: half the time a value taken from the
      // column; the expected answer is recomputed by scanning (a NaN picked
      // from the column is never found).
      if (uniform<int>(0, 1)) magic = want[uniform<size_t>(0, column.get_n_values() - 1)];
      for (size_t i = 0; i < column.get_n_values(); ++i) expect |= want[i] == magic;
    } else if (p.kernel == enums::Kernel::QueryMultiColumn) {
      magic = std::numeric_limits<T>::max();
    }
    if (p.patcher == P::Dummy || p.patcher == P::Stateless || p.patcher == P::Stateful ||
        p.patcher == P::None) {
      results.push_back(run_kernel<T>(column, want, p, expect, magic));
    } else {
      auto ext = column.create_extended_column();
      results.push_back(run_kernel<T>(ext, want, p, expect, magic));
      flsgpu::host::free_column(ext);
    }
    delete[] want;
  }
  flsgpu::host::free_column(column);
  return results;
}

}  // namespace

int run(int argc, char **argv) {
  if (argc != 14) throw std::invalid_argument("Wrong arg count.");
  int a = 0;
  Params p;
  p.data_type = enums::string_to_data_type(argv[++a]);
  p.kernel = enums::string_to_kernel(argv[++a]);
  p.n_vec = uint32_t(std::stoul(argv[++a]));
  p.n_val = uint32_t(std::stoul(argv[++a]));
  p.unpacker = enums::string_to_unpacker(argv[++a]);
  p.patcher = enums::string_to_patcher(argv[++a]);
  const unsigned v0 = unsigned(std::stoul(argv[++a])), v1 = unsigned(std::stoul(argv[++a]));
  p.vbw_min = vbw_t(v0);
  p.vbw_max = vbw_t(std::max(v0, v1));
  const unsigned e0 = unsigned(std::stoul(argv[++a])), e1 = unsigned(std::stoul(argv[++a]));
  p.ec_min = uint16_t(e0);
  p.ec_max = uint16_t(std::max(e0, e1));
  p.n_values = size_t(std::stoul(argv[++a])) * consts::VALUES_PER_VECTOR;
  p.n_samples = uint32_t(std::stoul(argv[++a]));
  p.print = static_cast<enums::Print>(std::stoul(argv[++a]));

  const bool debug = p.print != enums::Print::PrintNothing;
  int failed = 0;
  switch (p.data_type) {
    case enums::DataType::U32: failed = process(execute_ffor<uint32_t>(p), debug); break;
    case enums::DataType::U64: failed = process(execute_ffor<uint64_t>(p), debug); break;
    case enums::DataType::F32: failed = process(execute_alp<float>(p), debug); break;
    case enums::DataType::F64: failed = process(execute_alp<double>(p), debug); break;
  }
  if (p.print == enums::Print::PrintDebugExit0) return 0;
  return failed;
}

// A missing binding or bad argument throws std::invalid_argument, as in the
// reference; it is reported with exit status 100 instead of terminating.
int main(int argc, char **argv) {
  try {
    return run(argc, argv);
  } catch (const std::invalid_argument &e) {
    std::fprintf(stderr, "invalid_argument: %s\n", e.what());
    return 100;
  }
}
