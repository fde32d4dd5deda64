// flsgpu C++ host API — the reference's decoder API re-expressed over the
// C ABI of libflsgpu_hip.so (flsgpu.h, flsgpu_host.h), header-only.
//
// It keeps the reference's names, argument meaning and error behaviour so
// that a host program written against the reference compiles against this
// header and runs on gfx950:
//   - utils:: geometry helpers       src/flsgpu/utils.cuh:55-90
//   - consts::                        src/flsgpu/consts.cuh:9-29
//   - flsgpu::device:: column structs src/flsgpu/structs.cuh:33-84
//   - flsgpu::host:: column structs,  structs.cuh:89-331
//     copy_to_device, free_column     structs.cuh:334-400
//   - enums:: + string_to_*           src/engine/enums.cuh:6-70, enums.cu:5-100
//   - bindings::decompress_column,    src/generated-bindings/kernel-bindings.cuh:11-175
//     query_column, compute_column,   (bodies generated by generate-kernel-bindings.py)
//     query_multi_column              src/engine/multi-column-host-kernels.cuh:11-123
//
// What differs underneath: the reference's unpacker / n_vec / n_val choose
// among CUDA decoding strategies at compile time. Here they are checked
// against the combinations the reference provides (same
// std::invalid_argument when one is missing, generate-kernel-bindings.py:
// 207-209, 241-268, 355-357) and the decode runs the autotuned native
// variant (FLSGPU_VARIANT_AUTO). Every decoder produces the same bits, so the
// choice only affects speed. The FACT/FRAC tables are compiled into the
// kernels; copy_to_device still uploads them for code that reads
// device::ALPColumn::factors / fractions.
//
// Errors: a combination the reference does not provide, or an argument the
// library rejects, throws std::invalid_argument; a HIP runtime failure prints
// the message and calls exit(EXIT_FAILURE), as the reference's
// CUDA_SAFE_CALL does (src/flsgpu/host-utils.cuh:12-20).
#pragma once

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <random>
#include <stdexcept>
#include <string>
#include <tuple>
#include <type_traits>
#include <unordered_map>
#include <vector>

#include "flsgpu.h"
#include "flsgpu_host.h"

static_assert(sizeof(size_t) == sizeof(uint64_t), "size_t must be 64-bit");

using vbw_t = uint8_t;    // src/flsgpu/device-types.cuh:6-9
using lane_t = uint16_t;
using si_t = uint32_t;
using vi_t = uint32_t;

namespace consts {
constexpr int32_t REGISTER_WIDTH = 1024;
constexpr int32_t VALUES_PER_VECTOR = 1024;
// The reference's warp size is 32; a gfx950 wavefront has 64 threads.
constexpr int32_t THREADS_PER_WARP = 64;
constexpr unsigned MAX_UNPACK_N_VECS = 4;

template <class T>
struct as {
  static inline constexpr T MAGIC_NUMBER = 1;
};
template <>
struct as<float> {
  static inline constexpr float MAGIC_NUMBER = 0.3214f;
  static inline constexpr int32_t FACT_ARR_COUNT = 10;
  static inline constexpr int32_t FRAC_ARR_COUNT = 11;
};
template <>
struct as<double> {
  static inline constexpr double MAGIC_NUMBER = 0.3214;
  static inline constexpr int32_t FACT_ARR_COUNT = 19;
  static inline constexpr int32_t FRAC_ARR_COUNT = 21;
};
}  // namespace consts

namespace utils {
template <typename T>
struct same_width_uint {
  using type = std::conditional_t<sizeof(T) == 8, uint64_t,
                                  std::conditional_t<sizeof(T) == 4, uint32_t,
                                                     std::conditional_t<sizeof(T) == 2, uint16_t,
                                                                        uint8_t>>>;
};
template <typename T>
struct same_width_int {
  using type = std::make_signed_t<typename same_width_uint<T>::type>;
};
template <typename T>
constexpr int32_t sizeof_in_bits() {
  return int32_t(sizeof(T) * 8);
}
template <typename T>
constexpr int32_t get_n_lanes() {
  return consts::REGISTER_WIDTH / sizeof_in_bits<T>();
}
template <typename T>
constexpr int32_t get_values_per_lane() {
  return consts::VALUES_PER_VECTOR / get_n_lanes<T>();
}
template <typename T>
constexpr int32_t get_compressed_vector_size(int32_t value_bit_width) {
  return (consts::VALUES_PER_VECTOR * value_bit_width) / sizeof_in_bits<T>();
}
constexpr size_t get_n_vecs_from_size(const size_t size) {
  return (size + consts::VALUES_PER_VECTOR - 1) / consts::VALUES_PER_VECTOR;
}
// Low `count` bits set, count in 0..W (defined at count == W, unlike a
// plain shift).
template <typename T>
constexpr T h_set_first_n_bits(const int32_t count) {
  return count < sizeof_in_bits<T>() ? static_cast<T>((T{1} << count) - T{1})
                                     : static_cast<T>(~T{0});
}
template <typename T>
constexpr T set_first_n_bits(const int32_t count) {
  return h_set_first_n_bits<T>(count);
}
template <typename T>
T *copy_array(const T *in, const size_t n_elements) {
  T *out = new T[n_elements];
  if (n_elements) std::memcpy(out, in, sizeof(T) * n_elements);
  return out;
}
}  // namespace utils

namespace flsgpu {
namespace detail {
[[noreturn]] inline void fail_hip(const char *what) {
  std::fprintf(stderr, "flsgpu: %s: %s\n", what, flsgpu_last_error());
  std::exit(EXIT_FAILURE);
}
// EINVAL / EUNSUPPORTED -> std::invalid_argument, HIP failures -> exit.
inline void check(int rc, const char *what) {
  if (rc == FLSGPU_OK) return;
  if (rc == FLSGPU_EINVAL || rc == FLSGPU_EUNSUPPORTED)
    throw std::invalid_argument(std::string(what) + ": " + flsgpu_last_error());
  fail_hip(what);
}
template <typename T>
T *upload(const T *host, size_t n, size_t pad_elements = 0) {
  void *d = nullptr;
  check(flsgpu_upload(&d, host, n * sizeof(T), pad_elements * sizeof(T), nullptr), "upload");
  return static_cast<T *>(d);
}
inline void free_device_pointer(void *p) {
  if (p) check(flsgpu_free(p), "free");
}
// The reference's FACT_ARR / FRAC_ARR (src/alp/constants.hpp:49-55, 76-144).
template <typename T>
struct tables;
template <>
struct tables<float> {
  static constexpr int32_t fact[10] = {1,      10,      100,      1000,      10000,
                                       100000, 1000000, 10000000, 100000000, 1000000000};
  // Double literals converted to float, as the reference spells them.
  static constexpr float frac[11] = {1.0,      0.1,       0.01,       0.001,
                                     0.0001,   0.00001,   0.000001,   0.0000001,
                                     0.00000001, 0.000000001, 0.0000000001};
};
template <>
struct tables<double> {
  static constexpr int64_t fact[19] = {1LL,
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
  static constexpr double frac[21] = {1.0,   0.1,   0.01,  0.001, 1e-4,  1e-5,  1e-6,
                                      1e-7,  1e-8,  1e-9,  1e-10, 1e-11, 1e-12, 1e-13,
                                      1e-14, 1e-15, 1e-16, 1e-17, 1e-18, 1e-19, 1e-20};
};
}  // namespace detail

// ---------------------------------------------------------------------------
// Device columns (plain structs of device pointers)
// ---------------------------------------------------------------------------
namespace device {
template <typename T>
struct BPColumn {
  using UINT_T = typename utils::same_width_uint<T>::type;
  size_t n_values;
  size_t n_vecs;
  UINT_T *packed_array;
  vbw_t *bit_widths;
  size_t *vector_offsets;
};

template <typename T>
struct FFORColumn {
  using UINT_T = typename utils::same_width_uint<T>::type;
  size_t n_values;
  BPColumn<T> bp;
  UINT_T *bases;
};

template <typename T>
struct ALPColumn {
  using INT_T = typename utils::same_width_int<T>::type;
  using UINT_T = typename utils::same_width_uint<T>::type;
  size_t n_values;
  FFORColumn<UINT_T> ffor;
  INT_T *factors;
  T *fractions;
  uint8_t *factor_indices;
  uint8_t *fraction_indices;
  size_t n_exceptions;
  size_t *exceptions_offsets;
  T *exceptions;
  uint16_t *positions;
  uint16_t *counts;
};

template <typename T>
struct ALPExtendedColumn {
  using INT_T = typename utils::same_width_int<T>::type;
  using UINT_T = typename utils::same_width_uint<T>::type;
  size_t n_values;
  FFORColumn<UINT_T> ffor;
  INT_T *factors;
  T *fractions;
  uint8_t *factor_indices;
  uint8_t *fraction_indices;
  size_t n_exceptions;
  size_t *exceptions_offsets;
  T *exceptions;
  uint16_t *positions;
  uint16_t *offsets_counts;
};
}  // namespace device

// ---------------------------------------------------------------------------
// Host columns (new[]-allocated arrays; copy_to_device uploads them)
// ---------------------------------------------------------------------------
namespace host {
template <typename T>
struct BPColumn {
  using UINT_T = typename utils::same_width_uint<T>::type;
  using DeviceColumnT = device::BPColumn<T>;

  size_t n_values;
  size_t n_packed_values;

  size_t get_n_values() const { return n_values; }
  size_t get_n_vecs() const { return utils::get_n_vecs_from_size(n_values); }

  UINT_T *packed_array;
  vbw_t *bit_widths;
  size_t *vector_offsets;

  // Keeps the reference's over-read allowance after the packed words
  // (structs.cuh:104-109), which its branchless unpackers need.
  device::BPColumn<T> copy_to_device() const {
    const size_t pad = sizeof(T) * utils::get_n_lanes<T>() * 4;
    return device::BPColumn<T>{n_values, get_n_vecs(),
                               detail::upload(packed_array, n_packed_values, pad),
                               detail::upload(bit_widths, get_n_vecs()),
                               detail::upload(vector_offsets, get_n_vecs())};
  }
};

template <typename T>
struct FFORColumn {
  using UINT_T = typename utils::same_width_uint<T>::type;
  using DeviceColumnT = device::FFORColumn<T>;

  BPColumn<T> bp;
  UINT_T *bases;

  size_t get_n_values() const { return bp.n_values; }
  size_t get_n_vecs() const { return bp.get_n_vecs(); }

  device::FFORColumn<T> copy_to_device() const {
    return device::FFORColumn<T>{get_n_values(), bp.copy_to_device(),
                                 detail::upload(bases, bp.get_n_vecs())};
  }
};

template <typename T>
struct ALPExtendedColumn {
  using INT_T = typename utils::same_width_int<T>::type;
  using UINT_T = typename utils::same_width_uint<T>::type;
  using DeviceColumnT = device::ALPExtendedColumn<T>;

  FFORColumn<UINT_T> ffor;
  uint8_t *factor_indices;
  uint8_t *fraction_indices;
  size_t n_exceptions;
  size_t *exceptions_offsets;
  T *exceptions;
  uint16_t *positions;
  uint16_t *offsets_counts;
  size_t compressed_size_bytes_alp_extended;

  size_t get_n_values() const { return ffor.bp.n_values; }
  size_t get_n_vecs() const { return ffor.bp.get_n_vecs(); }
  double get_compression_ratio() const {
    return double(ffor.bp.n_values * sizeof(T)) / double(compressed_size_bytes_alp_extended);
  }

  // Exceptions and positions keep the reference's MAX_UNPACK_N_VECS
  // over-read allowance (structs.cuh:160-177).
  device::ALPExtendedColumn<T> copy_to_device() const {
    const size_t pad = consts::MAX_UNPACK_N_VECS;
    return device::ALPExtendedColumn<T>{
        get_n_values(),
        ffor.copy_to_device(),
        detail::upload(detail::tables<T>::fact, consts::as<T>::FACT_ARR_COUNT),
        detail::upload(detail::tables<T>::frac, consts::as<T>::FRAC_ARR_COUNT),
        detail::upload(factor_indices, get_n_vecs()),
        detail::upload(fraction_indices, get_n_vecs()),
        n_exceptions,
        detail::upload(exceptions_offsets, get_n_vecs()),
        detail::upload(exceptions, n_exceptions, pad),
        detail::upload(positions, n_exceptions, pad),
        detail::upload(offsets_counts, get_n_vecs() * utils::get_n_lanes<T>())};
  }
};

template <typename T>
struct ALPColumn {
  using INT_T = typename utils::same_width_int<T>::type;
  using UINT_T = typename utils::same_width_uint<T>::type;
  using DeviceColumnT = device::ALPColumn<T>;

  FFORColumn<UINT_T> ffor;
  uint8_t *factor_indices;
  uint8_t *fraction_indices;
  size_t n_exceptions;
  size_t *exceptions_offsets;
  T *exceptions;
  uint16_t *positions;
  uint16_t *counts;
  size_t compressed_size_bytes_alp;
  size_t compressed_size_bytes_alp_extended;

  size_t get_n_values() const { return ffor.bp.n_values; }
  size_t get_n_vecs() const { return ffor.bp.get_n_vecs(); }
  double get_compression_ratio() const {
    return double(ffor.bp.n_values * sizeof(T)) / double(compressed_size_bytes_alp);
  }

  device::ALPColumn<T> copy_to_device() const {
    return device::ALPColumn<T>{
        get_n_values(),
        ffor.copy_to_device(),
        detail::upload(detail::tables<T>::fact, consts::as<T>::FACT_ARR_COUNT),
        detail::upload(detail::tables<T>::frac, consts::as<T>::FRAC_ARR_COUNT),
        detail::upload(factor_indices, get_n_vecs()),
        detail::upload(fraction_indices, get_n_vecs()),
        n_exceptions,
        detail::upload(exceptions_offsets, get_n_vecs()),
        detail::upload(exceptions, n_exceptions),
        detail::upload(positions, n_exceptions),
        detail::upload(counts, get_n_vecs())};
  }

  // Lane-divided (G-ALP) exception layout, structs.cuh:231-306: per vector,
  // exceptions stably partitioned by lane (position % N_LANES) and a header
  // (count << 10) | offset per lane. malloc-allocated, as in the reference.
  std::tuple<T *, uint16_t *, uint16_t *> convert_exceptions_to_lane_divided_format() const {
    const size_t nv = get_n_vecs();
    T *e = static_cast<T *>(std::malloc(sizeof(T) * std::max<size_t>(n_exceptions, 1)));
    uint16_t *p = static_cast<uint16_t *>(std::malloc(sizeof(uint16_t) * std::max<size_t>(n_exceptions, 1)));
    uint16_t *oc = static_cast<uint16_t *>(
        std::malloc(sizeof(uint16_t) * std::max<size_t>(nv * utils::get_n_lanes<T>(), 1)));
    int rc;
    if constexpr (sizeof(T) == 4)
      rc = flsgpu_galp_convert_f32(nv, exceptions, positions, counts,
                                   reinterpret_cast<const uint64_t *>(exceptions_offsets), e, p, oc);
    else
      rc = flsgpu_galp_convert_f64(nv, exceptions, positions, counts,
                                   reinterpret_cast<const uint64_t *>(exceptions_offsets), e, p, oc);
    if (rc != FLSGPU_OK) {
      std::free(e);
      std::free(p);
      std::free(oc);
      throw std::invalid_argument(std::string("convert_exceptions_to_lane_divided_format: ") +
                                  flsgpu_encoder_last_error());
    }
    return std::make_tuple(e, p, oc);
  }

  ALPExtendedColumn<T> create_extended_column() const {
    auto [e, p, oc] = convert_exceptions_to_lane_divided_format();
    const size_t nv = get_n_vecs();
    // new[] copies so that free_column (delete[]) releases them.
    ALPExtendedColumn<T> out{
        FFORColumn<UINT_T>{
            BPColumn<UINT_T>{ffor.bp.n_values, ffor.bp.n_packed_values,
                             utils::copy_array(ffor.bp.packed_array, ffor.bp.n_packed_values),
                             utils::copy_array(ffor.bp.bit_widths, nv),
                             utils::copy_array(ffor.bp.vector_offsets, nv)},
            utils::copy_array(ffor.bases, nv)},
        utils::copy_array(factor_indices, nv),
        utils::copy_array(fraction_indices, nv),
        n_exceptions,
        utils::copy_array(exceptions_offsets, nv),
        utils::copy_array(e, n_exceptions),
        utils::copy_array(p, n_exceptions),
        utils::copy_array(oc, nv * utils::get_n_lanes<T>()),
        compressed_size_bytes_alp_extended};
    std::free(e);
    std::free(p);
    std::free(oc);
    return out;
  }
};

template <typename T>
void free_column(BPColumn<T> column) {
  delete[] column.packed_array;
  delete[] column.bit_widths;
  delete[] column.vector_offsets;
}
template <typename T>
void free_column(FFORColumn<T> column) {
  free_column(column.bp);
  delete[] column.bases;
}
template <typename T>
void free_column(ALPColumn<T> column) {
  free_column(column.ffor);
  delete[] column.factor_indices;
  delete[] column.fraction_indices;
  delete[] column.exceptions_offsets;
  delete[] column.exceptions;
  delete[] column.positions;
  delete[] column.counts;
}
template <typename T>
void free_column(ALPExtendedColumn<T> column) {
  free_column(column.ffor);
  delete[] column.factor_indices;
  delete[] column.fraction_indices;
  delete[] column.exceptions_offsets;
  delete[] column.exceptions;
  delete[] column.positions;
  delete[] column.offsets_counts;
}
template <typename T>
void free_column(device::BPColumn<T> column) {
  detail::free_device_pointer(column.packed_array);
  detail::free_device_pointer(column.bit_widths);
  detail::free_device_pointer(column.vector_offsets);
}
template <typename T>
void free_column(device::FFORColumn<T> column) {
  free_column(column.bp);
  detail::free_device_pointer(column.bases);
}
template <typename T>
void free_column(device::ALPColumn<T> column) {
  free_column(column.ffor);
  detail::free_device_pointer(column.factors);
  detail::free_device_pointer(column.fractions);
  detail::free_device_pointer(column.factor_indices);
  detail::free_device_pointer(column.fraction_indices);
  detail::free_device_pointer(column.exceptions_offsets);
  detail::free_device_pointer(column.exceptions);
  detail::free_device_pointer(column.positions);
  detail::free_device_pointer(column.counts);
}
template <typename T>
void free_column(device::ALPExtendedColumn<T> column) {
  free_column(column.ffor);
  detail::free_device_pointer(column.factors);
  detail::free_device_pointer(column.fractions);
  detail::free_device_pointer(column.factor_indices);
  detail::free_device_pointer(column.fraction_indices);
  detail::free_device_pointer(column.exceptions_offsets);
  detail::free_device_pointer(column.exceptions);
  detail::free_device_pointer(column.positions);
  detail::free_device_pointer(column.offsets_counts);
}
}  // namespace host
}  // namespace flsgpu

// ---------------------------------------------------------------------------
// Enums and their command-line spellings
// ---------------------------------------------------------------------------
namespace enums {
enum class DataType { U32, U64, F32, F64 };
enum class Kernel { Decompress, Query, QueryMultiColumn };
enum class Unpacker {
  Dummy,
  OldFls,
  SwitchCase,
  Stateless,
  StatelessBranchless,
  StatefulCache,
  StatefulLocal1,
  StatefulLocal2,
  StatefulLocal4,
  StatefulShared1,
  StatefulShared2,
  StatefulShared4,
  StatefulRegister1,
  StatefulRegister2,
  StatefulRegister4,
  StatefulRegisterBranchless1,
  StatefulRegisterBranchless2,
  StatefulRegisterBranchless4,
  StatefulBranchless,
};
// Same order as flsgpu_patcher in flsgpu.h.
enum class Patcher {
  None,
  Dummy,
  Stateless,
  Stateful,
  Naive,
  NaiveBranchless,
  PrefetchPosition,
  PrefetchAll,
  PrefetchAllBranchless,
};
enum class Print { PrintNothing, PrintDebug, PrintDebugExit0 };

namespace detail {
template <typename E>
E lookup(const std::unordered_map<std::string, E> &m, const std::string &s, const char *what) {
  auto it = m.find(s);
  if (it == m.end()) throw std::invalid_argument(std::string("Unknown ") + what + ": " + s);
  return it->second;
}
}  // namespace detail

inline DataType string_to_data_type(const std::string &str) {
  static const std::unordered_map<std::string, DataType> m = {
      {"u32", DataType::U32}, {"u64", DataType::U64}, {"f32", DataType::F32}, {"f64", DataType::F64}};
  return detail::lookup(m, str, "data type");
}
inline Kernel string_to_kernel(const std::string &str) {
  static const std::unordered_map<std::string, Kernel> m = {
      {"decompress", Kernel::Decompress},
      {"query", Kernel::Query},
      {"query-multi-column", Kernel::QueryMultiColumn}};
  return detail::lookup(m, str, "kernel type");
}
inline Unpacker string_to_unpacker(const std::string &str) {
  static const std::unordered_map<std::string, Unpacker> m = {
      {"dummy", Unpacker::Dummy},
      {"old-fls", Unpacker::OldFls},
      {"switch-case", Unpacker::SwitchCase},
      {"stateless", Unpacker::Stateless},
      {"stateless-branchless", Unpacker::StatelessBranchless},
      {"stateful-cache", Unpacker::StatefulCache},
      {"stateful-local-1", Unpacker::StatefulLocal1},
      {"stateful-local-2", Unpacker::StatefulLocal2},
      {"stateful-local-4", Unpacker::StatefulLocal4},
      {"stateful-shared-1", Unpacker::StatefulShared1},
      {"stateful-shared-2", Unpacker::StatefulShared2},
      {"stateful-shared-4", Unpacker::StatefulShared4},
      {"stateful-register-1", Unpacker::StatefulRegister1},
      {"stateful-register-2", Unpacker::StatefulRegister2},
      {"stateful-register-4", Unpacker::StatefulRegister4},
      {"stateful-register-branchless-1", Unpacker::StatefulRegisterBranchless1},
      {"stateful-register-branchless-2", Unpacker::StatefulRegisterBranchless2},
      {"stateful-register-branchless-4", Unpacker::StatefulRegisterBranchless4},
      {"stateful-branchless", Unpacker::StatefulBranchless}};
  return detail::lookup(m, str, "unpacker type");
}
inline Patcher string_to_patcher(const std::string &str) {
  static const std::unordered_map<std::string, Patcher> m = {
      {"none", Patcher::None},
      {"dummy", Patcher::Dummy},
      {"stateless", Patcher::Stateless},
      {"stateful", Patcher::Stateful},
      {"naive", Patcher::Naive},
      {"naive-branchless", Patcher::NaiveBranchless},
      {"prefetch-position", Patcher::PrefetchPosition},
      {"prefetch-all", Patcher::PrefetchAll},
      {"prefetch-all-branchless", Patcher::PrefetchAllBranchless}};
  return detail::lookup(m, str, "patcher type");
}
}  // namespace enums

// ---------------------------------------------------------------------------
// bindings:: — synchronous host entry points
// ---------------------------------------------------------------------------
namespace bindings {
namespace detail {
enum class Fn { Decompress, Query, Compute, MultiColumn };
enum class Enc { BP, FFOR, ALP, ALPExtended };

template <typename ColumnT>
struct column_traits;
template <typename T>
struct column_traits<flsgpu::device::BPColumn<T>> {
  static constexpr Enc enc = Enc::BP;
  static constexpr const char *name = "BP";
};
template <typename T>
struct column_traits<flsgpu::device::FFORColumn<T>> {
  static constexpr Enc enc = Enc::FFOR;
  static constexpr const char *name = "FFOR";
};
template <typename T>
struct column_traits<flsgpu::device::ALPColumn<T>> {
  static constexpr Enc enc = Enc::ALP;
  static constexpr const char *name = "ALP";
};
template <typename T>
struct column_traits<flsgpu::device::ALPExtendedColumn<T>> {
  static constexpr Enc enc = Enc::ALPExtended;
  static constexpr const char *name = "ALPExtended";
};
template <typename T>
struct column_traits<flsgpu::host::FFORColumn<T>> : column_traits<flsgpu::device::FFORColumn<T>> {};
template <typename T>
struct column_traits<flsgpu::host::ALPColumn<T>> : column_traits<flsgpu::device::ALPColumn<T>> {};
template <typename T>
struct column_traits<flsgpu::host::ALPExtendedColumn<T>>
    : column_traits<flsgpu::device::ALPExtendedColumn<T>> {};

inline const char *fn_name(Fn f) {
  switch (f) {
    case Fn::Decompress: return "decompress_column";
    case Fn::Query: return "query_column";
    case Fn::Compute: return "compute_column";
    default: return "query_multi_column";
  }
}
template <typename T>
constexpr const char *type_name() {
  if constexpr (std::is_same_v<T, uint32_t>) return "uint32_t";
  else if constexpr (std::is_same_v<T, uint64_t>) return "uint64_t";
  else if constexpr (std::is_same_v<T, float>) return "float";
  else return "double";
}

// The combinations the reference's generated bindings provide
// (generate-kernel-bindings.py:241-268 filters, :284-357 loops).
template <typename T>
bool provided(Fn fn, Enc enc, unsigned n_vec, unsigned n_val, enums::Unpacker u,
              enums::Patcher p) {
  using U = enums::Unpacker;
  using P = enums::Patcher;
  if (n_vec != 1 && n_vec != 4) return false;
  if (u == U::OldFls) {
    if (n_val != 32 || n_vec != 1) return false;
    if (!(std::is_same_v<T, uint32_t> || std::is_same_v<T, float>)) return false;
  } else if (n_val != 1) {
    return false;
  }
  if (u == U::SwitchCase &&
      (n_vec != 1 || n_val != 1 || !std::is_integral_v<T> || fn == Fn::Compute))
    return false;
  const bool alp = enc == Enc::ALP || enc == Enc::ALPExtended;
  if (alp && u == U::Dummy) return false;
  bool patcher_ok;
  switch (enc) {
    case Enc::BP:
    case Enc::FFOR: patcher_ok = p == P::None; break;
    case Enc::ALP: patcher_ok = p == P::Dummy || p == P::Stateless || p == P::Stateful; break;
    default:
      patcher_ok = p == P::Naive || p == P::NaiveBranchless || p == P::PrefetchPosition ||
                   p == P::PrefetchAll || p == P::PrefetchAllBranchless;
  }
  if (!patcher_ok) return false;
  if (fn == Fn::MultiColumn) {
    if (!(u == U::OldFls || u == U::StatefulRegisterBranchless2 || u == U::StatefulBranchless))
      return false;
    if (!(p == P::None || p == P::Stateful || p == P::Naive || p == P::NaiveBranchless ||
          p == P::PrefetchAll || p == P::PrefetchAllBranchless))
      return false;
    if (enc == Enc::BP) return false;
  }
  if (fn == Fn::Compute && enc != Enc::FFOR) return false;
  return true;
}

template <typename T, typename ColumnT>
void require(Fn fn, unsigned n_vec, unsigned n_val, enums::Unpacker u, enums::Patcher p) {
  if (!provided<T>(fn, column_traits<ColumnT>::enc, n_vec, n_val, u, p))
    throw std::invalid_argument(std::string("Could not find correct binding in ") + fn_name(fn) +
                                " " + column_traits<ColumnT>::name + "<" + type_name<T>() + ">");
}

template <typename T>
flsgpu_bp_column to_c(const flsgpu::device::BPColumn<T> &c) {
  return flsgpu_bp_column{c.n_values, c.n_vecs, c.packed_array, c.bit_widths,
                          reinterpret_cast<const uint64_t *>(c.vector_offsets)};
}
template <typename T>
flsgpu_ffor_column to_c(const flsgpu::device::FFORColumn<T> &c) {
  return flsgpu_ffor_column{to_c(c.bp), c.bases};
}
template <typename T>
flsgpu_alp_column to_c(const flsgpu::device::ALPColumn<T> &c) {
  return flsgpu_alp_column{to_c(c.ffor), c.factor_indices, c.fraction_indices, c.n_exceptions,
                           reinterpret_cast<const uint64_t *>(c.exceptions_offsets),
                           c.exceptions, c.positions, c.counts, nullptr};
}
template <typename T>
flsgpu_alp_column to_c(const flsgpu::device::ALPExtendedColumn<T> &c) {
  return flsgpu_alp_column{to_c(c.ffor), c.factor_indices, c.fraction_indices, c.n_exceptions,
                           reinterpret_cast<const uint64_t *>(c.exceptions_offsets),
                           c.exceptions, c.positions, nullptr, c.offsets_counts};
}

inline void sync() { flsgpu::detail::check(flsgpu_synchronize(nullptr), "synchronize"); }

// One launch of the native decoder for each column type.
inline int launch_decompress(const flsgpu::device::BPColumn<uint32_t> &c, uint32_t *o, int) {
  const auto d = to_c(c);
  return flsgpu_decompress_bp_u32(&d, o, FLSGPU_VARIANT_AUTO, nullptr);
}
inline int launch_decompress(const flsgpu::device::BPColumn<uint64_t> &c, uint64_t *o, int) {
  const auto d = to_c(c);
  return flsgpu_decompress_bp_u64(&d, o, FLSGPU_VARIANT_AUTO, nullptr);
}
inline int launch_decompress(const flsgpu::device::FFORColumn<uint32_t> &c, uint32_t *o, int) {
  const auto d = to_c(c);
  return flsgpu_decompress_ffor_u32(&d, o, FLSGPU_VARIANT_AUTO, nullptr);
}
inline int launch_decompress(const flsgpu::device::FFORColumn<uint64_t> &c, uint64_t *o, int) {
  const auto d = to_c(c);
  return flsgpu_decompress_ffor_u64(&d, o, FLSGPU_VARIANT_AUTO, nullptr);
}
inline int launch_decompress(const flsgpu::device::ALPColumn<float> &c, float *o, int p) {
  const auto d = to_c(c);
  return flsgpu_decompress_alp_f32(&d, o, p, FLSGPU_VARIANT_AUTO, nullptr);
}
inline int launch_decompress(const flsgpu::device::ALPColumn<double> &c, double *o, int p) {
  const auto d = to_c(c);
  return flsgpu_decompress_alp_f64(&d, o, p, FLSGPU_VARIANT_AUTO, nullptr);
}
inline int launch_decompress(const flsgpu::device::ALPExtendedColumn<float> &c, float *o, int p) {
  const auto d = to_c(c);
  return flsgpu_decompress_alp_ext_f32(&d, o, p, FLSGPU_VARIANT_AUTO, nullptr);
}
inline int launch_decompress(const flsgpu::device::ALPExtendedColumn<double> &c, double *o, int p) {
  const auto d = to_c(c);
  return flsgpu_decompress_alp_ext_f64(&d, o, p, FLSGPU_VARIANT_AUTO, nullptr);
}

inline int launch_query(const flsgpu::device::BPColumn<uint32_t> &c, uint32_t m, uint8_t *r, int) {
  const auto d = to_c(c);
  return flsgpu_query_any_bp_u32(&d, m, r, FLSGPU_VARIANT_AUTO, nullptr);
}
inline int launch_query(const flsgpu::device::BPColumn<uint64_t> &c, uint64_t m, uint8_t *r, int) {
  const auto d = to_c(c);
  return flsgpu_query_any_bp_u64(&d, m, r, FLSGPU_VARIANT_AUTO, nullptr);
}
inline int launch_query(const flsgpu::device::FFORColumn<uint32_t> &c, uint32_t m, uint8_t *r, int) {
  const auto d = to_c(c);
  return flsgpu_query_any_ffor_u32(&d, m, r, FLSGPU_VARIANT_AUTO, nullptr);
}
inline int launch_query(const flsgpu::device::FFORColumn<uint64_t> &c, uint64_t m, uint8_t *r, int) {
  const auto d = to_c(c);
  return flsgpu_query_any_ffor_u64(&d, m, r, FLSGPU_VARIANT_AUTO, nullptr);
}
inline int launch_query(const flsgpu::device::ALPColumn<float> &c, float m, uint8_t *r, int p) {
  const auto d = to_c(c);
  return flsgpu_query_any_alp_f32(&d, m, r, p, FLSGPU_VARIANT_AUTO, nullptr);
}
inline int launch_query(const flsgpu::device::ALPColumn<double> &c, double m, uint8_t *r, int p) {
  const auto d = to_c(c);
  return flsgpu_query_any_alp_f64(&d, m, r, p, FLSGPU_VARIANT_AUTO, nullptr);
}
inline int launch_query(const flsgpu::device::ALPExtendedColumn<float> &c, float m, uint8_t *r, int p) {
  const auto d = to_c(c);
  return flsgpu_query_any_alp_ext_f32(&d, m, r, p, FLSGPU_VARIANT_AUTO, nullptr);
}
inline int launch_query(const flsgpu::device::ALPExtendedColumn<double> &c, double m, uint8_t *r,
                        int p) {
  const auto d = to_c(c);
  return flsgpu_query_any_alp_ext_f64(&d, m, r, p, FLSGPU_VARIANT_AUTO, nullptr);
}

inline int launch_multi(const flsgpu::device::FFORColumn<uint32_t> *c, uint32_t n, uint32_t v,
                        uint8_t *r, int) {
  flsgpu_ffor_column d[10];
  for (uint32_t i = 0; i < n; ++i) d[i] = to_c(c[i]);
  return flsgpu_query_multi_column_ffor_u32(d, n, v, r, FLSGPU_VARIANT_AUTO, nullptr);
}
inline int launch_multi(const flsgpu::device::FFORColumn<uint64_t> *c, uint32_t n, uint64_t v,
                        uint8_t *r, int) {
  flsgpu_ffor_column d[10];
  for (uint32_t i = 0; i < n; ++i) d[i] = to_c(c[i]);
  return flsgpu_query_multi_column_ffor_u64(d, n, v, r, FLSGPU_VARIANT_AUTO, nullptr);
}
template <typename ColT>
int launch_multi_alp(const ColT *c, uint32_t n, float v, uint8_t *r, int p) {
  flsgpu_alp_column d[10];
  for (uint32_t i = 0; i < n; ++i) d[i] = to_c(c[i]);
  return flsgpu_query_multi_column_alp_f32(d, n, v, r, p, FLSGPU_VARIANT_AUTO, nullptr);
}
template <typename ColT>
int launch_multi_alp(const ColT *c, uint32_t n, double v, uint8_t *r, int p) {
  flsgpu_alp_column d[10];
  for (uint32_t i = 0; i < n; ++i) d[i] = to_c(c[i]);
  return flsgpu_query_multi_column_alp_f64(d, n, v, r, p, FLSGPU_VARIANT_AUTO, nullptr);
}
inline int launch_multi(const flsgpu::device::ALPColumn<float> *c, uint32_t n, float v, uint8_t *r,
                        int p) {
  return launch_multi_alp(c, n, v, r, p);
}
inline int launch_multi(const flsgpu::device::ALPColumn<double> *c, uint32_t n, double v,
                        uint8_t *r, int p) {
  return launch_multi_alp(c, n, v, r, p);
}
inline int launch_multi(const flsgpu::device::ALPExtendedColumn<float> *c, uint32_t n, float v,
                        uint8_t *r, int p) {
  return launch_multi_alp(c, n, v, r, p);
}
inline int launch_multi(const flsgpu::device::ALPExtendedColumn<double> *c, uint32_t n, double v,
                        uint8_t *r, int p) {
  return launch_multi_alp(c, n, v, r, p);
}

// Reads the one-byte result slot; frees it.
inline bool take_flag(uint8_t *d_flag) {
  uint8_t h = 0;
  flsgpu::detail::check(flsgpu_download(&h, d_flag, 1, nullptr), "download");
  flsgpu::detail::free_device_pointer(d_flag);
  return h != 0;
}
inline uint8_t *new_flag() {
  const uint8_t zero = 0;
  return flsgpu::detail::upload(&zero, 1);
}

// Random permutation of the vectors' bit widths with offsets rebuilt
// (src/engine/data.cuh shuffle_bit_widths): the packed words stay, so each
// copy decodes to different values while keeping the total size.
template <typename UT>
void shuffle_bit_widths(flsgpu::host::BPColumn<UT> &bp, std::mt19937_64 &rng) {
  const size_t nv = bp.get_n_vecs();
  std::shuffle(bp.bit_widths, bp.bit_widths + nv, rng);
  size_t off = 0;
  for (size_t v = 0; v < nv; ++v) {
    bp.vector_offsets[v] = off;
    off += size_t(bp.bit_widths[v]) * utils::get_n_lanes<UT>();
  }
}
inline std::mt19937_64 &shuffle_rng() {
  static std::mt19937_64 rng(0x5eed);
  return rng;
}
}  // namespace detail

// decompress_column: n_samples launches (synchronised), then the column is
// copied back into a new T[n_values] that the caller deletes
// (kernels::host::decompress_column, src/engine/kernels.cuh:94-112).
template <typename T, typename ColumnT>
T *decompress_column(const ColumnT column, const unsigned unpack_n_vectors,
                     const unsigned unpack_n_values, const enums::Unpacker unpacker,
                     const enums::Patcher patcher, const uint32_t n_samples) {
  detail::require<T, ColumnT>(detail::Fn::Decompress, unpack_n_vectors, unpack_n_values, unpacker,
                              patcher);
  T *d_out = flsgpu::detail::upload<T>(nullptr, 0, column.n_values);
  for (uint32_t i = 0; i < n_samples; ++i) {
    flsgpu::detail::check(detail::launch_decompress(column, d_out, int(patcher)),
                          "decompress_column");
    detail::sync();
  }
  T *out = new T[column.n_values];
  flsgpu::detail::check(flsgpu_download(out, d_out, column.n_values * sizeof(T), nullptr),
                        "download");
  flsgpu::detail::free_device_pointer(d_out);
  return out;
}

// query_column: "does any value equal magic_value" (kernels.cuh:114-133).
template <typename T, typename ColumnT>
bool query_column(const ColumnT column, const unsigned unpack_n_vectors,
                  const unsigned unpack_n_values, const enums::Unpacker unpacker,
                  const enums::Patcher patcher, const T magic_value, const uint32_t n_samples) {
  detail::require<T, ColumnT>(detail::Fn::Query, unpack_n_vectors, unpack_n_values, unpacker,
                              patcher);
  uint8_t *flag = detail::new_flag();
  for (uint32_t i = 0; i < n_samples; ++i) {
    flsgpu::detail::check(detail::launch_query(column, magic_value, flag, int(patcher)),
                          "query_column");
    detail::sync();
  }
  return detail::take_flag(flag);
}

// compute_column: decode + n_repetitions of ALU work per value, then
// "any == 1" (kernels.cuh:56-89, 135-153). The reference's generated
// bindings fix the repetitions at 10; here the argument is honoured.
template <typename T, typename ColumnT>
bool compute_column(const ColumnT column, const unsigned unpack_n_vectors,
                    const unsigned unpack_n_values, const enums::Unpacker unpacker,
                    const enums::Patcher patcher, const unsigned n_repetitions,
                    const uint32_t n_samples) {
  detail::require<T, ColumnT>(detail::Fn::Compute, unpack_n_vectors, unpack_n_values, unpacker,
                              patcher);
  static_assert(std::is_integral_v<T>, "compute_column decodes FFOR columns");
  uint8_t *flag = detail::new_flag();
  const auto d = detail::to_c(column);
  for (uint32_t i = 0; i < n_samples; ++i) {
    int rc;
    if constexpr (sizeof(T) == 4)
      rc = flsgpu_compute_ffor_u32(&d, n_repetitions, flag, FLSGPU_VARIANT_AUTO, nullptr);
    else
      rc = flsgpu_compute_ffor_u64(&d, n_repetitions, flag, FLSGPU_VARIANT_AUTO, nullptr);
    flsgpu::detail::check(rc, "compute_column");
    detail::sync();
  }
  return detail::take_flag(flag);
}

// query_multi_column (multi-column-host-kernels.cuh:11-123): takes a HOST
// column, shuffles its bit widths in place before each of 10 uploads, then
// queries arities 1..10, n_samples times each; the answer is whether any
// launch found a lane where every value of every column equals magic_value.
template <typename T, typename ColumnT>
bool query_multi_column(ColumnT column, const unsigned unpack_n_vectors,
                        const unsigned unpack_n_values, const enums::Unpacker unpacker,
                        const enums::Patcher patcher, const T magic_value,
                        const uint32_t n_samples) {
  detail::require<T, ColumnT>(detail::Fn::MultiColumn, unpack_n_vectors, unpack_n_values,
                              unpacker, patcher);
  using DeviceColumnT = typename ColumnT::DeviceColumnT;
  constexpr uint32_t MAX_N_COLS = 10;
  DeviceColumnT cols[MAX_N_COLS];
  for (uint32_t c = 0; c < MAX_N_COLS; ++c) {
    if constexpr (std::is_integral_v<T>)
      detail::shuffle_bit_widths(column.bp, detail::shuffle_rng());
    else
      detail::shuffle_bit_widths(column.ffor.bp, detail::shuffle_rng());
    cols[c] = column.copy_to_device();
  }
  bool found = false;
  for (uint32_t arity = 1; arity <= MAX_N_COLS; ++arity) {
    uint8_t *flag = detail::new_flag();
    for (uint32_t i = 0; i < n_samples; ++i) {
      flsgpu::detail::check(detail::launch_multi(cols, arity, magic_value, flag, int(patcher)),
                            "query_multi_column");
      detail::sync();
    }
    found |= detail::take_flag(flag);
  }
  for (uint32_t c = 0; c < MAX_N_COLS; ++c) flsgpu::host::free_column(cols[c]);
  return found;
}

// Non-template overloads of kernel-bindings.cuh:37-175.
#define FLSGPU_DECOMPRESS_OVERLOAD(T, COL)                                                    \
  inline T *decompress_column(const flsgpu::device::COL<T> column, const unsigned nvec,       \
                              const unsigned nval, const enums::Unpacker u,                   \
                              const enums::Patcher p, const uint32_t n_samples) {             \
    return decompress_column<T, flsgpu::device::COL<T>>(column, nvec, nval, u, p, n_samples); \
  }                                                                                           \
  inline bool query_column(const flsgpu::device::COL<T> column, const unsigned nvec,          \
                           const unsigned nval, const enums::Unpacker u,                      \
                           const enums::Patcher p, const T magic, const uint32_t n_samples) { \
    return query_column<T, flsgpu::device::COL<T>>(column, nvec, nval, u, p, magic, n_samples); \
  }
FLSGPU_DECOMPRESS_OVERLOAD(uint32_t, BPColumn)
FLSGPU_DECOMPRESS_OVERLOAD(uint64_t, BPColumn)
FLSGPU_DECOMPRESS_OVERLOAD(uint32_t, FFORColumn)
FLSGPU_DECOMPRESS_OVERLOAD(uint64_t, FFORColumn)
FLSGPU_DECOMPRESS_OVERLOAD(float, ALPColumn)
FLSGPU_DECOMPRESS_OVERLOAD(double, ALPColumn)
FLSGPU_DECOMPRESS_OVERLOAD(float, ALPExtendedColumn)
FLSGPU_DECOMPRESS_OVERLOAD(double, ALPExtendedColumn)
#undef FLSGPU_DECOMPRESS_OVERLOAD
}  // namespace bindings
