/*
 * flsgpu — MI355X (gfx950) FastLanes / ALP / ALP-RD decoder, C ABI.
 *
 * Plain C: device pointers, sizes and a hipStream_t; no C++ or torch types.
 * Every entry point is asynchronous on `stream` (NULL = the null stream),
 * returns FLSGPU_OK (0) or a negative status, and never calls exit().
 * flsgpu_last_error() returns a message for the calling thread's last failure.
 *
 * The column descriptors mirror the reference's device column structs field
 * by field (src/flsgpu/structs.cuh:33-84): a packed array of W-bit words,
 * per-vector bit widths (uint8), per-vector word offsets (uint64) and, for
 * FFOR, per-vector bases. A vector holds 1024 values; n_values must be
 * n_vecs * 1024 (the reference's format has full vectors only,
 * src/engine/data.cuh:206, src/alp/alp-bindings.cu:116-126).
 *
 * Unlike the reference (structs.cuh:104-109, 160-177), the kernels read no
 * word beyond a vector's own bit_width * N_LANES words, so no over-read
 * padding is needed after the packed array or the exception arrays.
 *
 * Reductions (query_any, filter_sum, compute, multi-column) are two launches
 * on `stream`: the work kernel leaves one partial per block in a scratch
 * area owned by the library, and a one-block kernel writes the final result
 * to the caller's device pointer. Each (device, stream) pair gets its own
 * 1 MiB area on its first reduction, so reductions on different streams may
 * run concurrently; calls on one stream are ordered by the stream. A
 * reduction captured into a hipGraph keeps the area of the stream it was
 * captured on: replays of that graph must not overlap reductions issued on
 * that stream. flsgpu_release_stream frees a stream's area.
 *
 * Exception lists: positions are 0..1023 within the vector and strictly
 * increasing within each vector's list (standard layout, ALP-RD) or within
 * each lane's list (G-ALP, whose lane l holds only positions p with
 * p % N_LANES == l), as the reference's encoder and G-ALP conversion write
 * them (src/alp/encoder.hpp:352-358, src/flsgpu/structs.cuh:231-306). The
 * decoders rely on it; flsgpu_column_upload and flsgpu_file_load reject
 * other lists with FLSGPU_EINVAL.
 */
#ifndef FLSGPU_H
#define FLSGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t *flsgpu_stream_t; /* == hipStream_t */

enum flsgpu_status {
  FLSGPU_OK = 0,
  FLSGPU_EINVAL = -1,       /* bad argument: null pointer, size, enum value */
  FLSGPU_EUNSUPPORTED = -2, /* variant / patcher / type combination not provided */
  FLSGPU_EHIP = -3,         /* a HIP runtime call failed */
  FLSGPU_ENOMEM = -4
};

/* Reference enums::Unpacker values (src/engine/enums.cuh:21-41), accepted as
 * the `unpacker` argument of the C++ shim; the C ABI takes a native variant. */
enum flsgpu_variant {
  FLSGPU_VARIANT_AUTO = 0,        /* autotuned table, else the default below */
  FLSGPU_VARIANT_LANE = 1,        /* one thread per FastLanes lane, full lane in registers */
  FLSGPU_VARIANT_LANE_NT = 2,     /* as LANE, non-temporal (streaming) stores */
  FLSGPU_VARIANT_HALF_LANE = 3,   /* two threads per lane, each half of the slots */
  FLSGPU_VARIANT_HALF_LANE_NT = 4,
  FLSGPU_VARIANT_QUARTER_LANE = 5,/* four threads per lane */
  FLSGPU_VARIANT_QUARTER_LANE_NT = 6,
  /* decode into LDS, then 16-byte non-temporal stores of each vector's slot
   * range; ALP exceptions are applied in LDS. 64-bit columns use the quarter
   * split for both. Needs a 16-byte aligned output (else the plain NT
   * variant of the same split runs). */
  FLSGPU_VARIANT_HALF_LANE_STAGED = 7,
  FLSGPU_VARIANT_QUARTER_LANE_STAGED = 8,
  /* staged, eight threads per lane (64-bit FFOR / ALP / ALP-RD and 32-bit
   * FFOR; 32-bit ALP and ALP-RD use their HALF_LANE_STAGED split): a smaller
   * LDS stage per wave, so more waves in flight per CU */
  FLSGPU_VARIANT_EIGHTH_LANE_STAGED = 9,
  FLSGPU_VARIANT_COUNT = 10
};

/* Reference enums::Patcher values (src/engine/enums.cuh:43-53). */
enum flsgpu_patcher {
  FLSGPU_PATCHER_NONE = 0,
  FLSGPU_PATCHER_DUMMY = 1,
  FLSGPU_PATCHER_STATELESS = 2,
  FLSGPU_PATCHER_STATEFUL = 3,
  FLSGPU_PATCHER_NAIVE = 4,
  FLSGPU_PATCHER_NAIVE_BRANCHLESS = 5,
  FLSGPU_PATCHER_PREFETCH_POSITION = 6,
  FLSGPU_PATCHER_PREFETCH_ALL = 7,
  FLSGPU_PATCHER_PREFETCH_ALL_BRANCHLESS = 8
};

/* device::BPColumn<T> (src/flsgpu/structs.cuh:33-41) */
typedef struct flsgpu_bp_column {
  uint64_t n_values;
  uint64_t n_vecs;
  const void *packed_array;        /* uint32_t or uint64_t words */
  const uint8_t *bit_widths;       /* [n_vecs] */
  const uint64_t *vector_offsets;  /* [n_vecs], in words */
} flsgpu_bp_column;

/* device::FFORColumn<T> (structs.cuh:43-48) */
typedef struct flsgpu_ffor_column {
  flsgpu_bp_column bp;
  const void *bases;               /* [n_vecs] W-bit bases */
} flsgpu_ffor_column;

/* device::ALPColumn<T> / ALPExtendedColumn<T> (structs.cuh:50-84). The
 * FACT/FRAC tables are compiled into the kernels instead of being uploaded
 * per column. Exactly one of counts (standard layout) and offsets_counts
 * (G-ALP lane-divided layout, (count << 10) | offset per lane) is used. */
typedef struct flsgpu_alp_column {
  flsgpu_ffor_column ffor;         /* unsigned words of the float's width */
  const uint8_t *factor_indices;   /* [n_vecs] */
  const uint8_t *fraction_indices; /* [n_vecs] */
  uint64_t n_exceptions;
  const uint64_t *exceptions_offsets; /* [n_vecs] */
  const void *exceptions;          /* float / double */
  const uint16_t *positions;       /* 0..1023 within the vector */
  const uint16_t *counts;          /* [n_vecs]            (ALPColumn) */
  const uint16_t *offsets_counts;  /* [n_vecs * N_LANES]  (ALPExtendedColumn) */
} flsgpu_alp_column;

/* ALP-RD column. The reference has no packed ALP-RD format (src/alp/rd.hpp
 * decodes unpacked parts only); this one follows the upstream-ALP convention
 * of rd.hpp:48-56: right parts bit-packed with FFOR base 0 at right_bit_width
 * (vector v at word v * right_bit_width * N_LANES), left parts as dictionary
 * indices bit-packed as uint16 at left_bit_width (vector v at
 * v * left_bit_width * 64 words), exceptions = original left parts. */
typedef struct flsgpu_alprd_column {
  uint64_t n_values;
  uint64_t n_vecs;
  uint32_t right_bit_width;        /* W-16 .. W-1 */
  uint32_t left_bit_width;         /* 1 .. 3 */
  uint32_t dictionary_size;        /* 1 .. 8 */
  uint16_t dictionary[8];
  const void *right_packed;        /* uint32_t / uint64_t words */
  const uint16_t *left_packed;
  const uint64_t *exceptions_offsets; /* [n_vecs] */
  const uint16_t *exceptions;      /* original left parts */
  const uint16_t *positions;
  const uint16_t *counts;          /* [n_vecs] */
} flsgpu_alprd_column;

/* Fused filter predicates for flsgpu_query_filter_sum_*. */
enum flsgpu_predicate {
  /* value in [lo, hi) */
  FLSGPU_PRED_RANGE = 0,
  /* FOR offset (value - vector base) below half of the vector's range:
   * offset < 2^(bit_width-1); no row passes at bit_width 0. */
  FLSGPU_PRED_OFFSET_LOWER_HALF = 1,
  /* every row: count and sum of the whole column */
  FLSGPU_PRED_ALL = 2
};

/* ---- library ------------------------------------------------------------ */
const char *flsgpu_last_error(void);
/* Optional; the first call on a device initialises it lazily. */
int flsgpu_init(int device);
/* Autotuner table (replaces the SQL best-unpacker query,
 * sql/select-best-unpacker-ffor.sql:1-7): the variant FLSGPU_VARIANT_AUTO
 * resolves to for an encoding (enum flsgpu_encoding below) and bit width;
 * bit_width 255 is the entry used when a call does not know the width.
 * flsgpu.autotune measures the table and loads it through this call. */
int flsgpu_set_tuning(uint32_t encoding, uint32_t bit_width, int variant);
int flsgpu_get_tuning(uint32_t encoding, uint32_t bit_width);

/* ---- device memory helpers (replace GPUArray / copy_to_device,
 *      src/flsgpu/host-utils.cuh:42-111, structs.cuh:103-229) --------------- */
int flsgpu_upload(void **d_ptr, const void *h_ptr, size_t bytes, size_t pad_bytes,
                  flsgpu_stream_t stream);
int flsgpu_download(void *h_ptr, const void *d_ptr, size_t bytes, flsgpu_stream_t stream);
int flsgpu_free(void *d_ptr);
/* Waits for `stream` (NULL: the whole device, as cudaDeviceSynchronize in
 * kernels::host::*, src/engine/kernels.cuh:106,128,148). */
int flsgpu_synchronize(flsgpu_stream_t stream);
/* Waits for `stream` and frees its reduction scratch area on the current
 * device (no reference counterpart; the reference has one default stream).
 * A later reduction on the stream allocates a new one. At most 256 streams
 * per device hold an area at once. */
int flsgpu_release_stream(flsgpu_stream_t stream);

/* ---- decompress: kernels::host::decompress_column (src/engine/kernels.cuh:94-112),
 *      bindings::decompress_column (src/generated-bindings/kernel-bindings.cuh:11-16,37-175) */
int flsgpu_decompress_bp_u32(const flsgpu_bp_column *col, uint32_t *out, int variant,
                             flsgpu_stream_t stream);
int flsgpu_decompress_bp_u64(const flsgpu_bp_column *col, uint64_t *out, int variant,
                             flsgpu_stream_t stream);
int flsgpu_decompress_ffor_u32(const flsgpu_ffor_column *col, uint32_t *out, int variant,
                               flsgpu_stream_t stream);
int flsgpu_decompress_ffor_u64(const flsgpu_ffor_column *col, uint64_t *out, int variant,
                               flsgpu_stream_t stream);
/* ALPColumn: patcher in {NONE, DUMMY, STATELESS, STATEFUL} (Dummy/None skip patching).
 * ALPExtendedColumn: patcher in {NAIVE, NAIVE_BRANCHLESS, PREFETCH_POSITION,
 * PREFETCH_ALL, PREFETCH_ALL_BRANCHLESS} (generate-kernel-bindings.py:355-357).
 * All patchers produce identical output; they map to the native patch paths. */
int flsgpu_decompress_alp_f32(const flsgpu_alp_column *col, float *out, int patcher, int variant,
                              flsgpu_stream_t stream);
int flsgpu_decompress_alp_f64(const flsgpu_alp_column *col, double *out, int patcher, int variant,
                              flsgpu_stream_t stream);
int flsgpu_decompress_alp_ext_f32(const flsgpu_alp_column *col, float *out, int patcher,
                                  int variant, flsgpu_stream_t stream);
int flsgpu_decompress_alp_ext_f64(const flsgpu_alp_column *col, double *out, int patcher,
                                  int variant, flsgpu_stream_t stream);
/* ALP-RD: no reference GPU counterpart (rd.hpp:146-172 is CPU-only, uncalled). */
int flsgpu_decompress_alprd_f32(const flsgpu_alprd_column *col, float *out, int variant,
                                flsgpu_stream_t stream);
int flsgpu_decompress_alprd_f64(const flsgpu_alprd_column *col, double *out, int variant,
                                flsgpu_stream_t stream);

/* ---- query "any value == magic": kernels::host::query_column
 *      (src/engine/kernels.cuh:114-133, device :34-54, MagicChecker
 *      src/engine/device-utils.cuh:122-143). *d_found (uint8) is written with
 *      0 or 1 by the kernel (the reference only writes on a hit). */
int flsgpu_query_any_bp_u32(const flsgpu_bp_column *col, uint32_t magic, uint8_t *d_found,
                            int variant, flsgpu_stream_t stream);
int flsgpu_query_any_bp_u64(const flsgpu_bp_column *col, uint64_t magic, uint8_t *d_found,
                            int variant, flsgpu_stream_t stream);
int flsgpu_query_any_ffor_u32(const flsgpu_ffor_column *col, uint32_t magic, uint8_t *d_found,
                              int variant, flsgpu_stream_t stream);
int flsgpu_query_any_ffor_u64(const flsgpu_ffor_column *col, uint64_t magic, uint8_t *d_found,
                              int variant, flsgpu_stream_t stream);
int flsgpu_query_any_alp_f32(const flsgpu_alp_column *col, float magic, uint8_t *d_found,
                             int patcher, int variant, flsgpu_stream_t stream);
int flsgpu_query_any_alp_f64(const flsgpu_alp_column *col, double magic, uint8_t *d_found,
                             int patcher, int variant, flsgpu_stream_t stream);
int flsgpu_query_any_alp_ext_f32(const flsgpu_alp_column *col, float magic, uint8_t *d_found,
                                 int patcher, int variant, flsgpu_stream_t stream);
int flsgpu_query_any_alp_ext_f64(const flsgpu_alp_column *col, double magic, uint8_t *d_found,
                                 int patcher, int variant, flsgpu_stream_t stream);

/* ---- fused decode + filter + sum (no reference counterpart: the reference
 *      query kernel only answers "any == magic"). d_result[0] = number of
 *      rows passing, d_result[1] = sum (mod 2^64) of the passing values.
 *      For a BP column the FOR offset of FLSGPU_PRED_OFFSET_LOWER_HALF is the
 *      value itself (base 0). */
int flsgpu_query_filter_sum_bp_u32(const flsgpu_bp_column *col, int predicate, uint64_t lo,
                                   uint64_t hi, uint64_t *d_result, int variant,
                                   flsgpu_stream_t stream);
int flsgpu_query_filter_sum_bp_u64(const flsgpu_bp_column *col, int predicate, uint64_t lo,
                                   uint64_t hi, uint64_t *d_result, int variant,
                                   flsgpu_stream_t stream);
int flsgpu_query_filter_sum_ffor_u32(const flsgpu_ffor_column *col, int predicate, uint64_t lo,
                                     uint64_t hi, uint64_t *d_result, int variant,
                                     flsgpu_stream_t stream);
int flsgpu_query_filter_sum_ffor_u64(const flsgpu_ffor_column *col, int predicate, uint64_t lo,
                                     uint64_t hi, uint64_t *d_result, int variant,
                                     flsgpu_stream_t stream);

/* The same fused filter + sum for 32-bit FFOR columns, allocating nothing.
 * No counterpart in the reference. For callers that cannot allocate device
 * memory (code under hipGraph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the block partials live in a static 32 KiB device array
 * of the library (2048 partial pairs; the grid is capped at that, so any
 * column size works), one copy per device, set up by flsgpu_init or the
 * library's first call on the device. Every call on a device shares that
 * device's array: calls that overlap on one device, on any streams, race;
 * calls ordered on one stream do not. */
int flsgpu_query_filter_sum_ffor_u32_noalloc(const flsgpu_ffor_column *col, int predicate,
                                             uint64_t lo, uint64_t hi, uint64_t *d_result,
                                             flsgpu_stream_t stream);

/* ---- synthetic ALU load: kernels::host::compute_column (kernels.cuh:135-153,
 *      device :56-89): per value n_repetitions of *=3, <<=3, +=3, >>=3, &=0,
 *      then "any == 1"; *d_result is written with 0 or 1. */
int flsgpu_compute_bp_u32(const flsgpu_bp_column *col, uint32_t n_repetitions,
                          uint8_t *d_result, int variant, flsgpu_stream_t stream);
int flsgpu_compute_bp_u64(const flsgpu_bp_column *col, uint32_t n_repetitions,
                          uint8_t *d_result, int variant, flsgpu_stream_t stream);
int flsgpu_compute_ffor_u32(const flsgpu_ffor_column *col, uint32_t n_repetitions,
                            uint8_t *d_result, int variant, flsgpu_stream_t stream);
int flsgpu_compute_ffor_u64(const flsgpu_ffor_column *col, uint32_t n_repetitions,
                            uint8_t *d_result, int variant, flsgpu_stream_t stream);

/* ---- multi-column query: multi_column::query_multi_column (arity 1..10,
 *      src/generated-bindings/multi-column-device-kernels.cuh:10-443). The
 *      answer (*d_found = 1) holds iff some (vector, lane) has every slot of
 *      that lane equal to `value` in every column, i.e. a whole FastLanes
 *      lane where all columns match; 0 otherwise. Columns must have equal
 *      n_vecs; n_cols in 1..10. */
int flsgpu_query_multi_column_ffor_u32(const flsgpu_ffor_column *cols, uint32_t n_cols,
                                       uint32_t value, uint8_t *d_found, int variant,
                                       flsgpu_stream_t stream);
int flsgpu_query_multi_column_ffor_u64(const flsgpu_ffor_column *cols, uint32_t n_cols,
                                       uint64_t value, uint8_t *d_found, int variant,
                                       flsgpu_stream_t stream);
int flsgpu_query_multi_column_alp_f32(const flsgpu_alp_column *cols, uint32_t n_cols,
                                      float value, uint8_t *d_found, int patcher, int variant,
                                      flsgpu_stream_t stream);
int flsgpu_query_multi_column_alp_f64(const flsgpu_alp_column *cols, uint32_t n_cols,
                                      double value, uint8_t *d_found, int patcher, int variant,
                                      flsgpu_stream_t stream);

/* ---- device-side encode (no reference counterpart: the reference encodes on
 *      the host only, src/alp/alp-bindings.cu:83-206, src/fls/ffor.cpp). The
 *      output bytes equal the host encoder's: flsgpu_alp_encode_with_params_*
 *      (encode_simdized + analyze_ffor + FastLanes FFOR, src/alp/encoder.hpp:
 *      104-115, 302-379) and, for integers, FFOR at base = vector minimum and
 *      width of (max - min). Asynchronous on `stream`; three launches.
 *
 *      The descriptor's arrays are caller-allocated device buffers written by
 *      the call, sized for the worst case: packed n_vecs * 1024 words;
 *      exceptions and positions n_vecs * 1024; offsets_counts n_vecs * N_LANES
 *      (G-ALP only); every per-vector array n_vecs. The call sets n_values and
 *      n_vecs in the descriptor. d_totals (device, 3 x u64) receives the packed
 *      words, the exception count and, for G-ALP, 1 if some lane holds more
 *      than 63 exceptions (the 16-bit header cannot encode it; the column is
 *      then unusable in that layout). d_workspace: flsgpu_encode_workspace_bytes
 *      (n_vecs) device bytes. exps / facs: per-vector exponent (<= 10 / 18) and
 *      factor (<= exponent, <= 9 for floats), device arrays; a vector whose
 *      pair is outside these ranges is encoded with (0, 0), which is lossless
 *      (every value that does not round-trip becomes an exception). */
size_t flsgpu_encode_workspace_bytes(uint64_t n_vecs);
int flsgpu_alp_encode_device_f32(const float *in, uint64_t n_vecs, const uint8_t *exps,
                                 const uint8_t *facs, int galp, flsgpu_alp_column *col,
                                 uint64_t *d_totals, void *d_workspace, flsgpu_stream_t stream);
int flsgpu_alp_encode_device_f64(const double *in, uint64_t n_vecs, const uint8_t *exps,
                                 const uint8_t *facs, int galp, flsgpu_alp_column *col,
                                 uint64_t *d_totals, void *d_workspace, flsgpu_stream_t stream);
int flsgpu_ffor_encode_device_u32(const uint32_t *in, uint64_t n_vecs, flsgpu_ffor_column *col,
                                  uint64_t *d_totals, void *d_workspace, flsgpu_stream_t stream);
int flsgpu_ffor_encode_device_u64(const uint64_t *in, uint64_t n_vecs, flsgpu_ffor_column *col,
                                  uint64_t *d_totals, void *d_workspace, flsgpu_stream_t stream);

/* ---- batched mixed-encoding decode (no reference counterpart: the
 *      reference launches and synchronises once per column,
 *      src/engine/kernels.cuh:94-112). One persistent launch decodes every
 *      item from a device-side task list that each wave walks on its own
 *      (task t to wave t mod the grid's waves); no state carries over from
 *      one launch to the next, so the launch can be replayed from a
 *      hipGraph. ---------------------------------------------------------- */
enum flsgpu_encoding {
  FLSGPU_ENC_BP_U32 = 0,    /* column: flsgpu_bp_column */
  FLSGPU_ENC_FFOR_U32 = 1,  /* column: flsgpu_ffor_column */
  FLSGPU_ENC_BP_U64 = 2,
  FLSGPU_ENC_FFOR_U64 = 3,
  FLSGPU_ENC_ALP_F32 = 4,   /* column: flsgpu_alp_column, standard layout */
  FLSGPU_ENC_ALP_F64 = 5,
  FLSGPU_ENC_ALP_EXT_F32 = 6, /* column: flsgpu_alp_column, G-ALP layout */
  FLSGPU_ENC_ALP_EXT_F64 = 7,
  FLSGPU_ENC_ALPRD_F32 = 8, /* column: flsgpu_alprd_column */
  FLSGPU_ENC_ALPRD_F64 = 9
};

typedef struct flsgpu_batch_item {
  int encoding;           /* enum flsgpu_encoding */
  const void *column;     /* host pointer to the descriptor (read at plan creation) */
  void *out;              /* device output of n_values elements */
  int patcher;            /* ALP items: a patching patcher of the layout (none/dummy are rejected) */
} flsgpu_batch_item;

typedef struct flsgpu_batch_plan flsgpu_batch_plan;

/* Copies the descriptors and a task list to the device (allocates).
 * `variant` is ignored: every encoding of one element width runs the batch
 * kernel's one LDS-staged shape. */
int flsgpu_batch_plan_create(const flsgpu_batch_item *items, uint32_t n_items, int variant,
                             flsgpu_batch_plan **plan);
uint64_t flsgpu_batch_plan_tasks(const flsgpu_batch_plan *plan);
/* One kernel launch for the whole batch. Launches of one plan write the same
 * outputs: they must not overlap. */
int flsgpu_batch_run(flsgpu_batch_plan *plan, flsgpu_stream_t stream);
/* Captures flsgpu_batch_run into a hipGraph on first use, then replays it. */
int flsgpu_batch_graph_launch(flsgpu_batch_plan *plan, flsgpu_stream_t stream);
int flsgpu_batch_plan_destroy(flsgpu_batch_plan *plan);

/* ---- column files (.flsc): one encoded column on disk. Stands in for the
 *      reference's byte storer (src/alp/storer.hpp:10-53) and its size model
 *      (src/alp/alp-bindings.cu:19-43), which only append arrays to memory,
 *      with a self-describing file: a 4 KiB header (magic "FLSGPUC1",
 *      version, encoding, sizes, ALP-RD scalars, section table) and the
 *      column's arrays at 256-byte aligned offsets. Loading validates every
 *      size and offset a decoder derives from the metadata (bit widths,
 *      vector and exception offsets, counts, G-ALP headers, positions, ALP
 *      table indices), so a corrupt file fails with FLSGPU_EINVAL instead of
 *      reaching a kernel. --------------------------------------------------- */
typedef struct flsgpu_file_info {
  int encoding;             /* enum flsgpu_encoding */
  uint64_t n_values;
  uint64_t n_vecs;
  uint64_t payload_bytes;   /* bytes of the column's arrays (device block size) */
  uint64_t file_bytes;
} flsgpu_file_info;

/* Writes a column whose descriptor (the struct named by `encoding` in enum
 * flsgpu_encoding) holds HOST pointers. Only the used part of each array is
 * stored (packed words up to the last vector's end, n_exceptions entries). */
int flsgpu_file_write(const char *path, int encoding, const void *host_column);
int flsgpu_file_stat(const char *path, flsgpu_file_info *info);
/* Reads the column into one new device block (*d_block; release with
 * flsgpu_free) and fills `column` (the descriptor type of `encoding`; the
 * file must hold that encoding) with device pointers into it. Staged through
 * two pinned buffers; returns after the copy has completed on `stream`. */
int flsgpu_file_load(const char *path, int encoding, void *column, void **d_block,
                     flsgpu_stream_t stream);
/* The same into host memory (*h_block; release with flsgpu_file_free_host). */
int flsgpu_file_read_host(const char *path, int encoding, void *column, void **h_block);
void flsgpu_file_free_host(void *h_block);

/* host::*Column::copy_to_device and free_column (src/flsgpu/structs.cuh:
 * 103-229, 334-400) for any encoding: copies a column whose descriptor holds
 * HOST pointers into one new device block (*d_block) and fills
 * `device_column` (the descriptor type of `encoding`) with device pointers
 * into it. The block keeps the reference's over-read padding after the packed
 * words (sizeof(T) * N_LANES * 4 elements) and the exception arrays (4
 * elements), zero-filled, for kernels written against the reference layout;
 * this library's kernels do not need it. Returns after the copies completed
 * on `stream`. flsgpu_column_free releases the block. */
int flsgpu_column_upload(int encoding, const void *host_column, void *device_column,
                         void **d_block, flsgpu_stream_t stream);
int flsgpu_column_free(void *d_block);

#ifdef __cplusplus
}
#endif
#endif /* FLSGPU_H */
