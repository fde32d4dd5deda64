// Column files (.flsc): one encoded column on disk, loadable into device
// memory with one allocation and large host-to-device copies. Plays the role
// of the reference's byte storer (src/alp/storer.hpp:10-53), which only
// appends arrays to a memory buffer, as a self-describing file with a
// section table, so a reader can validate every offset before a kernel
// touches the data.
//
// Layout (little-endian):
//   [0, 4096)     header: magic, version, encoding, sizes, ALP-RD scalars,
//                 section table {id, element bytes, element count, offset}
//   [4096, ...)   payload: the sections, each at a 256-byte aligned offset
//                 relative to the payload start
// Loading copies the payload unchanged into one device block, so section
// offsets are device offsets too.
#include <cstdio>
#include <cstring>
#include <memory>
#include <vector>

#include "common.hpp"

namespace flsgpu_hip {
namespace {

constexpr char MAGIC[8] = {'F', 'L', 'S', 'G', 'P', 'U', 'C', '1'};
constexpr uint32_t VERSION = 1;
constexpr uint64_t HEADER_BYTES = 4096;
constexpr uint64_t ALIGN = 256;
constexpr int MAX_SECTIONS = 16;
constexpr size_t STAGE_BYTES = size_t(32) << 20;

enum SectionId : uint32_t {
  S_PACKED = 1, S_BIT_WIDTHS, S_VECTOR_OFFSETS, S_BASES, S_FACTOR_INDICES, S_FRACTION_INDICES,
  S_EXCEPTIONS_OFFSETS, S_EXCEPTIONS, S_POSITIONS, S_COUNTS, S_OFFSETS_COUNTS, S_RIGHT_PACKED,
  S_LEFT_PACKED
};

struct Section {
  uint32_t id;
  uint32_t elem_bytes;
  uint64_t count;
  uint64_t offset;  // from the payload start
};

struct Header {
  char magic[8];
  uint32_t version;
  int32_t encoding;
  uint64_t n_values;
  uint64_t n_vecs;
  uint64_t payload_bytes;
  uint32_t n_sections;
  uint32_t rd_right_bit_width;
  uint32_t rd_left_bit_width;
  uint32_t rd_dictionary_size;
  uint16_t rd_dictionary[8];
  Section sections[MAX_SECTIONS];
};
static_assert(sizeof(Header) <= HEADER_BYTES, "header fits its block");

struct Kind {
  int width;     // bits of the element type
  bool ffor;     // has bases
  bool alp;      // ALP or G-ALP
  bool ext;      // G-ALP layout
  bool rd;       // ALP-RD
};

bool kind_of(int encoding, Kind *k) {
  switch (encoding) {
    case FLSGPU_ENC_BP_U32: *k = {32, false, false, false, false}; return true;
    case FLSGPU_ENC_FFOR_U32: *k = {32, true, false, false, false}; return true;
    case FLSGPU_ENC_BP_U64: *k = {64, false, false, false, false}; return true;
    case FLSGPU_ENC_FFOR_U64: *k = {64, true, false, false, false}; return true;
    case FLSGPU_ENC_ALP_F32: *k = {32, true, true, false, false}; return true;
    case FLSGPU_ENC_ALP_F64: *k = {64, true, true, false, false}; return true;
    case FLSGPU_ENC_ALP_EXT_F32: *k = {32, true, true, true, false}; return true;
    case FLSGPU_ENC_ALP_EXT_F64: *k = {64, true, true, true, false}; return true;
    case FLSGPU_ENC_ALPRD_F32: *k = {32, false, false, false, true}; return true;
    case FLSGPU_ENC_ALPRD_F64: *k = {64, false, false, false, true}; return true;
    default: return false;
  }
}

struct FileCloser {
  void operator()(FILE *f) const {
    if (f) fclose(f);
  }
};
using FilePtr = std::unique_ptr<FILE, FileCloser>;

uint64_t align_up(uint64_t x) { return (x + ALIGN - 1) / ALIGN * ALIGN; }

// ---- writing ---------------------------------------------------------------

struct Pending {
  Section s;
  const void *data;
};

struct Writer {
  Header h{};
  std::vector<Pending> parts;
  uint64_t cursor = 0;
  // Upload layout: the reference's over-read padding after the packed words
  // (sizeof(T) * N_LANES * 4 elements, src/flsgpu/structs.cuh:104-109) and
  // the exception arrays (MAX_UNPACK_N_VECS elements, :160,173-177). The
  // kernels here never read it; callers' own kernels written for the
  // reference layout may. Files carry no padding.
  bool pad = false;
  void add(uint32_t id, uint32_t elem_bytes, uint64_t count, const void *data) {
    uint64_t extra = 0;
    if (pad && (id == S_PACKED || id == S_RIGHT_PACKED)) extra = 512;  // (W/8) * (1024/W) * 4
    if (pad && (id == S_EXCEPTIONS || id == S_POSITIONS)) extra = 4;
    Pending p{{id, elem_bytes, count, cursor}, data};
    parts.push_back(p);
    cursor = align_up(cursor + uint64_t(elem_bytes) * (count + extra));
  }
};

// Words of a packed array that vectors with these widths and offsets use.
int packed_words(const uint8_t *bw, const uint64_t *offs, uint64_t n_vecs, int width,
                 uint64_t *words) {
  const uint64_t nl = 1024 / width;
  uint64_t end = 0;
  for (uint64_t v = 0; v < n_vecs; ++v) {
    if (bw[v] > width)
      return set_error(FLSGPU_EINVAL, "vector %llu: bit width %u above %d",
                       (unsigned long long)v, unsigned(bw[v]), width);
    const uint64_t e = offs[v] + uint64_t(bw[v]) * nl;
    if (e > end) end = e;
  }
  *words = end;
  return FLSGPU_OK;
}

int exception_total(const uint64_t *eo, const uint16_t *counts, uint64_t n_vecs, uint64_t *n) {
  uint64_t end = 0;
  for (uint64_t v = 0; v < n_vecs; ++v) {
    if (counts[v] > 1024)
      return set_error(FLSGPU_EINVAL, "vector %llu: %u exceptions", (unsigned long long)v,
                       unsigned(counts[v]));
    const uint64_t e = eo[v] + counts[v];
    if (e > end) end = e;
  }
  *n = end;
  return FLSGPU_OK;
}

int ext_exception_total(const uint64_t *eo, const uint16_t *oc, uint64_t n_vecs, int width,
                        uint64_t *n) {
  const uint64_t nl = 1024 / width;
  uint64_t end = 0;
  for (uint64_t v = 0; v < n_vecs; ++v)
    for (uint64_t l = 0; l < nl; ++l) {
      const uint32_t x = oc[v * nl + l];
      const uint64_t e = eo[v] + (x & 0x3FFu) + (x >> 10);
      if (e > end) end = e;
    }
  *n = end;
  return FLSGPU_OK;
}

// Exception positions must lie in 0..1023 and be strictly increasing within
// a vector's list (standard layout, ALP-RD) or within a lane's list (G-ALP,
// where lane l's list may only hold positions p with p % N_LANES == l). The
// decoders rely on it: the standard-layout patch takes an exception's index
// in the list as its rank among the vector's exception positions, and each
// position is written by one owner exactly once. The reference's CPU
// patcher (src/alp/decoder.hpp:137-145) would apply other lists last-wins,
// so such lists are rejected rather than decoded differently. Bounds of
// every list against n_exc must have been checked by the caller.
int check_positions(const Kind &k, uint64_t nv, const uint64_t *eo, const uint16_t *counts,
                    const uint16_t *oc, const uint16_t *pos) {
  const uint64_t nl = uint64_t(1024 / k.width);
  auto increasing = [&](uint64_t first, uint64_t n, int lane, uint64_t v) -> int {
    int prev = -1;
    for (uint64_t i = 0; i < n; ++i) {
      const int p = pos[first + i];
      if (p >= 1024 || p <= prev || (lane >= 0 && uint64_t(p) % nl != uint64_t(lane)))
        return set_error(FLSGPU_EINVAL,
                         "vector %llu: exception positions must be strictly increasing%s "
                         "(entry %llu, position %d)",
                         (unsigned long long)v, lane >= 0 ? " and in their lane" : "",
                         (unsigned long long)i, p);
      prev = p;
    }
    return FLSGPU_OK;
  };
  for (uint64_t v = 0; v < nv; ++v) {
    int rc;
    if (k.ext) {
      for (uint64_t l = 0; l < nl; ++l) {
        const uint32_t x = oc[v * nl + l];
        if ((rc = increasing(eo[v] + (x & 0x3FFu), x >> 10, int(l), v))) return rc;
      }
    } else if ((rc = increasing(eo[v], counts[v], -1, v))) {
      return rc;
    }
  }
  return FLSGPU_OK;
}

int plan_columns(int encoding, const void *column, Writer *w) {
  Kind k;
  if (!kind_of(encoding, &k)) return set_error(FLSGPU_EINVAL, "unknown encoding %d", encoding);
  if (!column) return set_error(FLSGPU_EINVAL, "null column");
  const uint32_t wb = uint32_t(k.width / 8);
  const uint64_t nl = uint64_t(1024 / k.width);
  Header &h = w->h;
  memcpy(h.magic, MAGIC, sizeof(MAGIC));
  h.version = VERSION;
  h.encoding = encoding;
  int rc;
  if (k.rd) {
    const auto *c = static_cast<const flsgpu_alprd_column *>(column);
    h.n_values = c->n_values;
    h.n_vecs = c->n_vecs;
    if (c->n_values != c->n_vecs * 1024) return set_error(FLSGPU_EINVAL, "n_values != n_vecs * 1024");
    if (c->dictionary_size < 1 || c->dictionary_size > 8 || c->left_bit_width < 1 ||
        c->left_bit_width > 3 || c->right_bit_width < uint32_t(k.width - 16) ||
        c->right_bit_width > uint32_t(k.width - 1))
      return set_error(FLSGPU_EINVAL, "ALP-RD parameters out of range");
    h.rd_right_bit_width = c->right_bit_width;
    h.rd_left_bit_width = c->left_bit_width;
    h.rd_dictionary_size = c->dictionary_size;
    memcpy(h.rd_dictionary, c->dictionary, sizeof(h.rd_dictionary));
    uint64_t n_exc = 0;
    if (c->n_vecs && (!c->exceptions_offsets || !c->counts))
      return set_error(FLSGPU_EINVAL, "null exception offsets or counts");
    if (c->n_vecs && (rc = exception_total(c->exceptions_offsets, c->counts, c->n_vecs, &n_exc)))
      return rc;
    if (n_exc && !c->positions) return set_error(FLSGPU_EINVAL, "null positions");
    if (c->n_vecs && (rc = check_positions(k, c->n_vecs, c->exceptions_offsets, c->counts,
                                           nullptr, c->positions)))
      return rc;
    w->add(S_RIGHT_PACKED, wb, c->n_vecs * c->right_bit_width * nl, c->right_packed);
    w->add(S_LEFT_PACKED, 2, c->n_vecs * c->left_bit_width * 64, c->left_packed);
    w->add(S_EXCEPTIONS_OFFSETS, 8, c->n_vecs, c->exceptions_offsets);
    w->add(S_EXCEPTIONS, 2, n_exc, c->exceptions);
    w->add(S_POSITIONS, 2, n_exc, c->positions);
    w->add(S_COUNTS, 2, c->n_vecs, c->counts);
  } else {
    const flsgpu_bp_column *bp;
    const flsgpu_alp_column *a = nullptr;
    const void *bases = nullptr;
    if (k.alp) {
      a = static_cast<const flsgpu_alp_column *>(column);
      bp = &a->ffor.bp;
      bases = a->ffor.bases;
    } else if (k.ffor) {
      const auto *f = static_cast<const flsgpu_ffor_column *>(column);
      bp = &f->bp;
      bases = f->bases;
    } else {
      bp = static_cast<const flsgpu_bp_column *>(column);
    }
    h.n_values = bp->n_values;
    h.n_vecs = bp->n_vecs;
    if (bp->n_values != bp->n_vecs * 1024) return set_error(FLSGPU_EINVAL, "n_values != n_vecs * 1024");
    uint64_t words = 0;
    if (bp->n_vecs &&
        (rc = packed_words(bp->bit_widths, bp->vector_offsets, bp->n_vecs, k.width, &words)))
      return rc;
    w->add(S_PACKED, wb, words, bp->packed_array);
    w->add(S_BIT_WIDTHS, 1, bp->n_vecs, bp->bit_widths);
    w->add(S_VECTOR_OFFSETS, 8, bp->n_vecs, bp->vector_offsets);
    if (k.ffor) w->add(S_BASES, wb, bp->n_vecs, bases);
    if (a) {
      uint64_t n_exc = 0;
      if (bp->n_vecs) {
        if (!a->exceptions_offsets || !(k.ext ? a->offsets_counts : a->counts))
          return set_error(FLSGPU_EINVAL, "null exception offsets or counts");
        rc = k.ext ? ext_exception_total(a->exceptions_offsets, a->offsets_counts, bp->n_vecs,
                                         k.width, &n_exc)
                   : exception_total(a->exceptions_offsets, a->counts, bp->n_vecs, &n_exc);
        if (rc) return rc;
        if (n_exc && !a->positions) return set_error(FLSGPU_EINVAL, "null positions");
        if ((rc = check_positions(k, bp->n_vecs, a->exceptions_offsets, a->counts,
                                  a->offsets_counts, a->positions)))
          return rc;
      }
      w->add(S_FACTOR_INDICES, 1, bp->n_vecs, a->factor_indices);
      w->add(S_FRACTION_INDICES, 1, bp->n_vecs, a->fraction_indices);
      w->add(S_EXCEPTIONS_OFFSETS, 8, bp->n_vecs, a->exceptions_offsets);
      w->add(S_EXCEPTIONS, wb, n_exc, a->exceptions);
      w->add(S_POSITIONS, 2, n_exc, a->positions);
      if (k.ext)
        w->add(S_OFFSETS_COUNTS, 2, bp->n_vecs * nl, a->offsets_counts);
      else
        w->add(S_COUNTS, 2, bp->n_vecs, a->counts);
    }
  }
  for (const Pending &p : w->parts)
    if (p.s.count && !p.data) return set_error(FLSGPU_EINVAL, "null array (section %u)", p.s.id);
  h.n_sections = uint32_t(w->parts.size());
  for (size_t i = 0; i < w->parts.size(); ++i) h.sections[i] = w->parts[i].s;
  h.payload_bytes = w->cursor;
  return FLSGPU_OK;
}

// ---- reading ---------------------------------------------------------------

struct Parsed {
  Header h;
  Kind k;
  uint64_t file_bytes;
};

const Section *find(const Header &h, uint32_t id) {
  for (uint32_t i = 0; i < h.n_sections; ++i)
    if (h.sections[i].id == id) return &h.sections[i];
  return nullptr;
}

int read_at(FILE *f, uint64_t pos, void *dst, uint64_t bytes) {
  if (fseeko(f, off_t(pos), SEEK_SET) != 0)
    return set_error(FLSGPU_EINVAL, "seek to %llu failed", (unsigned long long)pos);
  if (bytes && fread(dst, 1, bytes, f) != bytes)
    return set_error(FLSGPU_EINVAL, "short read at %llu", (unsigned long long)pos);
  return FLSGPU_OK;
}

template <typename X>
int read_section(FILE *f, const Header &h, uint32_t id, std::vector<X> *out) {
  const Section *s = find(h, id);
  if (!s) return set_error(FLSGPU_EINVAL, "section %u missing", id);
  out->resize(s->count);
  return read_at(f, HEADER_BYTES + s->offset, out->data(), s->count * sizeof(X));
}

// Checks the header and every offset a kernel derives from the metadata, so
// that a loaded column cannot make a decoder read outside its block.
int parse(FILE *f, Parsed *p) {
  if (fseeko(f, 0, SEEK_END) != 0) return set_error(FLSGPU_EINVAL, "seek failed");
  p->file_bytes = uint64_t(ftello(f));
  if (p->file_bytes < HEADER_BYTES) return set_error(FLSGPU_EINVAL, "file shorter than its header");
  Header &h = p->h;
  int rc = read_at(f, 0, &h, sizeof(Header));
  if (rc) return rc;
  if (memcmp(h.magic, MAGIC, sizeof(MAGIC)) != 0) return set_error(FLSGPU_EINVAL, "not a column file");
  if (h.version != VERSION) return set_error(FLSGPU_EINVAL, "column file version %u", h.version);
  if (!kind_of(h.encoding, &p->k)) return set_error(FLSGPU_EINVAL, "unknown encoding %d", h.encoding);
  const Kind &k = p->k;
  if (h.n_values != h.n_vecs * 1024 || h.n_vecs > (uint64_t(1) << 40))
    return set_error(FLSGPU_EINVAL, "bad sizes in header");
  if (h.n_sections > MAX_SECTIONS) return set_error(FLSGPU_EINVAL, "bad section count");
  if (HEADER_BYTES + h.payload_bytes > p->file_bytes)
    return set_error(FLSGPU_EINVAL, "file truncated: %llu of %llu bytes",
                     (unsigned long long)p->file_bytes,
                     (unsigned long long)(HEADER_BYTES + h.payload_bytes));
  for (uint32_t i = 0; i < h.n_sections; ++i) {
    const Section &s = h.sections[i];
    if (s.offset % ALIGN || s.elem_bytes == 0 || s.elem_bytes > 8 ||
        s.count > h.payload_bytes / s.elem_bytes ||
        s.offset > h.payload_bytes - s.count * s.elem_bytes)
      return set_error(FLSGPU_EINVAL, "section %u out of the payload", s.id);
  }
  const uint64_t nv = h.n_vecs, nl = uint64_t(1024 / k.width), wb = uint64_t(k.width / 8);
  auto need = [&](uint32_t id, uint64_t elem, uint64_t count, bool exact) -> int {
    const Section *s = find(h, id);
    if (!s) return set_error(FLSGPU_EINVAL, "section %u missing", id);
    if (s->elem_bytes != elem || (exact ? s->count != count : s->count < count))
      return set_error(FLSGPU_EINVAL, "section %u has the wrong size", id);
    return FLSGPU_OK;
  };
  if (k.rd) {
    if (h.rd_dictionary_size < 1 || h.rd_dictionary_size > 8 || h.rd_left_bit_width < 1 ||
        h.rd_left_bit_width > 3 || h.rd_right_bit_width < uint32_t(k.width - 16) ||
        h.rd_right_bit_width > uint32_t(k.width - 1))
      return set_error(FLSGPU_EINVAL, "ALP-RD parameters out of range");
    if ((rc = need(S_RIGHT_PACKED, wb, nv * h.rd_right_bit_width * nl, true)) ||
        (rc = need(S_LEFT_PACKED, 2, nv * h.rd_left_bit_width * 64, true)) ||
        (rc = need(S_EXCEPTIONS_OFFSETS, 8, nv, true)) || (rc = need(S_COUNTS, 2, nv, true)))
      return rc;
  } else {
    if ((rc = need(S_PACKED, wb, 0, false)) || (rc = need(S_BIT_WIDTHS, 1, nv, true)) ||
        (rc = need(S_VECTOR_OFFSETS, 8, nv, true)) ||
        (k.ffor && (rc = need(S_BASES, wb, nv, true))))
      return rc;
    std::vector<uint8_t> bw;
    std::vector<uint64_t> offs;
    if ((rc = read_section(f, h, S_BIT_WIDTHS, &bw)) ||
        (rc = read_section(f, h, S_VECTOR_OFFSETS, &offs)))
      return rc;
    const uint64_t words = find(h, S_PACKED)->count;
    for (uint64_t v = 0; v < nv; ++v)
      if (bw[v] > k.width || offs[v] > words || uint64_t(bw[v]) * nl > words - offs[v])
        return set_error(FLSGPU_EINVAL, "vector %llu reads outside the packed array",
                         (unsigned long long)v);
    if (k.alp) {
      if ((rc = need(S_FACTOR_INDICES, 1, nv, true)) ||
          (rc = need(S_FRACTION_INDICES, 1, nv, true)) ||
          (rc = need(S_EXCEPTIONS_OFFSETS, 8, nv, true)) ||
          (k.ext ? (rc = need(S_OFFSETS_COUNTS, 2, nv * nl, true))
                 : (rc = need(S_COUNTS, 2, nv, true))))
        return rc;
      std::vector<uint8_t> fi, ri;
      if ((rc = read_section(f, h, S_FACTOR_INDICES, &fi)) ||
          (rc = read_section(f, h, S_FRACTION_INDICES, &ri)))
        return rc;
      const unsigned max_f = k.width == 32 ? 10 : 18, max_e = k.width == 32 ? 10 : 20;
      for (uint64_t v = 0; v < nv; ++v)
        if (fi[v] > max_f || ri[v] > max_e)
          return set_error(FLSGPU_EINVAL, "vector %llu: ALP table index out of range",
                           (unsigned long long)v);
    }
  }
  if (k.alp || k.rd) {
    const Section *se = find(h, S_EXCEPTIONS), *sp = find(h, S_POSITIONS);
    if (!se || !sp || se->elem_bytes != (k.rd ? 2 : wb) || sp->elem_bytes != 2 ||
        se->count != sp->count)
      return set_error(FLSGPU_EINVAL, "exception sections malformed");
    std::vector<uint64_t> eo;
    std::vector<uint16_t> pos;
    if ((rc = read_section(f, h, S_EXCEPTIONS_OFFSETS, &eo)) ||
        (rc = read_section(f, h, S_POSITIONS, &pos)))
      return rc;
    const uint64_t n_exc = se->count;
    for (uint64_t i = 0; i < n_exc; ++i)
      if (pos[i] >= 1024)
        return set_error(FLSGPU_EINVAL, "exception %llu: position %u outside the vector",
                         (unsigned long long)i, unsigned(pos[i]));
    if (k.ext) {
      std::vector<uint16_t> oc;
      if ((rc = read_section(f, h, S_OFFSETS_COUNTS, &oc))) return rc;
      for (uint64_t v = 0; v < nv; ++v)
        for (uint64_t l = 0; l < nl; ++l) {
          const uint32_t x = oc[v * nl + l];
          if (eo[v] > n_exc || (x & 0x3FFu) + (x >> 10) > n_exc - eo[v])
            return set_error(FLSGPU_EINVAL, "vector %llu lane %llu: exceptions out of range",
                             (unsigned long long)v, (unsigned long long)l);
        }
      if ((rc = check_positions(k, nv, eo.data(), nullptr, oc.data(), pos.data()))) return rc;
    } else {
      std::vector<uint16_t> counts;
      if ((rc = read_section(f, h, S_COUNTS, &counts))) return rc;
      for (uint64_t v = 0; v < nv; ++v)
        if (counts[v] > 1024 || eo[v] > n_exc || counts[v] > n_exc - eo[v])
          return set_error(FLSGPU_EINVAL, "vector %llu: exceptions out of range",
                           (unsigned long long)v);
      if ((rc = check_positions(k, nv, eo.data(), counts.data(), nullptr, pos.data()))) return rc;
    }
  }
  return FLSGPU_OK;
}

// Fills the encoding's descriptor with pointers into `base` (host or device).
void describe(const Parsed &p, char *base, void *column) {
  const Header &h = p.h;
  auto at = [&](uint32_t id) -> void * {
    const Section *s = find(h, id);
    return s ? static_cast<void *>(base + s->offset) : nullptr;
  };
  if (p.k.rd) {
    auto *c = static_cast<flsgpu_alprd_column *>(column);
    memset(c, 0, sizeof(*c));
    c->n_values = h.n_values;
    c->n_vecs = h.n_vecs;
    c->right_bit_width = h.rd_right_bit_width;
    c->left_bit_width = h.rd_left_bit_width;
    c->dictionary_size = h.rd_dictionary_size;
    memcpy(c->dictionary, h.rd_dictionary, sizeof(c->dictionary));
    c->right_packed = at(S_RIGHT_PACKED);
    c->left_packed = static_cast<const uint16_t *>(at(S_LEFT_PACKED));
    c->exceptions_offsets = static_cast<const uint64_t *>(at(S_EXCEPTIONS_OFFSETS));
    c->exceptions = static_cast<const uint16_t *>(at(S_EXCEPTIONS));
    c->positions = static_cast<const uint16_t *>(at(S_POSITIONS));
    c->counts = static_cast<const uint16_t *>(at(S_COUNTS));
    return;
  }
  flsgpu_bp_column bp{h.n_values, h.n_vecs, at(S_PACKED),
                      static_cast<const uint8_t *>(at(S_BIT_WIDTHS)),
                      static_cast<const uint64_t *>(at(S_VECTOR_OFFSETS))};
  if (p.k.alp) {
    auto *c = static_cast<flsgpu_alp_column *>(column);
    memset(c, 0, sizeof(*c));
    c->ffor.bp = bp;
    c->ffor.bases = at(S_BASES);
    c->factor_indices = static_cast<const uint8_t *>(at(S_FACTOR_INDICES));
    c->fraction_indices = static_cast<const uint8_t *>(at(S_FRACTION_INDICES));
    c->n_exceptions = find(h, S_EXCEPTIONS)->count;
    c->exceptions_offsets = static_cast<const uint64_t *>(at(S_EXCEPTIONS_OFFSETS));
    c->exceptions = at(S_EXCEPTIONS);
    c->positions = static_cast<const uint16_t *>(at(S_POSITIONS));
    c->counts = p.k.ext ? nullptr : static_cast<const uint16_t *>(at(S_COUNTS));
    c->offsets_counts = p.k.ext ? static_cast<const uint16_t *>(at(S_OFFSETS_COUNTS)) : nullptr;
  } else if (p.k.ffor) {
    auto *c = static_cast<flsgpu_ffor_column *>(column);
    c->bp = bp;
    c->bases = at(S_BASES);
  } else {
    *static_cast<flsgpu_bp_column *>(column) = bp;
  }
}

int open_parsed(const char *path, int expected, FilePtr *f, Parsed *p) {
  if (!path) return set_error(FLSGPU_EINVAL, "null path");
  f->reset(fopen(path, "rb"));
  if (!*f) return set_error(FLSGPU_EINVAL, "cannot open %s", path);
  int rc = parse(f->get(), p);
  if (rc) return rc;
  if (expected >= 0 && p->h.encoding != expected)
    return set_error(FLSGPU_EINVAL, "%s holds encoding %d, not %d", path, p->h.encoding, expected);
  return FLSGPU_OK;
}

}  // namespace
}  // namespace flsgpu_hip

using namespace flsgpu_hip;

extern "C" {

int flsgpu_file_write(const char *path, int encoding, const void *host_column) {
  if (!path) return set_error(FLSGPU_EINVAL, "null path");
  Writer w;
  int rc = plan_columns(encoding, host_column, &w);
  if (rc) return rc;
  FilePtr f(fopen(path, "wb"));
  if (!f) return set_error(FLSGPU_EINVAL, "cannot create %s", path);
  std::vector<char> head(HEADER_BYTES, 0);
  memcpy(head.data(), &w.h, sizeof(Header));
  if (fwrite(head.data(), 1, head.size(), f.get()) != head.size())
    return set_error(FLSGPU_EINVAL, "write failed: %s", path);
  static const char zeros[ALIGN] = {};
  uint64_t pos = 0;
  for (const Pending &p : w.parts) {
    if (p.s.offset > pos && fwrite(zeros, 1, p.s.offset - pos, f.get()) != p.s.offset - pos)
      return set_error(FLSGPU_EINVAL, "write failed: %s", path);
    pos = p.s.offset;
    const uint64_t bytes = uint64_t(p.s.elem_bytes) * p.s.count;
    if (bytes && fwrite(p.data, 1, bytes, f.get()) != bytes)
      return set_error(FLSGPU_EINVAL, "write failed: %s", path);
    pos += bytes;
  }
  if (w.h.payload_bytes > pos && fwrite(zeros, 1, w.h.payload_bytes - pos, f.get()) !=
                                     w.h.payload_bytes - pos)
    return set_error(FLSGPU_EINVAL, "write failed: %s", path);
  if (fflush(f.get()) != 0) return set_error(FLSGPU_EINVAL, "write failed: %s", path);
  return FLSGPU_OK;
}

int flsgpu_column_upload(int encoding, const void *host_column, void *device_column,
                         void **d_block, flsgpu_stream_t stream) {
  if (!device_column || !d_block) return set_error(FLSGPU_EINVAL, "null argument");
  *d_block = nullptr;
  Writer w;
  w.pad = true;
  int rc = plan_columns(encoding, host_column, &w);
  if (rc) return rc;
  const uint64_t total = w.h.payload_bytes;
  void *dev = nullptr;
  if ((rc = check_hip(hipMalloc(&dev, total ? total : 1), "hipMalloc"))) return rc;
  char *base = static_cast<char *>(dev);
  rc = check_hip(hipMemsetAsync(dev, 0, total ? total : 1, stream), "hipMemsetAsync");
  for (size_t i = 0; !rc && i < w.parts.size(); ++i) {
    const Pending &p = w.parts[i];
    const uint64_t bytes = uint64_t(p.s.elem_bytes) * p.s.count;
    if (bytes)
      rc = check_hip(hipMemcpyAsync(base + p.s.offset, p.data, bytes, hipMemcpyHostToDevice, stream),
                     "hipMemcpyAsync H2D");
  }
  if (!rc) rc = check_hip(hipStreamSynchronize(stream), "hipStreamSynchronize");
  if (rc) {
    (void)hipFree(dev);
    return rc;
  }
  Parsed p;
  p.h = w.h;
  kind_of(encoding, &p.k);
  p.file_bytes = 0;
  describe(p, base, device_column);
  *d_block = dev;
  return FLSGPU_OK;
}

int flsgpu_column_free(void *d_block) {
  if (!d_block) return FLSGPU_OK;
  return check_hip(hipFree(d_block), "hipFree");
}

int flsgpu_file_stat(const char *path, flsgpu_file_info *info) {
  if (!info) return set_error(FLSGPU_EINVAL, "null info");
  FilePtr f;
  Parsed p;
  int rc = open_parsed(path, -1, &f, &p);
  if (rc) return rc;
  info->encoding = p.h.encoding;
  info->n_values = p.h.n_values;
  info->n_vecs = p.h.n_vecs;
  info->payload_bytes = p.h.payload_bytes;
  info->file_bytes = p.file_bytes;
  return FLSGPU_OK;
}

int flsgpu_file_read_host(const char *path, int encoding, void *column, void **h_block) {
  if (!column || !h_block) return set_error(FLSGPU_EINVAL, "null argument");
  *h_block = nullptr;
  FilePtr f;
  Parsed p;
  int rc = open_parsed(path, encoding, &f, &p);
  if (rc) return rc;
  char *block = static_cast<char *>(malloc(p.h.payload_bytes ? p.h.payload_bytes : 1));
  if (!block) return set_error(FLSGPU_ENOMEM, "host allocation of %llu bytes",
                               (unsigned long long)p.h.payload_bytes);
  if ((rc = read_at(f.get(), HEADER_BYTES, block, p.h.payload_bytes))) {
    free(block);
    return rc;
  }
  describe(p, block, column);
  *h_block = block;
  return FLSGPU_OK;
}

void flsgpu_file_free_host(void *h_block) { free(h_block); }

int flsgpu_file_load(const char *path, int encoding, void *column, void **d_block,
                     flsgpu_stream_t stream) {
  if (!column || !d_block) return set_error(FLSGPU_EINVAL, "null argument");
  *d_block = nullptr;
  FilePtr f;
  Parsed p;
  int rc = open_parsed(path, encoding, &f, &p);
  if (rc) return rc;
  const uint64_t total = p.h.payload_bytes;
  void *dev = nullptr;
  if ((rc = check_hip(hipMalloc(&dev, total ? total : 1), "hipMalloc"))) return rc;
  // two pinned staging buffers: the file read of one overlaps the copy of
  // the other
  void *stage[2] = {nullptr, nullptr};
  hipEvent_t done[2] = {nullptr, nullptr};
  auto cleanup = [&](int code) {
    for (int i = 0; i < 2; ++i) {
      if (done[i]) (void)hipEventDestroy(done[i]);
      if (stage[i]) (void)hipHostFree(stage[i]);
    }
    if (code) (void)hipFree(dev);
    return code;
  };
  const size_t chunk = size_t(total < STAGE_BYTES ? (total ? total : 1) : STAGE_BYTES);
  for (int i = 0; i < 2; ++i) {
    if ((rc = check_hip(hipHostMalloc(&stage[i], chunk, hipHostMallocDefault), "hipHostMalloc")) ||
        (rc = check_hip(hipEventCreateWithFlags(&done[i], hipEventDisableTiming), "hipEventCreate")))
      return cleanup(rc);
  }
  if (fseeko(f.get(), off_t(HEADER_BYTES), SEEK_SET) != 0)
    return cleanup(set_error(FLSGPU_EINVAL, "seek failed"));
  bool used[2] = {false, false};
  for (uint64_t pos = 0, i = 0; pos < total; pos += chunk, ++i) {
    const int b = int(i % 2);
    const size_t n = size_t(total - pos < chunk ? total - pos : chunk);
    if (used[b] && (rc = check_hip(hipEventSynchronize(done[b]), "hipEventSynchronize")))
      return cleanup(rc);
    if (fread(stage[b], 1, n, f.get()) != n)
      return cleanup(set_error(FLSGPU_EINVAL, "short read at %llu", (unsigned long long)pos));
    if ((rc = check_hip(hipMemcpyAsync(static_cast<char *>(dev) + pos, stage[b], n,
                                       hipMemcpyHostToDevice, stream), "hipMemcpyAsync H2D")) ||
        (rc = check_hip(hipEventRecord(done[b], stream), "hipEventRecord")))
      return cleanup(rc);
    used[b] = true;
  }
  if ((rc = check_hip(hipStreamSynchronize(stream), "hipStreamSynchronize"))) return cleanup(rc);
  describe(p, static_cast<char *>(dev), column);
  *d_block = dev;
  return cleanup(FLSGPU_OK);
}

}  // extern "C"
