// Reading the target sections of a Sabertooth index file (<db>.stidx,
// src/index.rs:1-47): the numeric residues, the sequence offsets and the
// identifiers, with the checks SeedIndex::open makes on them
// (src/index.rs:1654-1800), straight into a device database.
//
// Layout (little-endian): a 4 KiB header (magic, version, alphabet, counts,
// a table of up to 16 sections {kind, offset, len, hash}, a 64-bit checksum
// in its last 8 bytes), the sections at 64-byte aligned offsets, and a
// 16-byte footer ("SBTIDXOK", file length).  The file is memory-mapped; the
// k-mer sections (postings and lookup tables, the prefilter's) are not read.
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cstring>
#include <string>
#include <vector>

#include "../../include/sabertooth_amd.h"
#include "host_internal.h"

namespace {

using st::fail;

constexpr uint64_t kHeader = 4096, kFooter = 16;
constexpr int kMaxSections = 16;
// section kinds (SectionKind, src/index.rs:221-241)
enum : uint32_t { kResidues = 1, kSeqOffsets = 2, kIdBytes = 3, kIdOffsets = 4, kPostings = 5 };

// Hasher64 (src/index.rs:122-182): the index's integrity hash.  Words are
// read little-endian; each is folded in by xor and a multiply-xorshift, the
// tail word zero-padded, then the length is folded in and the state goes
// through SplitMix64's finaliser.  The format fixes it: the header and
// section checksums in a file are this function's values.
struct Hash64 {
  uint64_t s = 0x9E3779B97F4A7C15ull;
  uint64_t len = 0;
  void mix(uint64_t w) {
    const uint64_t x = (s ^ w) * 0xBF58476D1CE4E5B9ull;
    s = x ^ (x >> 31);
  }
  // one contiguous buffer (every use here hashes a whole region at once)
  uint64_t of(const uint8_t *p, uint64_t n) {
    len = n;
    uint64_t k = 0;
    for (; k + 8 <= n; k += 8) {
      uint64_t w;
      std::memcpy(&w, p + k, 8);
      mix(w);
    }
    if (k < n) {
      uint8_t b[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      std::memcpy(b, p + k, n - k);
      uint64_t w;
      std::memcpy(&w, b, 8);
      mix(w);
    }
    mix(len);
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
};

uint32_t rd32(const uint8_t *p) {
  uint32_t v;
  std::memcpy(&v, p, 4);
  return v;
}
uint64_t rd64(const uint8_t *p) {
  uint64_t v;
  std::memcpy(&v, p, 8);
  return v;
}

struct Sec {
  uint64_t off = 0, len = 0, hash = 0;
  bool present = false;
};

// A validated, mapped index file.
struct Mapped {
  const uint8_t *base = nullptr;
  uint64_t size = 0;
  st_stidx_info info{};
  Sec sec[6];  // by kind 1..5
  uint64_t num_postings = 0;
  uint32_t post_w = 8, pos_bits = 32;  // format 1: 8-byte (target << 32 | position) postings
  ~Mapped() {
    if (base) munmap((void *)base, size);
  }
  const uint8_t *at(uint32_t k) const { return base + sec[k].off; }
};

std::string who(const char *fn, const char *path) { return std::string(fn) + ": " + path + ": "; }

// SeedIndex::open's checks on the header, the footer and the four target
// sections (src/index.rs:536-640, 1654-1800); `verify` adds each section's
// checksum (SeedIndex::verify, src/index.rs:1872-1880).
int open_index(const char *fn, const char *path, bool verify, Mapped &m) {
  if (!path) return fail(ST_ERR_INVALID, std::string(fn) + ": null path");
  const int fd = ::open(path, O_RDONLY);
  if (fd < 0) return fail(ST_ERR_INVALID, who(fn, path) + "cannot open");
  struct stat sb;
  if (fstat(fd, &sb) != 0) {
    ::close(fd);
    return fail(ST_ERR_INVALID, who(fn, path) + "cannot stat");
  }
  m.size = (uint64_t)sb.st_size;
  if (m.size < kHeader + kFooter) {
    ::close(fd);
    return fail(ST_ERR_INVALID, who(fn, path) + "too small to be an index (incomplete build?)");
  }
  void *p = mmap(nullptr, m.size, PROT_READ, MAP_PRIVATE, fd, 0);
  ::close(fd);
  if (p == MAP_FAILED) return fail(ST_ERR_INVALID, who(fn, path) + "mmap failed");
  m.base = (const uint8_t *)p;
  const uint8_t *h = m.base;
  if (std::memcmp(h, "SBTIDX\0\x01", 8) != 0) return fail(ST_ERR_INVALID, who(fn, path) + "not a Sabertooth index (bad magic)");
  Hash64 hh;
  if (hh.of(h, kHeader - 8) != rd64(h + kHeader - 8))
    return fail(ST_ERR_INVALID, who(fn, path) + "index header checksum mismatch (corrupt file)");
  // header fields (encode_header, src/index.rs:475-532)
  m.info.version = rd32(h + 8);
  if (m.info.version < 1 || m.info.version > 2)
    return fail(ST_ERR_INVALID, who(fn, path) + "index format version not supported (reads 1..2)");
  m.info.alphabet = rd32(h + 16);
  uint64_t o = 8 + 4 * 4 + 64 + 4 + 4 + 8 + 8;  // magic, 4 u32, pattern, masked, window, locut, hicut
  m.info.num_seqs = rd64(h + o);
  m.info.total_residues = rd64(h + o + 8);
  m.info.max_seq_len = rd64(h + o + 16);
  m.num_postings = rd64(h + o + 24);
  o += 7 * 8;                       // num_seqs .. masked_residues
  o += 4 + 4 + 3 * 16 + 1024 + 8 + 32;  // source kind, pad, files, path, created, tool version
  const uint32_t ns = rd32(h + o);
  o += 8;
  if (ns > (uint32_t)kMaxSections) return fail(ST_ERR_INVALID, who(fn, path) + "bad section count");
  for (uint32_t i = 0; i < ns; i++) {
    const uint8_t *e = h + o + 32 * (uint64_t)i;
    const uint32_t k = rd32(e);
    if (k < 1 || k > 10) return fail(ST_ERR_INVALID, who(fn, path) + "unknown section kind");
    if (k <= kPostings) {
      m.sec[k].off = rd64(e + 8);
      m.sec[k].len = rd64(e + 16);
      m.sec[k].hash = rd64(e + 24);
      m.sec[k].present = true;
    }
  }
  // format 2 appends the posting encoding after the section table
  // (posting_width bytes, posting_pos_bits position bits)
  if (m.info.version >= 2) {
    const uint8_t *f2 = h + o + 32 * (uint64_t)kMaxSections;
    m.post_w = rd32(f2 + 4);
    m.pos_bits = rd32(f2 + 8);
    if (m.post_w < 1 || m.post_w > 8 || m.pos_bits > 32 || m.pos_bits > 8 * m.post_w)
      return fail(ST_ERR_INVALID, who(fn, path) + "bad posting encoding in header (corrupt index)");
  }
  // completion footer
  if (std::memcmp(m.base + m.size - kFooter, "SBTIDXOK", 8) != 0 || rd64(m.base + m.size - 8) != m.size)
    return fail(ST_ERR_INVALID, who(fn, path) + "missing completion footer (incomplete or truncated index)");
  if (m.info.alphabet != 1) return fail(ST_ERR_INVALID, who(fn, path) + "unsupported alphabet id");
  for (uint32_t k = kResidues; k <= kIdOffsets; k++) {
    const Sec &s = m.sec[k];
    if (!s.present) return fail(ST_ERR_INVALID, who(fn, path) + "a target section is missing");
    if (s.off < kHeader || s.len > m.size - kFooter || s.off > m.size - kFooter - s.len)
      return fail(ST_ERR_INVALID, who(fn, path) + "section out of bounds");
    if ((k == kSeqOffsets || k == kIdOffsets) && (s.off % 8 != 0 || s.len % 8 != 0))
      return fail(ST_ERR_INVALID, who(fn, path) + "misaligned u64 section");
  }
  const uint64_t n = m.info.num_seqs;
  if (n >= m.size / 8) return fail(ST_ERR_INVALID, who(fn, path) + "sequence count exceeds the file (corrupt index)");
  const uint64_t *so = (const uint64_t *)m.at(kSeqOffsets);
  const uint64_t *io = (const uint64_t *)m.at(kIdOffsets);
  // the end-value checks of SeedIndex::open (src/index.rs:1756-1768)
  if (m.sec[kResidues].len != m.info.total_residues || m.sec[kSeqOffsets].len != 8 * (n + 1) ||
      m.sec[kIdOffsets].len != 8 * (n + 1) || so[n] != m.info.total_residues || io[n] != m.sec[kIdBytes].len)
    return fail(ST_ERR_INVALID, who(fn, path) + "section sizes disagree with the header (corrupt index)");
  // offsets must not decrease (each record is a slice of its section)
  for (uint64_t i = 0; i < n; i++)
    if (so[i + 1] < so[i] || io[i + 1] < io[i])
      return fail(ST_ERR_INVALID, who(fn, path) + "decreasing offsets (corrupt index)");
  if (verify)
    for (uint32_t k = kResidues; k <= kIdOffsets; k++) {
      Hash64 hs;
      if (hs.of(m.at(k), m.sec[k].len) != m.sec[k].hash)
        return fail(ST_ERR_INVALID, who(fn, path) + "section checksum mismatch");
    }
  m.info.id_bytes = m.sec[kIdBytes].len;
  return ST_OK;
}

}  // namespace

extern "C" {

int st_stidx_info_read(const char *path, int verify, st_stidx_info *info) {
  if (!info) return fail(ST_ERR_INVALID, "st_stidx_info_read: null info");
  Mapped m;
  const int rc = open_index("st_stidx_info_read", path, verify != 0, m);
  if (rc) return rc;
  *info = m.info;
  return ST_OK;
}

int st_stidx_read_targets(const char *path, uint8_t *residues, uint64_t *offsets) {
  Mapped m;
  const int rc = open_index("st_stidx_read_targets", path, false, m);
  if (rc) return rc;
  if (residues) std::memcpy(residues, m.at(kResidues), m.sec[kResidues].len);
  if (offsets) std::memcpy(offsets, m.at(kSeqOffsets), m.sec[kSeqOffsets].len);
  return ST_OK;
}

int st_stidx_read_ids(const char *path, uint8_t *bytes, uint64_t *offsets) {
  Mapped m;
  const int rc = open_index("st_stidx_read_ids", path, false, m);
  if (rc) return rc;
  if (bytes) std::memcpy(bytes, m.at(kIdBytes), m.sec[kIdBytes].len);
  if (offsets) std::memcpy(offsets, m.at(kIdOffsets), m.sec[kIdOffsets].len);
  return ST_OK;
}

}  // extern "C"

namespace {

// The decoded posting section (SeedIndex::posting, src/index.rs:2294-2303:
// post_w little-endian bytes, the target above pos_bits, the position below).
int read_postings(const char *fn, const char *path, std::vector<uint64_t> &dec, uint64_t *count) {
  Mapped m;
  int rc = open_index(fn, path, false, m);
  if (rc) return rc;
  const Sec &s = m.sec[kPostings];
  if (!s.present || s.off < kHeader || s.len > m.size - kFooter || s.off > m.size - kFooter - s.len)
    return fail(ST_ERR_INVALID, who(fn, path) + "posting section missing or out of bounds");
  if (m.num_postings > s.len / m.post_w || s.len != m.num_postings * m.post_w)
    return fail(ST_ERR_INVALID, who(fn, path) + "posting count disagrees with the section");
  const uint64_t n = m.num_postings;
  if (count) *count = n;
  const uint64_t pmask = (1ull << m.pos_bits) - 1;
  dec.resize(n);
  const uint8_t *p = m.at(kPostings);
  for (uint64_t k = 0; k < n; k++) {
    uint64_t v = 0;
    std::memcpy(&v, p + k * m.post_w, m.post_w);
    const uint64_t tid = v >> m.pos_bits, pos = v & pmask;
    if (tid >= m.info.num_seqs || pos > 0xffffffffull)
      return fail(ST_ERR_INVALID, who(fn, path) + "posting outside the targets (corrupt index)");
    dec[k] = (tid << 32) | pos;
  }
  return ST_OK;
}

}  // namespace

extern "C" {

int st_stidx_read_postings(const char *path, uint64_t *postings, uint64_t cap, uint64_t *n) {
  if (!n) return fail(ST_ERR_INVALID, "st_stidx_read_postings: null count");
  std::vector<uint64_t> dec;
  const int rc = read_postings("st_stidx_read_postings", path, dec, n);
  if (rc) return rc;
  if (postings && cap >= dec.size()) std::memcpy(postings, dec.data(), dec.size() * 8);
  return ST_OK;
}

int st_postings_from_stidx(st_ctx *ctx, const char *path, st_postings **out) {
  if (!ctx || !out) return fail(ST_ERR_INVALID, "st_postings_from_stidx: null argument");
  *out = nullptr;
  std::vector<uint64_t> dec;
  const int rc = read_postings("st_postings_from_stidx", path, dec, nullptr);
  if (rc) return rc;
  return st_postings_create(ctx, dec.data(), dec.size(), out);
}

int st_db_create_from_stidx(st_ctx *ctx, const char *path, st_db **out) {
  if (!ctx || !out) return fail(ST_ERR_INVALID, "st_db_create_from_stidx: null argument");
  *out = nullptr;
  Mapped m;
  const int rc = open_index("st_db_create_from_stidx", path, false, m);
  if (rc) return rc;
  // the mapped sections go to the device as they lie in the file
  return st_db_create(ctx, m.at(kResidues), m.info.total_residues, (const uint64_t *)m.at(kSeqOffsets),
                      m.info.num_seqs, out);
}

}  // extern "C"
