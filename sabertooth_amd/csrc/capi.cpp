// Host side of the C ABI: device context, resident target arena, query
// profiles, work planning (length buckets, launch shapes) and the launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cctype>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <functional>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "../../include/sabertooth_amd.h"
#include "host_internal.h"

namespace st {
static thread_local std::string g_err;

int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}
}  // namespace st

using namespace st;

// ---- host inputs -----------------------------------------------------------------

namespace st {
// MMseqs2 blosum62.out (as shipped in the reference's src/data/blosum62.out):
// half-bit log-odds in residue-code order, background and lambda.
static const double kB62Back[21] = {0.07422, 0.02469, 0.05363, 0.05431, 0.04742, 0.07415, 0.02621,
                                    0.06792, 0.05815, 0.09891, 0.02499, 0.04465, 0.03854, 0.03426,
                                    0.05161, 0.05723, 0.05089, 0.07292, 0.01303, 0.03228, 0.00001};
static const double kB62Lambda = 0.34657;
static const double kB62Half[21][21] = {
#include "blosum62_halfbits.inc"
};
}  // namespace st

extern "C" {

const char *st_last_error(void) { return st::g_err.c_str(); }
const char *st_version(void) { return "sabertooth_amd 0.1.0 (gfx950)"; }

int st_blosum62(int8_t *score, double *p_back, float *r, double *lambda) {
  // SubstitutionMatrix::from_out_str (src/matrix.rs:56-175) on the embedded data
  double pb[21];
  std::memcpy(pb, kB62Back, sizeof pb);
  bool x_pos = false;
  for (int j = 0; j < 21; j++)
    if (kB62Half[20][j] > 0.0 || kB62Half[j][20] > 0.0) x_pos = true;
  if (!x_pos) {
    pb[20] = 1e-5;
    for (int i = 0; i < 20; i++) pb[i] *= 1.0 - pb[20];
  }
  double prob[21][21];
  for (int i = 0; i < 21; i++)
    for (int j = 0; j < 21; j++)
      prob[i][j] = std::exp(kB62Lambda * (double)kB62Half[i][j]) * pb[i] * pb[j];
  if (r)
    for (int a = 0; a < 20; a++)
      for (int b = 0; b < 20; b++) r[a * 20 + b] = (float)(prob[a][b] / pb[b]);
  if (score)
    for (int i = 0; i < 21; i++)
      for (int j = 0; j < 21; j++) {
        const double v = 2.0 * std::log2(prob[i][j] / (pb[i] * pb[j])) - 0.2;
        double rd = v < 0.0 ? v - 0.5 : v + 0.5;
        rd = std::min(127.0, std::max(-128.0, rd));
        score[i * 21 + j] = (int8_t)rd;
      }
  if (p_back) std::memcpy(p_back, pb, sizeof pb);
  if (lambda) *lambda = kB62Lambda;
  return ST_OK;
}

int st_aa2num(uint8_t table[256]) {
  if (!table) return fail(ST_ERR_INVALID, "st_aa2num: null table");
  static const char kLetters[] = "ACDEFGHIKLMNPQRSTVWYX";
  for (int i = 0; i < 256; i++) table[i] = 20;
  for (int k = 0; k < 21; k++) {
    const unsigned char c = (unsigned char)kLetters[k];
    table[c] = (uint8_t)k;
    table[c - 'A' + 'a'] = (uint8_t)k;
  }
  return ST_OK;
}

int st_seq_rows(const uint8_t *query, uint64_t m, int32_t *rows) {
  if ((!query || !rows) && m) return fail(ST_ERR_INVALID, "st_seq_rows: null pointer");
  int8_t mat[21 * 21];
  st_blosum62(mat, nullptr, nullptr, nullptr);
  for (uint64_t i = 0; i < m; i++) {
    const int qa = query[i] < 21 ? query[i] : 20;
    for (int r = 0; r < 21; r++) rows[i * 21 + r] = mat[qa * 21 + r];
  }
  return ST_OK;
}

int32_t st_size_adjusted_gate(int32_t gate, uint64_t db_residues) {
  // src/prefilter.rs:246-257
  const double x = std::log((double)std::max<uint64_t>(db_residues, 1) / 1.0e8);
  const double slope = x < 0.0 ? 2.5 : 1.0;
  const double v = std::round((double)gate + slope * x);  // half away from zero
  return std::min(60, std::max(20, (int32_t)v));
}

// ---- context -------------------------------------------------------------------------

int st_ctx_create(int device, st_ctx **out) {
  if (!out) return fail(ST_ERR_INVALID, "st_ctx_create: null out");
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
    return fail(ST_ERR_DEVICE, "st_ctx_create: no HIP device available");
  if (device < 0 || device >= n) return fail(ST_ERR_INVALID, "st_ctx_create: bad device ordinal");
  auto *c = new st_ctx();
  c->device = device;
  if (hipSetDevice(device) != hipSuccess) {
    delete c;
    return fail(ST_ERR_DEVICE, "st_ctx_create: hipSetDevice failed");
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->cus = prop.multiProcessorCount;
  bool streams = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) == hipSuccess;
  for (hipStream_t &s : c->slots) streams = streams && hipStreamCreateWithFlags(&s, hipStreamNonBlocking) == hipSuccess;
  if (!streams) {
    for (hipStream_t &s : c->slots)
      if (s) (void)hipStreamDestroy(s);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return fail(ST_ERR_DEVICE, "st_ctx_create: stream creation failed");
  }
  if (c->counter.alloc(4) != hipSuccess) {
    for (hipStream_t &s : c->slots) (void)hipStreamDestroy(s);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return fail(ST_ERR_NOMEM, "st_ctx_create: counter allocation failed");
  }
  if (int rc = noalloc_init(device)) {
    c->counter.release();
    for (hipStream_t &s : c->slots) (void)hipStreamDestroy(s);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return rc;
  }
  *out = c;
  return ST_OK;
}

static void span_work_free(SpanWork *w);

int st_ctx_destroy(st_ctx *ctx) {
  if (!ctx) return ST_OK;
  if (ctx->end_batch) st_batch_destroy(ctx->end_batch);
  ctx->end_batch = nullptr;
  span_work_free(ctx->spans);
  ctx->spans = nullptr;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  for (hipStream_t &s : ctx->slots) {
    (void)hipStreamSynchronize(s);
    (void)hipStreamDestroy(s);
  }
  ctx->counter.release();
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return ST_OK;
}

// ---- target database ------------------------------------------------------------------

int st_db_create(st_ctx *ctx, const uint8_t *residues, uint64_t n_residues, const uint64_t *offsets,
                 uint64_t n_targets, st_db **out) {
  if (!ctx || !out || !offsets || (!residues && n_residues))
    return fail(ST_ERR_INVALID, "st_db_create: null argument");
  *out = nullptr;
  if (offsets[0] != 0 && n_targets == 0) return fail(ST_ERR_INVALID, "st_db_create: bad offsets");
  std::vector<uint64_t> doff(n_targets + 1);
  std::vector<uint32_t> len(n_targets);
  uint64_t pos = 0;
  for (uint64_t t = 0; t < n_targets; t++) {
    if (offsets[t + 1] < offsets[t] || offsets[t + 1] > n_residues)
      return fail(ST_ERR_INVALID, "st_db_create: offsets must be non-decreasing and within the arena");
    const uint64_t l = offsets[t + 1] - offsets[t];
    if (l > 0x7fffffffull) return fail(ST_ERR_INVALID, "st_db_create: target longer than 2^31");
    len[t] = (uint32_t)l;
    doff[t] = pos;
    pos += (l + 15) & ~15ull;
  }
  doff[n_targets] = pos;
  // repack: 16-byte aligned starts, codes >= 21 -> 21, padding 21
  std::vector<uint8_t> host(pos + 64, (uint8_t)kSentinel);
  for (uint64_t t = 0; t < n_targets; t++) {
    const uint8_t *src = residues + offsets[t];
    uint8_t *dst = host.data() + doff[t];
    for (uint32_t k = 0; k < len[t]; k++) dst[k] = src[k] < kSentinel ? src[k] : (uint8_t)kSentinel;
  }
  std::lock_guard<std::mutex> lk(ctx->mu);
  ST_HIP(hipSetDevice(ctx->device));
  auto *db = new st_db();
  db->ctx = ctx;
  db->n_targets = n_targets;
  db->n_res = n_residues;
  db->hlen = len;
  for (uint32_t l : len) db->max_len = std::max(db->max_len, l);
  if (db->arena.upload(host.data(), host.size(), ctx->stream) != hipSuccess ||
      db->off.upload(doff.data(), doff.size(), ctx->stream) != hipSuccess ||
      db->len.upload(len.data(), len.size(), ctx->stream) != hipSuccess) {
    delete db;
    return fail(ST_ERR_NOMEM, "st_db_create: device allocation/upload failed");
  }
  if (n_targets == 0) (void)db->len.alloc(1);
  *out = db;
  return ST_OK;
}

int st_db_destroy(st_db *db) {
  delete db;
  return ST_OK;
}
uint64_t st_db_targets(const st_db *db) { return db ? db->n_targets : 0; }
uint64_t st_db_residues(const st_db *db) { return db ? db->n_res : 0; }

// ---- query profiles ---------------------------------------------------------------------

int st_qprof_create(st_ctx *ctx, const int32_t *rows, uint64_t m, st_qprof **out) {
  if (!ctx || !out || (!rows && m)) return fail(ST_ERR_INVALID, "st_qprof_create: null argument");
  if (m > 0x7fffffffull) return fail(ST_ERR_INVALID, "st_qprof_create: query too long");
  *out = nullptr;
  auto *qp = new st_qprof();
  qp->ctx = ctx;
  qp->m = (uint32_t)m;
  qp->rows.assign(rows, rows + m * 21);
  qp->ld = (uint32_t)std::max<uint64_t>(kMaxPassRows, (m + kMaxPassRows - 1) / kMaxPassRows * kMaxPassRows);
  bool fits = true;
  for (uint64_t k = 0; k < m * 21; k++)
    if (rows[k] < -32768 || rows[k] > 32767) fits = false;
  qp->fits16 = fits;
  bool fits8 = true;
  for (uint64_t k = 0; k < m * 21; k++)
    if (rows[k] < -128 || rows[k] > 127) fits8 = false;
  qp->fits8 = fits8;
  // code-major profiles [22][ld]: code 21 and rows >= m are zero
  std::vector<int32_t> p32((size_t)kCodes * qp->ld, 0);
  for (uint64_t i = 0; i < m; i++)
    for (int c = 0; c < 21; c++) p32[(size_t)c * qp->ld + i] = rows[i * 21 + c];
  std::lock_guard<std::mutex> lk(ctx->mu);
  if (hipSetDevice(ctx->device) != hipSuccess ||
      qp->prof32.upload(p32.data(), p32.size(), ctx->stream) != hipSuccess) {
    delete qp;
    return fail(ST_ERR_NOMEM, "st_qprof_create: device upload failed");
  }
  if (fits) {
    std::vector<int16_t> p16(p32.size());
    for (size_t k = 0; k < p32.size(); k++) p16[k] = (int16_t)p32[k];
    if (qp->prof16.upload(p16.data(), p16.size(), ctx->stream) != hipSuccess) {
      delete qp;
      return fail(ST_ERR_NOMEM, "st_qprof_create: device upload failed");
    }
  }
  if (fits8) {
    std::vector<int8_t> p8(p32.size());
    for (size_t k = 0; k < p32.size(); k++) p8[k] = (int8_t)p32[k];
    if (qp->prof8.upload(p8.data(), p8.size(), ctx->stream) != hipSuccess) {
      delete qp;
      return fail(ST_ERR_NOMEM, "st_qprof_create: device upload failed");
    }
  }
  qp->view.prof8 = qp->prof8.p;
  qp->view.prof16 = qp->prof16.p;
  qp->view.prof32 = qp->prof32.p;
  qp->view.m = qp->m;
  qp->view.ld = qp->ld;
  if (qp->qd.upload(&qp->view, 1, ctx->stream) != hipSuccess) {
    delete qp;
    return fail(ST_ERR_NOMEM, "st_qprof_create: device upload failed");
  }
  *out = qp;
  return ST_OK;
}

int st_qprof_destroy(st_qprof *qp) {
  delete qp;
  return ST_OK;
}
uint64_t st_qprof_len(const st_qprof *qp) { return qp ? qp->m : 0; }

}  // extern "C"


// ---- planning ------------------------------------------------------------------------------

namespace st {

Kind kind_for(const st_qprof *qp, int32_t open, int32_t ext) {
  if (ext < 0 || open < 0) return kI32N;
  if (qp->fits16 && open <= 32767 && ext <= 32767) return kPk16;
  return kI32F;
}

// ST_FORCE_R=8|12|16 and ST_FORCE_G=1..64 pin the rows per lane and the
// lanes per group (tuning runs and tests); read at every plan
static int forced_rows() {
  const char *e = std::getenv("ST_FORCE_R");
  const int v = e ? std::atoi(e) : 0;
  return (v == 8 || v == 12 || v == 16) ? v : 0;
}

static int forced_groups() {
  const char *e = std::getenv("ST_FORCE_G");
  const int v = e ? std::atoi(e) : 0;
  return (v >= 1 && v <= 64 && (v & (v - 1)) == 0) ? v : 0;
}

// Device time per executed cell of each launch shape (picoseconds, packed
// int16 kernel with the ungapped evidence and the gate, every height in one
// instance), measured on MI355X with uniform workloads where every unit has
// the shape (median over queries of 96, 300 and 1,200 residues); rows R = 8, 12, 16 by lanes per
// group G = 1..64 (R = 12: packed int16 only).  Narrow groups pay for their
// many distinct pass-boundary and residue streams (memory-pipeline bound),
// 64-lane groups for cross-row shuffles.
static const int kShapeRows[3] = {8, 12, 16};
static const double kShapePs[3][7] = {
    {1.98, 0.80, 0.286, 0.265, 0.283, 0.272, 0.40},   // R = 8
    {1.27, 0.48, 0.252, 0.233, 0.245, 0.289, 0.534},  // R = 12
    {0.88, 0.405, 0.236, 0.230, 0.237, 0.325, 0.665}, // R = 16
};

// The cost table in use: the compiled one, or the 21 values (rows R = 8,
// 12, 16; columns G = 1, 2, 4, ..., 64; '#' starts a comment) of the file
// that ST_SHAPE_TABLE names, as `python -m sabertooth_amd.autotune` writes
// it.  A file is read once per path; a malformed one is reported by
// shape_table_check and never used.
namespace {
struct ShapeTableState {
  std::mutex mu;
  std::string path;
  bool ok = false;
  std::string err;
  double ps[3][7];
};
ShapeTableState &shape_state() {
  static ShapeTableState s;
  return s;
}
// loads (or reuses) the file of `path`; called with the state's lock held
void load_shape_table(ShapeTableState &st, const std::string &path) {
  st.path = path;
  st.ok = false;
  FILE *f = std::fopen(path.c_str(), "r");
  if (!f) {
    st.err = "ST_SHAPE_TABLE: cannot read " + path;
    return;
  }
  std::string text;
  char buf[4096];
  size_t n;
  while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) text.append(buf, n);
  std::fclose(f);
  std::vector<double> v;
  size_t i = 0;
  while (i < text.size()) {
    if (text[i] == '#') {
      while (i < text.size() && text[i] != '\n') i++;
      continue;
    }
    if (std::isspace((unsigned char)text[i]) || text[i] == ',') {
      i++;
      continue;
    }
    char *end = nullptr;
    const double x = std::strtod(text.c_str() + i, &end);
    if (end == text.c_str() + i || !std::isfinite(x) || x <= 0.0) {
      st.err = "ST_SHAPE_TABLE: " + path + ": not a positive number at byte " + std::to_string(i);
      return;
    }
    v.push_back(x);
    i = (size_t)(end - text.c_str());
  }
  if (v.size() != 21) {
    st.err = "ST_SHAPE_TABLE: " + path + ": expected 21 values (3 heights x 7 group widths), found " +
             std::to_string(v.size());
    return;
  }
  for (int r = 0; r < 3; r++)
    for (int g = 0; g < 7; g++) st.ps[r][g] = v[(size_t)r * 7 + g];
  st.ok = true;
}
}  // namespace

// ST_OK, or ST_ERR_INVALID when ST_SHAPE_TABLE names a file that is not a
// valid table.
int shape_table_check() {
  const char *e = std::getenv("ST_SHAPE_TABLE");
  if (!e || !*e) return ST_OK;
  ShapeTableState &st = shape_state();
  std::lock_guard<std::mutex> lk(st.mu);
  if (st.path != e) load_shape_table(st, e);
  return st.ok ? ST_OK : fail(ST_ERR_INVALID, st.err);
}

// the table the planner uses now
static void shape_table(double (&ps)[3][7]);

PlanEnv plan_env() {
  PlanEnv e;
  e.fr = forced_rows();
  e.fg = forced_groups();
  shape_table(e.ps);
  return e;
}

static void shape_table(double (&ps)[3][7]) {
  std::memcpy(ps, kShapePs, sizeof ps);
  const char *e = std::getenv("ST_SHAPE_TABLE");
  if (!e || !*e) return;
  ShapeTableState &st = shape_state();
  std::lock_guard<std::mutex> lk(st.mu);
  if (st.path != e) load_shape_table(st, e);
  if (st.ok) std::memcpy(ps, st.ps, sizeof ps);
}

double unit_steps_cap(uint64_t batch_cells) {
  // half the mean work of a resident wave (256 CUs x 8 waves, 2048 cells per
  // packed step), at least 300 steps: a lone query spreads over many short
  // units (latency); in a large batch a unit is a small part of a wave's
  // work, so no step count is a tail and the cheapest shape wins (the
  // long-pair workload's units run 300,000 steps and more)
  const double mean = (double)batch_cells / 2048.0 / 2048.0;
  return std::max(300.0, mean / 2.0);
}

// Cut the sorted candidates of one query into wave units.  Each unit takes
// the launch shape (G, R) with the least modelled device time per useful
// cell: executed cells (rows padded to passes x G x R, columns to the unit's
// longest target plus the skew's G-1 steps, empty group slots) times the
// shape's measured cost, with a penalty on units longer than `cap` steps
// (normalised to R = 16), whose single wave would set a launch's tail.
bool plan_query(Plan &pl, std::vector<Unit> &pending, uint32_t qi, const st_qprof *qp, const uint32_t *tids,
                const uint64_t *outs, size_t count, Kind kind, const st_db *db, double cap, const PlanLimits *lim,
                const PlanEnv *env) {
  if (count == 0) return true;
  // stable length order: (length << 32 | index) keys; two 8-bit LSD radix
  // passes when every length is below 65536
  std::vector<uint64_t> key(count);
  uint32_t lmax = 0;
  for (size_t k = 0; k < count; k++) {
    const uint32_t l = db->hlen[tids[k]];
    lmax = std::max(lmax, l);
    key[k] = ((uint64_t)l << 32) | k;
  }
  if (count <= 32) {
    // short lists (a lone hit, a serving request's few targets): insertion
    // sort, the same stable order
    for (size_t k = 1; k < count; k++)
      for (size_t x = k; x > 0 && key[x - 1] > key[x]; x--) std::swap(key[x - 1], key[x]);
  } else if (lmax < 65536) {
    std::vector<uint64_t> tmp(count);
    for (int pass = 0; pass < 2; pass++) {
      const int sh = 32 + 8 * pass;
      size_t cnt[257] = {0};
      for (size_t k = 0; k < count; k++) cnt[((key[k] >> sh) & 0xff) + 1]++;
      for (int b = 0; b < 256; b++) cnt[b + 1] += cnt[b];
      for (size_t k = 0; k < count; k++) tmp[cnt[(key[k] >> sh) & 0xff]++] = key[k];
      key.swap(tmp);
    }
  } else {
    std::sort(key.begin(), key.end());
  }
  std::vector<uint64_t> csum(count + 1, 0);
  for (size_t k = 0; k < count; k++) csum[k + 1] = csum[k] + (key[k] >> 32);
  const int per = kind == kPk16 ? 2 : 1;
  const uint64_t m = qp->m;
  const PlanEnv own = env ? PlanEnv{} : plan_env();
  const PlanEnv &E = env ? *env : own;
  int fr = E.fr, fg = E.fg;
  const double (&shape_ps)[3][7] = E.ps;
  // the final pass of a packed int16 unit may be lower than the others (4,
  // 8 or 12 rows per lane): the kernels with every height take it; the
  // serving kernel (one height) and forced shapes keep it uniform
  const bool mixed_last = kind == kPk16 && !(lim && lim->only_R) && !fr;
  for (size_t s = 0; s < count;) {
    int bG = 16, bR = kind_unfloored(kind) ? 8 : 16, bRL = bR;
    double bc = 1e300;
    for (int ri = 0; ri < 3; ri++) {
      const int R = kShapeRows[ri];
      if (kind_unfloored(kind) && R != 8) continue;
      if (R == 12 && kind != kPk16) continue;  // the int32 kernels have R = 8, 16
      // a forced height the instance lacks (12 for the int32 kernels) is ignored
      if (fr && R != fr && !kind_unfloored(kind) && !(fr == 12 && kind != kPk16)) continue;
      for (int gi = 0; gi < 7; gi++) {
        const int G = 1 << gi;
        if (fg && G != fg) continue;
        const uint64_t GR = (uint64_t)G * R;
        // the pass's profile slice must fit 64 KB of LDS
        if ((uint64_t)kCodes * kind_lds_stride((int)GR, kind) * kind_entry_bytes(kind) > 65536) continue;
        const size_t e = std::min(count, s + (size_t)(64 / G * per));
        const uint64_t np = key[e - 1] >> 32;
        const uint64_t passes = std::max<uint64_t>(1, (m + GR - 1) / GR);
        // the passes' profile rows must lie inside the profile's padded rows
        if (passes * GR > qp->ld) continue;
        if (lim && ((lim->only_R && R != lim->only_R) || GR > (uint64_t)lim->max_GR ||
                    (passes > 1 && (uint64_t)(64 / G) * np > lim->max_half)))
          continue;
        // the lowest final-pass height that covers the query's last rows
        int RL = R;
        if (mixed_last) {
          const uint64_t rem = m - (passes - 1) * GR;  // 1 .. GR rows
          for (int h : {8, 12})
            if (h < R && (uint64_t)G * h >= rem) {
              RL = h;
              break;
            }
        }
        // device time: full passes at this shape's cost per executed cell,
        // the final pass at its own height's (a 4-row pass at the 8-row
        // cost scaled by its larger share of per-step overhead, 1.24)
        const double cols = 64.0 * per * (np + G - 1);
        const double ps_last = RL == R ? shape_ps[ri][gi]
                               : RL == 4 ? 1.24 * shape_ps[0][gi] : shape_ps[RL == 8 ? 0 : 1][gi];
        const double useful = std::max(1.0, (double)m * (double)(csum[e] - csum[s]));
        double c = cols * ((double)(passes - 1) * R * shape_ps[ri][gi] + (double)RL * ps_last) / useful;
        const double steps = (double)((passes - 1) * R + RL) * (np + G - 1) / 16.0;
        if (steps > cap) c *= steps / cap;
        if (c < bc) {
          bc = c;
          bG = G;
          bR = R;
          bRL = RL;
        }
      }
    }
    if (bc >= 1e300) {
      // a forced shape this kind cannot take (the int32 kernels' profile
      // slice is twice as large): plan this query's units freely
      if (fr || fg) {
        fr = fg = 0;
        continue;
      }
      return false;  // the limits admit no shape
    }
    const size_t e = std::min(count, s + (size_t)(64 / bG * per));
    const uint32_t GR = (uint32_t)(bG * bR);
    Unit u{};
    u.query = qi;
    u.pair0 = (uint32_t)pl.pairs.size();
    u.npairs = (uint32_t)(e - s);
    u.G = (uint32_t)bG;
    u.R = (uint32_t)bR;
    u.r_last = (uint32_t)bRL;
    u.passes = std::max<uint32_t>(1, (uint32_t)((m + GR - 1) / GR));
    for (size_t k = s; k < e; k++) {
      const uint32_t o = (uint32_t)(key[k] & 0xffffffffu);
      PairDev pd;
      pd.tid = tids[o];
      pd.out = (uint32_t)outs[o];
      pl.pairs.push_back(pd);
    }
    u.n_pad = (uint32_t)(key[e - 1] >> 32);
    pl.exec_cells += (uint64_t)64 * per * ((u.passes - 1) * bR + bRL) * (u.n_pad + bG - 1);
    pending.push_back(u);
    s = e;
  }
  return true;
}

// Window units (end-cell batches): hit k scores target tids[k] against the
// profile view qps[views[k]] and writes result row outs[k]; the views are
// windows of one source query, so they share gap costs and kind, and
// rep_query (any of them) stands for them in the launch order.  Hits are
// ordered by their rows' 64-row pass count, then target length, and cut
// into units of 64 / G hits at 8 rows per lane, G per unit by the least
// modelled device time per useful cell (as plan_query, with the step-cap
// penalty).  Unit.query carries bit 31 to mark the kind of unit.
bool plan_windows(Plan &pl, std::vector<Unit> &pending, uint32_t rep_query, const uint32_t *views,
                  const uint32_t *tids, const uint64_t *outs, size_t count, const st_qprof *const *qps,
                  const st_db *db, double cap, const PlanEnv *env) {
  if (count == 0) return true;
  if (rep_query >> 31) return false;
  constexpr int R = 8;
  std::vector<std::pair<uint64_t, uint32_t>> ord(count);
  for (size_t k = 0; k < count; k++) {
    const uint64_t m = qps[views[k]]->m, n = db->hlen[tids[k]];
    ord[k] = {(((m + 63) / 64) << 32) | n, (uint32_t)k};
  }
  std::sort(ord.begin(), ord.end());
  // each element's (m, n) in unit order: the shape search reads them often
  std::vector<std::pair<uint32_t, uint32_t>> mn(count);
  for (size_t x = 0; x < count; x++)
    mn[x] = {qps[views[ord[x].second]]->m, db->hlen[tids[ord[x].second]]};
  const PlanEnv own = env ? PlanEnv{} : plan_env();
  const PlanEnv &E = env ? *env : own;
  for (size_t s = 0; s < count;) {
    int bG = 64;
    double bc = 1e300;
    for (int gi = 0; gi < 7; gi++) {
      const int G = 1 << gi;
      if (E.fg && G != E.fg) continue;
      const size_t e = std::min(count, s + (size_t)(64 / G));
      uint64_t mmax = 0, np = 0;
      double useful = 0;
      for (size_t x = s; x < e; x++) {
        const uint64_t m = mn[x].first, n = mn[x].second;
        mmax = std::max(mmax, m);
        np = std::max(np, n);
        useful += (double)m * (double)n;
      }
      const uint64_t passes = std::max<uint64_t>(1, (mmax + (uint64_t)G * R - 1) / ((uint64_t)G * R));
      const double exec = 64.0 * R * (double)passes * (double)(np + G - 1);
      double c = exec * E.ps[0][gi] / std::max(1.0, useful);
      const double steps = (double)(passes * R) * (double)(np + G - 1) / 16.0;
      if (steps > cap) c *= steps / cap;
      if (c < bc) {
        bc = c;
        bG = G;
      }
    }
    const size_t e = std::min(count, s + (size_t)(64 / bG));
    Unit u{};
    u.query = rep_query | 0x80000000u;
    u.pair0 = (uint32_t)pl.pairs.size();
    u.npairs = (uint32_t)(e - s);
    u.G = (uint32_t)bG;
    u.R = R;
    u.r_last = R;
    uint64_t mmax = 0, np = 0;
    for (size_t x = s; x < e; x++) {
      const uint32_t k = ord[x].second;
      mmax = std::max<uint64_t>(mmax, mn[x].first);
      np = std::max<uint64_t>(np, mn[x].second);
      PairDev pd;
      pd.tid = tids[k];
      pd.out = (uint32_t)outs[k];
      pl.pairs.push_back(pd);
      pl.pair_q.push_back(views[k]);
    }
    u.n_pad = (uint32_t)np;
    u.passes = (uint32_t)std::max<uint64_t>(1, (mmax + (uint64_t)bG * R - 1) / ((uint64_t)bG * R));
    pl.exec_cells += (uint64_t)64 * R * u.passes * (u.n_pad + bG - 1);
    pending.push_back(u);
    s = e;
  }
  return true;
}

// Host worker threads shared by every planning call of the process, created
// once (up to 15, one fewer than the cores, max 16): run(n, f) calls f(t)
// for t in [0, n) on the workers and the calling thread and returns when all
// n have finished.  Concurrent callers queue their tasks in one FIFO; a
// caller runs queued tasks itself while it waits, so it never waits on
// workers that are busy elsewhere without making progress.
class WorkPool {
 public:
  static WorkPool &get() {
    static WorkPool *p = new WorkPool();  // never destroyed: workers outlive static teardown
    return *p;
  }
  unsigned width() const { return (unsigned)workers_.size() + 1; }
  template <class F>
  void run(unsigned n, F f) {
    struct Group {
      std::mutex mu;
      std::condition_variable cv;
      unsigned left;
    };
    auto g = std::make_shared<Group>();
    g->left = n;
    auto finish = [g] {
      std::lock_guard<std::mutex> lk(g->mu);
      if (--g->left == 0) g->cv.notify_all();
    };
    {
      std::lock_guard<std::mutex> lk(mu_);
      for (unsigned t = 1; t < n; t++)
        tasks_.push_back([f, t, finish]() mutable {
          f(t);
          finish();
        });
    }
    cv_.notify_all();
    f(0);
    finish();
    for (;;) {
      std::function<void()> task;
      {
        std::lock_guard<std::mutex> lk(mu_);
        if (!tasks_.empty()) {
          task = std::move(tasks_.front());
          tasks_.pop_front();
        }
      }
      if (task) {
        task();
        continue;
      }
      std::unique_lock<std::mutex> lk(g->mu);
      if (g->left == 0) return;
      g->cv.wait_for(lk, std::chrono::microseconds(200));
      if (g->left == 0) return;
    }
  }

 private:
  WorkPool() {
    const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    for (unsigned k = 0; k + 1 < hw; k++)
      workers_.emplace_back([this] {
        for (;;) {
          std::function<void()> task;
          {
            std::unique_lock<std::mutex> lk(mu_);
            cv_.wait(lk, [this] { return !tasks_.empty(); });
            task = std::move(tasks_.front());
            tasks_.pop_front();
          }
          task();
        }
      });
    for (auto &w : workers_) w.detach();
  }
  std::mutex mu_;
  std::condition_variable cv_;
  std::deque<std::function<void()>> tasks_;
  std::vector<std::thread> workers_;
};

// plan_query over the queries [0, nq): one(pl, pending, q) plans query q.
// Batches of many pairs are planned on the shared worker threads (one run
// per 2,048 pairs, at most one per core), each over a contiguous run of
// queries into a plan of its own; the runs are appended in query order with
// their pair indices shifted, which gives the plan one thread would make.
// Returns the first query whose planning failed, or nq.
template <class F>
static uint64_t plan_queries(Plan &pl, std::vector<Unit> &pending, uint64_t nq, uint64_t npairs, F one) {
  const uint64_t T = std::min<uint64_t>({(uint64_t)WorkPool::get().width(), nq, npairs / 2048 + 1});
  if (T <= 1) {
    for (uint64_t q = 0; q < nq; q++)
      if (!one(pl, pending, q)) return q;
    return nq;
  }
  std::vector<Plan> pls(T);
  std::vector<std::vector<Unit>> pend(T);
  std::vector<uint64_t> bad(T, nq);
  WorkPool::get().run((unsigned)T, [&](unsigned t) {
    for (uint64_t q = nq * t / T; q < nq * (t + 1) / T; q++)
      if (!one(pls[t], pend[t], q)) {
        bad[t] = q;
        return;
      }
  });
  for (uint64_t t = 0; t < T; t++)
    if (bad[t] < nq) return bad[t];
  for (uint64_t t = 0; t < T; t++) {
    const uint32_t off = (uint32_t)pl.pairs.size();
    pl.pairs.insert(pl.pairs.end(), pls[t].pairs.begin(), pls[t].pairs.end());
    pl.pair_q.insert(pl.pair_q.end(), pls[t].pair_q.begin(), pls[t].pair_q.end());
    pl.exec_cells += pls[t].exec_cells;
    for (Unit u : pend[t]) {
      u.pair0 += off;
      pending.push_back(u);
    }
  }
  return nq;
}

// dp_occupancy memoised per (kind, R, ung, LDS bytes): the runtime query is
// slow next to a serving micro-batch's planning.
static int occupancy_cached(int kind, int R, bool ung, size_t lds, bool end) {
  static std::mutex mu;
  static std::vector<std::pair<uint64_t, int>> memo;
  const uint64_t key = ((uint64_t)end << 48) | ((uint64_t)kind << 40) | ((uint64_t)R << 32) | ((uint64_t)ung << 31) | lds;
  std::lock_guard<std::mutex> lk(mu);
  for (const auto &e : memo)
    if (e.first == key) return e.second;
  const int n = dp_occupancy(kind, R, ung, lds, end);
  memo.emplace_back(key, n);
  return n;
}

static double unit_cost(const Unit &u) { return (double)((u.passes - 1) * u.R + u.r_last) * (u.n_pad + u.G - 1); }

// Group pending units into launches by (kind, R, gaps, LDS class), longest
// first, and lay out their pass-buffer regions.
// Stable LSD radix sort of (key, position) pairs by key, 8 bits a pass;
// passes whose byte is the same in every key are skipped.
static void radix_sort_keys(std::vector<std::pair<uint64_t, uint32_t>> &v) {
  std::vector<std::pair<uint64_t, uint32_t>> tmp(v.size());
  for (int sh = 0; sh < 64; sh += 8) {
    size_t cnt[257] = {0};
    for (const auto &x : v) cnt[((x.first >> sh) & 0xff) + 1]++;
    bool one = false;
    for (int d = 1; d <= 256; d++) one |= cnt[d] == v.size();
    if (one) continue;
    for (int d = 0; d < 256; d++) cnt[d + 1] += cnt[d];
    for (const auto &x : v) tmp[cnt[(x.first >> sh) & 0xff]++] = x;
    v.swap(tmp);
  }
}

void finish_plan(Plan &pl, std::vector<Unit> &pending, const std::vector<Kind> &qkind,
                        const std::vector<int32_t> &open, const std::vector<int32_t> &ext, int force_kind,
                        int cus, bool ung, bool end) {
  // a window unit (bit 31 of query) stages 64 lanes' rows, one part per group
  auto qi = [](const Unit &u) { return u.query & 0x7fffffffu; };
  auto gr_of = [](const Unit &u) { return (u.query >> 31) ? 64 * u.R : u.G * u.R; };
  auto kind_of = [&](const Unit &u) { return force_kind >= 0 ? force_kind : (int)qkind[qi(u)]; };
  auto big = [&](const Unit &u) { return gr_of(u) > 256 ? 1 : 0; };
  // packed int16 units of both heights share a launch (the kernel instance
  // with R = 0 takes each unit's own R); the int32 kernels are per R
  auto r_of = [&](const Unit &u) { return kind_of(u) == kPk16 ? 0 : (int)u.R; };
  // launch order: kind, R, wide-group launches first (their few long units
  // must start before the main launch fills the device), gap costs, then
  // the longest units first, ties in planning order; as one precomputed key
  // per unit (a comparator reading the query tables costs milliseconds for
  // tens of thousands of units)
  // the distinct gap costs (few: usually one), ranked
  std::vector<std::pair<int32_t, int32_t>> gaps;
  for (const Unit &u : pending) {
    const std::pair<int32_t, int32_t> g(open[qi(u)], ext[qi(u)]);
    if (gaps.empty() || gaps.back() != g) gaps.push_back(g);
  }
  std::sort(gaps.begin(), gaps.end());
  gaps.erase(std::unique(gaps.begin(), gaps.end()), gaps.end());
  // class = (kind, R, width class, gap rank), ranked; key = class rank in
  // the top 12 bits, then the cost descending (saturated at 2^52 - 1 cells
  // x steps), ties by position
  std::vector<uint64_t> cls(pending.size());
  for (size_t i = 0; i < pending.size(); i++) {
    const Unit &u = pending[i];
    const uint64_t g = (uint64_t)(std::lower_bound(gaps.begin(), gaps.end(), std::make_pair(open[qi(u)], ext[qi(u)])) -
                                  gaps.begin());
    cls[i] = ((uint64_t)kind_of(u) << 56) | ((uint64_t)r_of(u) << 48) | ((uint64_t)(1 - big(u)) << 40) | g;
  }
  std::vector<uint64_t> ranks;
  for (uint64_t c : cls)
    if (ranks.empty() || ranks.back() != c) ranks.push_back(c);
  std::sort(ranks.begin(), ranks.end());
  ranks.erase(std::unique(ranks.begin(), ranks.end()), ranks.end());
  std::vector<std::pair<uint64_t, uint32_t>> keys(pending.size());
  constexpr uint64_t kCostMax = (1ull << 52) - 1;
  if (ranks.size() <= 4096) {
    for (size_t i = 0; i < pending.size(); i++) {
      const uint64_t r = (uint64_t)(std::lower_bound(ranks.begin(), ranks.end(), cls[i]) - ranks.begin());
      const uint64_t c = std::min<uint64_t>((uint64_t)unit_cost(pending[i]), kCostMax);
      keys[i] = {(r << 52) | (kCostMax - c), (uint32_t)i};
    }
    radix_sort_keys(keys);
  } else {
    // more classes than the key holds (thousands of distinct gap costs)
    for (size_t i = 0; i < pending.size(); i++) keys[i] = {0, (uint32_t)i};
    std::sort(keys.begin(), keys.end(), [&](const std::pair<uint64_t, uint32_t> &x, const std::pair<uint64_t, uint32_t> &y) {
      if (cls[x.second] != cls[y.second]) return cls[x.second] < cls[y.second];
      const double cx = unit_cost(pending[x.second]), cy = unit_cost(pending[y.second]);
      return cx != cy ? cx > cy : x.second < y.second;
    });
  }
  {
    std::vector<Unit> sorted(pending.size());
    for (size_t i = 0; i < keys.size(); i++) sorted[i] = pending[keys[i].second];
    pending.swap(sorted);
  }
  for (size_t i = 0; i < pending.size();) {
    const Unit &u = pending[i];
    Launch L{};
    L.kind = kind_of(u);
    L.R = r_of(u);
    L.open = open[qi(u)];
    L.ext = ext[qi(u)];
    L.unit0 = (uint32_t)pl.units.size();
    size_t j = i;
    L.gr_max = 0;
    L.half = 0;
    while (j < pending.size()) {
      const Unit &v = pending[j];
      if (kind_of(v) != L.kind || r_of(v) != L.R || big(v) != big(u) || open[qi(v)] != L.open ||
          ext[qi(v)] != L.ext)
        break;
      L.gr_max = std::max(L.gr_max, (int)gr_of(v));
      if (v.passes > 1) L.half = std::max<uint64_t>(L.half, (uint64_t)(64 / v.G) * v.n_pad);
      pl.units.push_back(v);
      j++;
    }
    L.nunits = (uint32_t)(j - i);
    const size_t lds = dp_lds_bytes(kind_lds_stride(L.gr_max, L.kind), kind_entry_bytes(L.kind), ung && !end);
    const int occ = std::max(1, occupancy_cached(L.kind, L.R, ung, lds, end));
    L.grid = std::max<uint32_t>(1, std::min<uint32_t>(L.nunits, (uint32_t)(occ * cus)));
    L.slot = (int)(pl.launches.size() % kLaunchSlots);
    L.buf0 = pl.buf_elems;
    pl.buf_elems += (uint64_t)L.grid * (2 * L.half + kEdgeQuadsPerBlock);
    pl.launches.push_back(L);
    i = j;
  }
  pending.clear();
}

}  // namespace st

// A planned multi-query batch with every input resident on the device.
struct st_batch {
  st_ctx *ctx = nullptr;
  const st_db *db = nullptr;
  bool ung = true;
  bool end = false;  // score + end-cell batch (int32, no evidence)
  DevBuf<int32_t> d_qend, d_tend;
  uint64_t npairs = 0, cells = 0;
  std::vector<const st_qprof *> qps;
  std::vector<int32_t> open, ext, gate;
  std::vector<Kind> qkind;
  std::vector<uint32_t> cand_tids;
  std::vector<uint32_t> cand_q;
  Plan plan;
  std::vector<QueryDev> qv;  // the queries' device views, source of d_queries
  DevBuf<QueryDev> d_queries;
  DevBuf<int32_t> d_open, d_ext, d_gate;
  DevBuf<Unit> d_units;
  DevBuf<PairDev> d_pairs;
  DevBuf<uint32_t> d_pair_q;    // window units: each pair's profile view
  DevBuf<PairOut> d_pout;       // score-only batches
  DevBuf<uint8_t> d_res;        // evidence batches: st_pair_result rows
  DevBuf<uint32_t> d_sat;       // int16 pairs at the clamp: count, then rows
  // end-cell read-back staging, kept between refills of the batch object
  std::vector<PairOut> h_pout;
  std::vector<int32_t> h_qend, h_tend;
  // window units' early stop: each pair's known score, per result row
  bool has_stop = false;
  std::vector<int32_t> stop;
  DevBuf<int32_t> d_stop;
  DevBuf<uint4> d_passbuf;
  // exact int32 rescoring of saturated int16 pairs (built at run time)
  Plan resc;
  DevBuf<Unit> d_rs_units;
  DevBuf<PairDev> d_rs_pairs;
  DevBuf<uint4> d_rs_passbuf;
  uint64_t rescored = 0;
  // device time of the last run: main DP launches, int32 rescoring, and a
  // third phase kept at 0 (the gate runs inside the DP launches)
  hipEvent_t ev[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  hipEvent_t join[kLaunchSlots] = {nullptr, nullptr, nullptr, nullptr};
  double ms[3] = {0, 0, 0};
  // per-launch start/end events of the main plan and their last durations
  std::vector<hipEvent_t> lev;
  std::vector<double> launch_ms;
  // the main DP phase captured once (clamp-list clear + every DP launch,
  // forked over the launch streams) and replayed by st_batch_run_graph
  hipGraphExec_t gexec = nullptr;
  hipEvent_t gfork = nullptr;
  ~st_batch() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (gfork) (void)hipEventDestroy(gfork);
    for (hipEvent_t &e : ev)
      if (e) (void)hipEventDestroy(e);
    for (hipEvent_t &e : join)
      if (e) (void)hipEventDestroy(e);
    for (hipEvent_t &e : lev)
      if (e) (void)hipEventDestroy(e);
  }
};

namespace st {

// Issue a plan's launches on the context's launch streams, after `start` on
// the main stream, and make the main stream wait for all of them.  `timing`
// records per-launch events around the main plan's launches (off while the
// launches are being captured into a graph).
static int launch_plan(st_batch *b, const Plan &pl, const Unit *d_units, const PairDev *d_pairs, uint4 *d_buf,
                       hipEvent_t start, bool timing = true) {
  st_ctx *ctx = b->ctx;
  bool used[kLaunchSlots] = {false, false, false, false};
  for (const Launch &L : pl.launches) {
    hipStream_t s = ctx->slots[L.slot];
    if (!used[L.slot]) {
      ST_HIP(hipStreamWaitEvent(s, start, 0));
      used[L.slot] = true;
    }
    DpLaunch a{};
    a.units = d_units + L.unit0;
    a.nunits = L.nunits;
    a.counter = ctx->counter.p + L.slot;
    a.pairs = d_pairs;
    a.queries = b->d_queries.p;
    // window units occur in the main plan of end-cell batches only
    a.pair_q = &pl == &b->plan && !b->plan.pair_q.empty() ? b->d_pair_q.p : nullptr;
    a.arena = b->db->arena.p;
    a.toff = b->db->off.p;
    a.tlen = b->db->len.p;
    a.open = L.open;
    a.ext = L.ext;
    a.grid = L.grid;
    a.passbuf = d_buf + L.buf0;
    a.passbuf_half = L.half;
    a.out.pout = b->d_pout.p;
    a.out.res = (PairResultDev *)b->d_res.p;
    a.out.qgate = b->d_gate.p;
    // only the main plan's packed int16 pairs can reach the clamp
    a.out.sat = &pl == &b->plan ? b->d_sat.p : nullptr;
    a.out.sat_cap = (uint32_t)b->npairs;
    a.out.qend = b->d_qend.p;
    a.out.tend = b->d_tend.p;
    a.out.stop = &pl == &b->plan && b->has_stop ? b->d_stop.p : nullptr;
    a.lds_stride = kind_lds_stride(L.gr_max, L.kind);
    const bool timed = timing && &pl == &b->plan && b->lev.size() == 2 * pl.launches.size();
    const size_t li = (size_t)(&L - pl.launches.data());
    if (timed) ST_HIP(hipEventRecord(b->lev[2 * li], s));
    ST_HIP(launch_dp(a, L.kind, L.R, b->ung, s, b->end));
    if (timed) ST_HIP(hipEventRecord(b->lev[2 * li + 1], s));
  }
  for (int k = 0; k < kLaunchSlots; k++)
    if (used[k]) {
      ST_HIP(hipEventRecord(b->join[k], ctx->slots[k]));
      ST_HIP(hipStreamWaitEvent(ctx->stream, b->join[k], 0));
    }
  return ST_OK;
}

}  // namespace st

extern "C" {

// Plan a batch and put its tables on the device.  reuse (nullable): an
// earlier batch object of the same context to refill; its device buffers
// are kept where they are large enough.  win_off (nullable, end-cell
// batches): the queries are profile views, those in [win_off[s],
// win_off[s + 1]) windows of one source s, planned as window units.
static int batch_create_impl(st_ctx *ctx, const st_db *db, const st_qprof *const *queries, uint64_t nq,
                             const uint64_t *cand_off, const uint32_t *cand_tids, const int32_t *gap_open,
                             const int32_t *gap_extend, const int32_t *min_diag_score, bool ung,
                             st_batch **out, bool end = false, st_batch *reuse = nullptr,
                             const uint64_t *win_off = nullptr, uint64_t nwin = 0,
                             const int32_t *stop_scores = nullptr) {
  if (!ctx || !db || !out || (nq && (!queries || !cand_off || !gap_open || !gap_extend)))
    return fail(ST_ERR_INVALID, "st_batch_create: null argument");
  *out = nullptr;
  if (db->ctx != ctx) return fail(ST_ERR_INVALID, "st_batch_create: database of another context");
  if (int rc = shape_table_check()) return rc;
  const uint64_t np = nq ? cand_off[nq] : 0;
  if (nq && cand_off[0] != 0) return fail(ST_ERR_INVALID, "st_batch_create: cand_off[0] must be 0");
  if (np && !cand_tids) return fail(ST_ERR_INVALID, "st_batch_create: null candidate list");
  if (np > 0xffffffffull) return fail(ST_ERR_INVALID, "st_batch_create: more than 2^32 pairs");
  for (uint64_t q = 0; q < nq; q++) {
    if (!queries[q] || queries[q]->ctx != ctx) return fail(ST_ERR_INVALID, "st_batch_create: bad query handle");
    if (cand_off[q + 1] < cand_off[q]) return fail(ST_ERR_INVALID, "st_batch_create: cand_off must not decrease");
  }
  for (uint64_t k = 0; k < np; k++)
    if (cand_tids[k] >= db->n_targets) return fail(ST_ERR_INVALID, "st_batch_create: target id out of range");
  st_batch *b = reuse ? reuse : new st_batch();
  auto drop = [&]() {
    if (!reuse) delete b;
  };
  if (reuse && reuse->gexec) {
    (void)hipGraphExecDestroy(reuse->gexec);
    reuse->gexec = nullptr;
  }
  b->ctx = ctx;
  b->db = db;
  b->ung = ung;
  b->end = end;
  b->npairs = np;
  b->plan.clear();
  b->resc.clear();
  b->rescored = 0;
  b->qps.assign(queries, queries + nq);
  b->open.assign(gap_open, gap_open + nq);
  b->ext.assign(gap_extend, gap_extend + nq);
  if (min_diag_score)
    b->gate.assign(min_diag_score, min_diag_score + nq);
  else
    b->gate.assign(nq, 0);
  b->cand_tids.assign(cand_tids, cand_tids + np);
  b->cand_q.resize(np);
  b->has_stop = stop_scores != nullptr;
  if (stop_scores) b->stop.assign(stop_scores, stop_scores + np);
  b->qkind.resize(nq);
  b->qv.resize(nq);
  // ST_END_PK16=0 keeps every end-cell batch on the int32 kernels
  const char *e16 = std::getenv("ST_END_PK16");
  const bool end_pk16 = !(e16 && e16[0] == '0');
  // one pass over the queries (each handle is read once): cells, the
  // candidates' query, the precision class and the device view
  uint64_t cells = 0;
  for (uint64_t q = 0; q < nq; q++) {
    const st_qprof *qp = b->qps[q];
    b->qv[q] = qp->view;
    uint32_t nmax = 0;
    uint64_t tl = 0;
    for (uint64_t k = cand_off[q]; k < cand_off[q + 1]; k++) {
      const uint32_t n = db->hlen[cand_tids[k]];
      b->cand_q[k] = (uint32_t)q;
      tl += n;
      nmax = std::max(nmax, n);
    }
    cells += (uint64_t)qp->m * tl;
    b->qkind[q] = kind_for(qp, b->open[q], b->ext[q]);
    // the packed kernel's diagonal keys hold m + n - 1 <= kMaxDiagsPk16
    if (b->qkind[q] == kPk16 && (uint64_t)qp->m + nmax - 1 > kMaxDiagsPk16) b->qkind[q] = kI32F;
    // the end cell: the packed int16 kernel (two pairs per group, an int32
    // rescoring at the clamp) where the query has its own int16 profile and
    // the gaps allow it; the window units of profile views, and the rest,
    // on the int32 kernels over the int8 or int16 profile where there is one
    if (end) {
      const bool p8 = qp->fits8 && qp->view.prof8;
      const bool p16 = qp->fits16 && qp->view.prof16;
      if (!(b->qkind[q] == kPk16 && p16 && !win_off && end_pk16))
        b->qkind[q] = kind_unfloored(b->qkind[q]) ? (p8 ? kI32N8 : p16 ? kI32N16 : kI32N)
                                                  : (p8 ? kI32F8 : p16 ? kI32F16 : kI32F);
    }
  }
  b->cells = cells;
  std::lock_guard<std::mutex> lk(ctx->mu);
  if (hipSetDevice(ctx->device) != hipSuccess) {
    drop();
    return fail(ST_ERR_DEVICE, "st_batch_create: hipSetDevice failed");
  }
  std::vector<Unit> pending;
  const double cap = unit_steps_cap(cells);
  const PlanEnv env = plan_env();
  uint64_t bad = 0;
  if (win_off) {
    // one plan_windows per source over the hits of all its views
    bad = plan_queries(b->plan, pending, nwin, np, [&](Plan &pl, std::vector<Unit> &pend, uint64_t w) {
      const uint64_t v0 = win_off[w], v1 = win_off[w + 1];
      if (v0 >= v1) return true;
      std::vector<uint32_t> vs, ts;
      std::vector<uint64_t> os;
      for (uint64_t v = v0; v < v1; v++)
        for (uint64_t k = cand_off[v]; k < cand_off[v + 1]; k++) {
          vs.push_back((uint32_t)v);
          ts.push_back(cand_tids[k]);
          os.push_back(k);
        }
      return plan_windows(pl, pend, (uint32_t)v0, vs.data(), ts.data(), os.data(), vs.size(), b->qps.data(), db, cap,
                          &env);
    });
    if (bad < nwin) {
      drop();
      return fail(ST_ERR_INVALID, "st_batch_create: no window units for source " + std::to_string(bad));
    }
  } else {
    bad = plan_queries(b->plan, pending, nq, np, [&](Plan &pl, std::vector<Unit> &pend, uint64_t q) {
      const uint64_t s = cand_off[q], e = cand_off[q + 1];
      std::vector<uint64_t> outs(e - s);
      std::iota(outs.begin(), outs.end(), s);
      return plan_query(pl, pend, (uint32_t)q, b->qps[q], cand_tids + s, outs.data(), e - s, b->qkind[q], db, cap,
                        nullptr, &env);
    });
    if (bad < nq) {
      drop();
      return fail(ST_ERR_INVALID, "st_batch_create: no launch shape fits query " + std::to_string(bad));
    }
  }
  finish_plan(b->plan, pending, b->qkind, b->open, b->ext, -1, ctx->cus, ung, end);
  // useful and executed cells of every launch (roofline accounting)
  for (Launch &L : b->plan.launches) {
    L.useful_cells = L.exec_cells = 0;
    for (uint32_t u = L.unit0; u < L.unit0 + L.nunits; u++) {
      const Unit &U = b->plan.units[u];
      for (uint32_t k = 0; k < U.npairs; k++) {
        const uint32_t v = (U.query >> 31) ? b->plan.pair_q[U.pair0 + k] : U.query;
        L.useful_cells += (uint64_t)b->qv[v].m * db->hlen[b->plan.pairs[U.pair0 + k].tid];
      }
      L.exec_cells += (uint64_t)64 * (L.kind == kPk16 ? 2 : 1) * ((U.passes - 1) * U.R + U.r_last) * (U.n_pad + U.G - 1);
    }
  }
  // the tables are copied in stream order from the batch's own vectors
  // (unchanged until the batch is planned again or destroyed, after its runs
  // have synchronised); outputs and the pass buffer are grow-only
  hipStream_t st = ctx->stream;
  bool ok = b->d_queries.put(b->qv.data(), nq, st) == hipSuccess &&
            b->d_open.put(b->open.data(), nq, st) == hipSuccess &&
            b->d_ext.put(b->ext.data(), nq, st) == hipSuccess &&
            b->d_gate.put(b->gate.data(), nq, st) == hipSuccess &&
            b->d_units.put(b->plan.units.data(), b->plan.units.size(), st) == hipSuccess &&
            b->d_pairs.put(b->plan.pairs.data(), b->plan.pairs.size(), st) == hipSuccess &&
            b->d_pair_q.put(b->plan.pair_q.data(), b->plan.pair_q.size(), st) == hipSuccess &&
            b->d_pout.reserve(std::max<uint64_t>(np, 1)) == hipSuccess &&
            (!ung || b->d_res.reserve(std::max<uint64_t>(np, 1) * sizeof(st_pair_result)) == hipSuccess) &&
            b->d_sat.reserve(np + 1) == hipSuccess &&
            b->d_passbuf.reserve(std::max<uint64_t>(b->plan.buf_elems, 1)) == hipSuccess &&
            (!end || (b->d_qend.reserve(std::max<uint64_t>(np, 1)) == hipSuccess &&
                      b->d_tend.reserve(std::max<uint64_t>(np, 1)) == hipSuccess)) &&
            (!stop_scores || b->d_stop.put(b->stop.data(), np, st) == hipSuccess);
  if (!ok) {
    drop();
    return fail(ST_ERR_NOMEM, "st_batch_create: device allocation failed");
  }
  *out = b;
  return ST_OK;
}

int st_batch_create(st_ctx *ctx, const st_db *db, const st_qprof *const *queries, uint64_t nq,
                    const uint64_t *cand_off, const uint32_t *cand_tids, const int32_t *gap_open,
                    const int32_t *gap_extend, const int32_t *min_diag_score, st_batch **out) {
  return batch_create_impl(ctx, db, queries, nq, cand_off, cand_tids, gap_open, gap_extend, min_diag_score,
                           true, out);
}

// Capture the main DP phase into b->gexec: the clamp-list clear (a kernel;
// graphs here take no memset nodes) and every launch of the plan, forked
// from the main stream over the launch streams and joined back.
static int capture_dp_graph(st_batch *b) {
  st_ctx *ctx = b->ctx;
  if (!b->gfork) ST_HIP(hipEventCreateWithFlags(&b->gfork, hipEventDisableTiming));
  ST_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  int rc = ST_OK;
  if (clear_u32(b->d_sat.p, 1, ctx->stream) != hipSuccess || hipEventRecord(b->gfork, ctx->stream) != hipSuccess)
    rc = fail(ST_ERR_DEVICE, "st_batch_run_graph: capture failed");
  if (!rc) rc = launch_plan(b, b->plan, b->d_units.p, b->d_pairs.p, b->d_passbuf.p, b->gfork, false);
  // end the capture whatever happened, so the stream leaves capture mode
  hipGraph_t g = nullptr;
  const hipError_t ec = hipStreamEndCapture(ctx->stream, &g);
  if (!rc && (ec != hipSuccess || !g)) rc = fail(ST_ERR_DEVICE, "st_batch_run_graph: hipStreamEndCapture failed");
  if (!rc && hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0) != hipSuccess) {
    b->gexec = nullptr;
    rc = fail(ST_ERR_DEVICE, "st_batch_run_graph: hipGraphInstantiate failed");
  }
  if (g) (void)hipGraphDestroy(g);
  return rc;
}

static int batch_run_locked(st_batch *b, bool graph = false) {
  st_ctx *ctx = b->ctx;
  ST_HIP(hipSetDevice(ctx->device));
  for (hipEvent_t &e : b->ev)
    if (!e) ST_HIP(hipEventCreate(&e));
  for (hipEvent_t &e : b->join)
    if (!e) ST_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  if (b->lev.size() != 2 * b->plan.launches.size()) {
    // a refilled batch object may have another launch count
    for (hipEvent_t &e : b->lev)
      if (e) (void)hipEventDestroy(e);
    b->lev.assign(2 * b->plan.launches.size(), nullptr);
    for (hipEvent_t &e : b->lev) ST_HIP(hipEventCreate(&e));
  }
  int rc = ST_OK;
  if (graph) {
    if (!b->gexec && (rc = capture_dp_graph(b))) return rc;
    ST_HIP(hipEventRecord(b->ev[0], ctx->stream));
    ST_HIP(hipGraphLaunch(b->gexec, ctx->stream));
  } else {
    ST_HIP(clear_u32(b->d_sat.p, 1, ctx->stream));
    ST_HIP(hipEventRecord(b->ev[0], ctx->stream));
    rc = launch_plan(b, b->plan, b->d_units.p, b->d_pairs.p, b->d_passbuf.p, b->ev[0]);
    if (rc) return rc;
  }
  ST_HIP(hipEventRecord(b->ev[1], ctx->stream));
  // int16 pairs whose best reached 32767 are rescored in exact int32
  b->resc.clear();
  b->rescored = 0;
  ST_HIP(hipEventRecord(b->ev[2], ctx->stream));
  ST_HIP(hipEventRecord(b->ev[3], ctx->stream));
  bool any16 = false;
  for (const Launch &L : b->plan.launches) any16 |= L.kind == kPk16;
  if (any16) {
    // the DP launches listed the result rows of int16 pairs at the clamp
    uint32_t nsat = 0;
    ST_HIP(hipMemcpyAsync(&nsat, b->d_sat.p, sizeof nsat, hipMemcpyDeviceToHost, ctx->stream));
    ST_HIP(hipStreamSynchronize(ctx->stream));
    nsat = (uint32_t)std::min<uint64_t>(nsat, b->npairs);
    std::vector<uint32_t> rows(nsat);
    if (nsat) {
      ST_HIP(hipMemcpyAsync(rows.data(), b->d_sat.p + 1, nsat * sizeof(uint32_t), hipMemcpyDeviceToHost,
                            ctx->stream));
      ST_HIP(hipStreamSynchronize(ctx->stream));
      std::sort(rows.begin(), rows.end());
    }
    std::vector<std::vector<uint64_t>> perq(b->qps.size());
    for (uint32_t k : rows) perq[b->cand_q[k]].push_back(k);
    std::vector<Unit> pending;
    const PlanEnv env = plan_env();
    for (size_t q = 0; q < perq.size(); q++) {
      if (perq[q].empty()) continue;
      std::vector<uint32_t> tids;
      for (uint64_t o : perq[q]) tids.push_back(b->cand_tids[o]);
      b->rescored += perq[q].size();
      if (!plan_query(b->resc, pending, (uint32_t)q, b->qps[q], tids.data(), perq[q].data(), tids.size(), kI32F,
                      b->db, unit_steps_cap(0), nullptr, &env))
        return fail(ST_ERR_INVALID, "st_batch_run: no int32 launch shape fits query " + std::to_string(q));
    }
    finish_plan(b->resc, pending, b->qkind, b->open, b->ext, (int)kI32F, ctx->cus, b->ung, b->end);
    if (!b->resc.units.empty()) {
      if (b->d_rs_units.upload(b->resc.units.data(), b->resc.units.size(), ctx->stream) != hipSuccess ||
          b->d_rs_pairs.upload(b->resc.pairs.data(), b->resc.pairs.size(), ctx->stream) != hipSuccess ||
          b->d_rs_passbuf.alloc(std::max<uint64_t>(b->resc.buf_elems, 1)) != hipSuccess)
        return fail(ST_ERR_NOMEM, "rescoring upload failed");
      ST_HIP(hipEventRecord(b->ev[2], ctx->stream));
      rc = launch_plan(b, b->resc, b->d_rs_units.p, b->d_rs_pairs.p, b->d_rs_passbuf.p, b->ev[2]);
      if (rc) return rc;
      ST_HIP(hipEventRecord(b->ev[3], ctx->stream));
    }
  }
  ST_HIP(hipEventRecord(b->ev[4], ctx->stream));
  ST_HIP(hipEventRecord(b->ev[5], ctx->stream));
  ST_HIP(hipStreamSynchronize(ctx->stream));
  float t = 0;
  ST_HIP(hipEventElapsedTime(&t, b->ev[0], b->ev[1]));
  b->ms[0] = t;
  ST_HIP(hipEventElapsedTime(&t, b->ev[2], b->ev[3]));
  b->ms[1] = t;
  ST_HIP(hipEventElapsedTime(&t, b->ev[4], b->ev[5]));
  b->ms[2] = t;
  b->launch_ms.assign(b->plan.launches.size(), 0.0);
  // a replayed graph records no per-launch events
  for (size_t i = 0; !graph && i < b->plan.launches.size(); i++) {
    ST_HIP(hipEventElapsedTime(&t, b->lev[2 * i], b->lev[2 * i + 1]));
    b->launch_ms[i] = t;
  }
  return ST_OK;
}

int st_shape_table_status(void) { return shape_table_check(); }

int st_batch_run(st_batch *b) {
  if (!b) return fail(ST_ERR_INVALID, "st_batch_run: null batch");
  std::lock_guard<std::mutex> lk(b->ctx->mu);
  return batch_run_locked(b);
}

int st_batch_run_graph(st_batch *b) {
  if (!b) return fail(ST_ERR_INVALID, "st_batch_run_graph: null batch");
  std::lock_guard<std::mutex> lk(b->ctx->mu);
  return batch_run_locked(b, true);
}

int st_batch_results(st_batch *b, st_pair_result *out) {
  if (!b || (!out && b->npairs)) return fail(ST_ERR_INVALID, "st_batch_results: null argument");
  if (!b->npairs) return ST_OK;
  std::lock_guard<std::mutex> lk(b->ctx->mu);
  ST_HIP(hipSetDevice(b->ctx->device));
  if (b->ung) {
    ST_HIP(hipMemcpyAsync(out, b->d_res.p, b->npairs * sizeof(st_pair_result), hipMemcpyDeviceToHost,
                          b->ctx->stream));
    ST_HIP(hipStreamSynchronize(b->ctx->stream));
  } else {
    std::vector<PairOut> po(b->npairs);
    ST_HIP(hipMemcpyAsync(po.data(), b->d_pout.p, b->npairs * sizeof(PairOut), hipMemcpyDeviceToHost,
                          b->ctx->stream));
    ST_HIP(hipStreamSynchronize(b->ctx->stream));
    for (uint64_t k = 0; k < b->npairs; k++) {
      std::memset(&out[k], 0, sizeof(st_pair_result));
      out[k].swg = po[k].swg;
      out[k].kernel = po[k].flag == 2 ? ST_KERNEL_I32 : ST_KERNEL_I16;
    }
  }
  return ST_OK;
}

int st_batch_timing(const st_batch *b, double *ms3, uint64_t *rescored) {
  if (!b || !ms3) return fail(ST_ERR_INVALID, "st_batch_timing: null argument");
  for (int k = 0; k < 3; k++) ms3[k] = b->ms[k];
  if (rescored) *rescored = b->rescored;
  return ST_OK;
}

int st_batch_stats(const st_batch *b, uint64_t *exec_cells, uint64_t *units, uint64_t *launches) {
  if (!b) return fail(ST_ERR_INVALID, "st_batch_stats: null batch");
  if (exec_cells) *exec_cells = b->plan.exec_cells;
  if (units) *units = b->plan.units.size();
  if (launches) *launches = b->plan.launches.size();
  return ST_OK;
}

int st_batch_launches(const st_batch *b, st_launch_info *out, uint64_t cap, uint64_t *n) {
  if (!b || !n || (cap && !out)) return fail(ST_ERR_INVALID, "st_batch_launches: null argument");
  *n = b->plan.launches.size();
  for (uint64_t i = 0; i < std::min<uint64_t>(cap, *n); i++) {
    const Launch &L = b->plan.launches[i];
    st_launch_info &o = out[i];
    std::memset(&o, 0, sizeof o);
    o.kind = L.kind;
    o.rows_per_lane = L.R;
    o.units = L.nunits;
    o.grid = L.grid;
    o.useful_cells = L.useful_cells;
    o.exec_cells = L.exec_cells;
    o.ms = i < b->launch_ms.size() ? b->launch_ms[i] : 0.0;
  }
  return ST_OK;
}

uint64_t st_batch_cells(const st_batch *b) { return b ? b->cells : 0; }
uint64_t st_batch_pairs(const st_batch *b) { return b ? b->npairs : 0; }

int st_batch_destroy(st_batch *b) {
  if (b) {
    std::lock_guard<std::mutex> lk(b->ctx->mu);
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
  }
  delete b;
  return ST_OK;
}

// ---- single-query drop-ins ------------------------------------------------------------------

int st_sw_score_batch(st_ctx *ctx, const st_qprof *qp, const st_db *db, const uint32_t *tids, uint64_t count,
                      int32_t gap_open, int32_t gap_extend, int32_t *scores, uint8_t *kernel) {
  if (!ctx || !qp || !db || (count && (!tids || !scores)))
    return fail(ST_ERR_INVALID, "st_sw_score_batch: null argument");
  if (count == 0) return ST_OK;
  const uint64_t off[2] = {0, count};
  st_batch *b = nullptr;
  int rc = batch_create_impl(ctx, db, &qp, 1, off, tids, &gap_open, &gap_extend, nullptr, false, &b);
  if (rc) return rc;
  rc = st_batch_run(b);
  if (!rc) {
    std::vector<st_pair_result> res(count);
    rc = st_batch_results(b, res.data());
    if (!rc)
      for (uint64_t k = 0; k < count; k++) {
        scores[k] = res[k].swg;
        if (kernel) kernel[k] = res[k].kernel;
      }
  }
  st_batch_destroy(b);
  return rc;
}

static int score_end_multi(st_ctx *ctx, const st_db *db, const st_qprof *const *qps, uint64_t nq,
                           const uint64_t *off, const uint32_t *tids, const int32_t *open, const int32_t *ext,
                           st_score_end *out, const char *who, const uint64_t *win_off = nullptr,
                           uint64_t nwin = 0, std::vector<st_launch_info> *launches = nullptr,
                           double *phase_ms = nullptr, const int32_t *stop_scores = nullptr);

// host wall clock in milliseconds
static double wall_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int st_sw_score_end_batch(st_ctx *ctx, const st_qprof *qp, const st_db *db, const uint32_t *tids, uint64_t count,
                          int32_t gap_open, int32_t gap_extend, st_score_end *out) {
  if (!ctx || !qp || !db || (count && (!tids || !out)))
    return fail(ST_ERR_INVALID, "st_sw_score_end_batch: null argument");
  if (count == 0) return ST_OK;
  const uint64_t off[2] = {0, count};
  return score_end_multi(ctx, db, &qp, 1, off, tids, &gap_open, &gap_extend, out, "st_sw_score_end_batch");
}

// Scores and end cells of a multi-query candidate list (the int32
// end-tracking kernels), in candidate order.  win_off (nullable): the
// queries are profile views grouped by source, planned as window units.
static int score_end_multi(st_ctx *ctx, const st_db *db, const st_qprof *const *qps, uint64_t nq,
                           const uint64_t *off, const uint32_t *tids, const int32_t *open, const int32_t *ext,
                           st_score_end *out, const char *who, const uint64_t *win_off, uint64_t nwin,
                           std::vector<st_launch_info> *launches, double *phase_ms, const int32_t *stop_scores) {
  const uint64_t count = nq ? off[nq] : 0;
  if (count == 0) return ST_OK;
  // the context's end-cell batch object, refilled: its device buffers and
  // events outlive the call
  std::lock_guard<std::mutex> el(ctx->end_mu);
  const double t0 = wall_ms();
  st_batch *b = nullptr;
  int rc = batch_create_impl(ctx, db, qps, nq, off, tids, open, ext, nullptr, false, &b, true, ctx->end_batch,
                             win_off, nwin, stop_scores);
  if (rc) return rc;
  ctx->end_batch = b;
  const double t1 = wall_ms();
  rc = st_batch_run(b);
  const double t2 = wall_ms();
  if (!rc && launches) {
    // the run's launches, while the batch object is still this call's
    uint64_t n = 0;
    st_batch_launches(b, nullptr, 0, &n);
    const size_t at = launches->size();
    launches->resize(at + n);
    st_batch_launches(b, launches->data() + at, n, &n);
  }
  if (!rc) {
    std::lock_guard<std::mutex> lk(ctx->mu);
    std::vector<PairOut> &po = b->h_pout;
    std::vector<int32_t> &qe = b->h_qend, &te = b->h_tend;
    po.resize(count);
    qe.resize(count);
    te.resize(count);
    hipError_t e = hipMemcpyAsync(po.data(), b->d_pout.p, count * sizeof(PairOut), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(qe.data(), b->d_qend.p, count * 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(te.data(), b->d_tend.p, count * 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
      rc = fail(ST_ERR_DEVICE, std::string(who) + ": " + hipGetErrorString(e));
    } else {
      for (uint64_t k = 0; k < count; k++) {
        out[k].score = po[k].swg;
        out[k].reserved = 0;
        out[k].query_end = po[k].swg > 0 ? qe[k] : -1;
        out[k].target_end = po[k].swg > 0 ? te[k] : -1;
      }
    }
  }
  if (phase_ms) {
    phase_ms[0] = t1 - t0;
    phase_ms[1] = t2 - t1;
    phase_ms[2] = wall_ms() - t2;
  }
  return rc;
}

// The alignment spans' device buffers, kept by the context and grown as
// calls need (one call at a time, under the context's spans_mu): the
// row-reversed profiles, the reversed target prefixes as a database, and
// the tables of the kernels that write them.
// A reverse problem: query, end row, hit.
struct SpanRev {
  uint32_t q, e;
  uint64_t hit;
};

struct SpanWork {
  std::vector<st_qprof> views;  // the reversed prefixes' profile views, reused
  // host tables of a call, kept with their storage for the next call (tens
  // of thousands of entries each: fresh allocations cost page faults)
  std::vector<uint64_t> mir_off, doff, voff, win_off, rq_off, vq_off;
  std::vector<uint32_t> mir_ld, mir_es, chunk_job, chunk_first, rlen, vt;
  std::vector<MirrorJob> h_mj;
  std::vector<st_score_end> fw, rs;
  std::vector<SpanRev> rv;
  std::vector<RevPrefix> h_pre;
  std::vector<const st_qprof *> vq;
  std::vector<int32_t> vopen, vext, stop;
  std::vector<st_launch_info> launches;  // the last call's: forward, then reverse
  uint64_t n_forward = 0;
  double phase_ms[ST_SPAN_PHASES] = {};  // the last call's host phases
  DevBuf<uint8_t> mirror;
  DevBuf<RevPrefix> pre;
  DevBuf<MirrorJob> mj;
  DevBuf<uint32_t> cj, cf;
  st_db rdb;
};

static void span_work_free(SpanWork *w) { delete w; }

// ST_SPANS_STOP=0 runs every reverse problem over its whole rectangle
static bool spans_stop_enabled() {
  const char *e = std::getenv("ST_SPANS_STOP");
  return !(e && e[0] == '0');
}

int st_align_spans_batch(st_ctx *ctx, const st_db *db, const st_qprof *const *queries, uint64_t nq,
                         const uint64_t *hit_off, const uint32_t *hit_tids, const int32_t *gap_open,
                         const int32_t *gap_extend, st_span *out) {
  if (!ctx || !db || (nq && (!queries || !hit_off || !gap_open || !gap_extend)))
    return fail(ST_ERR_INVALID, "st_align_spans_batch: null argument");
  const uint64_t nh = nq ? hit_off[nq] : 0;
  if (nh == 0) return ST_OK;
  if (!out) return fail(ST_ERR_INVALID, "st_align_spans_batch: null output");
  std::lock_guard<std::mutex> sl(ctx->spans_mu);
  const double t_call = wall_ms();
  if (!ctx->spans) ctx->spans = new SpanWork();
  SpanWork &w = *ctx->spans;
  w.launches.clear();
  w.n_forward = 0;
  for (double &x : w.phase_ms) x = 0.0;
  // 0. row-reversed profiles of the queries with hits (int8 or int16 where
  // the query has such a profile: the end-cell kernels then stage fewer
  // bytes), built before the forward pass: they depend on the queries only.
  // A view's passes read at most its rows rounded up to a pass (<= 1024
  // rows) from its window start m - 1 - e: m + 1024 rows cover every view.
  std::vector<uint64_t> &mir_off = w.mir_off;  // bytes
  std::vector<uint32_t> &mir_ld = w.mir_ld, &mir_es = w.mir_es;
  std::vector<MirrorJob> &mj = w.h_mj;
  std::vector<uint32_t> &chunk_job = w.chunk_job, &chunk_first = w.chunk_first;
  mir_off.assign(nq + 1, 0);
  mir_ld.assign(nq, 0);
  mir_es.assign(nq, 4);
  mj.clear();
  chunk_job.clear();
  chunk_first.clear();
  for (uint64_t q = 0; q < nq; q++) {
    if (hit_off[q + 1] > hit_off[q] && queries[q]) {
      mir_ld[q] = (queries[q]->m + kMaxPassRows + 63) & ~63u;
      if (queries[q]->fits8 && queries[q]->prof8.p)
        mir_es[q] = 1;
      else if (queries[q]->fits16 && queries[q]->prof16.p)
        mir_es[q] = 2;
    }
    mir_off[q + 1] = mir_off[q] + (((uint64_t)kCodes * mir_ld[q] * mir_es[q] + 255) & ~255ull);
    if (mir_ld[q]) {
      chunk_first.push_back((uint32_t)chunk_job.size());
      const void *src = mir_es[q] == 1   ? (const void *)queries[q]->prof8.p
                        : mir_es[q] == 2 ? (const void *)queries[q]->prof16.p
                                         : (const void *)queries[q]->prof32.p;
      mj.push_back(MirrorJob{src, mir_off[q], queries[q]->ld, queries[q]->m, mir_ld[q], mir_es[q]});
      for (uint32_t c = 0; c < (mir_ld[q] + 255) / 256; c++) chunk_job.push_back((uint32_t)(mj.size() - 1));
    }
  }
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    ST_HIP(hipSetDevice(ctx->device));
    // grow-only buffers, uploads in stream order: the host tables are locals
    // of this call, which waits for the stream before they go
    hipStream_t st = ctx->stream;
    if (w.mirror.reserve(std::max<uint64_t>(mir_off[nq], 1)) != hipSuccess ||
        w.mj.put(mj.data(), mj.size(), st) != hipSuccess || w.cj.put(chunk_job.data(), chunk_job.size(), st) != hipSuccess ||
        w.cf.put(chunk_first.data(), chunk_first.size(), st) != hipSuccess)
      return fail(ST_ERR_NOMEM, "st_align_spans_batch: device allocation failed");
    ST_HIP(launch_mirror_prof(w.mj.p, w.cj.p, w.cf.p, (uint32_t)chunk_job.size(), w.mirror.p, st));
  }
  // every return from here waits for the stream: copies from this call's
  // host tables may still be queued on it
  struct StreamWait {
    hipStream_t s;
    ~StreamWait() { (void)hipStreamSynchronize(s); }
  } stream_wait{ctx->stream};
  // 1. forward pass: score and the column-major-first end cell
  std::vector<st_score_end> &fw = w.fw;
  fw.resize(nh);
  int rc = score_end_multi(ctx, db, queries, nq, hit_off, hit_tids, gap_open, gap_extend, fw.data(),
                           "st_align_spans_batch", nullptr, 0, &w.launches, w.phase_ms);
  if (rc) return rc;
  w.n_forward = w.launches.size();
  const double t_rev = wall_ms();
  for (uint64_t k = 0; k < nh; k++) {
    out[k].score = fw[k].score;
    out[k].reserved = 0;
    out[k].query_end = fw[k].query_end;
    out[k].target_end = fw[k].target_end;
    out[k].query_start = out[k].target_start = -1;
  }
  // 2. reverse pass over the hits with a positive score, one problem per
  // (query, end row): the reversed prefix rows[..=e] is the window of the
  // query's row-reversed profile that starts at row m - 1 - e
  // (query, end row) order: the hits come query by query, so each query's
  // run is sorted by end row on its own (stable: hit order within a row).
  // The tables are built on the planning threads, a run of queries each:
  // first the runs and their view counts, then the views, the reversed
  // target prefixes and the stop scores in place.
  std::vector<uint64_t> &rq_off = w.rq_off, &vq_off = w.vq_off;
  rq_off.assign(nq + 1, 0);
  for (uint64_t q = 0; q < nq; q++) {
    uint64_t cnt = 0;
    for (uint64_t k = hit_off[q]; k < hit_off[q + 1]; k++) cnt += fw[k].score > 0;
    rq_off[q + 1] = rq_off[q] + cnt;
  }
  const uint64_t nr = rq_off[nq];
  if (nr == 0) return ST_OK;
  if (nr > 0xffffffffull) return fail(ST_ERR_INVALID, "st_align_spans_batch: more than 2^32 hits");
  std::vector<SpanRev> &rv = w.rv;
  rv.resize(nr);
  vq_off.assign(nq + 1, 0);
  const unsigned T = (unsigned)std::min<uint64_t>({(uint64_t)WorkPool::get().width(), nq, nr / 2048 + 1});
  std::atomic<int> bad_end{0};
  WorkPool::get().run(T, [&](unsigned t) {
    for (uint64_t q = nq * t / T; q < nq * (t + 1) / T; q++) {
      uint64_t i = rq_off[q];
      const uint64_t r0 = i;
      for (uint64_t k = hit_off[q]; k < hit_off[q + 1]; k++)
        if (fw[k].score > 0) {
          // the gathers below read rows [0, query_end] and residues [0, target_end]
          if (fw[k].query_end < 0 || (uint64_t)fw[k].query_end >= queries[q]->m || fw[k].target_end < 0 ||
              (uint64_t)fw[k].target_end >= db->hlen[hit_tids[k]])
            bad_end.store(1, std::memory_order_relaxed);
          rv[i++] = SpanRev{(uint32_t)q, (uint32_t)std::max<int64_t>(fw[k].query_end, 0), k};
        }
      if (i - r0 <= 64) {
        // a short run: stable insertion sort, no allocation
        for (uint64_t a = r0 + 1; a < i; a++) {
          const SpanRev x = rv[a];
          uint64_t b = a;
          for (; b > r0 && rv[b - 1].e > x.e; b--) rv[b] = rv[b - 1];
          rv[b] = x;
        }
      } else {
        std::stable_sort(rv.begin() + r0, rv.begin() + i, [](const SpanRev &x, const SpanRev &y) { return x.e < y.e; });
      }
      uint64_t nv = 0;
      for (uint64_t a = r0; a < i; a++) nv += a == r0 || rv[a].e != rv[a - 1].e;
      vq_off[q + 1] = nv;
    }
  });
  if (bad_end.load()) return fail(ST_ERR_DEVICE, "st_align_spans_batch: end cell outside its problem");
  // one profile view per (query, end row): stride = the mirror's rows, rows
  // readable from the window's start = its extent (the planner keeps every
  // pass inside it); each query's views are consecutive and planned
  // together as one source, several views to a unit
  std::vector<uint64_t> &win_off = w.win_off;
  win_off.assign(1, 0);
  for (uint64_t q = 0; q < nq; q++) {
    vq_off[q + 1] += vq_off[q];
    if (vq_off[q + 1] > vq_off[q]) win_off.push_back(vq_off[q + 1]);
  }
  const uint64_t nviews = vq_off[nq];
  // st_qprof objects hold no allocations of their own here (views only)
  if (w.views.size() < nviews) w.views = std::vector<st_qprof>(nviews + nviews / 4);
  st_qprof *views = w.views.data();
  std::vector<const st_qprof *> &vq = w.vq;
  std::vector<uint64_t> &voff = w.voff;
  std::vector<int32_t> &vopen = w.vopen, &vext = w.vext, &stop = w.stop;
  std::vector<uint32_t> &vt = w.vt, &rlen = w.rlen;
  std::vector<RevPrefix> &pre = w.h_pre;
  vq.resize(nviews);
  voff.resize(nviews + 1);
  vopen.resize(nviews);
  vext.resize(nviews);
  vt.resize(nr);
  rlen.resize(nr);
  pre.resize(nr);
  stop.resize(nr);
  const uint8_t *mirror = w.mirror.p;
  WorkPool::get().run(T, [&](unsigned t) {
    for (uint64_t q = nq * t / T; q < nq * (t + 1) / T; q++) {
      uint64_t v = vq_off[q];
      const uint32_t o0 = queries[q]->m - 1, ld = mir_ld[q], es = mir_es[q];
      const uint8_t *mb = mirror + mir_off[q];
      for (uint64_t k = rq_off[q]; k < rq_off[q + 1]; k++) {
        const uint32_t e = rv[k].e;
        if (k == rq_off[q] || e != rv[k - 1].e) {
          const uint32_t o = o0 - e;
          st_qprof *vp = &views[v];
          vp->ctx = ctx;
          vp->m = e + 1;
          vp->ld = ld - o;
          vp->fits8 = es == 1;
          vp->fits16 = es <= 2;
          vp->view.prof8 = es == 1 ? (const int8_t *)mb + o : nullptr;
          vp->view.prof16 = es == 2 ? (const int16_t *)mb + o : nullptr;
          vp->view.prof32 = es == 4 ? (const int32_t *)mb + o : nullptr;
          vp->view.m = e + 1;
          vp->view.ld = ld;
          vq[v] = vp;
          vopen[v] = gap_open[q];
          vext[v] = gap_extend[q];
          voff[v] = k;
          v++;
        }
        const uint64_t h = rv[k].hit;
        vt[k] = (uint32_t)k;
        rlen[k] = (uint32_t)(fw[h].target_end + 1);
        pre[k] = RevPrefix{hit_tids[h], rlen[k]};
        // each reversed prefix is known to reach its hit's score: the
        // reverse pass stops once the first column holding it is complete
        stop[k] = fw[h].score;
      }
    }
  });
  voff[nviews] = nr;
  // the reversed target prefixes (database layout: 16-byte aligned starts,
  // code 21 padding)
  std::vector<uint64_t> &doff = w.doff;
  doff.resize(nr + 1);
  doff[0] = 0;
  uint32_t max_len = 0;
  for (uint64_t k = 0; k < nr; k++) {
    doff[k + 1] = doff[k] + ((rlen[k] + 15ull) & ~15ull);
    max_len = std::max(max_len, rlen[k]);
  }
  st_db *rdb = &ctx->spans->rdb;
  rdb->ctx = ctx;
  rdb->n_targets = nr;
  rdb->n_res = doff[nr];
  rdb->hlen = rlen;
  rdb->max_len = max_len;
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    ST_HIP(hipSetDevice(ctx->device));
    // in stream order, before the reverse pass's launches; the host tables
    // stay alive until this call has waited for the stream
    hipStream_t st = ctx->stream;
    if (w.pre.put(pre.data(), nr, st) != hipSuccess || rdb->off.put(doff.data(), nr + 1, st) != hipSuccess ||
        rdb->len.put(rlen.data(), nr, st) != hipSuccess || rdb->arena.reserve(doff[nr] + 64) != hipSuccess)
      return fail(ST_ERR_NOMEM, "st_align_spans_batch: device allocation failed");
    ST_HIP(launch_rev_prefix(db->arena.p, db->off.p, w.pre.p, (uint32_t)nr, rdb->off.p, rdb->arena.p, st));
  }
  std::vector<st_score_end> &rs = w.rs;
  rs.resize(nr);
  w.phase_ms[3] = wall_ms() - t_rev;
  rc = score_end_multi(ctx, rdb, vq.data(), vq.size(), voff.data(), vt.data(), vopen.data(), vext.data(),
                       rs.data(), "st_align_spans_batch", win_off.data(), win_off.size() - 1, &w.launches,
                       w.phase_ms + 4, spans_stop_enabled() ? stop.data() : nullptr);
  if (rc) return rc;
  for (uint64_t k = 0; k < nr; k++) {
    st_span &s = out[rv[k].hit];
    // with gap costs >= 0 the reversed prefixes hold the optimal alignment
    // and nothing better; with a negative cost the two passes can disagree
    // (src/align.rs:859-863), and the start stays -1 for the anchored search
    // of the traceback
    if (rs[k].score != s.score) continue;
    s.query_start = s.query_end - rs[k].query_end;
    s.target_start = s.target_end - rs[k].target_end;
  }
  w.phase_ms[7] = wall_ms() - t_call;
  return ST_OK;
}

int st_align_spans_phases(st_ctx *ctx, double ms[ST_SPAN_PHASES]) {
  if (!ctx || !ms) return fail(ST_ERR_INVALID, "st_align_spans_phases: null argument");
  std::lock_guard<std::mutex> sl(ctx->spans_mu);
  for (int k = 0; k < ST_SPAN_PHASES; k++) ms[k] = ctx->spans ? ctx->spans->phase_ms[k] : 0.0;
  return ST_OK;
}

int st_align_spans_launches(st_ctx *ctx, st_launch_info *out, uint64_t cap, uint64_t *n, uint64_t *n_forward) {
  if (!ctx || !n || (cap && !out)) return fail(ST_ERR_INVALID, "st_align_spans_launches: null argument");
  std::lock_guard<std::mutex> sl(ctx->spans_mu);
  const SpanWork *w = ctx->spans;
  *n = w ? w->launches.size() : 0;
  if (n_forward) *n_forward = w ? w->n_forward : 0;
  for (uint64_t i = 0; w && i < std::min<uint64_t>(cap, *n); i++) out[i] = w->launches[i];
  return ST_OK;
}

int st_ungapped_score_jobs(st_ctx *ctx, const st_qprof *qp, const st_db *db, const st_job *jobs, uint64_t njobs,
                           int32_t *scores) {
  if (!ctx || !qp || !db || (njobs && (!jobs || !scores)))
    return fail(ST_ERR_INVALID, "st_ungapped_score_jobs: null argument");
  if (njobs == 0) return ST_OK;
  for (uint64_t k = 0; k < njobs; k++) {
    const st_job &j = jobs[k];
    if (j.tid >= db->n_targets) return fail(ST_ERR_INVALID, "st_ungapped_score_jobs: target id out of range");
    if (j.hi > j.lo) {
      // the window must lie inside diag_range (src/ungapped.rs:305-309)
      const int64_t n = db->hlen[j.tid];
      const int64_t a = std::max<int64_t>(-(int64_t)j.diag, 0);
      const int64_t e = std::min<int64_t>((int64_t)qp->m, n - j.diag);
      if ((int64_t)j.lo < a || (int64_t)j.hi > e)
        return fail(ST_ERR_INVALID, "st_ungapped_score_jobs: job window outside its diagonal");
    }
  }
  std::lock_guard<std::mutex> lk(ctx->mu);
  ST_HIP(hipSetDevice(ctx->device));
  DevBuf<JobDev> dj;
  DevBuf<int32_t> ds;
  if (dj.upload((const JobDev *)jobs, njobs, ctx->stream) != hipSuccess || ds.alloc(njobs) != hipSuccess)
    return fail(ST_ERR_NOMEM, "st_ungapped_score_jobs: device allocation failed");
  ST_HIP(launch_ung_jobs(dj.p, njobs, qp->qd.p, qp->view.prof16 ? qp->m : 0, db->arena.p, db->off.p, ds.p, nullptr,
                         ctx->stream));
  ST_HIP(hipMemcpyAsync(scores, ds.p, njobs * 4, hipMemcpyDeviceToHost, ctx->stream));
  ST_HIP(hipStreamSynchronize(ctx->stream));
  return ST_OK;
}

}  // extern "C"

namespace st {

int score_and_gate_check(const char *fn, st_ctx *ctx, const st_qprof *qp, const st_db *db, const st_job *jobs,
                         uint64_t njobs, const st_target_rec *recs, uint64_t nrecs, const st_candidate *out,
                         uint64_t *n_out) {
  const std::string f(fn);
  if (!ctx || !qp || !db || !n_out || (njobs && !jobs) || (nrecs && (!recs || !out)))
    return fail(ST_ERR_INVALID, f + ": null argument");
  if (qp->ctx != ctx || db->ctx != ctx) return fail(ST_ERR_INVALID, f + ": handle of another context");
  *n_out = 0;
  for (uint64_t r = 0; r < nrecs; r++)
    if (recs[r].j1 < recs[r].j0 || recs[r].j1 > njobs || recs[r].tid >= db->n_targets)
      return fail(ST_ERR_INVALID, f + ": record out of range");
  // jobs must lie inside their diagonals (st_ungapped_score_jobs's checks)
  for (uint64_t k = 0; k < njobs; k++) {
    const st_job &j = jobs[k];
    if (j.tid >= db->n_targets) return fail(ST_ERR_INVALID, f + ": target id out of range");
    if (j.hi > j.lo) {
      const int64_t n = db->hlen[j.tid];
      const int64_t a = std::max<int64_t>(-(int64_t)j.diag, 0);
      const int64_t e = std::min<int64_t>((int64_t)qp->m, n - j.diag);
      if ((int64_t)j.lo < a || (int64_t)j.hi > e) return fail(ST_ERR_INVALID, f + ": job window outside its diagonal");
    }
  }
  return ST_OK;
}

}  // namespace st

extern "C" {

int st_score_and_gate(st_ctx *ctx, const st_qprof *qp, const st_db *db, const st_job *jobs, uint64_t njobs,
                      const st_target_rec *recs, uint64_t nrecs, int32_t gap_open, int32_t gap_extend,
                      int32_t min_diag_score, int regions, st_candidate *out, uint64_t *n_out,
                      uint64_t *chains_used, int32_t *job_scores) {
  if (int rc = score_and_gate_check("st_score_and_gate", ctx, qp, db, jobs, njobs, recs, nrecs, out, n_out)) return rc;
  if (njobs == 0 && nrecs == 0) return ST_OK;
  std::lock_guard<std::mutex> lk(ctx->mu);
  ST_HIP(hipSetDevice(ctx->device));
  DevBuf<JobDev> dj;
  DevBuf<int32_t> ds;
  DevBuf<int2> dseg;
  DevBuf<RecDev> dr;
  DevBuf<GateRes> dg;
  if (dj.upload((const JobDev *)jobs, njobs, ctx->stream) != hipSuccess || ds.alloc(njobs) != hipSuccess ||
      dseg.alloc(njobs) != hipSuccess || dr.upload((const RecDev *)recs, nrecs, ctx->stream) != hipSuccess ||
      dg.alloc(nrecs) != hipSuccess)
    return fail(ST_ERR_NOMEM, "st_score_and_gate: device allocation failed");
  // scores and segment bounds in one scan, then the gate on them
  ST_HIP(launch_ung_jobs(dj.p, njobs, qp->qd.p, qp->view.prof16 ? qp->m : 0, db->arena.p, db->off.p, ds.p, dseg.p,
                         ctx->stream));
  ST_HIP(launch_gate_recs(dr.p, nrecs, dj.p, ds.p, dseg.p, gap_open, gap_extend, min_diag_score, regions, dg.p,
                          ctx->stream));
  std::vector<GateRes> res(nrecs);
  if (job_scores && njobs)
    ST_HIP(hipMemcpyAsync(job_scores, ds.p, njobs * 4, hipMemcpyDeviceToHost, ctx->stream));
  if (nrecs)
    ST_HIP(hipMemcpyAsync(res.data(), dg.p, nrecs * sizeof(GateRes), hipMemcpyDeviceToHost, ctx->stream));
  ST_HIP(hipStreamSynchronize(ctx->stream));
  uint64_t nc = 0, chains = 0;
  for (uint64_t r = 0; r < nrecs; r++) {
    const GateRes &g = res[r];
    if (recs[r].j1 == recs[r].j0) continue;
    if (g.chain) chains++;
    if (!g.pass) continue;
    st_candidate c;
    c.target_id = recs[r].tid;
    c.hits = recs[r].hits;
    c.best_diag = g.best_diag;
    c.diag_score = g.score;
    c.t_lo = g.t_lo;
    c.t_hi = g.t_hi;
    out[nc++] = c;
  }
  *n_out = nc;
  if (chains_used) *chains_used += chains;
  return ST_OK;
}

}  // extern "C"

namespace st {

int score_multi(st_ctx *ctx, const st_db *db, const st_qprof *const *qps, uint64_t nq, const uint64_t *off,
                const uint32_t *tids, const int32_t *open, const int32_t *ext, int32_t *scores) {
  const uint64_t count = nq ? off[nq] : 0;
  if (count == 0) return ST_OK;
  st_batch *b = nullptr;
  int rc = batch_create_impl(ctx, db, qps, nq, off, tids, open, ext, nullptr, false, &b);
  if (rc) return rc;
  rc = st_batch_run(b);
  if (!rc) {
    std::vector<st_pair_result> res(count);
    rc = st_batch_results(b, res.data());
    if (!rc)
      for (uint64_t k = 0; k < count; k++) scores[k] = res[k].swg;
  }
  st_batch_destroy(b);
  return rc;
}
}  // namespace st
