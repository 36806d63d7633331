// Host-side internals shared by the C ABI translation units: kernel entry
// points, error reporting, device buffers, the context / database / profile
// objects and the work planner.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/sabertooth_amd.h"
#include "st_device.h"
#include "dp_launch.h"

namespace st {

// ---- kernel entry points (dp_kernels.hip, gate_kernels.hip) -----------------

// the kernels read the caller's st_job / st_target_rec buffers as JobDev /
// RecDev (st_device.h): the layouts must agree field by field
static_assert(sizeof(JobDev) == sizeof(st_job) && offsetof(JobDev, tid) == offsetof(st_job, tid) &&
                  offsetof(JobDev, diag) == offsetof(st_job, diag) && offsetof(JobDev, lo) == offsetof(st_job, lo) &&
                  offsetof(JobDev, hi) == offsetof(st_job, hi),
              "JobDev must have st_job's layout");
static_assert(sizeof(RecDev) == sizeof(st_target_rec) && offsetof(RecDev, tid) == offsetof(st_target_rec, tid) &&
                  offsetof(RecDev, hits) == offsetof(st_target_rec, hits) &&
                  offsetof(RecDev, j0) == offsetof(st_target_rec, j0) &&
                  offsetof(RecDev, j1) == offsetof(st_target_rec, j1),
              "RecDev must have st_target_rec's layout");
static_assert(sizeof(PairResultDev) == sizeof(st_pair_result) &&
                  offsetof(PairResultDev, swg) == offsetof(st_pair_result, swg) &&
                  offsetof(PairResultDev, ung) == offsetof(st_pair_result, ung) &&
                  offsetof(PairResultDev, best_diag) == offsetof(st_pair_result, best_diag) &&
                  offsetof(PairResultDev, gate_score) == offsetof(st_pair_result, gate_score) &&
                  offsetof(PairResultDev, pass) == offsetof(st_pair_result, pass) &&
                  offsetof(PairResultDev, kernel) == offsetof(st_pair_result, kernel),
              "PairResultDev must have st_pair_result's layout");
// stage_m: the longest query among the call's queries that has an int16
// profile (0 when none); runs of such queries are scored from LDS.  segs
// (nullable): each job's kadane_segment bounds, which the gate needs.
hipError_t launch_ung_jobs(const JobDev *jobs, uint64_t njobs, const QueryDev *qd, uint32_t stage_m,
                           const uint8_t *arena, const uint64_t *toff, int32_t *out, int2 *segs,
                           hipStream_t s);
hipError_t launch_gate_recs(const RecDev *recs, uint64_t nrecs, const JobDev *jobs,
                            const int32_t *scores, const int2 *segs, int32_t open, int32_t ext,
                            int32_t gate, int regions, GateRes *out, hipStream_t s);
// The same over jobs and records of several queries, sorted by query: job i
// of query jq[i], record i of query rq[i] with that query's gap costs and
// gate (open, ext, gate indexed by query; null: open1, ext1, gate1).  pass
// and chain (nullable): each record's flags as 0/1 words, for a compaction.
hipError_t launch_ung_jobs_multi(const JobDev *jobs, uint64_t njobs, const uint32_t *jq,
                                 const QueryDev *queries, uint32_t stage_m, const uint8_t *arena,
                                 const uint64_t *toff, int32_t *out, int2 *segs, hipStream_t s);
hipError_t launch_gate_recs_multi(const RecDev *recs, uint64_t nrecs, const JobDev *jobs,
                                  const int32_t *scores, const int2 *segs, const uint32_t *rq,
                                  const int32_t *open, const int32_t *ext, const int32_t *gate,
                                  int regions, GateRes *out, hipStream_t s, int32_t open1 = 0,
                                  int32_t ext1 = 0, int32_t gate1 = 0, uint32_t *pass = nullptr,
                                  uint32_t *chain = nullptr);

// ---- errors -------------------------------------------------------------------
// Record msg as this thread's last error (st_last_error) and return code.
int fail(int code, const std::string &msg);

#define ST_HIP(call)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess)                                                              \
      return fail(ST_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Device allocation owned by a host object.
template <class T>
struct DevBuf {
  T *p = nullptr;
  size_t n = 0;    // elements in use
  size_t cap = 0;  // elements allocated
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = cap = 0;
  }
  hipError_t alloc(size_t count) {
    release();
    n = count;
    if (count == 0) return hipSuccess;
    const hipError_t e = hipMalloc((void **)&p, count * sizeof(T));
    if (e == hipSuccess) cap = count;
    return e;
  }
  // grow-only: keeps the allocation when it holds count elements
  hipError_t reserve(size_t count) {
    if (p && count <= cap) {
      n = count;
      return hipSuccess;
    }
    return alloc(count);
  }
  hipError_t upload(const T *src, size_t count, hipStream_t s) {
    hipError_t e = alloc(count);
    if (e != hipSuccess || count == 0) return e;
    e = hipMemcpyAsync(p, src, count * sizeof(T), hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(s);
  }
  // grow-only upload in stream order, no wait: src must stay unchanged until
  // the stream has passed the copy
  hipError_t put(const T *src, size_t count, hipStream_t s) {
    hipError_t e = reserve(count);
    if (e != hipSuccess || count == 0) return e;
    return hipMemcpyAsync(p, src, count * sizeof(T), hipMemcpyHostToDevice, s);
  }
};

}  // namespace st

struct st_ctx {
  int device = 0;
  hipStream_t stream = nullptr;  // main stream: uploads, gate, results
  hipStream_t slots[4] = {nullptr, nullptr, nullptr, nullptr};  // concurrent DP launches
  std::mutex mu;
  int cus = 256;
  st::DevBuf<uint32_t> counter;  // one unit counter per launch slot
  // the end-cell batches (st_sw_score_end_batch, both passes of the
  // alignment spans) reuse one batch object and its grow-only device
  // buffers, one call at a time under end_mu; the alignment spans' own
  // buffers likewise under spans_mu
  std::mutex end_mu, spans_mu;
  st_batch *end_batch = nullptr;
  struct SpanWork *spans = nullptr;
};

struct st_db {
  st_ctx *ctx = nullptr;
  st::DevBuf<uint8_t> arena;
  st::DevBuf<uint64_t> off;
  st::DevBuf<uint32_t> len;
  std::vector<uint32_t> hlen;
  uint32_t max_len = 0;  // the longest target
  uint64_t n_targets = 0, n_res = 0;
};

// An index's posting lists on the device, decoded, and the grow-only device
// buffers of the seed-hit stage run against them (calls on one context are
// serialised, so one set serves every call).
struct st_postings {
  st_ctx *ctx = nullptr;
  st::DevBuf<uint64_t> post;  // (target << 32) | position
  uint64_t n = 0;
  mutable st::DevBuf<uint8_t> scratch[32];
  // events of the last chained seeds-to-candidates call: its start, the job
  // scan's start and end, the gate's end, its last kernel's end
  hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  mutable bool timed = false;
  ~st_postings() {
    for (hipEvent_t &e : ev)
      if (e) (void)hipEventDestroy(e);
  }
};

struct st_qprof {
  // what planning reads, first and together (profile views are built by
  // the ten thousand per call): the context, sizes, precision flags and the
  // device view
  st_ctx *ctx = nullptr;
  uint32_t m = 0, ld = 0;
  bool fits16 = false;
  bool fits8 = false;  // every value in [-128, 127]: prof8 holds them
  st::QueryDev view{};
  std::vector<int32_t> rows;
  st::DevBuf<int8_t> prof8;
  st::DevBuf<int16_t> prof16;
  st::DevBuf<int32_t> prof32;
  st::DevBuf<st::QueryDev> qd;  // a one-entry table for single-query calls
};

namespace st {

// st_score_and_gate's argument checks (handles, record and job bounds),
// shared by its variants; fn names the entry point in the message.
int score_and_gate_check(const char *fn, st_ctx *ctx, const st_qprof *qp, const st_db *db, const st_job *jobs,
                         uint64_t njobs, const st_target_rec *recs, uint64_t nrecs, const st_candidate *out,
                         uint64_t *n_out);

// Resolves, once per device and with that device current, the address of
// the fixed device scratch that st_score_and_gate_noalloc uses.
int noalloc_init(int device);

int dp_occupancy(int kind, int R, bool ung, size_t lds_bytes, bool end = false);

// Exact SWG scores of a multi-query candidate list (score-only batch, the
// int32 rescoring included), in candidate order: query q scores
// tids[off[q] .. off[q + 1]) with (open[q], ext[q]).
int score_multi(st_ctx *ctx, const st_db *db, const st_qprof *const *qps, uint64_t nq, const uint64_t *off,
                const uint32_t *tids, const int32_t *open, const int32_t *ext, int32_t *scores);

// Precision class of a query under given gaps.
// kI32F16 / kI32N16 and kI32F8 / kI32N8: the int32 recurrences over the
// query's int16 or int8 profile (end-cell batches of queries whose profile
// fits)
enum Kind { kPk16 = 0, kI32F = 1, kI32N = 2, kI32F16 = 3, kI32N16 = 4, kI32F8 = 5, kI32N8 = 6 };
// bytes per profile entry in LDS
inline int kind_entry_bytes(int k) {
  return (k == kI32F8 || k == kI32N8) ? 1 : (k == kPk16 || k == kI32F16 || k == kI32N16) ? 2 : 4;
}
// the LDS row stride (entries) of a launch whose passes hold at most gr_max
// rows: a whole number of dwords plus one (profile rows start on a dword,
// rows of consecutive codes fall in different banks)
inline int kind_lds_stride(int gr_max, int k) { return gr_max + (kind_entry_bytes(k) == 1 ? 4 : 2); }
// the unfloored recurrences (negative gap costs): 8 rows per lane only
inline bool kind_unfloored(int k) { return k == kI32N || k == kI32N16 || k == kI32N8; }

Kind kind_for(const st_qprof *qp, int32_t open, int32_t ext);

// ST_OK, or ST_ERR_INVALID (with the message) when ST_SHAPE_TABLE names a
// malformed launch-shape cost table.
int shape_table_check();

// Streams (and unit counters) for concurrent DP launches.
constexpr int kLaunchSlots = 4;

// One kernel launch: units [unit0, unit0 + nunits) of one precision kind, one
// R, one gap setting and one LDS class; buffer region [buf0, buf0 + grid *
// (2 half + edge)) of the batch's pass buffer.
struct Launch {
  int kind, R;
  int32_t open, ext;
  uint32_t unit0, nunits;
  int gr_max;
  uint64_t half, buf0;
  uint32_t grid;
  int slot;
  uint64_t useful_cells = 0, exec_cells = 0;  // filled by the batch planner
};

struct Plan {
  std::vector<Unit> units;
  std::vector<PairDev> pairs;
  std::vector<uint32_t> pair_q;  // window units: each pair's profile view (parallel to pairs)
  std::vector<Launch> launches;
  uint64_t buf_elems = 0;
  uint64_t exec_cells = 0;  // cells the launches execute, padding included
  // empty, keeping the vectors' storage for the next plan
  void clear() {
    units.clear();
    pairs.clear();
    pair_q.clear();
    launches.clear();
    buf_elems = exec_cells = 0;
  }
};

// Add the pairs of query qi (result rows outs[], target ids tids[]) to the
// pending units of `pl`: targets sorted by length, cut into wave units whose
// launch shape is chosen per unit; `cap` bounds a unit's DP steps.
// Shape limits of a kernel that cannot take every shape (the serving
// kernel: one R, bounded G*R and pass-boundary rows).
struct PlanLimits {
  int only_R = 0;              // 0: any
  int max_GR = 1 << 30;
  uint64_t max_half = ~0ull;   // (64 / G) * n_pad of multi-pass units
};

// The planner's environment: forced shapes (ST_FORCE_R, ST_FORCE_G) and the
// launch-shape cost table, read once per plan of many queries.
struct PlanEnv {
  int fr = 0, fg = 0;
  double ps[3][7];
};
PlanEnv plan_env();

// Returns false (and plans nothing further) when `lim` admits no shape for
// some unit.  env (nullable): the plan's environment, read here otherwise.
bool plan_query(Plan &pl, std::vector<Unit> &pending, uint32_t qi, const st_qprof *qp, const uint32_t *tids,
                const uint64_t *outs, size_t count, Kind kind, const st_db *db, double cap,
                const PlanLimits *lim = nullptr, const PlanEnv *env = nullptr);

// Unit step cap for a batch of `batch_cells` DP cells: short units when the
// batch is small (latency), growing with the batch without an upper bound.
double unit_steps_cap(uint64_t batch_cells);

// Group pending units into launches (kind, R, gaps, LDS class), longest
// first, size their grids for `cus` compute units and lay out their
// pass-buffer regions.  force_kind >= 0 overrides every query's kind.
void finish_plan(Plan &pl, std::vector<Unit> &pending, const std::vector<Kind> &qkind,
                 const std::vector<int32_t> &open, const std::vector<int32_t> &ext, int force_kind, int cus,
                 bool ung, bool end = false);

}  // namespace st
