// Serving runtime: the target database stays resident, requests (a query
// profile and its candidate list) queue up, and worker slots merge whatever
// is queued into micro-batches as DP-state memory and slots free up
// (continuous batching).  Each slot owns a HIP stream, a unit counter and
// grow-only device / pinned buffers, so its micro-batch's persistent DP
// launch runs beside the other slots' launches and new requests start while
// earlier ones are still on the device.
//
// The reference has no serving mode: its CLI scores fixed query batches
// (src/main.rs:255, Engine::search_batch src/search.rs:692-731).  Each
// request's outputs are exactly those of st_batch_create/st_batch_run for the
// same query and candidates: the kernels and the planner are shared, and no
// result depends on what else is in the micro-batch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <memory>
#include <numeric>
#include <thread>
#include <unordered_map>

#include "host_internal.h"
#include "serve_args.h"

namespace st {


namespace {

double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

// Device buffer that only grows (by at least half its size).
template <class T>
struct GrowBuf {
  DevBuf<T> b;
  hipError_t reserve(size_t n) {
    if (b.p && n <= b.n) return hipSuccess;
    return b.alloc(std::max<size_t>({n, b.n + b.n / 2, 1}));
  }
  T *p() const { return b.p; }
};

// Pinned host staging that only grows.
template <class T>
struct Pinned {
  T *p = nullptr;
  size_t cap = 0;
  Pinned() = default;
  Pinned(const Pinned &) = delete;
  Pinned &operator=(const Pinned &) = delete;
  ~Pinned() {
    if (p) (void)hipHostFree(p);
  }
  hipError_t reserve(size_t n) {
    if (p && n <= cap) return hipSuccess;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    cap = std::max<size_t>({n, cap + cap / 2, 1});
    return hipHostMalloc((void **)&p, cap * sizeof(T), hipHostMallocDefault);
  }
};

struct Req {
  uint64_t ticket = 0;
  const st_qprof *qp = nullptr;
  int32_t open = 0, ext = 0, gate = 0;
  std::vector<uint32_t> cands;
  st_pair_result *out = nullptr;
  uint64_t state_bytes = 0, cells = 0;
  double t_submit = 0, t_done = 0;
  int status = -1;  // -1 pending, else an ST_ status code
  std::string err;
};

// DP state a request holds on the device while it runs: per pair the raw
// outputs, diagonal keys, result row and pair slots; per target column the
// pass-boundary row (H, F, U, B words shared by two pairs, two halves).
// Device DP state of a request of n candidates with `cols` target residues
// in all: per-pair outputs and slots, and the pass-boundary rows.
uint64_t state_bytes_of(uint64_t n, uint64_t cols) {
  return n * (sizeof(PairOut) + 32 + sizeof(st_pair_result) + 2 * sizeof(PairDev) + 4) + cols * 16;
}

}  // namespace

// Concurrent launch streams of one slot: a micro-batch's launches of
// different kernel instances (rows per lane, precision, LDS class) run side
// by side.
constexpr int kSubStreams = 3;

struct Slot {
  st_server *srv = nullptr;
  int index = 0;
  hipStream_t stream = nullptr;
  hipStream_t sub[kSubStreams] = {nullptr, nullptr, nullptr};
  hipEvent_t fork = nullptr, join[kSubStreams] = {nullptr, nullptr, nullptr};
  hipEvent_t t_start = nullptr, t_end = nullptr;  // device time of a micro-batch
  DevBuf<uint32_t> counter;                       // one per sub-stream
  // inputs packed into one pinned block and one device block (one copy)
  Pinned<uint8_t> h_in;
  GrowBuf<uint8_t> d_in;
  GrowBuf<Unit> d_rs_units;
  GrowBuf<PairDev> d_rs_pairs;
  GrowBuf<st_pair_result> d_res;
  GrowBuf<uint4> d_buf, d_rs_buf;
  Pinned<Unit> h_units;
  Pinned<PairDev> h_pairs;
  Pinned<st_pair_result> h_res;
  std::thread th;
};

// Host-coherent pinned memory (fine-grained: the running kernel and the host
// see each other's writes without a kernel boundary).
template <class T>
struct Coherent {
  T *p = nullptr;  // host address
  T *d = nullptr;  // device address of the same memory
  size_t n = 0;
  Coherent() = default;
  Coherent(const Coherent &) = delete;
  Coherent &operator=(const Coherent &) = delete;
  ~Coherent() {
    if (p) (void)hipHostFree(p);
  }
  hipError_t alloc(size_t count) {
    n = count;
    hipError_t e = hipHostMalloc((void **)&p, count * sizeof(T), hipHostMallocCoherent | hipHostMallocMapped);
    if (e == hipSuccess) std::memset((void *)p, 0, count * sizeof(T));
    if (e == hipSuccess) e = hipHostGetDevicePointer((void **)&d, p, 0);
    return e;
  }
};

// Circular allocator of contiguous ranges, freed in any order and reclaimed
// in allocation order.
struct RangeRing {
  uint64_t cap = 0, wpos = 0;
  struct A {
    uint64_t off, len;
    bool done;
  };
  std::deque<A> live;
  bool alloc(uint64_t len, uint64_t *off) {
    len = std::max<uint64_t>(len, 1);
    if (len > cap) return false;
    if (live.empty()) wpos = 0;
    uint64_t o;
    if (live.empty()) {
      o = 0;
    } else {
      // live ranges lie in [oldest, wpos) unless they wrapped (wpos <= oldest)
      const uint64_t oldest = live.front().off;
      if (wpos > oldest) {
        if (wpos + len <= cap)
          o = wpos;
        else if (len <= oldest)
          o = 0;
        else
          return false;
      } else {
        if (wpos + len <= oldest)
          o = wpos;
        else
          return false;
      }
    }
    live.push_back(A{o, len, false});
    wpos = o + len;
    *off = o;
    return true;
  }
  void release(uint64_t off) {
    for (A &a : live)
      if (a.off == off && !a.done) {
        a.done = true;
        break;
      }
    while (!live.empty() && live.front().done) live.pop_front();
  }
};

// A request in the persistent kernel's queue.
struct Flight {
  std::shared_ptr<Req> req;
  uint32_t slot;       // completion counter
  uint64_t r0, n;      // result rows (plan-slot order)
  uint64_t u0, units;  // unit indices
  double t_pub;        // publication time
  std::vector<uint32_t> cand_of;  // plan slot -> candidate index
};

// The persistent serving engine: work queue, pools and the kernel's launches.
struct Persist {
  static constexpr uint64_t kRingUnits = 1ull << 20;
  static constexpr uint64_t kResultRows = 1ull << 23;
  static constexpr uint32_t kReqSlots = 1u << 16;
  static constexpr uint64_t kHalfQuads = 20480;  // pass-boundary half of one block
  uint64_t arena_bytes = 4ull << 30;
  Coherent<ServeUnit> ring;
  Coherent<uint64_t> ctl;  // [0] tail, [kClockWord] the device wall clock
  static constexpr int kClockWord = 8;  // its own 64-byte line
  // wave priority by request age: 1, 2, 3 from 0.6, 1.2, 2 ms (100 MHz ticks)
  uint32_t prio_ticks[3] = {60000, 120000, 200000};
  Coherent<uint32_t> stop;  // [0] stop, [4] malformed-unit flag
  Coherent<ServeRow> results;
  Coherent<uint32_t> remaining;
  DevBuf<uint64_t> dctl;  // the kernel's device control words
  DevBuf<uint4> passbuf;
  DevBuf<uint8_t> arena;  // profiles and pair slots of published requests
  Pinned<uint8_t> h_stage;  // a request's arena image
  uint64_t arena_used = 0;
  uint64_t inflight_cells = 0;  // guarded by the server mutex
  double close_started = 0;
  std::atomic<bool> draining{false};
  bool feeder_done = false;  // guarded by the server mutex
  bool broken = false;       // requests were failed; guarded by the server mutex
  hipStream_t kstream = nullptr, cstream = nullptr;
  hipEvent_t kdone = nullptr;
  uint32_t grid = 0;
  int lds_stride = 258;
  uint64_t published = 0;  // units published (host copy of the tail)
  uint64_t launches = 0;
  RangeRing rows;
  std::deque<uint64_t> unit_live;  // u0 of requests in flight, in order
  std::vector<uint32_t> free_slots;
  std::vector<Flight> flights;
  std::mutex kmu;  // launches
  std::thread feeder, completer;
  ~Persist() {
    if (kdone) (void)hipEventDestroy(kdone);
    if (kstream) (void)hipStreamDestroy(kstream);
    if (cstream) (void)hipStreamDestroy(cstream);
  }
};

}  // namespace st

using namespace st;

struct st_server {
  st_ctx *ctx = nullptr;
  const st_db *db = nullptr;
  st_server_config cfg{};
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::condition_variable cv_feed;  // the persistent path's feeder
  std::deque<std::shared_ptr<Req>> queue;
  std::unordered_map<uint64_t, std::shared_ptr<Req>> live;
  uint64_t next_ticket = 1;
  bool closing = false;
  uint64_t inflight_bytes = 0;
  std::vector<std::unique_ptr<Slot>> slots;
  st_server_stats stats{};
  // persistent path: requests it cannot take go to fb_queue (the slots)
  std::unique_ptr<Persist> pe;
  std::deque<std::shared_ptr<Req>> fb_queue;
};

namespace st {
namespace {

// Launch a plan's DP kernels: one launch runs on the slot's stream, several
// fan out over the slot's sub-streams (each with its own unit counter) and
// join back before the slot stream continues.
int launch_on(Slot &sl, const Plan &pl, const QueryDev *d_q, const Unit *d_units, const PairDev *d_pairs,
              uint4 *d_buf, st_pair_result *d_res, const int32_t *d_gate) {
  const st_db *db = sl.srv->db;
  const size_t nl = pl.launches.size();
  const bool fan = nl > 1;
  if (fan) {
    ST_HIP(hipEventRecord(sl.fork, sl.stream));
    for (int k = 0; k < kSubStreams && k < (int)nl; k++) ST_HIP(hipStreamWaitEvent(sl.sub[k], sl.fork, 0));
  }
  for (size_t i = 0; i < nl; i++) {
    const Launch &L = pl.launches[i];
    const int k = (int)(i % kSubStreams);
    DpLaunch a{};
    a.units = d_units + L.unit0;
    a.nunits = L.nunits;
    a.counter = sl.counter.p + (fan ? k : 0);
    a.pairs = d_pairs;
    a.queries = d_q;
    a.arena = db->arena.p;
    a.toff = db->off.p;
    a.tlen = db->len.p;
    a.open = L.open;
    a.ext = L.ext;
    a.grid = L.grid;
    a.passbuf = d_buf + L.buf0;
    a.passbuf_half = L.half;
    a.out = DpOut{};
    a.out.res = (PairResultDev *)d_res;  // the gate runs in the DP launches
    a.out.qgate = d_gate;
    a.lds_stride = kind_lds_stride(L.gr_max, L.kind);
    ST_HIP(launch_dp(a, L.kind, L.R, true, fan ? sl.sub[k] : sl.stream));
  }
  if (fan)
    for (int k = 0; k < kSubStreams && k < (int)nl; k++) {
      ST_HIP(hipEventRecord(sl.join[k], sl.sub[k]));
      ST_HIP(hipStreamWaitEvent(sl.stream, sl.join[k], 0));
    }
  return ST_OK;
}

template <class T>
hipError_t h2d(GrowBuf<T> &d, const T *h, size_t n, hipStream_t s) {
  hipError_t e = d.reserve(n);
  if (e != hipSuccess || n == 0) return e;
  return hipMemcpyAsync(d.p(), h, n * sizeof(T), hipMemcpyHostToDevice, s);
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Score one micro-batch on the slot's streams and fill every request's output.
int run_micro_batch(Slot &sl, std::vector<std::shared_ptr<Req>> &reqs) {
  const double t_begin = now_ms();
  st_server *srv = sl.srv;
  const st_db *db = srv->db;
  const size_t nq = reqs.size();
  std::vector<uint64_t> off(nq + 1, 0);
  for (size_t q = 0; q < nq; q++) off[q + 1] = off[q] + reqs[q]->cands.size();
  const uint64_t np = off[nq];
  std::vector<Kind> qkind(nq);
  std::vector<int32_t> open(nq), ext(nq);
  uint64_t batch_cells = 0;
  for (const auto &r : reqs) batch_cells += r->cells;
  const double cap = unit_steps_cap(batch_cells);
  Plan pl;
  pl.pairs.reserve(np);
  std::vector<Unit> pending;
  const PlanEnv env = plan_env();
  for (size_t q = 0; q < nq; q++) {
    const Req &r = *reqs[q];
    open[q] = r.open;
    ext[q] = r.ext;
    qkind[q] = kind_for(r.qp, r.open, r.ext);
    uint32_t nmax = 0;
    for (uint32_t t : r.cands) nmax = std::max(nmax, db->hlen[t]);
    if (qkind[q] == kPk16 && (uint64_t)r.qp->m + nmax - 1 > kMaxDiagsPk16) qkind[q] = kI32F;
    std::vector<uint64_t> outs(r.cands.size());
    std::iota(outs.begin(), outs.end(), off[q]);
    if (!plan_query(pl, pending, (uint32_t)q, r.qp, r.cands.data(), outs.data(), r.cands.size(), qkind[q], db, cap,
                    nullptr, &env))
      return fail(ST_ERR_INVALID, "serving micro-batch: no launch shape fits a request");
  }
  finish_plan(pl, pending, qkind, open, ext, -1, srv->ctx->cus, true);
  // packed inputs: query table, gaps (open, ext, gate), units, unit pair
  // slots
  const size_t o_q = 0, o_g = align256(o_q + nq * sizeof(QueryDev)),
               o_u = align256(o_g + 3 * nq * sizeof(int32_t)),
               o_p = align256(o_u + pl.units.size() * sizeof(Unit)),
               total = align256(o_p + pl.pairs.size() * sizeof(PairDev));
  ST_HIP(sl.h_in.reserve(total));
  uint8_t *h = sl.h_in.p;
  auto *hq = (QueryDev *)(h + o_q);
  auto *hg = (int32_t *)(h + o_g);
  for (size_t q = 0; q < nq; q++) {
    const Req &r = *reqs[q];
    hq[q] = r.qp->view;
    hg[q] = r.open;
    hg[nq + q] = r.ext;
    hg[2 * nq + q] = r.gate;
  }
  std::copy(pl.units.begin(), pl.units.end(), (Unit *)(h + o_u));
  std::copy(pl.pairs.begin(), pl.pairs.end(), (PairDev *)(h + o_p));
  const double t_planned = now_ms();
  hipStream_t s = sl.stream;
  ST_HIP(sl.d_in.reserve(total));
  ST_HIP(sl.d_res.reserve(np));
  ST_HIP(sl.d_buf.reserve(pl.buf_elems));
  ST_HIP(sl.h_res.reserve(np));
  uint8_t *d = sl.d_in.p();
  const auto *dq = (const QueryDev *)(d + o_q);
  const auto *dg = (const int32_t *)(d + o_g);
  ST_HIP(hipEventRecord(sl.t_start, s));
  ST_HIP(hipMemcpyAsync(d, h, total, hipMemcpyHostToDevice, s));
  int rc = launch_on(sl, pl, dq, (const Unit *)(d + o_u), (const PairDev *)(d + o_p), sl.d_buf.p(), sl.d_res.p(),
                     dg + 2 * nq);
  if (rc) return rc;
  if (np) ST_HIP(hipMemcpyAsync(sl.h_res.p, sl.d_res.p(), np * sizeof(st_pair_result), hipMemcpyDeviceToHost, s));
  ST_HIP(hipEventRecord(sl.t_end, s));
  ST_HIP(hipEventSynchronize(sl.t_end));
  float dev_ms = 0;
  ST_HIP(hipEventElapsedTime(&dev_ms, sl.t_start, sl.t_end));
  // int16 pairs whose best reached 32767 (an int16 result at the clamp):
  // exact int32 rescoring, whose launches rewrite those pairs' rows
  Plan rs;
  std::vector<Unit> rs_pending;
  uint64_t nres = 0;
  for (size_t q = 0; q < nq; q++) {
    std::vector<uint64_t> outs;
    std::vector<uint32_t> tids;
    for (uint64_t k = off[q]; k < off[q + 1]; k++)
      if (sl.h_res.p[k].kernel == ST_KERNEL_I16 && sl.h_res.p[k].swg >= 32767) {
        outs.push_back(k);
        tids.push_back(reqs[q]->cands[k - off[q]]);
      }
    if (outs.empty()) continue;
    nres += outs.size();
    if (!plan_query(rs, rs_pending, (uint32_t)q, reqs[q]->qp, tids.data(), outs.data(), tids.size(), kI32F, db,
                    unit_steps_cap(0), nullptr, &env))
      return fail(ST_ERR_INVALID, "serving micro-batch: no int32 launch shape fits a request");
  }
  if (nres) {
    finish_plan(rs, rs_pending, qkind, open, ext, (int)kI32F, srv->ctx->cus, true);
    ST_HIP(sl.h_units.reserve(rs.units.size()));
    ST_HIP(sl.h_pairs.reserve(rs.pairs.size()));
    std::copy(rs.units.begin(), rs.units.end(), sl.h_units.p);
    std::copy(rs.pairs.begin(), rs.pairs.end(), sl.h_pairs.p);
    ST_HIP(h2d(sl.d_rs_units, sl.h_units.p, rs.units.size(), s));
    ST_HIP(h2d(sl.d_rs_pairs, sl.h_pairs.p, rs.pairs.size(), s));
    ST_HIP(sl.d_rs_buf.reserve(rs.buf_elems));
    rc = launch_on(sl, rs, dq, sl.d_rs_units.p(), sl.d_rs_pairs.p(), sl.d_rs_buf.p(), sl.d_res.p(), dg + 2 * nq);
    if (rc) return rc;
    ST_HIP(hipMemcpyAsync(sl.h_res.p, sl.d_res.p(), np * sizeof(st_pair_result), hipMemcpyDeviceToHost, s));
    ST_HIP(hipEventRecord(sl.t_end, s));
    ST_HIP(hipEventSynchronize(sl.t_end));
  }
  for (size_t q = 0; q < nq; q++)
    if (reqs[q]->out && !reqs[q]->cands.empty())
      std::memcpy(reqs[q]->out, sl.h_res.p + off[q], reqs[q]->cands.size() * sizeof(st_pair_result));
  std::lock_guard<std::mutex> lk(srv->mu);
  srv->stats.rescored += nres;
  srv->stats.plan_ms += t_planned - t_begin;
  srv->stats.device_ms += dev_ms;
  return ST_OK;
}

void worker(Slot *sl) {
  st_server *srv = sl->srv;
  (void)hipSetDevice(srv->ctx->device);
  for (;;) {
    std::vector<std::shared_ptr<Req>> take;
    uint64_t bytes = 0;
    {
      std::unique_lock<std::mutex> lk(srv->mu);
      auto &Q = srv->pe ? srv->fb_queue : srv->queue;
      for (;;) {
        if (srv->closing && Q.empty() && (!srv->pe || srv->pe->feeder_done)) return;
        // admit FIFO while the DP-state budget allows; a request larger than
        // the whole budget runs alone once nothing else is in flight
        while (!Q.empty() && take.size() < srv->cfg.max_batch_queries) {
          const uint64_t need = Q.front()->state_bytes;
          const bool alone = take.empty() && srv->inflight_bytes == 0;
          if (!alone && srv->inflight_bytes + bytes + need > srv->cfg.state_budget_bytes) break;
          take.push_back(Q.front());
          Q.pop_front();
          bytes += need;
        }
        if (!take.empty()) break;
        srv->cv_work.wait(lk);
      }
      srv->inflight_bytes += bytes;
      srv->stats.peak_inflight_bytes = std::max(srv->stats.peak_inflight_bytes, srv->inflight_bytes);
    }
    const double t0 = now_ms();
    const int rc = run_micro_batch(*sl, take);
    const std::string err = rc ? std::string(st_last_error()) : std::string();
    const double t1 = now_ms();
    {
      std::lock_guard<std::mutex> lk(srv->mu);
      srv->inflight_bytes -= bytes;
      srv->stats.batches++;
      srv->stats.queries += take.size();
      srv->stats.max_batch_queries = std::max<uint64_t>(srv->stats.max_batch_queries, take.size());
      srv->stats.busy_ms += t1 - t0;
      for (auto &r : take) {
        srv->stats.pairs += r->cands.size();
        srv->stats.cells += r->cells;
        r->t_done = t1;
        r->status = rc;
        r->err = err;
      }
    }
    srv->cv_done.notify_all();
    srv->cv_work.notify_all();
  }
}

// ---- persistent serving kernel -------------------------------------------------

// Launch the serving kernel unless it is running.  Caller holds pe.kmu.
int ensure_running(st_server *srv) {
  Persist &pe = *srv->pe;
  if (hipEventQuery(pe.kdone) == hipErrorNotReady) return ST_OK;
  ServeArgs a;
  a.ring = pe.ring.d;
  a.ring_mask = Persist::kRingUnits - 1;
  a.tail = pe.ctl.d;
  a.stop = pe.stop.d;
  a.ctl = pe.dctl.p;
  a.results = pe.results.d;
  a.remaining = pe.remaining.d;
  a.error = pe.stop.d + 4;
  a.result_rows = Persist::kResultRows;
  a.req_slots = Persist::kReqSlots;
  a.sarena = pe.arena.p;
  a.sarena_bytes = pe.arena_bytes;
  a.arena = srv->db->arena.p;
  a.toff = srv->db->off.p;
  a.tlen = srv->db->len.p;
  a.n_targets = (uint32_t)srv->db->n_targets;
  a.passbuf = pe.passbuf.p;
  a.passbuf_half = Persist::kHalfQuads;
  a.idle_ticks = 500000;  // 5 ms of the 100 MHz wall clock
  a.clock = pe.ctl.d + Persist::kClockWord;
  for (int k = 0; k < 3; k++) a.prio_ticks[k] = pe.prio_ticks[k];
  a.epoch = (uint32_t)(pe.launches % 65535) + 1;
  a.zero = 0;
  // tickets restart at the first unit the previous launch did not finish
  ST_HIP(hipMemcpyAsync(pe.dctl.p + kCtlHead * kCtlStride, pe.dctl.p + kCtlDone * kCtlStride, sizeof(uint64_t),
                        hipMemcpyDeviceToDevice, pe.kstream));
  a.lds_stride = pe.lds_stride;
  ST_HIP(launch_serve(a, pe.grid, pe.kstream));
  ST_HIP(hipEventRecord(pe.kdone, pe.kstream));
  pe.launches++;
  {
    std::lock_guard<std::mutex> lk(srv->mu);
    srv->stats.kernel_launches++;
  }
  return ST_OK;
}

void complete(st_server *srv, const std::shared_ptr<Req> &r, int rc, const std::string &err) {
  std::lock_guard<std::mutex> lk(srv->mu);
  srv->stats.queries++;
  srv->stats.pairs += r->cands.size();
  srv->stats.cells += r->cells;
  r->t_done = now_ms();
  r->status = rc;
  r->err = err;
}

// Plan one request for the serving kernel and publish its units; false when
// the kernel cannot take it (the caller routes it to the slots).
bool publish(st_server *srv, const std::shared_ptr<Req> &r, int *rc) {
  Persist &pe = *srv->pe;
  const st_db *db = srv->db;
  const st_qprof *qp = r->qp;
  *rc = ST_OK;
  const uint64_t n = r->cands.size();
  if (kind_for(qp, r->open, r->ext) != kPk16 || !qp->prof16.p || qp->m == 0) return false;
  uint32_t nmax = 0;
  for (uint32_t t : r->cands) nmax = std::max(nmax, db->hlen[t]);
  if ((uint64_t)qp->m + nmax - 1 > kMaxDiagsPk16) return false;
  Plan pl;
  std::vector<Unit> units;
  std::vector<uint64_t> outs(n);
  std::iota(outs.begin(), outs.end(), 0);
  PlanLimits lim;
  lim.only_R = 8;
  lim.max_GR = 256;
  lim.max_half = Persist::kHalfQuads;
  uint64_t load;
  {
    std::lock_guard<std::mutex> lk(srv->mu);
    load = pe.inflight_cells + r->cells;
  }
  // unit size from the cells in flight: small units when the kernel is
  // nearly idle (latency), larger ones under load (fewer claims)
  if (!plan_query(pl, units, 0, qp, r->cands.data(), outs.data(), n, kPk16, db, unit_steps_cap(load), &lim))
    return false;
  // longest units first: they set the request's latency
  std::stable_sort(units.begin(), units.end(), [](const Unit &a, const Unit &b) {
    return (uint64_t)a.passes * (a.n_pad + a.G - 1) > (uint64_t)b.passes * (b.n_pad + b.G - 1);
  });
  uint32_t ld = std::max<uint32_t>(qp->m, 8);
  for (const Unit &u : units) ld = std::max<uint32_t>(ld, u.passes * u.G * u.R);
  ld = (ld + 7) & ~7u;
  if (ld > qp->ld) return false;
  const uint64_t b16 = ((uint64_t)kCodes * ld * 2 + 255) & ~255ull, b32 = ((uint64_t)kCodes * ld * 4 + 255) & ~255ull;
  const uint64_t bp = (pl.pairs.size() * sizeof(PairDev) + 255) & ~255ull;
  const uint64_t need = b16 + b32 + bp;
  if (need > pe.arena_bytes / 2 || units.size() > Persist::kRingUnits / 2 || n > Persist::kResultRows / 2)
    return false;
  // wait for room: rows, units, a counter slot, the arena (draining the
  // kernel when the arena must restart) and the DP-state budget
  uint64_t r0 = 0;
  uint32_t slot = 0;
  {
    std::unique_lock<std::mutex> lk(srv->mu);
    for (;;) {
      if (pe.broken) {
        *rc = fail(ST_ERR_DEVICE, "serving kernel: earlier requests did not finish");
        return true;
      }
      const bool arena_ok = pe.arena_used + need <= pe.arena_bytes;
      if (!arena_ok && pe.flights.empty()) {
        // every launch-lifetime address was used: restart the kernel
        pe.draining = true;
        lk.unlock();
        {
          std::lock_guard<std::mutex> kl(pe.kmu);
          __atomic_store_n(pe.stop.p, 1u, __ATOMIC_RELEASE);
          (void)hipEventSynchronize(pe.kdone);
          __atomic_store_n(pe.stop.p, 0u, __ATOMIC_RELEASE);
        }
        lk.lock();
        pe.arena_used = 0;
        pe.draining = false;
        continue;
      }
      const uint64_t oldest = pe.unit_live.empty() ? pe.published : pe.unit_live.front();
      const bool units_ok = pe.published + units.size() - oldest <= Persist::kRingUnits;
      const bool budget_ok = pe.flights.empty() || srv->inflight_bytes + r->state_bytes <= srv->cfg.state_budget_bytes;
      if (arena_ok && units_ok && budget_ok && !pe.free_slots.empty() && pe.rows.alloc(n ? n : 1, &r0)) {
        slot = pe.free_slots.back();
        pe.free_slots.pop_back();
        srv->inflight_bytes += r->state_bytes;
        pe.inflight_cells += r->cells;
        srv->stats.peak_inflight_bytes = std::max(srv->stats.peak_inflight_bytes, srv->inflight_bytes);
        break;
      }
      srv->cv_work.wait_for(lk, std::chrono::milliseconds(1));
    }
  }
  const double t0 = now_ms();
  // the query's profiles (code-major [kCodes][ld], zero beyond row m and for
  // code 21) and pair slots (result rows r0 + k), built on the host and
  // copied in one transfer to arena addresses fresh for the running launch
  uint8_t *dst = pe.arena.p + pe.arena_used;
  pe.arena_used += need;
  int16_t *p16 = (int16_t *)dst;
  int32_t *p32 = (int32_t *)(dst + b16);
  PairDev *pp = (PairDev *)(dst + b16 + b32);
  hipError_t e = pe.h_stage.reserve(need);
  if (e == hipSuccess) {
    uint8_t *hs = pe.h_stage.p;
    std::memset(hs, 0, b16 + b32);
    int16_t *s16 = (int16_t *)hs;
    int32_t *s32 = (int32_t *)(hs + b16);
    const int32_t *rows = qp->rows.data();
    for (int c = 0; c < 21; c++)
      for (uint32_t i = 0; i < qp->m; i++) {
        const int32_t v = rows[(size_t)i * 21 + c];
        s16[(size_t)c * ld + i] = (int16_t)v;
        s32[(size_t)c * ld + i] = v;
      }
    // slot k's result row is r0 + k: a unit's rows are consecutive
    PairDev *sp = (PairDev *)(hs + b16 + b32);
    for (size_t k = 0; k < pl.pairs.size(); k++) {
      sp[k] = pl.pairs[k];
      sp[k].out = (uint32_t)(r0 + k);
    }
    e = hipMemcpyAsync(dst, hs, need, hipMemcpyHostToDevice, pe.cstream);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(pe.cstream);
  if (e != hipSuccess) {
    *rc = fail(ST_ERR_DEVICE, std::string("serving upload: ") + hipGetErrorString(e));
    std::lock_guard<std::mutex> lk(srv->mu);
    pe.rows.release(r0);
    pe.free_slots.push_back(slot);
    srv->inflight_bytes -= r->state_bytes;
    pe.inflight_cells -= r->cells;
    return true;
  }
  // unit descriptors, stamped with the device clock the poller last wrote
  const uint32_t t_pub = (uint32_t)__atomic_load_n(pe.ctl.p + Persist::kClockWord, __ATOMIC_RELAXED);
  const uint64_t u0 = pe.published;
  for (size_t i = 0; i < units.size(); i++) {
    const Unit &u = units[i];
    ServeUnit su{};
    su.prof16 = p16;
    su.prof32 = p32;
    su.pairs = pp + u.pair0;
    su.m = qp->m;
    su.ld = ld;
    su.npairs = u.npairs;
    su.G = u.G;
    su.n_pad = u.n_pad;
    su.passes = u.passes;
    su.open = r->open;
    su.ext = r->ext;
    su.t_pub = t_pub;
    su.req = slot;
    pe.ring.p[(u0 + i) & (Persist::kRingUnits - 1)] = su;
  }
  __atomic_store_n(pe.remaining.p + slot, (uint32_t)units.size(), __ATOMIC_RELAXED);
  {
    std::lock_guard<std::mutex> lk(srv->mu);
    const double tp = now_ms();
    std::vector<uint32_t> cand_of(pl.pairs.size());
    for (size_t k = 0; k < pl.pairs.size(); k++) cand_of[k] = pl.pairs[k].out;
    pe.flights.push_back(Flight{r, slot, r0, n, u0, units.size(), tp, std::move(cand_of)});
    srv->stats.published++;
    srv->stats.wait_ms += tp - r->t_submit;
    pe.unit_live.push_back(u0);
    pe.published = u0 + units.size();
    srv->stats.plan_ms += now_ms() - t0;
  }
  std::atomic_thread_fence(std::memory_order_release);
  __atomic_store_n(pe.ctl.p, u0 + units.size(), __ATOMIC_RELEASE);
  std::lock_guard<std::mutex> kl(pe.kmu);
  *rc = ensure_running(srv);
  return true;
}

void feeder(st_server *srv) {
  (void)hipSetDevice(srv->ctx->device);
  for (;;) {
    std::shared_ptr<Req> r;
    {
      std::unique_lock<std::mutex> lk(srv->mu);
      srv->cv_feed.wait(lk, [&] { return srv->closing || !srv->queue.empty(); });
      if (srv->queue.empty()) {  // closing
        srv->pe->feeder_done = true;
        lk.unlock();
        srv->cv_work.notify_all();
        return;
      }
      r = srv->queue.front();
      srv->queue.pop_front();
    }
    if (r->cands.empty()) {
      complete(srv, r, ST_OK, "");
      srv->cv_done.notify_all();
      continue;
    }
    int rc = ST_OK;
    if (!publish(srv, r, &rc)) {
      std::lock_guard<std::mutex> lk(srv->mu);
      srv->fb_queue.push_back(r);
    } else if (rc) {
      complete(srv, r, rc, st_last_error());
      srv->cv_done.notify_all();
    }
    srv->cv_work.notify_all();
  }
}

// Exact int32 rescoring of results at the int16 clamp (the batch path).
int rescore_clamped(st_server *srv, const Req &r, st_pair_result *rows) {
  std::vector<uint32_t> tids;
  std::vector<uint64_t> idx;
  for (uint64_t k = 0; k < r.cands.size(); k++)
    if (rows[k].kernel == ST_KERNEL_I16 && rows[k].swg >= 32767) {
      tids.push_back(r.cands[k]);
      idx.push_back(k);
    }
  if (tids.empty()) return ST_OK;
  const uint64_t off[2] = {0, tids.size()};
  st_batch *b = nullptr;
  int rc = st_batch_create(srv->ctx, srv->db, &r.qp, 1, off, tids.data(), &r.open, &r.ext, &r.gate, &b);
  if (rc) return rc;
  rc = st_batch_run(b);
  std::vector<st_pair_result> res(tids.size());
  if (!rc) rc = st_batch_results(b, res.data());
  st_batch_destroy(b);
  if (rc) return rc;
  for (size_t i = 0; i < idx.size(); i++) rows[idx[i]] = res[i];
  std::lock_guard<std::mutex> lk(srv->mu);
  srv->stats.rescored += idx.size();
  return ST_OK;
}

void completer(st_server *srv) {
  (void)hipSetDevice(srv->ctx->device);
  Persist &pe = *srv->pe;
  for (;;) {
    std::vector<Flight> done;
    bool busy, quit;
    {
      std::lock_guard<std::mutex> lk(srv->mu);
      for (size_t i = 0; i < pe.flights.size();) {
        if (__atomic_load_n(pe.remaining.p + pe.flights[i].slot, __ATOMIC_ACQUIRE) == 0) {
          done.push_back(std::move(pe.flights[i]));
          pe.flights[i] = std::move(pe.flights.back());
          pe.flights.pop_back();
        } else {
          i++;
        }
      }
      busy = !pe.flights.empty();
      quit = srv->closing && pe.feeder_done && pe.flights.empty() && done.empty();
      // a malformed descriptor (reported by the kernel) or a server closing
      // with requests that no longer progress fails those requests
      const bool broken = __atomic_load_n(pe.stop.p + 4, __ATOMIC_ACQUIRE) != 0;
      if (srv->closing && pe.feeder_done && busy) {
        if (pe.close_started == 0) pe.close_started = now_ms();
      }
      if (busy && (broken || (pe.close_started > 0 && now_ms() - pe.close_started > 10000.0))) {
        for (Flight &f : pe.flights) {
          f.req->t_done = now_ms();
          f.req->status = ST_ERR_DEVICE;
          f.req->err = broken ? "serving kernel: malformed work unit" : "serving kernel: requests did not finish";
        }
        pe.flights.clear();
        pe.broken = true;
        busy = false;
        quit = srv->closing && pe.feeder_done;
        srv->cv_done.notify_all();
      }
    }
    if (quit) return;
    for (const Flight &f : done) {
      Req &r = *f.req;
      for (uint64_t k = 0; k < f.n; k++) {
        const ServeRow row = pe.results.p[f.r0 + k];
        st_pair_result &o = r.out[f.cand_of[k]];
        o.swg = row.swg;
        o.ung = row.ung;
        o.best_diag = row.best_diag;
        o.gate_score = row.gate_score;
        o.pass = row.gate_score >= r.gate;
        o.kernel = ST_KERNEL_I16;
        o.reserved[0] = o.reserved[1] = 0;
      }
      const int rc = rescore_clamped(srv, r, r.out);
      {
        std::lock_guard<std::mutex> lk(srv->mu);
        pe.rows.release(f.r0);
        for (auto it = pe.unit_live.begin(); it != pe.unit_live.end(); ++it)
          if (*it == f.u0) {
            pe.unit_live.erase(it);
            break;
          }
        pe.free_slots.push_back(f.slot);
        srv->inflight_bytes -= r.state_bytes;
        pe.inflight_cells -= r.cells;
      }
      {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->stats.run_ms += now_ms() - f.t_pub;
      }
      complete(srv, f.req, rc, rc ? std::string(st_last_error()) : std::string());
    }
    if (!done.empty()) {
      srv->cv_done.notify_all();
      srv->cv_work.notify_all();
    }
    if (busy) {
      // a launch that went idle and ended while units were being published
      std::lock_guard<std::mutex> kl(pe.kmu);
      if (!pe.draining && hipEventQuery(pe.kdone) != hipErrorNotReady) (void)ensure_running(srv);
    }
    std::this_thread::sleep_for(std::chrono::microseconds(busy ? 10 : 200));
  }
}

int setup_persistent(st_server *s) {
  auto pe = std::make_unique<Persist>();
  pe->rows.cap = Persist::kResultRows;
  // ST_SERVE_PRIO_US="a,b,c": a unit's wave priority rises to 1, 2, 3 once
  // its request is a, b, c microseconds old; "0": priorities off
  if (const char *e = std::getenv("ST_SERVE_PRIO_US")) {
    unsigned v[3] = {0, 0, 0};
    if (std::sscanf(e, "%u,%u,%u", &v[0], &v[1], &v[2]) == 3 && v[0] > 0 && v[0] <= v[1] && v[1] <= v[2] &&
        v[2] < 40000000u)
      for (int k = 0; k < 3; k++) pe->prio_ticks[k] = v[k] * 100u;
    else if (std::strcmp(e, "0") == 0)
      pe->prio_ticks[0] = pe->prio_ticks[1] = pe->prio_ticks[2] = 0;
  }
  if (s->cfg.arena_mib) pe->arena_bytes = (uint64_t)s->cfg.arena_mib << 20;
  const size_t lds = (size_t)kCodes * pe->lds_stride * 2;
  const int occ = serve_occupancy(lds);
  // leave one block per CU to the other launches (rescoring, int32 requests)
  pe->grid = (uint32_t)std::max(1, occ - 1) * (uint32_t)s->ctx->cus;
  bool ok = pe->ring.alloc(Persist::kRingUnits) == hipSuccess && pe->ctl.alloc(2 * Persist::kClockWord) == hipSuccess &&
            pe->stop.alloc(8) == hipSuccess && pe->results.alloc(Persist::kResultRows) == hipSuccess &&
            pe->remaining.alloc(Persist::kReqSlots) == hipSuccess &&
            pe->dctl.alloc((size_t)kCtlWords * kCtlStride) == hipSuccess &&
            pe->passbuf.alloc((size_t)pe->grid * (2 * Persist::kHalfQuads + kEdgeQuadsPerBlock)) == hipSuccess &&
            pe->arena.alloc(pe->arena_bytes) == hipSuccess &&
            hipMemset(pe->dctl.p, 0, (size_t)kCtlWords * kCtlStride * sizeof(uint64_t)) == hipSuccess &&
            hipStreamCreateWithFlags(&pe->kstream, hipStreamNonBlocking) == hipSuccess &&
            hipStreamCreateWithFlags(&pe->cstream, hipStreamNonBlocking) == hipSuccess &&
            hipEventCreateWithFlags(&pe->kdone, hipEventDisableTiming) == hipSuccess &&
            hipDeviceSynchronize() == hipSuccess;
  if (!ok) return fail(ST_ERR_NOMEM, "st_server_create: persistent serving buffers could not be allocated");
  for (uint32_t k = 0; k < Persist::kReqSlots; k++) pe->free_slots.push_back(Persist::kReqSlots - 1 - k);
  s->pe = std::move(pe);
  return ST_OK;
}

}  // namespace
}  // namespace st

namespace {

// Wait for and release every slot's streams and events (threads joined).
void release_slots(st_server *s) {
  for (auto &sl : s->slots) {
    if (sl->stream) (void)hipStreamSynchronize(sl->stream);
    for (hipStream_t &x : sl->sub)
      if (x) {
        (void)hipStreamSynchronize(x);
        (void)hipStreamDestroy(x);
      }
    for (hipEvent_t &e : sl->join)
      if (e) (void)hipEventDestroy(e);
    if (sl->fork) (void)hipEventDestroy(sl->fork);
    if (sl->t_start) (void)hipEventDestroy(sl->t_start);
    if (sl->t_end) (void)hipEventDestroy(sl->t_end);
    if (sl->stream) (void)hipStreamDestroy(sl->stream);
  }
  s->slots.clear();
}

}  // namespace

extern "C" {

int st_server_create(st_ctx *ctx, const st_db *db, const st_server_config *cfg, st_server **out) {
  if (!ctx || !db || !out) return fail(ST_ERR_INVALID, "st_server_create: null argument");
  if (db->ctx != ctx) return fail(ST_ERR_INVALID, "st_server_create: database of another context");
  *out = nullptr;
  if (int rc = shape_table_check()) return rc;
  auto *s = new st_server();
  s->ctx = ctx;
  s->db = db;
  s->cfg.workers = 4;
  s->cfg.max_batch_queries = 64;
  s->cfg.state_budget_bytes = 4ull << 30;
  if (cfg) {
    if (cfg->workers) s->cfg.workers = cfg->workers;
    if (cfg->max_batch_queries) s->cfg.max_batch_queries = cfg->max_batch_queries;
    if (cfg->state_budget_bytes) s->cfg.state_budget_bytes = cfg->state_budget_bytes;
  }
  if (s->cfg.workers > 16) {
    delete s;
    return fail(ST_ERR_INVALID, "st_server_create: at most 16 workers");
  }
  if (hipSetDevice(ctx->device) != hipSuccess) {
    delete s;
    return fail(ST_ERR_DEVICE, "st_server_create: hipSetDevice failed");
  }
  for (uint32_t k = 0; k < s->cfg.workers; k++) {
    auto sl = std::make_unique<Slot>();
    sl->srv = s;
    sl->index = (int)k;
    bool ok = hipStreamCreateWithFlags(&sl->stream, hipStreamNonBlocking) == hipSuccess &&
              hipEventCreateWithFlags(&sl->fork, hipEventDisableTiming) == hipSuccess &&
              hipEventCreate(&sl->t_start) == hipSuccess && hipEventCreate(&sl->t_end) == hipSuccess &&
              sl->counter.alloc(kSubStreams) == hipSuccess;
    for (int k = 0; ok && k < kSubStreams; k++)
      ok = hipStreamCreateWithFlags(&sl->sub[k], hipStreamNonBlocking) == hipSuccess &&
           hipEventCreateWithFlags(&sl->join[k], hipEventDisableTiming) == hipSuccess;
    s->slots.push_back(std::move(sl));
    if (!ok) {
      release_slots(s);
      delete s;
      return fail(ST_ERR_DEVICE, "st_server_create: stream/event/counter creation failed");
    }
  }
  if (cfg && cfg->persistent) {
    s->cfg.persistent = 1;
    s->cfg.arena_mib = cfg->arena_mib;
    const int rc = setup_persistent(s);
    if (rc) {
      release_slots(s);
      delete s;
      return rc;
    }
  }
  for (auto &sl : s->slots) sl->th = std::thread(worker, sl.get());
  if (s->pe) {
    s->pe->feeder = std::thread(feeder, s);
    s->pe->completer = std::thread(completer, s);
  }
  *out = s;
  return ST_OK;
}

int st_server_submit(st_server *s, const st_request *req, uint64_t *ticket) {
  if (!s || !req || !ticket || !req->query || (req->n_cands && (!req->cand_tids || !req->out)))
    return fail(ST_ERR_INVALID, "st_server_submit: null argument");
  if (req->query->ctx != s->ctx) return fail(ST_ERR_INVALID, "st_server_submit: query of another context");
  auto r = std::make_shared<Req>();
  r->qp = req->query;
  r->open = req->gap_open;
  r->ext = req->gap_extend;
  r->gate = req->min_diag_score;
  r->out = req->out;
  r->cands.assign(req->cand_tids, req->cand_tids + req->n_cands);
  // one pass over the candidates: validation and the target residues
  uint64_t cols = 0;
  const uint32_t *hl = s->db->hlen.data();
  const uint64_t nt = s->db->n_targets;
  for (uint32_t t : r->cands) {
    if (t >= nt) return fail(ST_ERR_INVALID, "st_server_submit: target id out of range");
    cols += hl[t];
  }
  r->cells = (uint64_t)r->qp->m * cols;
  r->state_bytes = state_bytes_of(r->cands.size(), cols);
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->closing) return fail(ST_ERR_CLOSED, "st_server_submit: server is closing");
    r->ticket = s->next_ticket++;
    r->t_submit = now_ms();
    s->live[r->ticket] = r;
    s->queue.push_back(r);
  }
  if (s->pe)
    s->cv_feed.notify_one();
  else
    s->cv_work.notify_one();
  *ticket = r->ticket;
  return ST_OK;
}

int st_server_poll(st_server *s, uint64_t ticket, int *done) {
  if (!s || !done) return fail(ST_ERR_INVALID, "st_server_poll: null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->live.find(ticket);
  if (it == s->live.end()) return fail(ST_ERR_INVALID, "st_server_poll: unknown ticket");
  *done = it->second->status >= 0;
  return ST_OK;
}

int st_server_wait(st_server *s, uint64_t ticket, double timeout_ms, double *latency_ms) {
  if (!s) return fail(ST_ERR_INVALID, "st_server_wait: null server");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->live.find(ticket);
  if (it == s->live.end()) return fail(ST_ERR_INVALID, "st_server_wait: unknown ticket");
  std::shared_ptr<Req> r = it->second;
  auto ready = [&] { return r->status >= 0; };
  if (timeout_ms < 0) {
    s->cv_done.wait(lk, ready);
  } else if (!s->cv_done.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), ready)) {
    return fail(ST_ERR_TIMEOUT, "st_server_wait: timed out");
  }
  s->live.erase(ticket);
  if (latency_ms) *latency_ms = r->t_done - r->t_submit;
  if (r->status != ST_OK) return fail(r->status, r->err);
  return ST_OK;
}

int st_server_run_trace(st_server *s, const st_request *reqs, const double *arrival_ms, uint64_t n,
                        double *latency_ms) {
  if (!s || (n && (!reqs || !arrival_ms || !latency_ms)))
    return fail(ST_ERR_INVALID, "st_server_run_trace: null argument");
  for (uint64_t k = 1; k < n; k++)
    if (arrival_ms[k] < arrival_ms[k - 1])
      return fail(ST_ERR_INVALID, "st_server_run_trace: arrival times must not decrease");
  std::vector<uint64_t> tickets(n);
  const auto t0 = std::chrono::steady_clock::now();
  int rc = ST_OK;
  uint64_t submitted = 0;
  for (uint64_t k = 0; k < n; k++) {
    std::this_thread::sleep_until(t0 + std::chrono::duration_cast<std::chrono::steady_clock::duration>(
                                           std::chrono::duration<double, std::milli>(arrival_ms[k])));
    rc = st_server_submit(s, &reqs[k], &tickets[k]);
    if (rc) break;
    submitted++;
  }
  int wrc = ST_OK;
  for (uint64_t k = 0; k < submitted; k++) {
    const int r = st_server_wait(s, tickets[k], -1.0, &latency_ms[k]);
    if (r && !wrc) wrc = r;
  }
  return rc ? rc : wrc;
}

int st_server_stats_get(st_server *s, st_server_stats *out) {
  if (!s || !out) return fail(ST_ERR_INVALID, "st_server_stats_get: null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  *out = s->stats;
  return ST_OK;
}

int st_server_destroy(st_server *s) {
  if (!s) return ST_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->closing = true;
  }
  s->cv_work.notify_all();
  s->cv_feed.notify_all();
  if (s->pe) {
    if (s->pe->feeder.joinable()) s->pe->feeder.join();
    if (s->pe->completer.joinable()) s->pe->completer.join();
  }
  s->cv_work.notify_all();
  for (auto &sl : s->slots)
    if (sl->th.joinable()) sl->th.join();
  if (s->pe) {
    (void)hipSetDevice(s->ctx->device);
    __atomic_store_n(s->pe->stop.p, 1u, __ATOMIC_RELEASE);
    (void)hipStreamSynchronize(s->pe->kstream);
    s->pe.reset();
  }
  (void)hipSetDevice(s->ctx->device);
  release_slots(s);
  delete s;
  return ST_OK;
}

}  // extern "C"
