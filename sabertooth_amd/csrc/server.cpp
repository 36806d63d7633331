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
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <numeric>
#include <thread>
#include <unordered_map>

#include "host_internal.h"

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
uint64_t state_bytes_of(const Req &r, const st_db *db) {
  uint64_t cols = 0;
  for (uint32_t t : r.cands) cols += db->hlen[t];
  return (uint64_t)r.cands.size() * (sizeof(PairOut) + 32 + sizeof(st_pair_result) + 2 * sizeof(PairDev) + 4) +
         cols * 16;
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
  GrowBuf<PairOut> d_pout;
  GrowBuf<uint64_t> d_top4;
  GrowBuf<st_pair_result> d_res;
  GrowBuf<uint4> d_buf, d_rs_buf;
  Pinned<Unit> h_units;
  Pinned<PairDev> h_pairs;
  Pinned<st_pair_result> h_res;
  std::thread th;
};

}  // namespace st

using namespace st;

struct st_server {
  st_ctx *ctx = nullptr;
  const st_db *db = nullptr;
  st_server_config cfg{};
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<std::shared_ptr<Req>> queue;
  std::unordered_map<uint64_t, std::shared_ptr<Req>> live;
  uint64_t next_ticket = 1;
  bool closing = false;
  uint64_t inflight_bytes = 0;
  std::vector<std::unique_ptr<Slot>> slots;
  st_server_stats stats{};
};

namespace st {
namespace {

// Launch a plan's DP kernels: one launch runs on the slot's stream, several
// fan out over the slot's sub-streams (each with its own unit counter) and
// join back before the slot stream continues.
int launch_on(Slot &sl, const Plan &pl, const QueryDev *d_q, const Unit *d_units, const PairDev *d_pairs,
              uint4 *d_buf, PairOut *d_pout, uint64_t *d_top4) {
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
    DpLaunch a;
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
    a.pout = d_pout;
    a.top4 = d_top4;
    a.qend = a.tend = nullptr;
    a.lds_stride = L.gr_max + 2;
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
    plan_query(pl, pending, (uint32_t)q, r.qp, r.cands.data(), outs.data(), r.cands.size(), qkind[q], db, cap);
  }
  finish_plan(pl, pending, qkind, open, ext, -1, srv->ctx->cus, true);
  // packed inputs: query table, gaps (open, ext, gate), units, unit pair
  // slots, candidate-order pairs, candidate query index
  const size_t o_q = 0, o_g = align256(o_q + nq * sizeof(QueryDev)),
               o_u = align256(o_g + 3 * nq * sizeof(int32_t)),
               o_p = align256(o_u + pl.units.size() * sizeof(Unit)),
               o_a = align256(o_p + pl.pairs.size() * sizeof(PairDev)), o_c = align256(o_a + np * sizeof(PairDev)),
               total = align256(o_c + np * sizeof(uint32_t));
  ST_HIP(sl.h_in.reserve(total));
  uint8_t *h = sl.h_in.p;
  auto *hq = (QueryDev *)(h + o_q);
  auto *hg = (int32_t *)(h + o_g);
  auto *ha = (PairDev *)(h + o_a);
  auto *hc = (uint32_t *)(h + o_c);
  for (size_t q = 0; q < nq; q++) {
    const Req &r = *reqs[q];
    hq[q] = r.qp->view;
    hg[q] = r.open;
    hg[nq + q] = r.ext;
    hg[2 * nq + q] = r.gate;
    for (size_t k = 0; k < r.cands.size(); k++) {
      ha[off[q] + k].tid = r.cands[k];
      ha[off[q] + k].out = (uint32_t)(off[q] + k);
      hc[off[q] + k] = (uint32_t)q;
    }
  }
  std::copy(pl.units.begin(), pl.units.end(), (Unit *)(h + o_u));
  std::copy(pl.pairs.begin(), pl.pairs.end(), (PairDev *)(h + o_p));
  const double t_planned = now_ms();
  hipStream_t s = sl.stream;
  ST_HIP(sl.d_in.reserve(total));
  ST_HIP(sl.d_pout.reserve(np));
  ST_HIP(sl.d_top4.reserve(np * 4));
  ST_HIP(sl.d_res.reserve(np));
  ST_HIP(sl.d_buf.reserve(pl.buf_elems));
  ST_HIP(sl.h_res.reserve(np));
  uint8_t *d = sl.d_in.p();
  const auto *dq = (const QueryDev *)(d + o_q);
  const auto *dg = (const int32_t *)(d + o_g);
  ST_HIP(hipEventRecord(sl.t_start, s));
  ST_HIP(hipMemcpyAsync(d, h, total, hipMemcpyHostToDevice, s));
  int rc = launch_on(sl, pl, dq, (const Unit *)(d + o_u), (const PairDev *)(d + o_p), sl.d_buf.p(), sl.d_pout.p(),
                     sl.d_top4.p());
  if (rc) return rc;
  PairGateArgs ga;
  ga.pairs = (const PairDev *)(d + o_a);
  ga.npairs = np;
  ga.pair_query = (const uint32_t *)(d + o_c);
  ga.queries = dq;
  ga.arena = db->arena.p;
  ga.toff = db->off.p;
  ga.tlen = db->len.p;
  ga.top4 = sl.d_top4.p();
  ga.open = dg;
  ga.ext = dg + nq;
  ga.gate = dg + 2 * nq;
  ga.pout = sl.d_pout.p();
  ga.res = sl.d_res.p();
  if (np) {
    ST_HIP(launch_gate_pairs(ga, s));
    ST_HIP(hipMemcpyAsync(sl.h_res.p, sl.d_res.p(), np * sizeof(st_pair_result), hipMemcpyDeviceToHost, s));
  }
  ST_HIP(hipEventRecord(sl.t_end, s));
  ST_HIP(hipEventSynchronize(sl.t_end));
  float dev_ms = 0;
  ST_HIP(hipEventElapsedTime(&dev_ms, sl.t_start, sl.t_end));
  // int16 pairs whose best reached 32767 (an int16 result at the clamp):
  // exact int32 rescoring, then the gate again over the micro-batch
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
    plan_query(rs, rs_pending, (uint32_t)q, reqs[q]->qp, tids.data(), outs.data(), tids.size(), kI32F, db,
               unit_steps_cap(0));
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
    rc = launch_on(sl, rs, dq, sl.d_rs_units.p(), sl.d_rs_pairs.p(), sl.d_rs_buf.p(), sl.d_pout.p(), sl.d_top4.p());
    if (rc) return rc;
    ST_HIP(launch_gate_pairs(ga, s));
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
      for (;;) {
        if (srv->closing && srv->queue.empty()) return;
        // admit FIFO while the DP-state budget allows; a request larger than
        // the whole budget runs alone once nothing else is in flight
        while (!srv->queue.empty() && take.size() < srv->cfg.max_batch_queries) {
          const uint64_t need = srv->queue.front()->state_bytes;
          const bool alone = take.empty() && srv->inflight_bytes == 0;
          if (!alone && srv->inflight_bytes + bytes + need > srv->cfg.state_budget_bytes) break;
          take.push_back(srv->queue.front());
          srv->queue.pop_front();
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
  for (auto &sl : s->slots) sl->th = std::thread(worker, sl.get());
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
  for (uint32_t t : r->cands) {
    if (t >= s->db->n_targets) return fail(ST_ERR_INVALID, "st_server_submit: target id out of range");
    r->cells += (uint64_t)r->qp->m * s->db->hlen[t];
  }
  r->state_bytes = state_bytes_of(*r, s->db);
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->closing) return fail(ST_ERR_CLOSED, "st_server_submit: server is closing");
    r->ticket = s->next_ticket++;
    r->t_submit = now_ms();
    s->live[r->ticket] = r;
    s->queue.push_back(r);
  }
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
  for (auto &sl : s->slots)
    if (sl->th.joinable()) sl->th.join();
  (void)hipSetDevice(s->ctx->device);
  release_slots(s);
  delete s;
  return ST_OK;
}

}  // extern "C"
