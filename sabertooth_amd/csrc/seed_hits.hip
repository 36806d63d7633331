// The seed-hit stage of prefilter_query in segment-evidence mode
// (src/prefilter.rs:999-1086, 1097-1225, 1412-1492) on the device: the
// posting lists of a query's seeds become hits, the hits are grouped by
// target and diagonal, and each target's best-supported diagonals become the
// ungapped jobs and target records that score_and_gate consumes.
//
// A hit packs (target, diagonal + m, query position) into one 64-bit key,
// most significant first, as the reference does; one device radix sort of
// the whole query's keys replaces its per-block sorts (the blocks only bound
// host memory: every block's targets are contiguous and ascending, so the
// concatenated output is the same).  The sort covers the key above the query
// position, so each (target, diagonal) run is contiguous with its positions
// unordered; the emit pass scans a chosen run for its smallest and largest
// position.  Runs and targets are found with boundary flags and scans;
// one thread per target keeps its top diagonals (hits descending, diagonal
// ascending, the order top_diagonals' insertion keeps) and emits its jobs.
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_reduce.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/sabertooth_amd.h"
#include "host_internal.h"

namespace st {
namespace {

constexpr int kMaxJobsPerTarget = 16;  // evaluate_block_kadane takes at most 16 runs

struct SeedDev {
  uint32_t qpos, reserved;
  uint64_t cur, end;
};

// Key fields, most significant first: query (several queries per call),
// target, diagonal + M (M = the longest query), query position.
struct Layout {
  uint32_t qbits, dbits, sh_t, tbits;
  uint64_t m;  // M
};

__device__ __forceinline__ uint64_t td_of(uint64_t k, uint32_t qbits) { return k >> qbits; }

// One wave per seed: its postings [cur, end) -> keys at off[s] ..
__global__ __launch_bounds__(256) void expand_kernel(const SeedDev *__restrict__ seeds, uint64_t nseeds,
                                                     const uint64_t *__restrict__ off,
                                                     const uint64_t *__restrict__ post, Layout lay,
                                                     uint64_t n_targets, uint64_t *__restrict__ keys,
                                                     uint32_t *__restrict__ bad) {
  const uint64_t s = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (s >= nseeds) return;
  const SeedDev sd = seeds[s];
  const uint64_t o = off[s];
  const uint64_t qhi = (uint64_t)sd.reserved << (lay.sh_t + lay.tbits);  // the seed's query
  for (uint64_t k = sd.cur + lane; k < sd.end; k += 64) {
    const uint64_t p = post[k];
    const uint64_t tid = p >> 32, tpos = p & 0xffffffffull;
    if (tid >= n_targets || tpos + lay.m - sd.qpos >= (1ull << lay.dbits)) {
      *bad = 1;
      keys[o + (k - sd.cur)] = 0;
      continue;
    }
    keys[o + (k - sd.cur)] = qhi | (tid << lay.sh_t) | ((tpos + lay.m - sd.qpos) << lay.qbits) | sd.qpos;
  }
}

// run_start[i]: key i begins a (target, diagonal) run
__global__ void run_flags_kernel(const uint64_t *__restrict__ keys, uint64_t n, uint32_t qbits,
                                 uint32_t *__restrict__ flag) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  flag[i] = (i == 0 || td_of(keys[i], qbits) != td_of(keys[i - 1], qbits)) ? 1u : 0u;
}

// run r's first key index; tflag[r]: run r begins a target
__global__ void run_first_kernel(const uint64_t *__restrict__ keys, uint64_t n, uint32_t qbits, uint32_t sh_t,
                                 const uint32_t *__restrict__ flag, const uint32_t *__restrict__ idx,
                                 uint32_t *__restrict__ first, uint32_t *__restrict__ tflag) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !flag[i]) return;
  const uint32_t r = idx[i];
  first[r] = (uint32_t)i;
  tflag[r] = (i == 0 || (keys[i] >> sh_t) != (keys[i - 1] >> sh_t)) ? 1u : 0u;
}

__global__ void target_first_kernel(uint64_t nruns, const uint32_t *__restrict__ tflag,
                                    const uint32_t *__restrict__ tidx, uint32_t *__restrict__ tfirst) {
  const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nruns || !tflag[r]) return;
  tfirst[tidx[r]] = (uint32_t)r;
}

// One thread per target: its top diagonals (top_diagonals, src/prefilter.rs:
// 1194-1225) and the gates min_hits / diag_hits (src/prefilter.rs:1445-1465).
__global__ void target_top_kernel(const uint64_t *__restrict__ keys, uint64_t nkeys, uint64_t ntg,
                                  uint64_t nruns, const uint32_t *__restrict__ rfirst,
                                  const uint32_t *__restrict__ tfirst, uint32_t kmax, uint32_t min_hits,
                                  uint32_t diag_hits, uint32_t *__restrict__ njobs,
                                  uint32_t *__restrict__ top_runs) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntg) return;
  const uint32_t r0 = tfirst[t];
  const uint32_t r1 = t + 1 < ntg ? tfirst[t + 1] : (uint32_t)nruns;
  const uint64_t h0 = rfirst[r0];
  const uint64_t h1 = r1 < nruns ? rfirst[r1] : nkeys;
  uint32_t nj = 0;
  if (h1 - h0 >= min_hits) {
    uint32_t cnt[kMaxJobsPerTarget], run[kMaxJobsPerTarget];
    uint32_t ntop = 0;
    for (uint32_t r = r0; r < r1; r++) {
      const uint32_t c = (uint32_t)((r + 1 < nruns ? rfirst[r + 1] : nkeys) - rfirst[r]);
      // the first held run with a smaller count (runs arrive by diagonal
      // ascending: equal counts keep the earlier diagonal first)
      uint32_t pos = ntop;
      for (uint32_t k = 0; k < ntop; k++)
        if (c > cnt[k]) {
          pos = k;
          break;
        }
      if (pos < kmax) {
        const uint32_t last = ntop < kmax ? ntop : kmax - 1;
        for (uint32_t k = last; k > pos; k--) {
          cnt[k] = cnt[k - 1];
          run[k] = run[k - 1];
        }
        cnt[pos] = c;
        run[pos] = r;
        if (ntop < kmax) ntop++;
      }
    }
    if (ntop > 0 && cnt[0] >= diag_hits) {
      nj = ntop;
      // slot-major ([slot][target]): a wave's stores of one slot are contiguous
      for (uint32_t k = 0; k < ntop; k++) top_runs[k * ntg + t] = run[k];
    }
  }
  njobs[t] = nj;
}

// The jobs and the record of each kept target (src/prefilter.rs:1466-1487).
__global__ void emit_kernel(const uint64_t *__restrict__ keys, uint64_t nkeys, uint64_t ntg, uint64_t nruns,
                            const uint32_t *__restrict__ rfirst, const uint32_t *__restrict__ tfirst,
                            const uint32_t *__restrict__ njobs, const uint32_t *__restrict__ joff,
                            const uint32_t *__restrict__ roff, const uint32_t *__restrict__ top_runs,
                            const uint32_t *__restrict__ tlen, const uint64_t *__restrict__ m_of, Layout lay,
                            uint32_t diag_window, uint32_t span, st_job *__restrict__ jobs,
                            st_target_rec *__restrict__ recs, uint32_t *__restrict__ rec_q,
                            uint32_t *__restrict__ job_q) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntg || njobs[t] == 0) return;
  const uint32_t r0 = tfirst[t];
  const uint32_t r1 = t + 1 < ntg ? tfirst[t + 1] : (uint32_t)nruns;
  const uint64_t h0 = rfirst[r0];
  const uint64_t h1 = r1 < nruns ? rfirst[r1] : nkeys;
  const uint32_t tid = (uint32_t)((keys[h0] >> lay.sh_t) & ((1ull << lay.tbits) - 1));
  const uint32_t q = (uint32_t)(lay.sh_t + lay.tbits >= 64 ? 0 : keys[h0] >> (lay.sh_t + lay.tbits));
  const int64_t n = tlen[tid], m = (int64_t)m_of[q], M = (int64_t)lay.m;
  const uint64_t qmask = (1ull << lay.qbits) - 1;
  const uint32_t j0 = joff[t];
  for (uint32_t k = 0; k < njobs[t]; k++) {
    const uint32_t r = top_runs[k * ntg + t];
    const uint64_t a = rfirst[r], b = (r + 1 < nruns ? rfirst[r + 1] : nkeys) - 1;
    const int64_t diag = (int64_t)((keys[a] >> lay.qbits) & ((1ull << lay.dbits) - 1)) - M;
    // the sort leaves positions inside a run unordered: scan the run
    int64_t qmin = (int64_t)(keys[a] & qmask), qmax = qmin;
    for (uint64_t h = a + 1; h <= b; h++) {
      const int64_t p = (int64_t)(keys[h] & qmask);
      qmin = p < qmin ? p : qmin;
      qmax = p > qmax ? p : qmax;
    }
    // diag_range (src/ungapped.rs:38-48): query positions whose target
    // position d + i lies in [0, n)
    int64_t lo = diag < 0 ? -diag : 0;
    int64_t hi = n - diag < m ? n - diag : m;
    if (hi <= lo) lo = hi = 0;
    if (diag_window > 0) {
      const int64_t wl = qmin - (int64_t)diag_window;
      lo = lo > wl ? lo : (wl > 0 ? wl : 0);
      const int64_t wh = qmax + (int64_t)span + (int64_t)diag_window;
      hi = hi < wh ? hi : wh;
    }
    st_job j;
    j.tid = tid;
    j.diag = (int32_t)diag;
    j.lo = (uint32_t)lo;
    j.hi = (uint32_t)(hi > lo ? hi : lo);
    jobs[j0 + k] = j;
    if (job_q) job_q[j0 + k] = q;
  }
  st_target_rec rc;
  rc.tid = tid;
  rc.hits = (uint32_t)(h1 - h0);
  rc.j0 = j0;
  rc.j1 = j0 + njobs[t];
  recs[roff[t]] = rc;
  rec_q[roff[t]] = q;
}

__global__ void cand_emit_kernel(const GateRes *__restrict__ g, const st_target_rec *__restrict__ recs,
                                 const uint32_t *__restrict__ rq, const uint32_t *__restrict__ pass,
                                 const uint32_t *__restrict__ idx, uint64_t n, st_candidate *__restrict__ out,
                                 uint32_t *__restrict__ cq) {
  const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n || !pass[r]) return;
  st_candidate c;
  c.target_id = recs[r].tid;
  c.hits = recs[r].hits;
  c.best_diag = g[r].best_diag;
  c.diag_score = g[r].score;
  c.t_lo = g[r].t_lo;
  c.t_hi = g[r].t_hi;
  out[idx[r]] = c;
  cq[idx[r]] = rq[r];
}

__global__ void flag_nonzero_kernel(const uint32_t *__restrict__ v, uint64_t n, uint32_t *__restrict__ f) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) f[i] = v[i] ? 1u : 0u;
}

uint32_t bits_for(uint64_t x) { return x ? 64 - __builtin_clzll(x) : 0; }

// n elements of T in a grow-only byte buffer (nullptr when it cannot grow)
template <class T>
T *grow(DevBuf<uint8_t> &b, uint64_t n) {
  const uint64_t bytes = std::max<uint64_t>(n, 1) * sizeof(T);
  if (b.n < bytes && b.alloc(bytes) != hipSuccess) return nullptr;
  return (T *)b.p;
}

unsigned grid_of(uint64_t n, unsigned b = 256) { return (unsigned)((n + b - 1) / b); }

// exclusive scan of n words into out; returns out[n-1] + in[n-1] (the total)
hipError_t scan_total(const uint32_t *in, uint32_t *out, uint64_t n, DevBuf<uint8_t> &tmp, hipStream_t s,
                      uint64_t *total) {
  size_t tb = 0;
  hipError_t e = rocprim::exclusive_scan(nullptr, tb, in, out, 0u, (size_t)n, rocprim::plus<uint32_t>(), s);
  if (e != hipSuccess) return e;
  if (tmp.n < tb && (e = tmp.alloc(tb)) != hipSuccess) return e;
  if ((e = rocprim::exclusive_scan(tmp.p, tb, in, out, 0u, (size_t)n, rocprim::plus<uint32_t>(), s)) != hipSuccess)
    return e;
  uint32_t a = 0, b = 0;
  if ((e = hipMemcpyAsync(&a, out + n - 1, 4, hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
  if ((e = hipMemcpyAsync(&b, in + n - 1, 4, hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
  *total = (uint64_t)a + b;
  return hipSuccess;
}

}  // namespace
}  // namespace st

using namespace st;

extern "C" {

int st_postings_create(st_ctx *ctx, const uint64_t *postings, uint64_t n, st_postings **out) {
  if (!ctx || !out || (n && !postings)) return fail(ST_ERR_INVALID, "st_postings_create: null argument");
  *out = nullptr;
  std::lock_guard<std::mutex> lk(ctx->mu);
  ST_HIP(hipSetDevice(ctx->device));
  auto *p = new st_postings();
  p->ctx = ctx;
  p->n = n;
  const hipError_t e = n ? p->post.upload(postings, n, ctx->stream) : p->post.alloc(1);
  if (e != hipSuccess) {
    delete p;
    return fail(ST_ERR_NOMEM, "st_postings_create: device allocation failed");
  }
  for (hipEvent_t &ev : p->ev)
    if (hipEventCreate(&ev) != hipSuccess) {
      delete p;
      return fail(ST_ERR_DEVICE, "st_postings_create: event creation failed");
    }
  *out = p;
  return ST_OK;
}

int st_postings_destroy(st_postings *p) {
  delete p;
  return ST_OK;
}

uint64_t st_postings_count(const st_postings *p) { return p ? p->n : 0; }

// The seed-hit stage over nq queries, left on the device: query q's seeds
// are seeds[seed_off[q] .. seed_off[q+1]) against its m[q] rows.  Jobs,
// records (j0 / j1 index all jobs) and each record's query stay in the
// postings' scratch, in (query, target) order.
struct SeedStage {
  uint64_t njobs = 0, nrecs = 0;
  st_job *jobs = nullptr;
  st_target_rec *recs = nullptr;
  uint32_t *rq = nullptr;
  uint32_t *jq = nullptr;  // each job's query (when asked for)
};

static int seed_stage_check(const std::string &f, st_ctx *ctx, const st_db *db, const st_postings *post, uint64_t nq,
                            const uint64_t *m, const uint64_t *seed_off, const st_seed *seeds,
                            const st_hit_params *hp) {
  if (!ctx || !db || !post || !hp || (nq && (!m || !seed_off))) return fail(ST_ERR_INVALID, f + ": null argument");
  if (db->ctx != ctx || post->ctx != ctx) return fail(ST_ERR_INVALID, f + ": handle of another context");
  if (hp->max_diagonals < 1) return fail(ST_ERR_INVALID, f + ": max_diagonals must be at least 1");
  if (nq > 0xffffffffull) return fail(ST_ERR_INVALID, f + ": more than 2^32 queries");
  if (nq && seed_off[0] != 0) return fail(ST_ERR_INVALID, f + ": seed_off[0] must be 0");
  if (nq && seed_off[nq] && !seeds) return fail(ST_ERR_INVALID, f + ": null seeds");
  return ST_OK;
}

// ctx->mu held by the caller; arguments checked by seed_stage_check
static int seed_stage(const std::string &f, st_ctx *ctx, const st_db *db, const st_postings *post, uint64_t nq,
                      const uint64_t *m, const uint64_t *seed_off, const st_seed *seeds, const st_hit_params *hp,
                      SeedStage &out, bool want_jq = false) {
  out = SeedStage();
  if (nq == 0 || db->n_targets == 0) return ST_OK;
  const uint64_t nseeds = seed_off[nq];
  // hits per seed -> key offsets; each seed learns its query
  std::vector<SeedDev> sv(nseeds);
  std::vector<uint64_t> off(nseeds + 1, 0);
  uint64_t M = 0;
  for (uint64_t q = 0; q < nq; q++) {
    if (seed_off[q + 1] < seed_off[q]) return fail(ST_ERR_INVALID, f + ": seed_off must not decrease");
    if (seed_off[q + 1] > seed_off[q]) M = std::max(M, m[q]);
    for (uint64_t k = seed_off[q]; k < seed_off[q + 1]; k++) {
      const st_seed &sd = seeds[k];
      if (sd.cur > sd.end || sd.end > post->n || sd.qpos >= m[q])
        return fail(ST_ERR_INVALID, f + ": seed outside the postings or the query");
      sv[k] = SeedDev{sd.qpos, (uint32_t)q, sd.cur, sd.end};
      off[k + 1] = off[k] + (sd.end - sd.cur);
    }
  }
  const uint64_t nh = off[nseeds];
  if (nh == 0) return ST_OK;
  // the sort, scans and keys index items in 32 bits
  if (nh > 0x7fffffffull) return fail(ST_ERR_INVALID, f + ": more than 2^31 - 1 seed hits");
  const uint32_t maxlen = db->max_len;
  Layout lay;
  lay.m = M;
  lay.qbits = bits_for(M - 1);
  lay.dbits = bits_for(M + maxlen);
  lay.sh_t = lay.qbits + lay.dbits;
  lay.tbits = bits_for(db->n_targets - 1);
  const uint32_t end_bit = lay.sh_t + lay.tbits + bits_for(nq - 1);
  if (end_bit > 64) return fail(ST_ERR_INVALID, f + ": queries, targets and database too large for a 64-bit hit key");
  ST_HIP(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  DevBuf<uint8_t> *sc = post->scratch;
  DevBuf<uint8_t> &tmp = sc[18];
  SeedDev *d_seeds = grow<SeedDev>(sc[0], nseeds);
  uint64_t *d_off = grow<uint64_t>(sc[1], nseeds + nq);
  uint64_t *k0 = grow<uint64_t>(sc[2], nh), *k1 = grow<uint64_t>(sc[3], nh);
  uint32_t *flag = grow<uint32_t>(sc[4], nh), *idx = grow<uint32_t>(sc[5], nh), *bad = grow<uint32_t>(sc[6], 1);
  if (!d_seeds || !d_off || !k0 || !k1 || !flag || !idx || !bad)
    return fail(ST_ERR_NOMEM, f + ": device allocation failed");
  uint64_t *d_m = d_off + nseeds;  // the queries' lengths after the key offsets
  ST_HIP(hipMemcpyAsync(d_seeds, sv.data(), nseeds * sizeof(SeedDev), hipMemcpyHostToDevice, s));
  ST_HIP(hipMemcpyAsync(d_off, off.data(), nseeds * sizeof(uint64_t), hipMemcpyHostToDevice, s));
  ST_HIP(hipMemcpyAsync(d_m, m, nq * sizeof(uint64_t), hipMemcpyHostToDevice, s));
  ST_HIP(clear_u32(bad, 1, s));
  hipLaunchKernelGGL(expand_kernel, dim3((unsigned)((nseeds + 3) / 4)), dim3(256), 0, s, d_seeds, nseeds, d_off,
                     post->post.p, lay, db->n_targets, k0, bad);
  ST_HIP(hipGetLastError());
  // the flag is read after the run scan's synchronisation: a bad hit's key
  // is 0, which the sort and the run passes only move
  uint32_t h_bad = 0;
  ST_HIP(hipMemcpyAsync(&h_bad, bad, 4, hipMemcpyDeviceToHost, s));
  // sort on (query, target, diagonal) only: runs come out contiguous, and
  // the query positions below them need no radix passes (emit takes each
  // chosen run's smallest and largest position by scanning it)
  size_t tb = 0;
  const int begin_bit = (int)lay.qbits;
  ST_HIP(rocprim::radix_sort_keys(nullptr, tb, k0, k1, (uint32_t)nh, (unsigned)begin_bit, (unsigned)end_bit, s));
  if (!grow<uint8_t>(tmp, tb)) return fail(ST_ERR_NOMEM, f + ": device allocation failed");
  ST_HIP(rocprim::radix_sort_keys(tmp.p, tb, k0, k1, (uint32_t)nh, (unsigned)begin_bit, (unsigned)end_bit, s));
  const uint64_t *keys = k1;
  // runs
  hipLaunchKernelGGL(run_flags_kernel, dim3(grid_of(nh)), dim3(256), 0, s, keys, nh, lay.qbits, flag);
  ST_HIP(hipGetLastError());
  uint64_t nruns = 0;
  ST_HIP(scan_total(flag, idx, nh, tmp, s, &nruns));
  if (h_bad) return fail(ST_ERR_INVALID, f + ": a posting names a target outside the database");
  uint32_t *rfirst = grow<uint32_t>(sc[7], nruns), *tflag = grow<uint32_t>(sc[8], nruns),
           *tidx = grow<uint32_t>(sc[9], nruns);
  if (!rfirst || !tflag || !tidx) return fail(ST_ERR_NOMEM, f + ": device allocation failed");
  hipLaunchKernelGGL(run_first_kernel, dim3(grid_of(nh)), dim3(256), 0, s, keys, nh, lay.qbits, lay.sh_t, flag,
                     idx, rfirst, tflag);
  ST_HIP(hipGetLastError());
  // (query, target) groups
  uint64_t ntg = 0;
  ST_HIP(scan_total(tflag, tidx, nruns, tmp, s, &ntg));
  // a target's jobs are the first 16 of its top max_diagonals diagonals
  // (src/prefilter.rs:1466); the top list's order is total (hits descending,
  // then arrival), so its first 16 entries are the top-16 list itself
  const uint32_t kmax = std::min<uint32_t>(hp->max_diagonals, (uint32_t)kMaxJobsPerTarget);
  uint32_t *tfirst = grow<uint32_t>(sc[10], ntg), *nj = grow<uint32_t>(sc[11], ntg),
           *joff = grow<uint32_t>(sc[12], ntg), *tkeep = grow<uint32_t>(sc[13], ntg),
           *roff = grow<uint32_t>(sc[14], ntg), *top_runs = grow<uint32_t>(sc[15], ntg * kmax);
  if (!tfirst || !nj || !joff || !tkeep || !roff || !top_runs) return fail(ST_ERR_NOMEM, f + ": device allocation failed");
  hipLaunchKernelGGL(target_first_kernel, dim3(grid_of(nruns)), dim3(256), 0, s, nruns, tflag, tidx, tfirst);
  ST_HIP(hipGetLastError());
  hipLaunchKernelGGL(target_top_kernel, dim3(grid_of(ntg)), dim3(256), 0, s, keys, nh, ntg, nruns, rfirst, tfirst,
                     kmax, hp->min_hits, hp->diag_hits, nj, top_runs);
  ST_HIP(hipGetLastError());
  uint64_t tot_jobs = 0, tot_recs = 0;
  ST_HIP(scan_total(nj, joff, ntg, tmp, s, &tot_jobs));
  hipLaunchKernelGGL(flag_nonzero_kernel, dim3(grid_of(ntg)), dim3(256), 0, s, nj, ntg, tkeep);
  ST_HIP(hipGetLastError());
  ST_HIP(scan_total(tkeep, roff, ntg, tmp, s, &tot_recs));
  if (tot_recs == 0) return ST_OK;
  st_job *d_jobs = grow<st_job>(sc[16], tot_jobs);
  st_target_rec *d_recs = grow<st_target_rec>(sc[17], tot_recs);
  uint32_t *d_rq = grow<uint32_t>(sc[19], tot_recs);
  uint32_t *d_jq = want_jq ? grow<uint32_t>(sc[20], tot_jobs) : nullptr;
  if (!d_jobs || !d_recs || !d_rq || (want_jq && !d_jq)) return fail(ST_ERR_NOMEM, f + ": device allocation failed");
  hipLaunchKernelGGL(emit_kernel, dim3(grid_of(ntg)), dim3(256), 0, s, keys, nh, ntg, nruns, rfirst, tfirst, nj,
                     joff, roff, top_runs, db->len.p, d_m, lay, hp->diag_window, hp->span, d_jobs, d_recs, d_rq, d_jq);
  ST_HIP(hipGetLastError());
  out.njobs = tot_jobs;
  out.nrecs = tot_recs;
  out.jobs = d_jobs;
  out.recs = d_recs;
  out.rq = d_rq;
  out.jq = d_jq;
  return ST_OK;
}

// The stage's jobs and records copied out in query order; job_off /
// rec_off (nq + 1) delimit each query's slices, and a record's j0 / j1
// index its query's jobs.
static int seed_hits_impl(const char *fn, st_ctx *ctx, const st_db *db, const st_postings *post, uint64_t nq,
                          const uint64_t *m, const uint64_t *seed_off, const st_seed *seeds,
                          const st_hit_params *hp, st_job *jobs, uint64_t jobs_cap, uint64_t *njobs,
                          st_target_rec *recs, uint64_t recs_cap, uint64_t *nrecs, uint64_t *job_off,
                          uint64_t *rec_off) {
  const std::string f(fn);
  if (!njobs || !nrecs) return fail(ST_ERR_INVALID, f + ": null argument");
  *njobs = *nrecs = 0;
  if (job_off) std::fill(job_off, job_off + nq + 1, 0);
  if (rec_off) std::fill(rec_off, rec_off + nq + 1, 0);
  if (int rc = seed_stage_check(f, ctx, db, post, nq, m, seed_off, seeds, hp)) return rc;
  std::lock_guard<std::mutex> lk(ctx->mu);
  SeedStage st;
  if (int rc = seed_stage(f, ctx, db, post, nq, m, seed_off, seeds, hp, st)) return rc;
  *njobs = st.njobs;
  *nrecs = st.nrecs;
  if (st.njobs > jobs_cap || st.nrecs > recs_cap || (st.njobs && !jobs) || (st.nrecs && !recs))
    return fail(ST_ERR_CAPACITY, f + ": output buffers too small (*njobs, *nrecs hold the sizes)");
  if (st.nrecs == 0) return ST_OK;
  hipStream_t s = ctx->stream;
  std::vector<uint32_t> rq(st.nrecs);
  ST_HIP(hipMemcpyAsync(jobs, st.jobs, st.njobs * sizeof(st_job), hipMemcpyDeviceToHost, s));
  ST_HIP(hipMemcpyAsync(recs, st.recs, st.nrecs * sizeof(st_target_rec), hipMemcpyDeviceToHost, s));
  ST_HIP(hipMemcpyAsync(rq.data(), st.rq, st.nrecs * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  ST_HIP(hipStreamSynchronize(s));
  // per-query slices; each record's jobs indexed within its query's slice
  std::vector<uint64_t> ro(nq + 1, 0), jo(nq + 1, 0);
  for (uint64_t k = 0; k < st.nrecs; k++) {
    ro[rq[k] + 1]++;
    jo[rq[k] + 1] += recs[k].j1 - recs[k].j0;
  }
  for (uint64_t q = 0; q < nq; q++) {
    ro[q + 1] += ro[q];
    jo[q + 1] += jo[q];
  }
  for (uint64_t k = 0; k < st.nrecs; k++) {
    recs[k].j0 -= (uint32_t)jo[rq[k]];
    recs[k].j1 -= (uint32_t)jo[rq[k]];
  }
  if (job_off) std::copy(jo.begin(), jo.end(), job_off);
  if (rec_off) std::copy(ro.begin(), ro.end(), rec_off);
  return ST_OK;
}

int st_seed_hits_to_candidates_batch(st_ctx *ctx, const st_db *db, const st_postings *post, uint64_t nq,
                                     const st_qprof *const *queries, const uint64_t *seed_off,
                                     const st_seed *seeds, const st_hit_params *hp, const int32_t *gap_open,
                                     const int32_t *gap_extend, const int32_t *min_diag_score, int regions,
                                     st_candidate *out, uint64_t cap, uint64_t *n_out, uint64_t *cand_off,
                                     uint64_t *chains_used) {
  const std::string f("st_seed_hits_to_candidates_batch");
  if (!n_out || (nq && (!queries || !gap_open || !gap_extend || !min_diag_score || !cand_off)))
    return fail(ST_ERR_INVALID, f + ": null argument");
  *n_out = 0;
  if (cand_off) std::fill(cand_off, cand_off + nq + 1, 0);
  std::vector<uint64_t> m(nq);
  std::vector<QueryDev> qv(nq);
  uint32_t stage_m = 0;  // the longest query with an int16 profile
  for (uint64_t q = 0; q < nq; q++) {
    if (!queries[q] || queries[q]->ctx != ctx) return fail(ST_ERR_INVALID, f + ": bad query handle");
    m[q] = queries[q]->m;
    qv[q] = queries[q]->view;
    if (qv[q].prof16) stage_m = std::max(stage_m, qv[q].m);
  }
  if (int rc = seed_stage_check(f, ctx, db, post, nq, m.data(), seed_off, seeds, hp)) return rc;
  std::lock_guard<std::mutex> lk(ctx->mu);
  post->timed = false;
  ST_HIP(hipSetDevice(ctx->device));
  ST_HIP(hipEventRecord(post->ev[0], ctx->stream));
  SeedStage st;
  if (int rc = seed_stage(f, ctx, db, post, nq, m.data(), seed_off, seeds, hp, st, true)) return rc;
  if (st.nrecs == 0) return ST_OK;
  hipStream_t s = ctx->stream;
  DevBuf<uint8_t> *sc = post->scratch;
  const uint32_t *jq = st.jq;
  int32_t *scores = grow<int32_t>(sc[21], st.njobs);
  GateRes *gres = grow<GateRes>(sc[22], st.nrecs);
  uint32_t *pass = grow<uint32_t>(sc[23], st.nrecs), *pidx = grow<uint32_t>(sc[24], st.nrecs);
  uint32_t *chain = grow<uint32_t>(sc[25], st.nrecs), *nchain_d = grow<uint32_t>(sc[26], 1);
  st_candidate *d_c = grow<st_candidate>(sc[27], st.nrecs);
  uint32_t *d_cq = grow<uint32_t>(sc[28], st.nrecs);
  QueryDev *d_q = grow<QueryDev>(sc[29], nq);
  int32_t *prm = grow<int32_t>(sc[30], 3 * nq);  // open, extend, gate per query
  int2 *segs = grow<int2>(sc[31], st.njobs);
  if (!jq || !scores || !gres || !pass || !pidx || !chain || !nchain_d || !d_c || !d_cq || !d_q || !prm || !segs)
    return fail(ST_ERR_NOMEM, f + ": device allocation failed");
  std::vector<int32_t> hprm(3 * nq);
  std::copy(gap_open, gap_open + nq, hprm.begin());
  std::copy(gap_extend, gap_extend + nq, hprm.begin() + nq);
  std::copy(min_diag_score, min_diag_score + nq, hprm.begin() + 2 * nq);
  ST_HIP(hipMemcpyAsync(d_q, qv.data(), nq * sizeof(QueryDev), hipMemcpyHostToDevice, s));
  ST_HIP(hipMemcpyAsync(prm, hprm.data(), 3 * nq * 4, hipMemcpyHostToDevice, s));
  // score every job against its query, then gate every record; the gate
  // writes the pass and chain flags the compaction scans
  ST_HIP(hipEventRecord(post->ev[1], s));
  ST_HIP(launch_ung_jobs_multi((const JobDev *)st.jobs, st.njobs, jq, d_q, stage_m, db->arena.p, db->off.p, scores,
                               segs, s));
  ST_HIP(hipEventRecord(post->ev[2], s));
  ST_HIP(launch_gate_recs_multi((const RecDev *)st.recs, st.nrecs, (const JobDev *)st.jobs, scores, segs, st.rq, prm,
                                prm + nq, prm + 2 * nq, regions, gres, s, 0, 0, 0, pass, chain));
  ST_HIP(hipEventRecord(post->ev[3], s));
  // the passing records, compacted in (query, target) order
  // the chain count is a sum; its copy completes with the pass scan's
  // synchronisation
  DevBuf<uint8_t> &tmp = sc[18];
  size_t tb = 0;
  ST_HIP(rocprim::reduce(nullptr, tb, chain, nchain_d, 0u, (size_t)st.nrecs, rocprim::plus<uint32_t>(), s));
  if (!grow<uint8_t>(tmp, tb)) return fail(ST_ERR_NOMEM, f + ": device allocation failed");
  ST_HIP(rocprim::reduce(tmp.p, tb, chain, nchain_d, 0u, (size_t)st.nrecs, rocprim::plus<uint32_t>(), s));
  uint32_t nchain = 0;
  ST_HIP(hipMemcpyAsync(&nchain, nchain_d, 4, hipMemcpyDeviceToHost, s));
  uint64_t nc = 0;
  ST_HIP(scan_total(pass, pidx, st.nrecs, tmp, s, &nc));
  ST_HIP(hipEventRecord(post->ev[4], s));
  post->timed = true;
  *n_out = nc;
  if (nc > cap || (nc && !out))
    return fail(ST_ERR_CAPACITY, f + ": output buffer too small (*n_out holds the size)");
  // counted only by a call that delivers its candidates: a caller that
  // retries with a larger buffer counts each chain once
  if (chains_used) *chains_used += nchain;
  if (nc == 0) return ST_OK;
  hipLaunchKernelGGL(cand_emit_kernel, dim3(grid_of(st.nrecs)), dim3(256), 0, s, gres, st.recs, st.rq, pass, pidx,
                     st.nrecs, d_c, d_cq);
  ST_HIP(hipGetLastError());
  std::vector<uint32_t> cq(nc);
  ST_HIP(hipMemcpyAsync(out, d_c, nc * sizeof(st_candidate), hipMemcpyDeviceToHost, s));
  ST_HIP(hipMemcpyAsync(cq.data(), d_cq, nc * 4, hipMemcpyDeviceToHost, s));
  ST_HIP(hipStreamSynchronize(s));
  for (uint32_t q : cq) cand_off[q + 1]++;
  for (uint64_t q = 0; q < nq; q++) cand_off[q + 1] += cand_off[q];
  return ST_OK;
}

int st_seed_stage_timing(const st_postings *post, double *ms) {
  if (!post || !ms) return fail(ST_ERR_INVALID, "st_seed_stage_timing: null argument");
  ms[0] = ms[1] = ms[2] = 0.0;
  std::lock_guard<std::mutex> lk(post->ctx->mu);
  if (!post->timed) return ST_OK;
  ST_HIP(hipSetDevice(post->ctx->device));
  ST_HIP(hipEventSynchronize(post->ev[4]));
  float a = 0, b = 0, c = 0;
  ST_HIP(hipEventElapsedTime(&a, post->ev[0], post->ev[4]));
  ST_HIP(hipEventElapsedTime(&b, post->ev[1], post->ev[2]));
  ST_HIP(hipEventElapsedTime(&c, post->ev[2], post->ev[3]));
  ms[0] = a;
  ms[1] = b;
  ms[2] = c;
  return ST_OK;
}

int st_seed_hits_to_jobs(st_ctx *ctx, const st_db *db, const st_postings *post, uint64_t m,
                         const st_seed *seeds, uint64_t nseeds, const st_hit_params *hp, st_job *jobs,
                         uint64_t jobs_cap, uint64_t *njobs, st_target_rec *recs, uint64_t recs_cap,
                         uint64_t *nrecs) {
  if (m == 0) {
    if (!njobs || !nrecs) return fail(ST_ERR_INVALID, "st_seed_hits_to_jobs: null argument");
    *njobs = *nrecs = 0;
    return ST_OK;
  }
  const uint64_t so[2] = {0, nseeds};
  return seed_hits_impl("st_seed_hits_to_jobs", ctx, db, post, 1, &m, so, seeds, hp, jobs, jobs_cap, njobs, recs,
                        recs_cap, nrecs, nullptr, nullptr);
}

int st_seed_hits_to_jobs_batch(st_ctx *ctx, const st_db *db, const st_postings *post, uint64_t nq,
                               const uint64_t *m, const uint64_t *seed_off, const st_seed *seeds,
                               const st_hit_params *hp, st_job *jobs, uint64_t jobs_cap, uint64_t *njobs,
                               st_target_rec *recs, uint64_t recs_cap, uint64_t *nrecs, uint64_t *job_off,
                               uint64_t *rec_off) {
  if (nq && (!job_off || !rec_off)) return fail(ST_ERR_INVALID, "st_seed_hits_to_jobs_batch: null offsets");
  return seed_hits_impl("st_seed_hits_to_jobs_batch", ctx, db, post, nq, m, seed_off, seeds, hp, jobs, jobs_cap,
                        njobs, recs, recs_cap, nrecs, job_off, rec_off);
}

}  // extern "C"
