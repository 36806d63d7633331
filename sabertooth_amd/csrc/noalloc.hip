// st_score_and_gate without device allocation, for callers that cannot
// allocate (code under HIP graph capture, tight serving loops, harnesses
// that forbid hipMalloc).  It has no counterpart in the reference.
//
// Every buffer beyond the caller's inputs and outputs lives in one fixed
// device array, `scratch`, defined below: the staged jobs and records, the
// job scores and kadane_segment bounds, the gate results, the compacted
// candidates and two counters.  Its size bounds one call to
// kNoAllocMaxJobs jobs and kNoAllocMaxRecs records (the reference scores
// EV_CHUNK = 16,384 jobs per score_and_gate call, src/prefilter.rs:1381).
//
// The array's name is only a host-side stand-in: host code uses the device
// address that hipGetSymbolAddress returns, resolved once per device with
// that device current (noalloc_init, called by st_ctx_create) and kept per
// device.  Each GPU holds its own copy until the process ends.  Calls that
// share a device share its array, so they are serialised by a per-device
// lock; every kernel writes what the next one reads, so nothing needs
// clearing between calls.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/sabertooth_amd.h"
#include "host_internal.h"

namespace st {
namespace {

constexpr uint64_t kNoAllocMaxJobs = 1u << 17;
constexpr uint64_t kNoAllocMaxRecs = 1u << 17;
constexpr int kMaxDevices = 64;

constexpr uint64_t align256(uint64_t x) { return (x + 255) & ~255ull; }

// byte offsets of the scratch regions
constexpr uint64_t kOffJobs = 0;
constexpr uint64_t kOffScores = align256(kOffJobs + kNoAllocMaxJobs * sizeof(JobDev));
constexpr uint64_t kOffSegs = align256(kOffScores + kNoAllocMaxJobs * sizeof(int32_t));
constexpr uint64_t kOffRecs = align256(kOffSegs + kNoAllocMaxJobs * sizeof(int2));
constexpr uint64_t kOffGate = align256(kOffRecs + kNoAllocMaxRecs * sizeof(RecDev));
constexpr uint64_t kOffCands = align256(kOffGate + kNoAllocMaxRecs * sizeof(GateRes));
constexpr uint64_t kOffCounts = align256(kOffCands + kNoAllocMaxRecs * sizeof(st_candidate));
constexpr uint64_t kScratchBytes = align256(kOffCounts + 2 * sizeof(uint32_t));

static __device__ __align__(256) uint8_t scratch[kScratchBytes];

std::mutex g_init_mu;
void *g_scratch[kMaxDevices];       // each device's address of `scratch`
std::mutex g_dev_mu[kMaxDevices];   // one call at a time per device

constexpr int kCompactThreads = 1024;

// The passing records' candidates in record order (score_and_gate's push
// order) and the chain count; one block: thread t takes a contiguous chunk
// of records, a block scan of the chunk counts places each candidate.
// counts[0] = candidates, counts[1] = records whose decision used a chain.
__global__ __launch_bounds__(kCompactThreads) void compact_kernel(const GateRes *__restrict__ g,
                                                                  const RecDev *__restrict__ recs, uint32_t nrecs,
                                                                  st_candidate *__restrict__ out,
                                                                  uint32_t *__restrict__ counts) {
  __shared__ uint32_t s_pass[kCompactThreads], s_chain[kCompactThreads];
  const uint32_t t = threadIdx.x;
  const uint32_t per = (nrecs + kCompactThreads - 1) / kCompactThreads;
  const uint32_t a = min(t * per, nrecs), b = min(a + per, nrecs);
  uint32_t np = 0, nc = 0;
  for (uint32_t r = a; r < b; r++) {
    if (recs[r].j1 == recs[r].j0) continue;  // a record without jobs is skipped
    np += g[r].pass;
    nc += g[r].chain;
  }
  s_pass[t] = np;
  s_chain[t] = nc;
  __syncthreads();
  for (uint32_t d = 1; d < kCompactThreads; d <<= 1) {
    const uint32_t vp = t >= d ? s_pass[t - d] : 0, vc = t >= d ? s_chain[t - d] : 0;
    __syncthreads();
    s_pass[t] += vp;
    s_chain[t] += vc;
    __syncthreads();
  }
  uint32_t pos = s_pass[t] - np;
  for (uint32_t r = a; r < b; r++) {
    if (recs[r].j1 == recs[r].j0 || !g[r].pass) continue;
    st_candidate c;
    c.target_id = recs[r].tid;
    c.hits = recs[r].hits;
    c.best_diag = g[r].best_diag;
    c.diag_score = g[r].score;
    c.t_lo = g[r].t_lo;
    c.t_hi = g[r].t_hi;
    out[pos++] = c;
  }
  if (t == kCompactThreads - 1) {
    counts[0] = s_pass[t];
    counts[1] = s_chain[t];
  }
}

}  // namespace

int noalloc_init(int device) {
  if (device < 0 || device >= kMaxDevices) return fail(ST_ERR_INVALID, "st_ctx_create: device ordinal above 63");
  std::lock_guard<std::mutex> lk(g_init_mu);
  if (g_scratch[device]) return ST_OK;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return fail(ST_ERR_DEVICE, std::string("st_ctx_create: hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_scratch[device] = p;
  return ST_OK;
}

}  // namespace st

using namespace st;

extern "C" {

int st_score_and_gate_noalloc(st_ctx *ctx, const st_qprof *qp, const st_db *db, const st_job *jobs, uint64_t njobs,
                              const st_target_rec *recs, uint64_t nrecs, int32_t gap_open, int32_t gap_extend,
                              int32_t min_diag_score, int regions, st_candidate *out, uint64_t *n_out,
                              uint64_t *chains_used, int32_t *job_scores) {
  const char *fn = "st_score_and_gate_noalloc";
  if (njobs > kNoAllocMaxJobs || nrecs > kNoAllocMaxRecs)
    return fail(ST_ERR_INVALID, std::string(fn) + ": at most 131072 jobs and 131072 records per call (" +
                                    std::to_string(njobs) + " jobs, " + std::to_string(nrecs) + " records given)");
  if (int rc = score_and_gate_check(fn, ctx, qp, db, jobs, njobs, recs, nrecs, out, n_out)) return rc;
  if (njobs == 0 && nrecs == 0) return ST_OK;
  std::lock_guard<std::mutex> lk(ctx->mu);
  std::lock_guard<std::mutex> dl(g_dev_mu[ctx->device]);
  ST_HIP(hipSetDevice(ctx->device));
  uint8_t *base = (uint8_t *)g_scratch[ctx->device];
  if (!base) return fail(ST_ERR_DEVICE, std::string(fn) + ": device scratch not resolved");
  JobDev *d_jobs = (JobDev *)(base + kOffJobs);
  int32_t *d_scores = (int32_t *)(base + kOffScores);
  int2 *d_segs = (int2 *)(base + kOffSegs);
  RecDev *d_recs = (RecDev *)(base + kOffRecs);
  GateRes *d_gate = (GateRes *)(base + kOffGate);
  st_candidate *d_cands = (st_candidate *)(base + kOffCands);
  uint32_t *d_counts = (uint32_t *)(base + kOffCounts);
  hipStream_t s = ctx->stream;
  if (njobs) ST_HIP(hipMemcpyAsync(d_jobs, jobs, njobs * sizeof(JobDev), hipMemcpyHostToDevice, s));
  if (nrecs) ST_HIP(hipMemcpyAsync(d_recs, recs, nrecs * sizeof(RecDev), hipMemcpyHostToDevice, s));
  // scores and segment bounds in one scan, the gate on them, then the
  // compaction, all in the scratch
  ST_HIP(launch_ung_jobs(d_jobs, njobs, qp->qd.p, qp->view.prof16 ? qp->m : 0, db->arena.p, db->off.p, d_scores,
                         d_segs, s));
  ST_HIP(launch_gate_recs(d_recs, nrecs, d_jobs, d_scores, d_segs, gap_open, gap_extend, min_diag_score, regions,
                          d_gate, s));
  hipLaunchKernelGGL(compact_kernel, dim3(1), dim3(kCompactThreads), 0, s, d_gate, d_recs, (uint32_t)nrecs, d_cands,
                     d_counts);
  ST_HIP(hipGetLastError());
  uint32_t counts[2] = {0, 0};
  ST_HIP(hipMemcpyAsync(counts, d_counts, sizeof counts, hipMemcpyDeviceToHost, s));
  if (job_scores && njobs) ST_HIP(hipMemcpyAsync(job_scores, d_scores, njobs * 4, hipMemcpyDeviceToHost, s));
  ST_HIP(hipStreamSynchronize(s));
  if (counts[0] > nrecs) return fail(ST_ERR_DEVICE, std::string(fn) + ": candidate count exceeds the records");
  if (counts[0]) {
    ST_HIP(hipMemcpyAsync(out, d_cands, counts[0] * sizeof(st_candidate), hipMemcpyDeviceToHost, s));
    ST_HIP(hipStreamSynchronize(s));
  }
  *n_out = counts[0];
  if (chains_used) *chains_used += counts[1];
  return ST_OK;
}

}  // extern "C"
