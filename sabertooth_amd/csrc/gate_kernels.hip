// Ungapped job scoring and the chain/gate rule on gfx950 (the job-list
// drop-ins; the all-pairs evidence and gate run inside the DP kernels).
//
//  * ung_jobs_kernel: kadane_score of explicit (target, diagonal, [lo, hi))
//    jobs (ungapped::score_batch, src/ungapped.rs:285-386; kadane_score,
//    src/ungapped.rs:53-62).  One lane per job; a lane streams its diagonal's
//    residues with 16-byte loads (the arena is 16-byte aligned per target).
//  * gate_recs_kernel: score_and_gate's per-target chain and gate
//    (src/prefilter.rs:1577-1657) over already scored jobs, on the
//    kadane_segment bounds the job scan wrote (it reads no residues).
//
// Jobs come sorted by query.  A block takes a run of them; when the run
// belongs to one query whose profile has an int16 copy of at most kStageRows
// rows, the block first copies that profile into LDS and every score lookup
// reads LDS.  Lanes of a wave look up unrelated (code, row) cells, so from
// global memory each lookup is a gather of its own cache line; from LDS it is
// a banked read.  Other runs read the int32 profile in global memory; both
// hold the same values.
#include <hip/hip_runtime.h>

#include <cstdlib>

#include "gate_core.h"

namespace st {

// Rows of a query profile a block can stage in LDS: 22 int16 rows of up to
// 1,112 entries (48,928 bytes), three blocks per compute unit.
constexpr int kStageRows = 1112;

struct ProfLds {
  const int16_t *p;
  int S;
  __device__ int32_t operator()(uint32_t code, int q) const { return p[code * S + q]; }
};

// Copy rows [0, S) of every code of qd's int16 profile into lds (S = m
// rounded up to 8; rows >= m are zero in the source).  Returns false, with
// nothing copied, when the query has no int16 profile or more than cap rows.
// Every thread of the block computes the same answer.
__device__ __forceinline__ bool stage_prof(const QueryDev qd, int16_t *lds, int cap, int &S) {
  if (!qd.prof16 || qd.m == 0 || (int)qd.m > cap) return false;
  S = (int)((qd.m + 7) & ~7u);
  const int vpr = S / 8;  // 16-byte vectors per code row
  for (int k = threadIdx.x; k < kCodes * vpr; k += blockDim.x) {
    const int c = k / vpr, v = k - c * vpr;
    *(uint4 *)(lds + c * S + v * 8) = *(const uint4 *)(qd.prof16 + (size_t)c * qd.ld + v * 8);
  }
  return true;
}

// Whether a block of the run [base, last] stages its query's profile: the
// run belongs to one query with an int16 profile of at most cap rows.
__device__ __forceinline__ bool stages(const uint32_t *qidx, const QueryDev *queries, uint64_t base, uint64_t last,
                                       int cap) {
  const uint32_t qa = qidx ? qidx[base] : 0, qb = qidx ? qidx[last] : 0;
  return qa == qb && queries[qa].prof16 && queries[qa].m > 0 && (int)queries[qa].m <= cap;
}

// kadane_score of one job; target codes >= 21 were mapped to 21 at upload.
// kSeg: also kadane_segment's bounds (src/ungapped.rs:69-88) in *seg.  The
// two scans restart alike (kadane_score's max(cur, 0) + s is s when
// cur <= 0), so one pass gives both and the same best score.
template <bool kSeg, class Acc>
__device__ __forceinline__ int32_t ung_job(const JobDev jb, const Acc &acc, const uint8_t *__restrict__ arena,
                                           const uint64_t *__restrict__ toff, int2 *seg) {
  const int lo = (int)jb.lo, hi = (int)(jb.hi > jb.lo ? jb.hi : jb.lo);
  const uint8_t *t = arena + toff[jb.tid];
  int32_t best = 0, cur = 0;
  int bs = 0, be = 0, cs = lo;
  int q = lo;
  if (q < hi) {
    // align the target pointer to 16 bytes and stream whole vectors
    const int64_t p0 = (int64_t)q + jb.diag;
    const int64_t pa = p0 & ~(int64_t)15;
    int skip = (int)(p0 - pa);
    uint4 v = *(const uint4 *)(t + pa);
    for (int64_t pos = pa; q < hi; pos += 16) {
      // the next vector's load is issued before this one's lookups; it is
      // made only when the scan reaches that vector
      const uint4 vn = q + 16 - skip < hi ? *(const uint4 *)(t + pos + 16) : v;
      const uint32_t w[4] = {v.x, v.y, v.z, v.w};
      int32_t s[16];
#pragma unroll
      for (int b = 0; b < 16; b++) {
        const int qq = q + b - skip;
        s[b] = (b >= skip && qq < hi) ? acc((w[b >> 2] >> ((b & 3) * 8)) & 0xffu, qq) : 0;
      }
      v = vn;
#pragma unroll
      for (int b = 0; b < 16; b++) {
        if (b >= skip && q < hi) {
          if (kSeg) {
            if (cur <= 0) {
              cur = s[b];
              cs = q;
            } else {
              cur += s[b];
            }
            if (cur > best) {
              best = cur;
              bs = cs;
              be = q;
            }
          } else {
            cur = (cur > 0 ? cur : 0) + s[b];
            best = best > cur ? best : cur;
          }
          q++;
        }
      }
      skip = 0;
    }
  }
  if (kSeg) *seg = make_int2(bs, be);
  return best;
}

constexpr int kJobsPerLane = 4;
constexpr int kJobBlock = 512;  // threads per job-scan block unless ST_JOB_BLOCK says otherwise

// Jobs [kThreads * kJobsPerLane * blockIdx.x, ...) of the list; job i belongs to
// query jq[i] of `queries` (jq null: every job to queries[0]).  kSeg: each
// job's kadane_segment bounds go to segs[i].
template <bool kSeg, int kThreads>
__global__ __launch_bounds__(kThreads) void ung_jobs_kernel(const JobDev *__restrict__ jobs, uint64_t njobs,
                                                       const uint32_t *__restrict__ jq,
                                                       const QueryDev *__restrict__ queries,
                                                       const uint8_t *__restrict__ arena,
                                                       const uint64_t *__restrict__ toff, int32_t *__restrict__ out,
                                                       int2 *__restrict__ segs, int cap) {
  extern __shared__ __align__(16) int16_t lds[];
  const uint64_t base = (uint64_t)blockIdx.x * (kThreads * kJobsPerLane);
  if (base >= njobs) return;
  const uint64_t last = (base + kThreads * kJobsPerLane < njobs ? base + kThreads * kJobsPerLane : njobs) - 1;
  const QueryDev qd = queries[jq ? jq[base] : 0];
  int S = 0;
  const bool staged = stages(jq, queries, base, last, cap) && stage_prof(qd, lds, cap, S);
  __syncthreads();
  for (int k = 0; k < kJobsPerLane; k++) {
    const uint64_t i = base + (uint64_t)k * kThreads + threadIdx.x;
    if (i >= njobs) break;
    if (staged) {
      out[i] = ung_job<kSeg>(jobs[i], ProfLds{lds, S}, arena, toff, segs + i);
    } else {
      const QueryDev q = queries[jq ? jq[i] : 0];
      out[i] = ung_job<kSeg>(jobs[i], ProfGlobal{q.prof32, q.ld}, arena, toff, segs + i);
    }
  }
}

// score_and_gate's chain and gate of one target record (up to 16 jobs) on
// its jobs' scores and kadane_segment bounds
__device__ __forceinline__ GateRes gate_rec(const RecDev rc, const JobDev *__restrict__ jobs,
                                            const int32_t *__restrict__ scores, const int2 *__restrict__ segs,
                                            int32_t open, int32_t ext, int32_t min_diag_score, int regions) {
  int ns = (int)(rc.j1 - rc.j0);
  if (ns > 16) ns = 16;
  if (ns == 0) return GateRes{0, 0, 0, 0, 0, 0, 0, 0, 0};
  int32_t sc[16], dg[16], lo[16], hi[16];
  for (int k = 0; k < ns; k++) {
    const JobDev jb = jobs[rc.j0 + k];
    sc[k] = scores[rc.j0 + k];
    dg[k] = jb.diag;
    lo[k] = (int32_t)jb.lo;
    hi[k] = (int32_t)(jb.hi > jb.lo ? jb.hi : jb.lo);
  }
  // the lookups are never made: every segment comes from segs
  return gate_eval(ProfGlobal{nullptr, 0}, nullptr, sc, dg, lo, hi, ns, open, ext, min_diag_score, regions != 0,
                   segs + rc.j0);
}

// Gap costs and gate: per query (arrays indexed by the record's query) or,
// with null arrays, one setting for every record.
struct GateParams {
  const int32_t *open, *ext, *gate;
  int32_t open1, ext1, gate1;
};

// One lane per record; record i belongs to query rq[i] (rq null: query 0).
__global__ __launch_bounds__(256) void gate_recs_kernel(const RecDev *__restrict__ recs, uint64_t nrecs,
                                                        const JobDev *__restrict__ jobs,
                                                        const int32_t *__restrict__ scores,
                                                        const int2 *__restrict__ segs,
                                                        const uint32_t *__restrict__ rq, GateParams gp,
                                                        int regions, GateRes *__restrict__ out,
                                                        uint32_t *__restrict__ pass, uint32_t *__restrict__ chain) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nrecs) return;
  const uint32_t q = rq ? rq[i] : 0;
  const int32_t open = gp.open ? gp.open[q] : gp.open1, ext = gp.ext ? gp.ext[q] : gp.ext1,
                gate = gp.gate ? gp.gate[q] : gp.gate1;
  const GateRes r = gate_rec(recs[i], jobs, scores, segs, open, ext, gate, regions);
  out[i] = r;
  // a record without jobs neither passes nor chains (score_and_gate skips it)
  if (pass) {
    pass[i] = r.pass;
    chain[i] = r.chain;
  }
}

// Staged rows for a launch whose longest query with an int16 profile has
// stage_m rows (0: none), and the dynamic LDS that takes.
// ST_STAGE_ROWS (8 .. kStageRows, read once) lowers the limit: less LDS per
// block, more blocks per compute unit.
static uint32_t stage_rows_limit() {
  static const uint32_t lim = [] {
    const char *e = std::getenv("ST_STAGE_ROWS");
    const long v = e ? std::strtol(e, nullptr, 10) : kStageRows;
    return (uint32_t)(v >= 8 && v <= kStageRows ? v & ~7L : kStageRows);
  }();
  return lim;
}

static int stage_cap(uint32_t stage_m) {
  const uint32_t r = (stage_m + 7) & ~7u, lim = stage_rows_limit();
  return (int)(r < lim ? r : lim);
}

// ST_JOB_BLOCK (256, 512 or 1024 threads, read once): the job scan's block.
// A block's staged profile serves all its waves, so larger blocks put more
// waves on a compute unit for the same LDS.
static int job_block() {
  static const int b = [] {
    const char *e = std::getenv("ST_JOB_BLOCK");
    const long v = e ? std::strtol(e, nullptr, 10) : kJobBlock;
    return (int)(v == 256 || v == 512 || v == 1024 ? v : kJobBlock);
  }();
  return b;
}

template <int kThreads>
static void launch_ung(const JobDev *jobs, uint64_t njobs, const uint32_t *jq, const QueryDev *queries, int cap,
                       const uint8_t *arena, const uint64_t *toff, int32_t *out, int2 *segs, hipStream_t s) {
  const dim3 grid((uint32_t)((njobs + kThreads * kJobsPerLane - 1) / (kThreads * kJobsPerLane)));
  const size_t lds = (size_t)kCodes * cap * 2;
  if (segs)
    hipLaunchKernelGGL((ung_jobs_kernel<true, kThreads>), grid, dim3(kThreads), lds, s, jobs, njobs, jq, queries,
                       arena, toff, out, segs, cap);
  else
    hipLaunchKernelGGL((ung_jobs_kernel<false, kThreads>), grid, dim3(kThreads), lds, s, jobs, njobs, jq, queries,
                       arena, toff, out, segs, cap);
}

hipError_t launch_ung_jobs_multi(const JobDev *jobs, uint64_t njobs, const uint32_t *jq, const QueryDev *queries,
                                 uint32_t stage_m, const uint8_t *arena, const uint64_t *toff, int32_t *out,
                                 int2 *segs, hipStream_t s) {
  if (njobs == 0) return hipSuccess;
  const int cap = stage_cap(stage_m);
  switch (job_block()) {
    case 1024: launch_ung<1024>(jobs, njobs, jq, queries, cap, arena, toff, out, segs, s); break;
    case 512: launch_ung<512>(jobs, njobs, jq, queries, cap, arena, toff, out, segs, s); break;
    default: launch_ung<256>(jobs, njobs, jq, queries, cap, arena, toff, out, segs, s); break;
  }
  return hipGetLastError();
}

hipError_t launch_ung_jobs(const JobDev *jobs, uint64_t njobs, const QueryDev *qd, uint32_t stage_m,
                           const uint8_t *arena, const uint64_t *toff, int32_t *out, int2 *segs, hipStream_t s) {
  return launch_ung_jobs_multi(jobs, njobs, nullptr, qd, stage_m, arena, toff, out, segs, s);
}

hipError_t launch_gate_recs_multi(const RecDev *recs, uint64_t nrecs, const JobDev *jobs, const int32_t *scores,
                                  const int2 *segs, const uint32_t *rq, const int32_t *open, const int32_t *ext,
                                  const int32_t *gate, int regions, GateRes *out, hipStream_t s, int32_t open1,
                                  int32_t ext1, int32_t gate1, uint32_t *pass, uint32_t *chain) {
  if (nrecs == 0) return hipSuccess;
  hipLaunchKernelGGL(gate_recs_kernel, dim3((uint32_t)((nrecs + 255) / 256)), dim3(256), 0, s, recs, nrecs, jobs,
                     scores, segs, rq, GateParams{open, ext, gate, open1, ext1, gate1}, regions, out, pass, chain);
  return hipGetLastError();
}

hipError_t launch_gate_recs(const RecDev *recs, uint64_t nrecs, const JobDev *jobs, const int32_t *scores,
                            const int2 *segs, int32_t open, int32_t ext, int32_t gate, int regions, GateRes *out,
                            hipStream_t s) {
  return launch_gate_recs_multi(recs, nrecs, jobs, scores, segs, nullptr, nullptr, nullptr, nullptr, regions, out, s,
                                open, ext, gate, nullptr, nullptr);
}

}  // namespace st
