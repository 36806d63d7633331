// Ungapped job scoring, the chain/gate rule and the all-diagonal reduction on
// gfx950.
//
//  * ung_jobs_kernel: kadane_score of explicit (target, diagonal, [lo, hi))
//    jobs (ungapped::score_batch, src/ungapped.rs:285-386; kadane_score,
//    src/ungapped.rs:53-62).  One lane per job; a lane streams its diagonal's
//    residues with 16-byte loads (the arena is 16-byte aligned per target).
//  * gate_recs_kernel: score_and_gate's per-target chain and gate
//    (src/prefilter.rs:1577-1657) over already scored jobs.
//  * gate_pairs_kernel: the seedless evidence of a (query, candidate) pair
//    from its 4 best diagonals (keys left by the DP kernel): best diagonal
//    and the same chain/gate rule over their whole ranges.
//
// kadane_segment scans read 16 residues per load and look scores up either
// in an LDS copy of the query's int16 profile (when a block's pairs share one
// query that fits) or in the int32 profile in HBM.
#include <hip/hip_runtime.h>

#include "gate_core.h"

namespace st {

struct JobDev {
  uint32_t tid;
  int32_t diag;
  uint32_t lo, hi;
};
struct RecDev {
  uint32_t tid, hits, j0, j1;
};

// kadane_score of each job; codes >= 21 were mapped to 21 at upload
__global__ __launch_bounds__(256) void ung_jobs_kernel(const JobDev *__restrict__ jobs, uint64_t njobs,
                                                       const QueryDev *__restrict__ qd,
                                                       const uint8_t *__restrict__ arena,
                                                       const uint64_t *__restrict__ toff,
                                                       int32_t *__restrict__ out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= njobs) return;
  const JobDev jb = jobs[i];
  const int32_t *prof = qd->prof32;
  const uint32_t ld = qd->ld;
  const int lo = (int)jb.lo, hi = (int)(jb.hi > jb.lo ? jb.hi : jb.lo);
  const uint8_t *t = arena + toff[jb.tid];
  int32_t best = 0, cur = 0;
  int q = lo;
  if (q < hi) {
    // align the target pointer to 16 bytes and stream whole vectors
    const int64_t p0 = (int64_t)q + jb.diag;
    const int64_t pa = p0 & ~(int64_t)15;
    int skip = (int)(p0 - pa);
    for (int64_t pos = pa; q < hi; pos += 16) {
      const uint4 v = *(const uint4 *)(t + pos);
      const uint32_t w[4] = {v.x, v.y, v.z, v.w};
      int32_t s[16];
#pragma unroll
      for (int b = 0; b < 16; b++) {
        const int qq = q + b - skip;
        s[b] = (b >= skip && qq < hi) ? prof[(uint64_t)((w[b >> 2] >> ((b & 3) * 8)) & 0xffu) * ld + qq] : 0;
      }
#pragma unroll
      for (int b = 0; b < 16; b++) {
        if (b >= skip && q < hi) {
          cur = (cur > 0 ? cur : 0) + s[b];
          best = best > cur ? best : cur;
          q++;
        }
      }
      skip = 0;
    }
  }
  out[i] = best;
}

__global__ __launch_bounds__(128) void gate_recs_kernel(const RecDev *__restrict__ recs, uint64_t nrecs,
                                                        const JobDev *__restrict__ jobs,
                                                        const int32_t *__restrict__ scores,
                                                        const QueryDev *__restrict__ qd,
                                                        const uint8_t *__restrict__ arena,
                                                        const uint64_t *__restrict__ toff,
                                                        int32_t open, int32_t ext, int32_t min_diag_score,
                                                        int regions, GateRes *__restrict__ out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nrecs) return;
  const RecDev rc = recs[i];
  int ns = (int)(rc.j1 - rc.j0);
  if (ns > 16) ns = 16;
  int32_t sc[16], dg[16], lo[16], hi[16];
  for (int k = 0; k < ns; k++) {
    const JobDev jb = jobs[rc.j0 + k];
    sc[k] = scores[rc.j0 + k];
    dg[k] = jb.diag;
    lo[k] = (int32_t)jb.lo;
    hi[k] = (int32_t)(jb.hi > jb.lo ? jb.hi : jb.lo);
  }
  const uint8_t *t = arena + toff[rc.tid];
  GateRes r;
  if (ns == 0) {
    r = GateRes{0, 0, 0, 0, 0, 0, 0, 0, 0};
  } else {
    r = gate_eval(ProfGlobal{qd->prof32, qd->ld}, t, sc, dg, lo, hi, ns, open, ext, min_diag_score,
                  regions != 0);
  }
  out[i] = r;
}

// All-diagonal evidence of each pair: the DP kernel leaves the 4 best
// diagonals as sorted keys (score desc, diagonal asc); the best one is the
// pair's ungapped score and best diagonal, and the gate runs over all four.
struct PairGateArgs {
  const PairDev *pairs;
  uint64_t npairs;
  const uint32_t *pair_query;  // query index of each pair slot
  const QueryDev *queries;
  const uint8_t *arena;
  const uint64_t *toff;
  const uint32_t *tlen;
  const uint64_t *top4;              // 4 keys per result row
  const int32_t *open, *ext, *gate;  // per query
  PairOut *pout;
  void *res;  // st_pair_result rows (by PairDev::out)
};

// query rows held in LDS by gate_pairs_kernel
constexpr int kGateLdsRows = 512;

__global__ __launch_bounds__(64) void gate_pairs_kernel(PairGateArgs a) {
  __shared__ int16_t lp[kCodes * kGateLdsRows];
  const uint64_t base = (uint64_t)blockIdx.x * 64;
  const uint64_t i = base + threadIdx.x;
  const uint64_t last = base + 63 < a.npairs ? base + 63 : a.npairs - 1;
  const uint32_t q0 = a.pair_query[base];
  const QueryDev Q = a.queries[q0];
  // one query for the whole block, short enough, with an int16 profile
  const bool uni = a.pair_query[last] == q0 && Q.prof16 != nullptr && Q.m <= (uint32_t)kGateLdsRows;
  if (uni) {
    const int m = (int)Q.m;
    for (int idx = threadIdx.x; idx < kCodes * m; idx += 64) {
      const int c = idx / m, r = idx - c * m;
      lp[c * m + r] = Q.prof16[(uint64_t)c * Q.ld + r];
    }
  }
  __syncthreads();
  if (i >= a.npairs) return;
  const PairDev pd = a.pairs[i];
  const uint32_t qi = a.pair_query[i];
  const QueryDev q = uni ? Q : a.queries[qi];
  const int m = (int)q.m;
  const int n = (int)a.tlen[pd.tid];
  int32_t ts[4], td[4], lo[4], hi[4];
  int nt = 0;
  for (int k = 0; k < 4; k++) {
    const uint64_t key = a.top4[(uint64_t)pd.out * 4 + k];
    if (key == 0) break;
    ts[k] = (int32_t)(key >> 32);
    td[k] = (int32_t)(0xFFFFFFFFu - (uint32_t)key) - (m - 1);
    // whole diagonal (diag_range, src/ungapped.rs:38-46)
    const int d = td[k];
    const int a0 = -d > 0 ? -d : 0;
    const int b0 = m < n - d ? m : n - d;
    lo[k] = a0;
    hi[k] = b0 > a0 ? b0 : a0;
    nt++;
  }
  const uint8_t *t = a.arena + a.toff[pd.tid];
  const GateRes r = uni ? gate_eval(ProfLds{lp, m}, t, ts, td, lo, hi, nt, a.open[qi], a.ext[qi], INT32_MIN, false)
                        : gate_eval(ProfGlobal{q.prof32, q.ld}, t, ts, td, lo, hi, nt, a.open[qi], a.ext[qi],
                                    INT32_MIN, false);
  const int32_t g = a.gate[qi];
  PairResultDev o;
  const PairOut po = a.pout[pd.out];
  o.swg = po.swg;
  o.ung = nt ? ts[0] : 0;
  o.best_diag = nt ? td[0] : 0;
  o.gate_score = nt ? r.score : 0;
  o.pass = (nt ? r.score : 0) >= g;
  o.kernel = po.flag == 2 ? 1 : 0;  // ST_KERNEL_I32 when the int32 pass produced swg
  o.r0 = o.r1 = 0;
  ((PairResultDev *)a.res)[pd.out] = o;
}

hipError_t launch_ung_jobs(const JobDev *jobs, uint64_t njobs, const QueryDev *qd, const uint8_t *arena,
                           const uint64_t *toff, int32_t *out, hipStream_t s) {
  if (njobs == 0) return hipSuccess;
  const uint64_t blocks = (njobs + 255) / 256;
  hipLaunchKernelGGL(ung_jobs_kernel, dim3((uint32_t)blocks), dim3(256), 0, s, jobs, njobs, qd, arena, toff, out);
  return hipGetLastError();
}

hipError_t launch_gate_recs(const RecDev *recs, uint64_t nrecs, const JobDev *jobs, const int32_t *scores,
                            const QueryDev *qd, const uint8_t *arena, const uint64_t *toff, int32_t open,
                            int32_t ext, int32_t gate, int regions, GateRes *out, hipStream_t s) {
  if (nrecs == 0) return hipSuccess;
  const uint64_t blocks = (nrecs + 127) / 128;
  hipLaunchKernelGGL(gate_recs_kernel, dim3((uint32_t)blocks), dim3(128), 0, s, recs, nrecs, jobs, scores, qd,
                     arena, toff, open, ext, gate, regions, out);
  return hipGetLastError();
}

hipError_t launch_gate_pairs(const PairGateArgs &a, hipStream_t s) {
  if (a.npairs == 0) return hipSuccess;
  const uint64_t blocks = (a.npairs + 63) / 64;
  hipLaunchKernelGGL(gate_pairs_kernel, dim3((uint32_t)blocks), dim3(64), 0, s, a);
  return hipGetLastError();
}

}  // namespace st
