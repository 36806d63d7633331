// Ungapped job scoring and the chain/gate rule on gfx950 (the job-list
// drop-ins; the all-pairs evidence and gate run inside the DP kernels).
//
//  * ung_jobs_kernel: kadane_score of explicit (target, diagonal, [lo, hi))
//    jobs (ungapped::score_batch, src/ungapped.rs:285-386; kadane_score,
//    src/ungapped.rs:53-62).  One lane per job; a lane streams its diagonal's
//    residues with 16-byte loads (the arena is 16-byte aligned per target).
//  * gate_recs_kernel: score_and_gate's per-target chain and gate
//    (src/prefilter.rs:1577-1657) over already scored jobs.
//
// kadane_segment scans read 16 residues per load and look scores up in the
// query's int32 profile.
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

// kadane_score of one job against profile (prof, ld); target codes >= 21
// were mapped to 21 at upload
__device__ __forceinline__ int32_t ung_job(const JobDev jb, const int32_t *__restrict__ prof, uint32_t ld,
                                           const uint8_t *__restrict__ arena, const uint64_t *__restrict__ toff) {
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
  return best;
}

__global__ __launch_bounds__(256) void ung_jobs_kernel(const JobDev *__restrict__ jobs, uint64_t njobs,
                                                       const QueryDev *__restrict__ qd,
                                                       const uint8_t *__restrict__ arena,
                                                       const uint64_t *__restrict__ toff,
                                                       int32_t *__restrict__ out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= njobs) return;
  out[i] = ung_job(jobs[i], qd->prof32, qd->ld, arena, toff);
}

// the same for jobs of several queries: job i belongs to query jq[i]
__global__ __launch_bounds__(256) void ung_jobs_multi_kernel(const JobDev *__restrict__ jobs, uint64_t njobs,
                                                             const uint32_t *__restrict__ jq,
                                                             const QueryDev *__restrict__ queries,
                                                             const uint8_t *__restrict__ arena,
                                                             const uint64_t *__restrict__ toff,
                                                             int32_t *__restrict__ out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= njobs) return;
  const QueryDev q = queries[jq[i]];
  out[i] = ung_job(jobs[i], q.prof32, q.ld, arena, toff);
}

// score_and_gate's chain and gate of one target record (up to 16 jobs)
__device__ __forceinline__ GateRes gate_rec(const RecDev rc, const JobDev *__restrict__ jobs,
                                            const int32_t *__restrict__ scores, const int32_t *prof, uint32_t ld,
                                            const uint8_t *__restrict__ arena, const uint64_t *__restrict__ toff,
                                            int32_t open, int32_t ext, int32_t min_diag_score, int regions) {
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
  if (ns == 0) return GateRes{0, 0, 0, 0, 0, 0, 0, 0, 0};
  return gate_eval(ProfGlobal{prof, ld}, t, sc, dg, lo, hi, ns, open, ext, min_diag_score, regions != 0);
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
  out[i] = gate_rec(recs[i], jobs, scores, qd->prof32, qd->ld, arena, toff, open, ext, min_diag_score, regions);
}

// the same for records of several queries: record i belongs to query rq[i],
// with that query's gap costs and gate
__global__ __launch_bounds__(128) void gate_recs_multi_kernel(
    const RecDev *__restrict__ recs, uint64_t nrecs, const JobDev *__restrict__ jobs,
    const int32_t *__restrict__ scores, const uint32_t *__restrict__ rq, const QueryDev *__restrict__ queries,
    const uint8_t *__restrict__ arena, const uint64_t *__restrict__ toff, const int32_t *__restrict__ open,
    const int32_t *__restrict__ ext, const int32_t *__restrict__ gate, int regions, GateRes *__restrict__ out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nrecs) return;
  const uint32_t q = rq[i];
  const QueryDev qd = queries[q];
  out[i] = gate_rec(recs[i], jobs, scores, qd.prof32, qd.ld, arena, toff, open[q], ext[q], gate[q], regions);
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

}  // namespace st

namespace st {

hipError_t launch_ung_jobs_multi(const JobDev *jobs, uint64_t njobs, const uint32_t *jq, const QueryDev *queries,
                                 const uint8_t *arena, const uint64_t *toff, int32_t *out, hipStream_t s) {
  if (njobs == 0) return hipSuccess;
  hipLaunchKernelGGL(ung_jobs_multi_kernel, dim3((uint32_t)((njobs + 255) / 256)), dim3(256), 0, s, jobs, njobs, jq,
                     queries, arena, toff, out);
  return hipGetLastError();
}

hipError_t launch_gate_recs_multi(const RecDev *recs, uint64_t nrecs, const JobDev *jobs, const int32_t *scores,
                                  const uint32_t *rq, const QueryDev *queries, const uint8_t *arena,
                                  const uint64_t *toff, const int32_t *open, const int32_t *ext,
                                  const int32_t *gate, int regions, GateRes *out, hipStream_t s) {
  if (nrecs == 0) return hipSuccess;
  hipLaunchKernelGGL(gate_recs_multi_kernel, dim3((uint32_t)((nrecs + 127) / 128)), dim3(128), 0, s, recs, nrecs,
                     jobs, scores, rq, queries, arena, toff, open, ext, gate, regions, out);
  return hipGetLastError();
}

}  // namespace st
