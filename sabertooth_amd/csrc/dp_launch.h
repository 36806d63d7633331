// Arguments of one DP kernel launch, shared by the kernels and the host code
// that plans and issues them.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "st_device.h"

namespace st {

// Where a DP launch puts its per-pair outputs.
struct DpOut {
  PairOut *pout;         // score-only launches: score and precision flag per result row
  PairResultDev *res;    // evidence launches: the final result row (the gate runs in the kernel)
  const int32_t *qgate;  // evidence launches: min_diag_score per query
  uint32_t *sat;         // packed int16 pairs at the clamp: sat[0] = count, sat[1 + k] = their result rows
  uint32_t sat_cap;      // entries after sat[0]
  int32_t *qend, *tend;  // score + end launches: end cell per result row
  // window units (end-cell launches of floored kernels): the score each
  // pair is known to reach, per result row; its pass stops once the first
  // column holding it is complete (null: no early stop)
  const int32_t *stop;
};

struct DpLaunch {
  const Unit *units;
  uint32_t nunits;
  uint32_t *counter;
  const PairDev *pairs;
  const QueryDev *queries;
  const uint32_t *pair_q;  // window units (end-cell kernels): each pair slot's profile view
  const uint8_t *arena;
  const uint64_t *toff;
  const uint32_t *tlen;
  int32_t open, ext;
  uint4 *passbuf;
  uint64_t passbuf_half;
  DpOut out;
  uint32_t grid;
  int lds_stride;  // >= largest G*R of the launch + 2, even
};
// Dynamic LDS of a DP launch: the pass's profile slice (kCodes rows of
// lds_stride entries), and for the evidence kernels at least the gate's
// scratch (GateLds, 6 KB), which reuses it after a unit's passes.
inline size_t dp_lds_bytes(int lds_stride, size_t entry_bytes, bool ung) {
  const size_t prof = (size_t)kCodes * (size_t)lds_stride * entry_bytes;
  return ung && prof < 6144 ? 6144 : prof;
}

// end: the int32 score + end-cell instances (no evidence)
hipError_t launch_dp(const DpLaunch &a, int kind, int R, bool ung, hipStream_t s, bool end = false);
// n words at p set to 0 by a kernel (not a memset: safe under graph capture)
hipError_t clear_u32(uint32_t *p, uint32_t n, hipStream_t s);

// Reverse score pass inputs (span_kernels.hip).  A query's profile (int8,
// int16 or int32 entries: esize 1, 2 or 4) with its m rows reversed into the bytes
// dst + dst_off ([kCodes][dst_ld] entries, dst_ld >= m + 1024); one launch
// for many queries: chunk_job / chunk_first map each 256-row chunk of a
// mirror to its job.
struct MirrorJob {
  const void *src;
  uint64_t dst_off;
  uint32_t src_ld, m, dst_ld, esize;
};
hipError_t launch_mirror_prof(const MirrorJob *jobs, const uint32_t *chunk_job, const uint32_t *chunk_first,
                              uint32_t nchunks, uint8_t *dst, hipStream_t s);
// Target tid's first len residues, reversed.
struct RevPrefix {
  uint32_t tid;
  uint32_t len;
};
// n reversed prefixes into dst at doff[0 .. n] (database arena layout).
hipError_t launch_rev_prefix(const uint8_t *arena, const uint64_t *toff, const RevPrefix *pre, uint32_t n,
                             const uint64_t *doff, uint8_t *dst, hipStream_t s);

// Profiles of simulated pairs (sim_kernels.hip): sample s's query codes
// q[s*m .. s*m + m) and (nullable) bias give its [kCodes][ld] int16 and int32
// profiles at p16 / p32 + s * kCodes * ld; mat is the kCodes x kCodes score
// table (code 21's row and column zero).
hipError_t launch_sim_prof(const uint8_t *q, const int32_t *bias, const int8_t *mat, uint32_t m, uint32_t ld,
                           uint32_t samples, int16_t *p16, int32_t *p32, hipStream_t s);

}  // namespace st
