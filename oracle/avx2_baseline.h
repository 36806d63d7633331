/*
 * CPU baseline — TEST/BENCH INFRASTRUCTURE ONLY (never linked by the product).
 *
 * A C++ restatement of the reference's AVX2 gapped-score cascade, timed on
 * the host cores beside the GPU: the biased-u8 Farrar striped pass
 * (StripedU8::score_avx2, src/simd.rs:418-591, 32 lanes), the signed-i16
 * striped pass with lazy F (Striped::score_avx2, src/simd.rs:715-897, 16
 * lanes) when the u8 headroom is exceeded, and the exact scalar Gotoh
 * (src/simd.rs:342-390) when i16 saturates; dispatch as sw_score_with
 * (src/simd.rs:252-292).  Profiles are built as StripedU8::build /
 * Striped::build (src/simd.rs:435-472, 730-754).
 */
#ifndef ST_AVX2_BASELINE_H
#define ST_AVX2_BASELINE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif
/* Score npairs pairs: pair k = query pair_q[k] (rows[pair_q[k]], length
 * qlen[pair_q[k]]) against target arena[off[t]..off[t]+len[t]) with
 * t = pair_t[k].  scores[k] receives the score and kernels[k] (nullable)
 * 0 = u8, 1 = i16, 2 = scalar.  Runs on `threads` OpenMP threads (queries'
 * profiles are built inside the timed call, as the reference builds one
 * QueryProfile per query).  Returns the wall seconds of the scoring. */
double avx2_score_pairs(uint32_t nq, const int32_t *const *rows, const uint32_t *qlen,
                        uint64_t npairs, const uint32_t *pair_q, const uint32_t *pair_t,
                        const uint8_t *arena, const uint64_t *off, const uint32_t *len,
                        int32_t open, int32_t ext, int threads, int32_t *scores,
                        uint8_t *kernels);

/* The ungapped evidence stage restated from the reference's AVX2 path:
 * score_and_gate (src/prefilter.rs:1498-1660) over EV_CHUNK = 16384-job
 * chunks (src/prefilter.rs:1381, 1488-1490): jobs counting-sorted by window
 * start, scored 32 lanes at a time by score_batch (src/ungapped.rs:285-386:
 * lane-major copy, 16 x 16 byte transposes, kadane_avx2's biased-u8
 * shuffle-lookup Kadane, exact recompute above the u8 limit; DiagProfile
 * src/ungapped.rs:188-255), then each target chained and gated with
 * on-demand kadane_segment bounds.  Query q has rows[q] (qlen[q] rows), jobs
 * jobs[job_off[q] .. job_off[q+1]) and records recs[rec_off[q] ..
 * rec_off[q+1]) (each record's j0/j1 index its query's jobs), gaps
 * (open[q], ext[q]) and gate[q].  Candidates of query q go to
 * cands[rec_off[q] ..], their count to n_cand[q], chains to chains[q];
 * job_scores (nullable) gets every job's score.  Queries run on `threads`
 * OpenMP threads.  Returns the wall seconds, or -1 when a query's records
 * do not index its jobs in order. */
typedef struct {
  uint32_t tid;
  int32_t diag;
  uint32_t lo, hi;
} avx2_job;
typedef struct {
  uint32_t tid, hits, j0, j1;
} avx2_rec;
typedef struct {
  uint32_t target_id, hits;
  int64_t best_diag;
  int32_t diag_score;
  uint32_t t_lo, t_hi;
} avx2_cand;
double avx2_score_and_gate(uint32_t nq, const int32_t *const *rows, const uint32_t *qlen,
                           const uint64_t *job_off, const avx2_job *jobs, const uint64_t *rec_off,
                           const avx2_rec *recs, const uint8_t *arena, const uint64_t *off,
                           const int32_t *open, const int32_t *ext, const int32_t *gate, int regions,
                           int threads, uint64_t *n_cand, avx2_cand *cands, uint64_t *chains,
                           int32_t *job_scores);
#ifdef __cplusplus
}
#endif
#endif
