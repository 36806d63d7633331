/*
 * sabertooth_amd — MI355X (gfx950) implementation of Sabertooth's
 * post-prefilter scoring path behind a plain C ABI.
 *
 * The path: ungapped diagonal evidence (Kadane per diagonal), the gate (best
 * segment or co-linear chain of two vs min_diag_score) and the exact
 * Smith-Waterman-Gotoh local score, for plain-sequence and PSSM score rows.
 * All results are bit-identical to the reference CPU kernels.
 *
 * Conventions (identical to the reference):
 *   - residue codes 0..20 (ACDEFGHIKLMNPQRSTVWYX); target codes >= 21 score 0
 *     (src/alphabet.rs:36, src/align.rs:180-187);
 *   - score rows are int32[m][21], row i = scores of query column i against
 *     target codes 0..20 (ScoreTable, src/align.rs:104-192);
 *   - a gap of length L costs open + (L-1)*extend (GapCosts, src/align.rs:196-223).
 *
 * Ownership: every pointer argument is borrowed for the duration of the call
 * only; outputs are caller-owned buffers.  Handles are opaque.  Every entry
 * point returns ST_OK or an error code; st_last_error() gives the message of
 * the calling thread's last failure.  One context serialises its own calls
 * internally, so callers may share it between threads (as rayon workers do).
 *
 * Citations are file:line in the reference (raw-lab/sabertooth v0.4.0).
 */
#ifndef SABERTOOTH_AMD_H
#define SABERTOOTH_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status ------------------------------------------------------------ */
#define ST_OK 0
#define ST_ERR_INVALID 1   /* bad argument (null, out of range, mismatched) */
#define ST_ERR_DEVICE 2    /* a HIP runtime call failed / no usable GPU */
#define ST_ERR_NOMEM 3     /* device or host allocation failed */
#define ST_ERR_CAPACITY 4  /* admission refused: DP-state memory exhausted */
#define ST_ERR_CLOSED 5    /* the serving queue was closed */
#define ST_ERR_TIMEOUT 6   /* a bounded wait ended before the request finished */

/* Message of the calling thread's last error ("" when none). */
const char *st_last_error(void);
/* Library version string. */
const char *st_version(void);

/* Which precision produced a score (mirrors the reference's Kernel enum,
 * src/simd.rs:38-49, for the scalar_fallbacks counter, src/search.rs:532-580). */
#define ST_KERNEL_I16 0 /* packed int16 pass, exact (no saturation) */
#define ST_KERNEL_I32 1 /* exact int32 pass (saturated int16, or wide rows) */

/* ---- context ----------------------------------------------------------- */
typedef struct st_ctx st_ctx;

/* Open device `device` (HIP ordinal).  Fails with ST_ERR_DEVICE when no GPU. */
int st_ctx_create(int device, st_ctx **out);
int st_ctx_destroy(st_ctx *ctx);

/* ---- host inputs --------------------------------------------------------- */

/* Integer BLOSUM62 at bit factor 2, bias -0.2, as the reference reconstructs
 * it from MMseqs2's blosum62.out (SubstitutionMatrix::blosum62,
 * src/matrix.rs:50-175).  score: int8[21*21] row-major; p_back: f64[21]
 * (after the X rescale); r: f32[20*20] = P(a|b); any output may be NULL. */
int st_blosum62(int8_t *score, double *p_back, float *r, double *lambda);

/* ASCII -> residue code (build_aa2num, src/alphabet.rs:41-50). */
int st_aa2num(uint8_t table[256]);

/* Score rows of a plain sequence query (ScoreTable::from_scorer(SeqScorer),
 * src/align.rs:115-131): rows[i*21 + r] = score[q[i]][r]. */
int st_seq_rows(const uint8_t *query, uint64_t m, int32_t *rows);

/* ---- resident target database -------------------------------------------- */
typedef struct st_db st_db;

/* Upload the residue arena and offsets (the .stidx residues / seq_offsets
 * sections, src/index.rs:10-13; SeedIndex::seq, src/index.rs:1975-1979).
 * Target t is residues[offsets[t] .. offsets[t+1]).  The device copy is
 * repacked: each target starts on a 16-byte boundary and is padded with the
 * zero-score code 21. */
int st_db_create(st_ctx *ctx, const uint8_t *residues, uint64_t n_residues,
                 const uint64_t *offsets, uint64_t n_targets, st_db **out);
int st_db_destroy(st_db *db);
uint64_t st_db_targets(const st_db *db);
uint64_t st_db_residues(const st_db *db);

/* The target sections of a Sabertooth index file (<db>.stidx,
 * src/index.rs:1-47): numeric residues, sequence offsets and identifiers.
 * Every call maps the file and makes SeedIndex::open's checks on what it
 * reads (src/index.rs:1654-1800): magic, the header's Hasher64 checksum,
 * format version 1..2, the completion footer, the alphabet, section bounds
 * and alignment, and the residue / offset / identifier end values; the
 * k-mer sections (the prefilter's) are not read. */
typedef struct {
  uint32_t version;        /* format version */
  uint32_t alphabet;       /* 1: ACDEFGHIKLMNPQRSTVWY + X */
  uint64_t num_seqs;
  uint64_t total_residues;
  uint64_t max_seq_len;
  uint64_t id_bytes;       /* bytes of the identifier section */
} st_stidx_info;
/* Header figures; verify != 0 also checks the four target sections'
 * checksums (SeedIndex::verify, src/index.rs:1872-1880). */
int st_stidx_info_read(const char *path, int verify, st_stidx_info *info);
/* residues: total_residues bytes; offsets: num_seqs + 1 (either NULL). */
int st_stidx_read_targets(const char *path, uint8_t *residues, uint64_t *offsets);
/* Identifier bytes (id_bytes) and offsets (num_seqs + 1; identifier i is
 * bytes[offsets[i] .. offsets[i+1]), SeedIndex::id src/index.rs:2031-2035). */
int st_stidx_read_ids(const char *path, uint8_t *bytes, uint64_t *offsets);
/* st_db_create on the index's residue and offset sections, read from the
 * mapped file (the reference searches these sections in place,
 * SeedIndex::seq src/index.rs:1975-1979). */
int st_db_create_from_stidx(st_ctx *ctx, const char *path, st_db **out);

/* ---- query profile ------------------------------------------------------- */
typedef struct st_qprof st_qprof;

/* QueryProfile::from_rows (src/simd.rs:172-210): rows int32[m][21].  The
 * device holds an int16 profile when every row value fits int16, and always
 * an int32 one. */
int st_qprof_create(st_ctx *ctx, const int32_t *rows, uint64_t m,
                    st_qprof **out);
int st_qprof_destroy(st_qprof *qp);
uint64_t st_qprof_len(const st_qprof *qp);

/* ---- SWG score ------------------------------------------------------------ */

/* Exact optimal SWG local score of the query against each listed target:
 * the batched form of QueryProfile::sw_score_with (src/simd.rs:252-292) as
 * the gapped stage calls it (src/search.rs:532-580).  scores[i] for tids[i];
 * kernel (nullable) receives ST_KERNEL_I16 / ST_KERNEL_I32 per pair. */
int st_sw_score_batch(st_ctx *ctx, const st_qprof *qp, const st_db *db,
                      const uint32_t *tids, uint64_t count, int32_t gap_open,
                      int32_t gap_extend, int32_t *scores, uint8_t *kernel);

/* A score and its end cell (0-based; query_end = target_end = -1 when the
 * score is 0). */
typedef struct {
  int32_t score;
  int32_t reserved;
  int64_t query_end;
  int64_t target_end;
} st_score_end;

/* sw_score_end_with (src/simd.rs:294-326) for target ids tids[0..count):
 * the exact score and the first optimum in column-major order (smallest
 * target position, then smallest query position; src/simd.rs:977-998),
 * computed by the int32 end-tracking DP kernel. */
int st_sw_score_end_batch(st_ctx *ctx, const st_qprof *qp, const st_db *db,
                          const uint32_t *tids, uint64_t count, int32_t gap_open,
                          int32_t gap_extend, st_score_end *out);

/* The span of a local alignment: score, start and end cells (0-based,
 * inclusive; all four -1 when the score is 0; the two start cells -1 when
 * only the end is known). */
typedef struct {
  int32_t score;
  int32_t reserved;
  int64_t query_start, query_end;
  int64_t target_start, target_end;
} st_span;

/* The two score passes of align_canonical (src/align.rs:847-866) for many
 * hits at once, on the device: query q's hits are the targets
 * hit_tids[hit_off[q] .. hit_off[q + 1]), scored with (gap_open[q],
 * gap_extend[q]).  Per hit, out[k] gets the score and its end cell (the
 * first optimum in column-major order, as st_sw_score_end_batch), then the
 * start from the same rule on the reversed prefixes rows[..=query_end] and
 * target[..=target_end] (start = end - the reversed problem's end).  With a
 * negative gap cost the reversed problem can score differently; the start
 * then stays -1, as align_canonical then searches it from the end
 * (src/align.rs:859-863), which st_align_span does.  Every
 * query's reversed prefix is a window of one row-reversed copy of its
 * profile; the reversed target prefixes are gathered on the device.  With
 * the span known, st_align_span (sabertooth_host.h) runs only the banded
 * traceback. */
int st_align_spans_batch(st_ctx *ctx, const st_db *db, const st_qprof *const *queries,
                         uint64_t nq, const uint64_t *hit_off, const uint32_t *hit_tids,
                         const int32_t *gap_open, const int32_t *gap_extend, st_span *out);

/* ---- ungapped evidence and gate -------------------------------------------- */

/* One diagonal to score: target, diagonal (target_pos - query_pos) and the
 * query positions [lo, hi) (Job, src/prefilter.rs:1362-1369; Lane,
 * src/ungapped.rs:265-270). */
typedef struct {
  uint32_t tid;
  int32_t diag;
  uint32_t lo;
  uint32_t hi;
} st_job;

/* One target's jobs jobs[j0 .. j1) (TargetRec, src/prefilter.rs:1372-1377). */
typedef struct {
  uint32_t tid;
  uint32_t hits;
  uint32_t j0;
  uint32_t j1;
} st_target_rec;

/* ---- seed hits -> ungapped jobs (the prefilter's step before) -------------- */

/* An index's posting lists resident on the device, each posting decoded to
 * (target << 32) | position (SeedIndex::posting, src/index.rs:2294-2303),
 * in the index's order. */
typedef struct st_postings st_postings;
int st_postings_create(st_ctx *ctx, const uint64_t *postings, uint64_t n, st_postings **out);
/* The posting section of an index file (<db>.stidx, formats 1 and 2: fixed
 * 8-byte or packed posting_width-byte entries), decoded and uploaded. */
int st_postings_from_stidx(st_ctx *ctx, const char *path, st_postings **out);
/* The same decoded postings into a host buffer (*n = count; written when
 * cap >= *n). */
int st_stidx_read_postings(const char *path, uint64_t *postings, uint64_t cap, uint64_t *n);
int st_postings_destroy(st_postings *p);
uint64_t st_postings_count(const st_postings *p);

/* One seed of a query (Seed, src/prefilter.rs:704-708): query position and
 * its posting range [cur, end). */
typedef struct {
  uint32_t qpos;
  uint32_t reserved;
  uint64_t cur;
  uint64_t end;
} st_seed;

/* The PrefilterParams fields the seed-hit stage reads
 * (src/prefilter.rs:263-303), and the seed pattern's span. */
typedef struct {
  uint32_t max_diagonals; /* diagonals per target, >= 1 (default 4); jobs: the first 16 */
  uint32_t min_hits;      /* seed hits per target (default 1) */
  uint32_t diag_hits;     /* seed hits on the best diagonal (default 1) */
  uint32_t diag_window;   /* residues around the seed hits, 0 = whole diagonal (default 48) */
  uint32_t span;          /* SeedPattern::span */
  uint32_t reserved;
} st_hit_params;

/* The seed-hit stage of prefilter_query with segment evidence (xdrop = 0):
 * every seed's postings become hits (target, diagonal = position - qpos,
 * qpos) (src/prefilter.rs:999-1086); hits are grouped by target and
 * diagonal; a target with at least min_hits hits whose best diagonal has at
 * least diag_hits gets its top max_diagonals diagonals (hits descending,
 * diagonal ascending; top_diagonals src/prefilter.rs:1194-1225), at most 16
 * of them (src/prefilter.rs:1466), as jobs over
 * diag_range, narrowed to [qmin - diag_window, qmax + span + diag_window)
 * when diag_window > 0, and one record (src/prefilter.rs:1412-1492).
 * Targets come in ascending order, as the reference pushes them; jobs and
 * records are what st_score_and_gate takes.  *njobs / *nrecs receive the
 * sizes; when they exceed jobs_cap / recs_cap nothing is written and
 * ST_ERR_CAPACITY is returned (caps of 0 query the sizes). */
int st_seed_hits_to_jobs(st_ctx *ctx, const st_db *db, const st_postings *post, uint64_t m,
                         const st_seed *seeds, uint64_t nseeds, const st_hit_params *params,
                         st_job *jobs, uint64_t jobs_cap, uint64_t *njobs, st_target_rec *recs,
                         uint64_t recs_cap, uint64_t *nrecs);

/* st_seed_hits_to_jobs for nq queries in one call (one device sort of all
 * their hits): query q has m[q] rows and the seeds seeds[seed_off[q] ..
 * seed_off[q + 1]).  Jobs and records are concatenated in query order;
 * job_off and rec_off (nq + 1 entries each) delimit query q's slices, and a
 * record's j0 / j1 index its own query's slice, so each slice equals the
 * single-query call's output. */
int st_seed_hits_to_jobs_batch(st_ctx *ctx, const st_db *db, const st_postings *post, uint64_t nq,
                               const uint64_t *m, const uint64_t *seed_off, const st_seed *seeds,
                               const st_hit_params *params, st_job *jobs, uint64_t jobs_cap,
                               uint64_t *njobs, st_target_rec *recs, uint64_t recs_cap,
                               uint64_t *nrecs, uint64_t *job_off, uint64_t *rec_off);

/* Candidate (src/prefilter.rs:370-381). */
typedef struct {
  uint32_t target_id;
  uint32_t hits;
  int64_t best_diag;
  int32_t diag_score;
  uint32_t t_lo;
  uint32_t t_hi;
} st_candidate;

/* Exact kadane_score of every job (ungapped::score_batch,
 * src/ungapped.rs:285-386, with lanes = jobs). */
int st_ungapped_score_jobs(st_ctx *ctx, const st_qprof *qp, const st_db *db,
                           const st_job *jobs, uint64_t njobs, int32_t *scores);

/* Score the jobs, then chain and gate every target record in order, writing
 * the passing candidates in record order (score_and_gate,
 * src/prefilter.rs:1498-1660).  *n_out = number written (<= nrecs);
 * chains_used (nullable) is incremented like PrefilterStats::chains_used;
 * job_scores (nullable) receives the per-job scores. */
int st_score_and_gate(st_ctx *ctx, const st_qprof *qp, const st_db *db,
                      const st_job *jobs, uint64_t njobs,
                      const st_target_rec *recs, uint64_t nrecs,
                      int32_t gap_open, int32_t gap_extend,
                      int32_t min_diag_score, int regions, st_candidate *out,
                      uint64_t *n_out, uint64_t *chains_used,
                      int32_t *job_scores);

/* st_score_and_gate with no device allocation: no counterpart in the
 * reference.  For callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, checks and results as st_score_and_gate; the staged jobs
 * and records, scores, segment bounds, gate results and candidates live in
 * one fixed device array of the library (about 12.5 MiB per GPU, held until
 * the process ends), so a call takes at most 131072 jobs and 131072 records
 * (ST_ERR_INVALID beyond).  Calls on one GPU share that array and are
 * serialised by the library; overlapping calls from several processes do not
 * share it. */
int st_score_and_gate_noalloc(st_ctx *ctx, const st_qprof *qp, const st_db *db,
                              const st_job *jobs, uint64_t njobs,
                              const st_target_rec *recs, uint64_t nrecs,
                              int32_t gap_open, int32_t gap_extend,
                              int32_t min_diag_score, int regions, st_candidate *out,
                              uint64_t *n_out, uint64_t *chains_used,
                              int32_t *job_scores);

/* The seed-hit stage and score_and_gate chained on the device for nq
 * queries in one call (evaluate_block_kadane's jobs scored and gated
 * without leaving the device, src/prefilter.rs:1412-1492, 1498-1660): query
 * q (handle queries[q]) has the seeds seeds[seed_off[q] .. seed_off[q+1])
 * and is gated with (gap_open[q], gap_extend[q], min_diag_score[q]).  The
 * candidates that pass, in each query's target order (score_and_gate's
 * push order), go to out; cand_off (nq + 1) delimits each query's.  Equal
 * to st_seed_hits_to_jobs_batch followed by st_score_and_gate per query.
 * *n_out = the count; ST_ERR_CAPACITY when it exceeds cap (nothing written
 * but *n_out); chains_used (nullable) is accumulated by a call that succeeds,
 * so a retry with a buffer of *n_out entries counts every chain once. */
int st_seed_hits_to_candidates_batch(st_ctx *ctx, const st_db *db, const st_postings *post,
                                     uint64_t nq, const st_qprof *const *queries,
                                     const uint64_t *seed_off, const st_seed *seeds,
                                     const st_hit_params *params, const int32_t *gap_open,
                                     const int32_t *gap_extend, const int32_t *min_diag_score,
                                     int regions, st_candidate *out, uint64_t cap, uint64_t *n_out,
                                     uint64_t *cand_off, uint64_t *chains_used);

/* Device times (ms, HIP events on the context's stream) of the last
 * st_seed_hits_to_candidates_batch call that reached the scoring step on
 * these postings: ms[0] its device window (the seed upload to the pass
 * scan), ms[1] the ungapped job scan, ms[2] the gate; all 0 when none did.
 * No counterpart in the reference (a measurement hook). */
int st_seed_stage_timing(const st_postings *post, double *ms);

/* size_adjusted_gate (src/prefilter.rs:246-257). */
int32_t st_size_adjusted_gate(int32_t gate, uint64_t db_residues);

/* ---- batched (query, candidate list) scoring ------------------------------ */

/* Per-pair result of the all-pairs path: SWG score, ungapped best-diagonal
 * score over every diagonal, the smallest diagonal reaching it, the gate
 * score (best single or chain of the top-4 diagonals) and its decision. */
typedef struct {
  int32_t swg;
  int32_t ung;
  int32_t best_diag;
  int32_t gate_score;
  uint8_t pass;
  uint8_t kernel; /* ST_KERNEL_I16 / ST_KERNEL_I32 */
  uint8_t reserved[2];
} st_pair_result;

typedef struct st_batch st_batch;

/* Plan a batch: query q scores targets cand_tids[cand_off[q] .. cand_off[q+1])
 * with gaps (gap_open[q], gap_extend[q]) and gate min_diag_score[q].  All
 * inputs are copied to the device; st_batch_run then runs without any host
 * transfer.  Autotuned launch shapes are chosen here. */
int st_batch_create(st_ctx *ctx, const st_db *db, const st_qprof *const *queries,
                    uint64_t nq, const uint64_t *cand_off,
                    const uint32_t *cand_tids, const int32_t *gap_open,
                    const int32_t *gap_extend, const int32_t *min_diag_score,
                    st_batch **out);
/* Run every kernel of the batch on the device and wait for it. */
int st_batch_run(st_batch *b);
/* The planner's launch-shape cost table (device picoseconds per executed
 * cell for rows per lane 8, 12, 16 by lanes per group 1..64) is compiled in,
 * or read from the file the environment variable ST_SHAPE_TABLE names, as
 * `python -m sabertooth_amd.autotune` measures and writes it (21 numbers,
 * '#' comments; a path is read once per process).  ST_OK when the variable
 * is unset or names a valid table;
 * otherwise ST_ERR_INVALID with the reason, which st_batch_create and
 * st_server_create then return too.  No counterpart in the reference. */
int st_shape_table_status(void);
/* st_batch_run with the main DP phase (clearing the clamp list, then every
 * DP launch of the plan forked over the context's launch streams) captured
 * into a HIP graph by the first call and replayed by every later call, one
 * hipGraphLaunch per run.  The int32 rescoring of clamped pairs depends on
 * how many pairs clamped and runs eagerly after the replay.  Results and
 * st_batch_timing equal st_batch_run's; st_launch_info.ms reads 0 after it.
 * No counterpart in the reference, whose CPU kernels have no launch cost. */
int st_batch_run_graph(st_batch *b);
/* Copy results (cand_off[nq] entries, in candidate order) to the host. */
int st_batch_results(st_batch *b, st_pair_result *out);
/* Device time of the last st_batch_run, from HIP events on the batch's
 * stream: ms3[0] the main DP launches, ms3[1] the int32 rescoring of
 * saturated pairs, ms3[2] the gate; *rescored (nullable) = pairs rescored. */
int st_batch_timing(const st_batch *b, double *ms3, uint64_t *rescored);
/* Plan statistics: DP cells the kernels execute including padding (rows to
 * whole passes, columns to a unit's longest target, the skew's idle steps,
 * empty group slots), work units and kernel launches. */
int st_batch_stats(const st_batch *b, uint64_t *exec_cells, uint64_t *units, uint64_t *launches);
/* One DP kernel launch of a batch's plan: precision (0 packed int16, 1
 * int32 floored, 2 int32 unfloored, 3 and 4 the same over the query's int16
 * profile, 5 and 6 over its int8 profile: end-cell launches), rows per lane
 * (8 or 16; 0: each unit's own,
 * 8, 12 or 16 — the packed int16 launches), units, grid, useful and
 * executed cells, and its device time in the last st_batch_run (HIP events
 * around the launch on its stream). */
typedef struct {
  int32_t kind;
  int32_t rows_per_lane;
  uint64_t units;
  uint64_t grid;
  uint64_t useful_cells;
  uint64_t exec_cells;
  double ms;
} st_launch_info;
/* Writes up to cap launches; *n = the plan's launch count. */
int st_batch_launches(const st_batch *b, st_launch_info *out, uint64_t cap, uint64_t *n);
/* The DP launches of the last st_align_spans_batch call on ctx: its forward
 * launches first (*n_forward of them), then its reverse launches; up to cap
 * are written, *n = all of them.  A diagnostic with no counterpart in the
 * reference, for the executed/useful cells and device time of each pass. */
int st_align_spans_launches(st_ctx *ctx, st_launch_info *out, uint64_t cap, uint64_t *n,
                            uint64_t *n_forward);
/* Host wall-clock phases of the last st_align_spans_batch call on ctx, in
 * milliseconds: ms[0] forward planning and uploads, [1] forward device run,
 * [2] forward read-back, [3] reverse preparation (the end rows sorted, the
 * reversed target prefixes gathered, the profile views), [4] reverse
 * planning and uploads, [5] reverse device run, [6] reverse read-back and
 * the spans, [7] the whole call.  A diagnostic with no counterpart in the
 * reference. */
#define ST_SPAN_PHASES 8
int st_align_spans_phases(st_ctx *ctx, double ms[ST_SPAN_PHASES]);
/* DP cells of the batch (sum of m*n) and pairs. */
uint64_t st_batch_cells(const st_batch *b);
uint64_t st_batch_pairs(const st_batch *b);
int st_batch_destroy(st_batch *b);

/* ---- serving: resident database, continuous batching --------------------- */

/* The reference has no serving mode; its CLI scores fixed query batches
 * (src/main.rs:255, Engine::search_batch src/search.rs:692-731).  A server
 * keeps the database resident.  With `persistent`, each admitted request is
 * planned into work units that are appended to the queue of one running
 * persistent kernel (scoring and gate fused per unit), so requests merge into
 * the work already on the device.  Otherwise (and for requests the
 * persistent kernel cannot take) it runs `workers` slots, each with its own
 * stream and device buffers.  A free slot takes every queued request (FIFO,
 * at most max_batch_queries) whose DP state fits the budget beside what is
 * already running, scores them as one micro-batch and completes them;
 * requests that arrive meanwhile wait for the next free slot, and the slots'
 * launches overlap on the device.  Every request's results equal those of
 * st_batch_create / st_batch_run on the same query and candidates.  In the
 * persistent kernel a unit's wave raises its priority as its request ages
 * (1, 2, 3 from 0.6, 1.2, 2 ms after publication), so older requests win the
 * SIMDs' issue slots; the environment variable ST_SERVE_PRIO_US="a,b,c"
 * (microseconds) sets other thresholds and "0" turns it off. */
typedef struct st_server st_server;

typedef struct {
  uint32_t workers;           /* concurrent micro-batches (slots); 0 = 4, max 16 */
  uint32_t max_batch_queries; /* requests merged into one micro-batch; 0 = 64 */
  uint64_t state_budget_bytes;/* device DP state admitted at once; 0 = 4 GiB */
  uint32_t persistent;        /* 1: merge requests into one persistent serving
                                 kernel through a device work queue; requests it
                                 cannot take (int32 precision, very long
                                 targets) use the micro-batch slots */
  uint32_t arena_mib;         /* persistent: device arena for the published
                                 requests' profiles and pair slots, MiB; the
                                 kernel restarts (after draining) when it is
                                 used up; 0 = 4096 */
} st_server_config;

/* One request: `out` receives n_cands results in candidate order and must
 * stay valid (with `query`) until st_server_wait returns for its ticket. */
typedef struct {
  const st_qprof *query;
  int32_t gap_open;
  int32_t gap_extend;
  int32_t min_diag_score;
  int32_t reserved;
  const uint32_t *cand_tids;
  uint64_t n_cands;
  st_pair_result *out;
} st_request;

typedef struct {
  uint64_t batches;             /* micro-batches run */
  uint64_t queries;             /* requests completed */
  uint64_t pairs;
  uint64_t cells;               /* sum of m*n over completed requests */
  uint64_t rescored;            /* pairs rescored in int32 */
  uint64_t max_batch_queries;   /* largest micro-batch */
  uint64_t peak_inflight_bytes; /* largest admitted DP state */
  double busy_ms;               /* summed slot time spent on micro-batches */
  double plan_ms;               /* of which host planning and packing */
  double device_ms;             /* of which device time (uploads, kernels, results) */
  /* persistent serving kernel */
  uint64_t kernel_launches;     /* launches of the serving kernel */
  uint64_t published;           /* requests it took */
  double wait_ms;               /* summed submit -> publication time */
  double run_ms;                /* summed publication -> completion time */
} st_server_stats;

/* cfg may be NULL (defaults). */
int st_server_create(st_ctx *ctx, const st_db *db, const st_server_config *cfg,
                     st_server **out);
/* Queue a request (candidate ids are copied); *ticket identifies it. */
int st_server_submit(st_server *s, const st_request *req, uint64_t *ticket);
/* *done = 1 once the request's results are in its out buffer. */
int st_server_poll(st_server *s, uint64_t ticket, int *done);
/* Wait for a request (timeout_ms < 0: no limit; ST_ERR_TIMEOUT otherwise);
 * *latency_ms (nullable) = submit to completion.  Retires the ticket. */
int st_server_wait(st_server *s, uint64_t ticket, double timeout_ms, double *latency_ms);
/* Open-loop replay: submit reqs[k] at arrival_ms[k] (non-decreasing, ms
 * from the call), then wait for all; latency_ms[k] = submit to completion. */
int st_server_run_trace(st_server *s, const st_request *reqs, const double *arrival_ms,
                        uint64_t n, double *latency_ms);
int st_server_stats_get(st_server *s, st_server_stats *out);
/* Finish queued requests, stop the workers and free the server. */
int st_server_destroy(st_server *s);

#ifdef __cplusplus
}
#endif
#endif
