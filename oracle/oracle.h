/*
 * CPU oracle for the post-prefilter scoring path — TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference's scalar algorithms.  It is the
 * checker of the HIP product path: only tests/, the smoke check and the
 * bench's cpu_baseline leg may load it.  Nothing in the product library links
 * or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative
 * to the raw-lab/sabertooth v0.4.0 source tree).
 *
 * Conventions (as in the reference): query length m, target length n, score
 * rows are int32[m][21] (column 20 = target X), target codes >= 21 score 0.
 */
#ifndef ST_ORACLE_H
#define ST_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_NPROF 21 /* alphabet size incl. X (src/alphabet.rs:20) */
#define OR_NAA 20   /* real amino acids (src/alphabet.rs:17) */

/* ---- alphabet / matrix ------------------------------------------------ */

/* ASCII -> residue code table (src/alphabet.rs:36-50). */
void or_aa2num(uint8_t table[256]);

/* BLOSUM62 reconstructed with bit factor 2, bias -0.2
 * (src/matrix.rs:50-175 over src/data/blosum62.out).  Any output may be NULL. */
void or_blosum62(int8_t score[OR_NPROF * OR_NPROF], double p_back[OR_NPROF],
                 float r[OR_NAA * OR_NAA], double *lambda);

/* Rows of a plain sequence query: rows[i][r] = score[q[i]][r]
 * (ScoreTable::from_scorer over SeqScorer, src/align.rs:115-131, 61-79). */
void or_seq_rows(const uint8_t *q, size_t m, int32_t *rows);

/* ---- Smith-Waterman-Gotoh ---------------------------------------------- */

/* Exact i32 Gotoh with O(n) memory and the column-major-first end cell
 * (scalar_gotoh_score_end, src/simd.rs:347-390).  Returns the score; *qend /
 * *tend receive the end cell, or -1 when the best score is 0. */
int32_t or_gotoh_score_end(const int32_t *rows, size_t m, const uint8_t *t,
                           size_t n, int32_t open, int32_t ext, int64_t *qend,
                           int64_t *tend);

/* Full three-matrix textbook reference (src/simd.rs:1124-1142); also returns
 * the first optimum in column-major order (src/simd.rs:1167-1195). */
int32_t or_gotoh_textbook(const int32_t *rows, size_t m, const uint8_t *t,
                          size_t n, int32_t open, int32_t ext, int64_t *qend,
                          int64_t *tend);

/* ---- ungapped evidence ------------------------------------------------- */

/* diag_range (src/ungapped.rs:38-46). */
void or_diag_range(size_t m, size_t n, int64_t d, size_t *lo, size_t *hi);
/* kadane_score (src/ungapped.rs:53-62). */
int32_t or_kadane_score(const int32_t *rows, const uint8_t *t, int64_t d,
                        size_t lo, size_t hi);
/* kadane_segment (src/ungapped.rs:69-88): returns score, writes the inclusive
 * query range of the first maximum. */
int32_t or_kadane_segment(const int32_t *rows, const uint8_t *t, int64_t d,
                          size_t lo, size_t hi, size_t *qs, size_t *qe);

/* ---- gate -------------------------------------------------------------- */

/* Job / TargetRec / Candidate (src/prefilter.rs:1362-1377, 370-381). */
typedef struct {
  uint32_t tid;
  int32_t diag;
  uint32_t lo, hi;
} or_job;
typedef struct {
  uint32_t tid, hits, j0, j1;
} or_target_rec;
typedef struct {
  uint32_t target_id, hits;
  int64_t best_diag;
  int32_t diag_score;
  uint32_t t_lo, t_hi;
} or_candidate;

/* Score the jobs (kadane_score each, src/ungapped.rs:285-386 semantics), then
 * chain / gate / emit candidates per target in record order
 * (score_and_gate, src/prefilter.rs:1498-1660).  `arena`/`offsets` hold the
 * targets (offsets[tid]..offsets[tid+1]).  Returns the candidate count;
 * *chains_used is incremented like PrefilterStats::chains_used.  `scores`
 * (njobs) receives the per-job Kadane scores when not NULL. */
size_t or_score_and_gate(const int32_t *rows, size_t m, const uint8_t *arena,
                         const uint64_t *offsets, const or_job *jobs,
                         size_t njobs, const or_target_rec *recs, size_t nrecs,
                         int32_t open, int32_t ext, int32_t min_diag_score,
                         int regions, or_candidate *out, uint64_t *chains_used,
                         int32_t *scores);

/* size_adjusted_gate (src/prefilter.rs:246-257). */
int32_t or_size_adjusted_gate(int32_t gate, uint64_t db_residues);

/* ---- seedless all-diagonal evidence for candidate lists ----------------
 * UNG(q,t) = max over every diagonal of kadane_score over its whole range;
 * best_diag = smallest diagonal reaching it; the gate runs score_and_gate's
 * chain rule on the top-4 diagonals (score desc, diagonal asc).  Writes
 * best, best_diag, gate score (best or chain) and pass flag. */
void or_ung_pair(const int32_t *rows, size_t m, const uint8_t *t, size_t n,
                 int32_t open, int32_t ext, int32_t min_diag_score,
                 int32_t *ung, int32_t *best_diag, int32_t *gate_score,
                 int32_t *pass);

/* ---- seeded synthetic data -------------------------------------------- */

/* SplitMix64 (src/evalue.rs:52-75). */
uint64_t or_splitmix_next(uint64_t *state);
/* BackgroundSampler::sequence (src/evalue.rs:85-114): len residues drawn from
 * the BLOSUM62 background using SplitMix64 state *state. */
void or_background_sequence(uint64_t *state, size_t len, uint8_t *out);

/* ---- host-side rows (host_oracle.c) ----------------------------------- */

/* Profile::from_numeric_msa (src/profile.rs:90-182): msa holds nrows rows of
 * `stride` codes (0..19, 20 = X, 0xFF = gap), the first qlen are match
 * columns.  pssm: qlen x 20 int8; prob (qlen x 20), neff (qlen), cons (qlen)
 * may be NULL.  Returns 0, or -1 for an empty MSA / zero columns. */
int or_pssm_from_msa(const uint8_t *msa, size_t nrows, size_t stride, size_t qlen, float pca, float pcb,
                     float score_bias, int8_t *pssm, float *prob, float *neff, uint8_t *cons);

/* gumbel_mu_given_lambda, fit_gumbel_ml (src/evalue.rs:130-182) and
 * fit_calibration (src/evalue.rs:576-668) over cells sorted by (m, n):
 * out6 = {lambda, K, alpha/lambda, beta, ln K rms, 0}, cell_out 8 doubles
 * per cell {m, n, samples, mean, sd, lambda_ml, mu_ml, mu}; -1 = None. */
double or_gumbel_mu_given_lambda(const double *xs, size_t n, double lam);
int or_fit_gumbel_ml(const double *xs, size_t n, double *lam, double *mu);
int or_fit_calibration(const uint64_t *cm, const uint64_t *cn, const uint64_t *off, const double *xs, size_t nc,
                       uint64_t lmin, double *out6, double *cell_out);

#ifdef __cplusplus
}
#endif
#endif
