/*
 * sabertooth_amd — host-side rows of the scoring path (C ABI).
 *
 * The reference computes these on the host around its scoring kernels; they
 * are restated here in C++ (no GPU involved) so a caller gets the same
 * inputs and outputs as the reference's Rust: the PSSM builder of profile
 * queries (and their per-profile statistics, whose random-target scores run
 * on the GPU), the row modifiers of sequence queries (composition-bias
 * correction, low-complexity masking), the exact traceback of reported hits,
 * the hit statistics (E-values, bit scores, search spaces), the gapped stage
 * of a query's search and its .m8 rows.  Floating-point code follows the reference's operation order and
 * libm calls; the functions are deterministic and thread-safe.
 *
 * Status codes and st_last_error() are those of sabertooth_amd.h.
 */
#ifndef SABERTOOTH_HOST_H
#define SABERTOOTH_HOST_H

#include <stddef.h>
#include <stdint.h>

#include "sabertooth_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- PSSM builder ---------------------------------------------------------- */

/* Numeric MSA gap code (src/alphabet.rs:33). */
#define ST_MSA_GAP 0xFF

/* parse_msa (src/profile.rs:372-449): aligned FASTA or a3m text -> *n_rows
 * rows of *query_len codes (0..19, 20 = X, ST_MSA_GAP), row-major into rows
 * when cap >= n_rows * query_len (call with rows = NULL for the sizes). */
int st_parse_msa(const char *text, uint64_t len, uint8_t *rows, uint64_t cap,
                 uint64_t *n_rows, uint64_t *query_len);

/* Profile::from_numeric_msa (src/profile.rs:90-182) with
 * PseudoCountParams (pca, pcb, score_bias) (src/profile.rs:47-66; defaults
 * 1.0, 1.5, 0.0): rows of `stride` codes, the first query_len are match
 * columns.  pssm: query_len x 20 int8 in residue-code order; prob
 * (query_len x 20 f32), neff (query_len f32), consensus (query_len) may be
 * NULL. */
int st_pssm_from_msa(const uint8_t *msa, uint64_t n_rows, uint64_t stride,
                     uint64_t query_len, float pca, float pcb, float score_bias,
                     int8_t *pssm, float *prob, float *neff, uint8_t *consensus);

/* Score rows of a profile query (ProfileScorer, src/align.rs:82-99):
 * rows[i][r] = pssm[i][r] for r < 20, rows[i][20] = 0 (target X). */
int st_pssm_rows(const int8_t *pssm, uint64_t query_len, int32_t *rows);

/* ---- sequence-query rows (src/compbias.rs, src/mask.rs, src/search.rs) ---- */

/* local_bias_correction_mode (src/compbias.rs:38-70): per-position additive
 * corrections (window 40), rounded half away from zero; penalize_only clips
 * positive corrections to 0.  Residue codes 0..20. */
int st_local_bias_correction(const uint8_t *query, uint64_t m, int penalize_only, int32_t *out);
/* low_complexity_mask_into (src/mask.rs:70-134): mask[i] = 1 for masked
 * positions; MaskParams (window, locut, hicut) = (12, 2.2, 2.5) for
 * seeding, (15, 2.3, 2.3) for scoring (src/mask.rs:33-66). */
int st_low_complexity_mask(const uint8_t *seq, uint64_t n, uint32_t window, double locut,
                           double hicut, uint8_t *mask, uint64_t *n_masked);
/* Rows of a sequence query as search_query builds them
 * (src/search.rs:432-463): BLOSUM62 rows, plus the composition-bias
 * correction of each column when comp_bias, then the columns the mask
 * (mask_window > 0) marks replaced by the matrix X row. */
int st_query_rows(const uint8_t *query, uint64_t m, int comp_bias, int penalize_only,
                  uint32_t mask_window, double mask_locut, double mask_hicut, int32_t *rows,
                  uint64_t *n_masked);

/* ---- exact traceback of reported hits (src/align.rs) ---------------------- */

/* Alignment (src/align.rs:224-248); coordinates 0-based inclusive. */
typedef struct {
  int32_t raw_score;
  int32_t reserved;
  uint64_t query_start, query_end;
  uint64_t target_start, target_end;
  uint64_t aln_len;    /* aligned pairs + gap columns */
  uint64_t identities; /* pairs of identical real residues */
  uint64_t mismatches;
  uint64_t gaps;       /* gap columns */
  uint64_t gap_opens;  /* gap runs (BLAST gapopen) */
} st_alignment;

/* scalar_gotoh_score_end (src/simd.rs:347-390): exact score and the first
 * optimum in column-major order (smallest target position, then smallest
 * query position); *qend = *tend = -1 when the score is 0. */
int st_sw_score_end_host(const int32_t *rows, uint64_t m, const uint8_t *t, uint64_t n,
                         int32_t gap_open, int32_t gap_extend, int32_t *score, int64_t *qend,
                         int64_t *tend);
/* align_canonical (src/align.rs:840-882) of rows [m][21] (qres: the query's
 * residues for identity counts, NULL for none) against t.  `known` (NULL:
 * computed here) is the forward score and end cell, e.g. from
 * st_sw_score_end_batch.  *found = 0 when no positive alignment exists, and
 * (as the reference's release build returns None, src/align.rs:859-878) when
 * no start or no band reproduces the score, which negative gap costs or a
 * wrong `known` can cause.  The
 * CIGAR (M, I = target residue against a query gap, D) is written
 * NUL-terminated up to cigar_cap bytes; *cigar_len receives its length.
 * st_score_end is declared in sabertooth_amd.h. */
int st_align_canonical(const int32_t *rows, const uint8_t *qres, uint64_t m, const uint8_t *t,
                       uint64_t n, int32_t gap_open, int32_t gap_extend, uint32_t band_radius,
                       const st_score_end *known, st_alignment *out, char *cigar,
                       uint64_t cigar_cap, uint64_t *cigar_len, int *found);
/* align_canonical's traceback (src/align.rs:869-881) for a hit whose span
 * (score, start and end cells) is known, e.g. from st_align_spans_batch:
 * the anchored banded traceback from start to end, band doubled until the
 * score is reproduced; a span whose start cells are -1 (end only) first
 * searches the start from the end, as align_canonical's fallback does
 * (src/align.rs:859-863).  Equal to st_align_canonical on the same problem.
 * *found = 0 when span->score <= 0 or no alignment of its score joins its
 * start and end; ST_ERR_INVALID when the span lies outside the problem.
 * st_span is declared in sabertooth_amd.h. */
int st_align_span(const int32_t *rows, const uint8_t *qres, uint64_t m, const uint8_t *t,
                  uint64_t n, int32_t gap_open, int32_t gap_extend, uint32_t band_radius,
                  const st_span *span, st_alignment *out, char *cigar, uint64_t cigar_cap,
                  uint64_t *cigar_len, int *found);
/* rescore_alignment (src/align.rs:1064-1135): recompute an alignment's
 * score and counts from its CIGAR; ST_ERR_INVALID on any disagreement. */
int st_rescore_alignment(const int32_t *rows, const uint8_t *qres, uint64_t m, const uint8_t *t,
                         uint64_t n, int32_t gap_open, int32_t gap_extend,
                         const st_alignment *aln, const char *cigar, int32_t *score);

/* ---- hit statistics (src/evalue.rs) ------------------------------------ */

/* GappedParams + LengthModel (src/evalue.rs:240-255). */
typedef struct {
  double lambda;
  double k;
  int32_t has_length; /* 1: BLAST length adjustment with (alpha_d_lambda, beta) */
  int32_t reserved;
  double alpha_d_lambda;
  double beta;
} st_gapped_params;

/* EValueParams (src/evalue.rs:311-329) bound to a database size. */
typedef struct {
  double lambda;
  double k;
  double log_k;
  double db_residues; /* global database residues (>= 1) */
  double db_seqs;     /* global database sequences (>= 1) */
  int32_t has_length;
  int32_t reserved;
  double alpha_d_lambda;
  double beta;
} st_evalue_params;

/* GumbelFit (src/evalue.rs:121-127). */
typedef struct {
  double lambda;
  double mu;
  uint64_t n;
} st_gumbel;

/* SEQUENCE_TABLE lookup (src/evalue.rs:274-310): gaps 11/1 with or without
 * composition-bias correction; ST_ERR_INVALID for other configurations (the
 * reference's EValueParams::sequence error, src/evalue.rs:347-366). */
int st_evalue_table(int32_t gap_open, int32_t gap_extend, int comp_bias, st_gapped_params *out);
/* EValueParams::from_params (src/evalue.rs:333-343). */
int st_evalue_params_from(const st_gapped_params *p, uint64_t db_residues, uint64_t db_seqs,
                          st_evalue_params *out);
/* EValueParams::for_profile (src/evalue.rs:391-403): diagnostic ln2/8. */
int st_evalue_params_profile(uint64_t db_residues, st_evalue_params *out);
/* length_adjustment (src/evalue.rs:193-237). */
double st_length_adjustment(double k, double alpha_d_lambda, double beta, double m,
                            double n_res, double n_seqs);
/* search_space (src/evalue.rs:427-438). */
double st_search_space(const st_evalue_params *p, uint64_t query_len);
/* bit_score (src/evalue.rs:420-423). */
double st_bit_score(const st_evalue_params *p, int32_t raw);
/* evalue / evalue_in_space (src/evalue.rs:441-450). */
double st_evalue(const st_evalue_params *p, int32_t raw, uint64_t query_len);
double st_evalue_in_space(const st_evalue_params *p, int32_t raw, double space);
/* min_score_for (src/evalue.rs:453-461). */
int32_t st_min_score_for(const st_evalue_params *p, double max_evalue, double space);
/* gumbel_mu_given_lambda / fit_gumbel_ml (src/evalue.rs:131-182);
 * st_fit_gumbel_ml returns ST_ERR_INVALID where the reference returns None. */
double st_gumbel_mu_given_lambda(const double *xs, uint64_t n, double lambda);
int st_fit_gumbel_ml(const double *xs, uint64_t n, st_gumbel *out);

/* CalibrationCell (src/evalue.rs:507-519): one simulated (m, n) cell. */
typedef struct {
  uint64_t m, n, samples;
  double mean, sd;        /* sample mean and standard deviation (n - 1) */
  double lambda_ml, mu_ml; /* the cell's own maximum-likelihood Gumbel */
  double mu;              /* its location under the shared lambda, + 1/2 */
} st_calibration_cell;

/* fit_calibration (src/evalue.rs:576-668): cell c holds the scores
 * scores[score_off[c] .. score_off[c + 1]) of (cell_m[c], cell_n[c]) pairs.
 * A shared lambda by golden-section maximum likelihood over the cells whose
 * shorter side is >= lambda_min_len (all cells when none is), each cell's
 * location under it, and the length model (alpha/lambda, beta) on the
 * reference's grid minimising the variance of ln K across cells;
 * K = e^(mean ln K).  cells (nullable, ncells entries) receives the cells in
 * (m, n) order, ln_k_rms (nullable) the root-mean-square spread of ln K.
 * ST_ERR_INVALID where the reference returns None. */
int st_fit_calibration(const uint64_t *cell_m, const uint64_t *cell_n, const uint64_t *score_off,
                       const double *scores, uint64_t ncells, uint64_t lambda_min_len,
                       st_gapped_params *out, st_calibration_cell *cells, double *ln_k_rms);
/* solve_lambda (src/evalue.rs:466-498): ungapped lambda of the integer
 * BLOSUM62 (diagnostic). */
double st_ungapped_lambda(void);

/* ---- gapped stage, hits and .m8 rows (src/search.rs) --------------------- */

/* Hit (src/search.rs:162-186) without the identifier strings. */
typedef struct {
  uint32_t target;   /* target ordinal */
  int32_t raw_score;
  double pident;     /* 100 * identities / aln_len */
  uint64_t aln_len, mismatches, gap_opens;
  uint64_t q_start, q_end, t_start, t_end; /* 0-based inclusive */
  double evalue;
  double bit_score;
  uint64_t qlen, tlen;
} st_hit;

/* SearchStats counters of the gapped stage (src/search.rs:565-666). */
typedef struct {
  uint64_t scored;            /* gapped_scored */
  uint64_t passed_score;      /* passed_score */
  uint64_t tracebacks;
  uint64_t score_mismatches;  /* traceback score != gapped score (expected 0) */
  uint64_t rejected_coverage;
  uint64_t hits;
} st_search_counters;

/* hit_order (src/search.rs:191-197): <0, 0, >0 as a sorts before b. */
int st_hit_order(const st_hit *a, const st_hit *b);
/* The pass filter and order of the gapped stage (src/search.rs:537-580):
 * order[0 .. *n_pass) = indices with score >= min_score and score > 0,
 * by score descending then target ascending. */
int st_gapped_stage(const int32_t *scores, const uint32_t *tids, uint64_t n, int32_t min_score,
                    uint64_t *order, uint64_t *n_pass);
/* One query's search after scoring (src/search.rs:532-666): min_score from
 * the E-value cut-off and the query's search space, the gapped stage,
 * deferred canonical traceback until max_hits (0 = unlimited) hits pass the
 * coverage filters, E-values and bit scores, hits in hit_order.  arena /
 * offsets: the database residues on the host.  Writes up to cap hits (and
 * their CIGARs, cigar_stride bytes each, when cigars is non-NULL);
 * *n_hits = all hits; counters (nullable) are accumulated. */
int st_search_hits(const int32_t *rows, const uint8_t *qres, uint64_t m, const uint8_t *arena,
                   const uint64_t *offsets, uint64_t n_targets, const uint32_t *cand_tids,
                   const int32_t *scores, uint64_t n_cands, int32_t gap_open, int32_t gap_extend,
                   const st_evalue_params *ev, double max_evalue, uint64_t max_hits,
                   double min_query_cov, double min_target_cov, uint32_t band_radius, st_hit *hits,
                   char *cigars, uint64_t cigar_stride, uint64_t cap, uint64_t *n_hits,
                   st_search_counters *counters);
/* st_search_hits with each candidate's span from st_align_spans_batch
 * (spans[c] for cand_tids[c]; only the candidates that pass the gapped
 * stage are read, and their span score must equal scores[c]): every
 * traceback is st_align_span's, so the host runs no score pass.  The hits
 * equal st_search_hits'. */
int st_search_hits_spans(const int32_t *rows, const uint8_t *qres, uint64_t m, const uint8_t *arena,
                         const uint64_t *offsets, uint64_t n_targets, const uint32_t *cand_tids,
                         const int32_t *scores, const st_span *spans, uint64_t n_cands,
                         int32_t gap_open, int32_t gap_extend, const st_evalue_params *ev,
                         double max_evalue, uint64_t max_hits, double min_query_cov,
                         double min_target_cov, uint32_t band_radius, st_hit *hits, char *cigars,
                         uint64_t cigar_stride, uint64_t cap, uint64_t *n_hits,
                         st_search_counters *counters);
/* Hit::to_m8 (src/search.rs:201-216) with Rust's number formatting: pident
 * {:.3}, E-value {:.2e} (exponent without '+' or padding), bits {:.1}. */
int st_format_m8(const char *query_id, const char *target_id, const st_hit *h, char *buf,
                 uint64_t cap, uint64_t *len);
/* A candidate trace (SABERTOOTH_TRACE_CANDIDATES, src/search.rs:358-385;
 * lines written at src/search.rs:522-530): "query \t target \t prefilter
 * score \t DP cells" per candidate.  Targets resolve against the database's
 * identifiers (id_bytes / id_off as st_stidx_read_ids gives them); queries
 * are numbered in order of first appearance and each query's candidates
 * keep their trace order.  *n_queries and *n_cands are always set; every
 * other output is written when non-NULL (qid_bytes when qid_cap holds them,
 * qid_off and cand_off n_queries + 1 entries, the candidate arrays n_cands),
 * so a first call with NULL outputs gives the sizes.  A malformed line or an
 * unknown target is ST_ERR_INVALID naming the line. */
int st_parse_candidate_trace(const char *text, uint64_t len, const uint8_t *id_bytes,
                             const uint64_t *id_off, uint64_t n_targets, uint64_t *n_queries,
                             uint64_t *n_cands, char *qid_bytes, uint64_t qid_cap, uint64_t *qid_off,
                             uint64_t *cand_off, uint32_t *cand_tids, int32_t *diag_scores,
                             uint64_t *cells);
/* The same lines for one query's candidates (the trace's writer side), so a
 * device prefilter's candidates can be compared with a traced reference run
 * line by line. */
int st_format_candidate_trace(const char *query_id, const char *const *target_ids,
                              const int32_t *diag_scores, const uint64_t *cells, uint64_t n,
                              char *buf, uint64_t cap, uint64_t *len);
/* Hit::extra_cols (src/search.rs:219-230): raw, qlen, tlen, qcov and tcov
 * ({:.3}, span / max(length, 1)), CIGAR; tab separated, no newline. */
int st_format_m8_extra(const st_hit *h, const char *cigar, char *buf, uint64_t cap, uint64_t *len);
/* A block of output lines as the reference's writer emits them
 * (write_header / write_hits, src/main.rs:667-687): the header line when
 * `header`, then one '\n'-terminated .m8 line per hit with the extra
 * columns appended when `extra` (cigars[i] per hit).  *len = the block's
 * length; up to cap - 1 bytes and a NUL are written to buf (call with
 * buf = NULL for the length), so a caller streams blocks to its output. */
int st_format_m8_block(const char *query_id, const char *const *target_ids, const st_hit *hits,
                       const char *const *cigars, uint64_t n, int extra, int header, char *buf,
                       uint64_t cap, uint64_t *len);

/* ---- per-profile statistics (src/evalue.rs:714-749) ----------------------- */

/* The random targets of calibrate_profile: sample i draws its length from
 * lengths[] and its residues from the BLOSUM62 background with
 * SplitMix64(stream_seed(seed, 0xBEEF, i)) (src/evalue.rs:52-114, 727-731).
 * offsets[0 .. samples] always; residues into arena when it holds
 * offsets[samples] bytes (cap). */
int st_calibration_targets(const uint64_t *lengths, uint64_t n_lengths, uint64_t samples,
                           uint64_t seed, uint8_t *arena, uint64_t cap, uint64_t *offsets);

/* calibrate_profile (src/evalue.rs:714-749): `samples` (reference: 2000)
 * random targets from st_calibration_targets scored against the profile on
 * the GPU with the given gaps, a maximum-likelihood Gumbel fit, lambda and
 * K = e^(lambda (mu + 1/2)) / (m * mean length); out->has_length = 0; fit
 * (nullable) receives the Gumbel.  ST_ERR_INVALID where the reference
 * returns None. */
int st_calibrate_profile(st_ctx *ctx, const st_qprof *qp, int32_t gap_open, int32_t gap_extend,
                         const uint64_t *lengths, uint64_t n_lengths, uint64_t samples, uint64_t seed,
                         st_gapped_params *out, st_gumbel *fit);

/* ---- sequence statistics by simulation (src/evalue.rs:537-700) ------------ */

/* simulate_scores (src/evalue.rs:537-565): sample i of cell cell_id draws a
 * query of m and then a target of n residues from the BLOSUM62 background
 * with SplitMix64(stream_seed(seed, cell_id, i)); out[i] = the exact SWG
 * score of the pair with the given gaps, the query's rows BLOSUM62 rows
 * plus, when comp_bias, its composition-bias correction
 * (local_bias_correction, src/compbias.rs:32-34).  The scores are computed
 * on the GPU, every query's profile built on the device. */
int st_simulate_scores(st_ctx *ctx, int32_t gap_open, int32_t gap_extend, int comp_bias, uint64_t m,
                       uint64_t n, uint64_t samples, uint64_t seed, uint64_t cell_id, double *out);

/* calibrate_sequence (src/evalue.rs:673-700): one cell of `samples` pairs
 * per length l = lengths[c] (m = n = l, cell id c) by st_simulate_scores,
 * then st_fit_calibration with lambda_min_len 200.  The reference's
 * SEQUENCE_TABLE came from lengths {60, 100, 150, 220, 320, 480, 700,
 * 1000}, 20000 samples, seed 24301.  cells (nullable) receives n_lengths
 * cells, ln_k_rms (nullable) the spread of ln K. */
int st_calibrate_sequence(st_ctx *ctx, int32_t gap_open, int32_t gap_extend, int comp_bias,
                          const uint64_t *lengths, uint64_t n_lengths, uint64_t samples, uint64_t seed,
                          st_gapped_params *out, st_calibration_cell *cells, double *ln_k_rms);

#ifdef __cplusplus
}
#endif
#endif
