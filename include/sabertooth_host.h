/*
 * sabertooth_amd — host-side rows of the scoring path (C ABI).
 *
 * The reference computes these on the host around its scoring kernels; they
 * are restated here in C++ (no GPU involved) so a caller gets the same
 * inputs and outputs as the reference's Rust: the PSSM builder of profile
 * queries, the row modifiers of sequence queries (composition-bias
 * correction, low-complexity masking) and the hit statistics (E-values, bit
 * scores, search spaces).  Floating-point code follows the reference's operation order and
 * libm calls; the functions are deterministic and thread-safe.
 *
 * Status codes and st_last_error() are those of sabertooth_amd.h.
 */
#ifndef SABERTOOTH_HOST_H
#define SABERTOOTH_HOST_H

#include <stddef.h>
#include <stdint.h>

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
/* solve_lambda (src/evalue.rs:466-498): ungapped lambda of the integer
 * BLOSUM62 (diagnostic). */
double st_ungapped_lambda(void);

#ifdef __cplusplus
}
#endif
#endif
