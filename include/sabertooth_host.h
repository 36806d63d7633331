/*
 * sabertooth_amd — host-side rows of the scoring path (C ABI).
 *
 * The reference computes these on the host around its scoring kernels; they
 * are restated here in C++ (no GPU involved) so a caller gets the same
 * inputs and outputs as the reference's Rust: the PSSM builder of profile
 * queries.  Floating-point code follows the reference's operation order and
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

#ifdef __cplusplus
}
#endif
#endif
