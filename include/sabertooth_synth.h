/*
 * Seeded synthetic protein workloads for benchmarks and tests (host only).
 * Exported by libsabertooth_amd.so next to the scoring ABI; it has no
 * counterpart in the reference's scoring interface.  Residues follow the
 * reference's BackgroundSampler (src/evalue.rs:52-114: SplitMix64 over the
 * BLOSUM62 background); homolog substitutions draw from R[a][b] = P(a|b)
 * (src/matrix.rs:145-151).  Every function advances *state.
 */
#ifndef SABERTOOTH_SYNTH_H
#define SABERTOOTH_SYNTH_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif
/* n background residues. */
void sy_background(uint64_t *state, uint64_t n, uint8_t *out);
/* count log-normal lengths round(exp(ln(median) + sigma*z)) clipped to [lo, hi]. */
void sy_lognormal(uint64_t *state, uint64_t count, double median, double sigma,
                  uint32_t lo, uint32_t hi, uint32_t *out);
/* count uniform integers in [lo, hi]. */
void sy_uniform(uint64_t *state, uint64_t count, uint32_t lo, uint32_t hi, uint32_t *out);
/* A homolog of src[0..m) at the given identity with indel runs, cut or
 * extended with background residues to exactly out_len; returns out_len. */
uint64_t sy_homolog(uint64_t *state, const uint8_t *src, uint64_t m, double identity,
                    double indel_rate, uint32_t indel_max, uint64_t out_len, uint8_t *out);
/* n ids uniform in [0, range). */
void sy_pick(uint64_t *state, uint64_t n, uint64_t range, uint32_t *out);
/* One a3m member row around `root`: per column the root residue with
 * probability identity (else a P(a|b) substitute), gap runs of 1..10 starting
 * with probability gap_rate, lowercase insertions of 1..5 residues after a
 * column with probability insert_rate.  Returns the text length (written up
 * to cap bytes, no newline). */
uint64_t sy_msa_member(uint64_t *state, const uint8_t *root, uint64_t m, double identity,
                       double gap_rate, double insert_rate, char *out, uint64_t cap);
#ifdef __cplusplus
}
#endif
#endif
