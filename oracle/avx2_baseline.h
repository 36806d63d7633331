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
#ifdef __cplusplus
}
#endif
#endif
