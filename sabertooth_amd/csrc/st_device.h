// Shared device/host structures of the scoring path (gfx950).
#pragma once
#include <stdint.h>

namespace st {

// Zero-score residue code: every target code >= 21 scores 0
// (src/align.rs:180-187).  The device arena and the profiles use it as the
// padding code.
constexpr int kSentinel = 21;
// Profile codes held on the device: 0..20 plus the zero row 21.
constexpr int kCodes = 22;
// Largest group-rows product G*R of one pass (G <= 64 lanes, R <= 16 rows).
constexpr int kMaxPassRows = 1024;
// Per-block right-edge parking area of the DP kernels, in 16-byte units
// (64 lanes x 16 words); it follows the two pass-boundary halves.
constexpr uint64_t kEdgeQuadsPerBlock = 64 * 16 * 4 / 16;
// Largest m + n - 1 whose diagonal index fits the 17-bit field of the packed
// int16 kernel's top-4 keys.
constexpr uint32_t kMaxDiagsPk16 = 131071;

// One wave-sized work unit of the DP kernels: a query and npairs consecutive
// pair slots (targets sorted by length), scored by groups of G lanes.
struct Unit {
  uint32_t query;   // index into the QueryDev table
  uint32_t pair0;   // first pair slot
  uint32_t npairs;  // slots in this unit
  uint32_t G;       // lanes per group (1, 2, 4, ..., 64)
  uint32_t n_pad;   // columns: longest target of the unit
  uint32_t passes;  // ceil(m / (G*R))
  uint32_t R;       // rows per lane (the kernel instance)
  uint32_t r_last;  // rows per lane of the final pass (4, 8, 12 or R; <= R):
                    // the query's last rows without a whole pass of padding
};

// Device view of a query profile.
struct QueryDev {
  const int16_t *prof16;  // [kCodes][ld], rows >= m and code 21 are zero
  const int32_t *prof32;  // [kCodes][ld]
  uint32_t m;
  uint32_t ld;            // leading dimension, multiple of kMaxPassRows
  const int8_t *prof8;    // [kCodes][ld] where every value fits int8, else null
};

// One (query, target) pair slot of a batch.
struct PairDev {
  uint32_t tid;  // target id
  uint32_t out;  // index of its result row (and of its 4 top-diagonal keys)
};

// Per-pair raw outputs of the DP kernels.
struct PairOut {
  int32_t swg;
  int32_t flag;  // 0: exact int16; 1: int16 saturated (rescore); 2: exact int32
};

// Per-pair result row (the C ABI's st_pair_result).
struct PairResultDev {
  int32_t swg, ung, best_diag, gate_score;
  uint8_t pass, kernel, r0, r1;
};

// One unit of the serving kernel's work queue: a request's query (profile
// copies in the serving arena), npairs consecutive slots of the pair pool,
// the launch shape and the request's gaps and gate.
struct ServeUnit {
  const int16_t *prof16;   // profile copies in the serving arena
  const int32_t *prof32;
  const PairDev *pairs;    // the unit's pair slots in the serving arena
  uint32_t m, ld;
  uint32_t npairs;
  uint32_t G, n_pad, passes;
  int32_t open, ext;
  uint32_t t_pub;          // low 32 bits of the device wall clock (100 MHz) when published
  uint32_t req;            // request slot: its completion counter
};
static_assert(sizeof(ServeUnit) == 64, "one 64-byte descriptor per unit");

// Gate outcome of one target record or pair: the best single segment, its
// diagonal, the best single-or-chain score and its target range
// (Candidate, src/prefilter.rs:370-381).
struct GateRes {
  int32_t score;      // best single or chain
  int32_t best;       // best single segment
  int32_t best_diag;  // diagonal of the best single segment
  uint32_t t_lo, t_hi;
  uint8_t pass, chain, pad0, pad1;
};

// A job and a target record as the kernels read them: the C ABI's st_job
// and st_target_rec (layouts checked where both are visible).
struct JobDev {
  uint32_t tid;
  int32_t diag;
  uint32_t lo, hi;
};
struct RecDev {
  uint32_t tid, hits, j0, j1;
};

// Top-diagonal key: best ungapped score of a diagonal in the high word, and
// 0xFFFFFFFF - (d + m - 1) in the low word, so a larger key means a higher
// score, then a smaller diagonal d; 0 = no diagonal.
__host__ __device__ inline uint64_t diag_key(int32_t score, uint32_t x) {
  return ((uint64_t)(uint32_t)score << 32) | (uint64_t)(0xFFFFFFFFu - x);
}

}  // namespace st
