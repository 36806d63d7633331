// Smith-Waterman-Gotoh score + all-diagonal ungapped evidence on gfx950.
//
// Recurrence (scalar_gotoh_score_end, src/simd.rs:347-390; the textbook
// form at src/simd.rs:1124-1142), i = query row, j = target column:
//   E[i][j] = max(H[i][j-1] - open, E[i][j-1] - ext)
//   F[i][j] = max(H[i-1][j] - open, F[i-1][j] - ext)
//   H[i][j] = max(H[i-1][j-1] + s(i, t_j), E, F, 0),   score = max H
// and, fused on the same cells, Kadane along every diagonal
// (kadane_score, src/ungapped.rs:53-62):
//   U[i][j] = max(U[i-1][j-1], 0) + s(i, t_j),  B[i][j] = max(B[i-1][j-1], U)
// so B at a diagonal's last cell is that diagonal's best ungapped segment.
// Each pair keeps the 4 best diagonals (score desc, diagonal asc) as sorted
// keys; the gate kernel finishes from those.
//
// Mapping.  A work unit is one wave64 and one query.  Lanes form groups of G
// (1..64 lanes); a group scores two targets at once in
// the int16 halves of 32-bit registers (Pk16) or one target in int32
// (I32F/I32N).  Lane l of a group owns R consecutive query rows of the
// current pass (G*R rows per pass) with H/E/U/B for those rows in VGPRs; the
// group walks the target columns with lane l skewed l steps behind lane l-1,
// so the row boundary (H, F, U, B of the lane's last row) moves to the next
// lane by one DPP row_shr:1 per step (a cross-row lane shuffle for groups of
// 32 or 64 lanes).  Between passes the last lane's
// boundary row goes through a per-block global buffer (prefetched one column
// ahead).  The query profile of the pass (22 codes x G*R rows) is staged in
// LDS; target codes stream into the group's first lane as dwords fetched
// eight columns ahead and move down the group with the boundary.  Targets are
// padded with the zero-score code 21, which never raises any H, E, F or B, so
// padded columns, padded rows and the skew's idle steps need no masking.
//
// Diagonal ends: a diagonal ends on the last padded column (right edge; each
// lane parks its R values in LDS and folds them into its top-4 lists after
// the pass) or on the last padded row (bottom; the group's last lane parks
// one value per step in LDS and the whole group folds them every
// kBottomChunk steps).  The group's lists merge by butterfly at the end.
//
// Exactness of the int16 pass: E and F are kept floored at 0 (exact for
// ext >= 0, since H >= 0 absorbs every negative E/F), which lets the u16
// saturating subtract do "minus penalty, floor 0" in one op; H + s uses the
// signed saturating add; a pair whose best reaches 32767 is flagged and
// rescored by the exact int32 kernel.  U <= H cell by cell, so an unsaturated
// H also means exact U and B.
#include <hip/hip_runtime.h>

#include "dp_core.h"
#include "dp_launch.h"
#include "gate_core.h"

namespace st {

static_assert(kGateLdsBytes <= 6144, "the gate scratch must fit the evidence launches' minimum LDS");

// The final pass of a unit at a lower height RL < R (4, 8 or 12 rows per
// lane): the query's last rows without a whole pass of padding rows.
template <class P, int R, bool UNG, bool END>
__device__ __forceinline__ void run_last_lower(int RL, const GroupCtx<P> &c, const typename P::L *lds, int stride,
                                               int base, bool first, const uint4 *rd, uint4 *wr, typename P::T e0,
                                               typename P::T vopen, typename P::T vext, typename P::T &best,
                                               typename P::K (&top)[P::kPer][4], typename P::T *edge, int lane,
                                               int (&bj)[2], int (&bi)[2]) {
  if constexpr (R > 4) {
    if (RL == 4) {
      run_pass_sel<P, 4, UNG, END>(UNG, c, lds, stride, base, first, true, rd, wr, e0, vopen, vext, best, top, edge,
                                   lane, bj, bi);
      return;
    }
  }
  if constexpr (R > 8) {
    if (RL == 8) {
      run_pass_sel<P, 8, UNG, END>(UNG, c, lds, stride, base, first, true, rd, wr, e0, vopen, vext, best, top, edge,
                                   lane, bj, bi);
      return;
    }
  }
  if constexpr (R > 12) {
    if (RL == 12) {
      run_pass_sel<P, 12, UNG, END>(UNG, c, lds, stride, base, first, true, rd, wr, e0, vopen, vext, best, top,
                                    edge, lane, bj, bi);
      return;
    }
  }
  run_pass_sel<P, R, UNG, END>(UNG, c, lds, stride, base, first, true, rd, wr, e0, vopen, vext, best, top, edge,
                               lane, bj, bi);
}

// One unit: every pass of its groups, then the merged outputs.
template <class P, int R, bool UNG, bool END>
__device__ __forceinline__ void dp_unit(const Unit &u, const QueryDev *__restrict__ queries,
                                        const PairDev *__restrict__ pairs, const uint8_t *__restrict__ arena,
                                        const uint64_t *__restrict__ toff, const uint32_t *__restrict__ tlen,
                                        typename P::L *lds, int lds_stride, uint4 *myb, typename P::T *edge,
                                        uint64_t passbuf_half, typename P::T e0, typename P::T vopen,
                                        typename P::T vext, int32_t gap_open, int32_t gap_ext, const DpOut &o,
                                        int lane) {
  using T = typename P::T;
  using L = typename P::L;
  using K = typename P::K;
  constexpr int kPer = P::kPer;
  const QueryDev q = queries[u.query];
  const L *prof = prof_of<P>(q);
  GroupCtx<P> c;
  c.G = (int)u.G;
  c.l = lane & (c.G - 1);
  c.g = lane / c.G;
  const int GR = c.G * R;
  c.n_pad = (int)u.n_pad;
  c.m = (int)q.m;
  // the final pass may be lower (r_last rows per lane); packed int16 only
  const int RL = (P::kPer == 2 && u.r_last != 0 && (int)u.r_last < R) ? (int)u.r_last : R;
  c.m_pad = ((int)u.passes - 1) * GR + c.G * RL;

  // this group's pair slots
  uint32_t outi[2] = {0, 0};
  bool valid[2] = {false, false};
  c.len[0] = c.len[1] = 0;
  c.tp[0] = c.tp[1] = arena;
#pragma unroll
  for (int x = 0; x < kPer; x++) {
    const uint32_t slot = (uint32_t)(c.g * kPer + x);
    if (slot < u.npairs) {
      const PairDev pd = pairs[u.pair0 + slot];
      valid[x] = true;
      c.len[x] = (int)tlen[pd.tid];
      c.tp[x] = arena + toff[pd.tid];
      outi[x] = pd.out;
    }
  }
#pragma unroll
  for (int x = 0; x < 2; x++) c.has[x] = valid[x] && c.len[x] > 0 && c.m > 0;

  K top[kPer][4];
#pragma unroll
  for (int x = 0; x < kPer; x++)
#pragma unroll
    for (int k = 0; k < 4; k++) top[x][k] = 0;

  T best = 0;
  int bj[2] = {0x7fffffff, 0x7fffffff}, bi[2] = {0, 0};
  for (int p = 0; p < (int)u.passes; p++) {
    const int base = p * GR;
    const bool final_pass = p == (int)u.passes - 1;
    const int GRp = final_pass ? c.G * RL : GR;
    // stage the profile rows [base, base + GRp) of all codes into LDS
    __syncthreads();
    // int16 / int8 rows a dword at a time where the rows start on a dword
    // (every query's own profile; a window of a row-reversed profile may not)
    constexpr int kPerDw = 4 / (int)sizeof(L);
    if (sizeof(L) < 4 && ((((uintptr_t)prof) | ((uintptr_t)q.ld * sizeof(L))) & 3) == 0) {
      const int dw = GRp / kPerDw;
      for (int idx = lane; idx < kCodes * dw; idx += 64) {
        const int cc = idx / dw, r = idx - cc * dw;
        const uint32_t v = *(const uint32_t *)(prof + (uint64_t)cc * q.ld + base + kPerDw * r);
        *(uint32_t *)(lds + cc * lds_stride + kPerDw * r) = v;
      }
    } else {
      for (int idx = lane; idx < kCodes * GRp; idx += 64) {
        const int cc = idx / GRp, r = idx - cc * GRp;
        lds[cc * lds_stride + r] = prof[(uint64_t)cc * q.ld + base + r];
      }
    }
    __syncthreads();
    const uint4 *rd = myb + (uint64_t)((p + 1) & 1) * passbuf_half + (uint64_t)c.g * c.n_pad;
    uint4 *wr = myb + (uint64_t)(p & 1) * passbuf_half + (uint64_t)c.g * c.n_pad;
    bool lower = false;
    if constexpr (P::kPer == 2) {
      if (final_pass && RL != R) {
        run_last_lower<P, R, UNG, END>(RL, c, lds, lds_stride, base, p == 0, rd, wr, e0, vopen, vext, best, top,
                                       edge, lane, bj, bi);
        lower = true;
      }
    }
    if (!lower)
      run_pass_sel<P, R, UNG, END>(UNG && final_pass, c, lds, lds_stride, base, p == 0, final_pass, rd, wr, e0,
                                   vopen, vext, best, top, edge, lane, bj, bi);
  }
  // best and top-4 over the group's lanes
  for (int off = 1; off < c.G; off <<= 1) {
    const T o = (T)__shfl_xor((int)best, off);
    if (END) end_merge<P>(o, best, bj, bi, off);
    best = P::max(best, o);
    if (UNG) {
#pragma unroll
      for (int x = 0; x < kPer; x++) {
        K other[4];
#pragma unroll
        for (int k = 0; k < 4; k++) other[k] = P::shfl_xor(top[x][k], off);
#pragma unroll
        for (int k = 0; k < 4; k++) insert4<K>(top[x], other[k]);
      }
    }
  }
  if (UNG) {
    // the gate of each pair slot s on lane s mod 64, from the merged keys of
    // its group (every lane of a group holds the merge), then the pair's
    // final result row
    for (int s0 = 0; s0 < (int)u.npairs; s0 += 64) {
      const int s = s0 + lane;
      const int x = s % kPer;
      const int src = ((s / kPer) * c.G) & 63;
      const T bs = (T)__shfl((int)best, src);
      uint64_t key[4];
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const K k0 = P::shfl(top[0][k], src);
        const K k1 = kPer == 2 ? P::shfl(top[kPer - 1][k], src) : k0;
        key[k] = P::key64(x ? k1 : k0);
      }
      const bool act = s < (int)u.npairs;
      const PairDev pd = act ? pairs[u.pair0 + s] : PairDev{0, 0};
      PairResultDev r;
      r.swg = P::half(bs, x);
      // the profile region of LDS is free until the next unit's first pass
      gate_wave(ProfGlobal{q.prof32, q.ld}, act, arena + toff[pd.tid], c.m, act ? (int)tlen[pd.tid] : 0, key,
                gap_open, gap_ext, lane, (GateLds *)lds, r.ung, r.best_diag, r.gate_score);
      if (act) {
        r.pass = r.gate_score >= o.qgate[u.query];
        r.kernel = kPer == 2 ? 0 : 1;  // ST_KERNEL_I16 / ST_KERNEL_I32
        r.r0 = r.r1 = 0;
        o.res[pd.out] = r;
        if (kPer == 2 && r.swg >= 32767 && o.sat) {
          // at the int16 clamp: listed for exact int32 rescoring
          const uint32_t k = atomicAdd(o.sat, 1u);
          if (k < o.sat_cap) o.sat[1 + k] = pd.out;
        }
      }
    }
  } else if (c.l == 0) {
#pragma unroll
    for (int x = 0; x < kPer; x++)
      if (valid[x]) {
        PairOut po;
        po.swg = P::half(best, x);
        // 0: exact int16; 1: int16 saturated (rescore); 2: exact int32
        po.flag = kPer == 2 ? (po.swg >= 32767 ? 1 : 0) : 2;
        o.pout[outi[x]] = po;
        if (po.flag == 1 && o.sat) {
          const uint32_t k = atomicAdd(o.sat, 1u);
          if (k < o.sat_cap) o.sat[1 + k] = outi[x];
        }
        if (END) {
          o.qend[outi[x]] = po.swg > 0 ? bi[x] : -1;
          o.tend[outi[x]] = po.swg > 0 ? bj[x] : -1;
        }
      }
  }
}

// A window unit (end-cell kernels, one pair per group): group g scores pair
// slot g against its own profile view queries[pair_q[slot]], staged into its
// own part of the LDS slice (rows [g * G * R, (g + 1) * G * R) of every
// code), so one wave holds 64 / G pairs of different views (the reversed
// prefixes of several hits).  Rows past a view's m are staged as zeros,
// like the zero rows that follow every view of a row-reversed profile; the
// unit runs the passes of its tallest view.
template <class P, int R>
__device__ __forceinline__ void dp_unit_win(const Unit &u, const QueryDev *__restrict__ queries,
                                            const uint32_t *__restrict__ pair_q, const PairDev *__restrict__ pairs,
                                            const uint8_t *__restrict__ arena, const uint64_t *__restrict__ toff,
                                            const uint32_t *__restrict__ tlen, typename P::L *lds, int lds_stride,
                                            uint4 *myb, typename P::T *edge, uint64_t passbuf_half,
                                            typename P::T e0, typename P::T vopen, typename P::T vext,
                                            const DpOut &o, int lane) {
  using T = typename P::T;
  using L = typename P::L;
  using K = typename P::K;
  GroupCtx<P> c;
  c.G = (int)u.G;
  c.l = lane & (c.G - 1);
  c.g = lane / c.G;
  const int GR = c.G * R;
  c.n_pad = (int)u.n_pad;
  const bool valid = (uint32_t)c.g < u.npairs;
  const L *prof = nullptr;
  uint32_t ld = 0, outi = 0;
  c.m = 0;
  c.len[0] = c.len[1] = 0;
  c.tp[0] = c.tp[1] = arena;
  if (valid) {
    const PairDev pd = pairs[u.pair0 + c.g];
    const QueryDev q = queries[pair_q[u.pair0 + c.g]];
    prof = prof_of<P>(q);
    ld = q.ld;
    c.m = (int)q.m;
    c.len[0] = (int)tlen[pd.tid];
    c.tp[0] = arena + toff[pd.tid];
    outi = pd.out;
  }
  c.m_pad = (int)u.passes * GR;
  c.has[0] = valid && c.len[0] > 0 && c.m > 0;
  c.has[1] = false;
  K top[1][4] = {{0, 0, 0, 0}};
  T best = 0;
  int bj[2] = {0x7fffffff, 0x7fffffff}, bi[2] = {0, 0};
  L *const mine = lds + c.g * GR;
  // early stop (floored kernels, when the launch gives each pair's known
  // score): a group without a pair is done at once; a pair's limit starts
  // unbounded
  const bool stop = P::kFloor && o.stop != nullptr;
  const int stop_at = valid && stop ? o.stop[outi] : (1 << 30);
  int climit = valid ? (1 << 30) : -c.G;
  for (int p = 0; p < (int)u.passes; p++) {
    const int base = p * GR;
    // this group's rows [base, base + GR) of its view, zeros past its m
    __syncthreads();
    for (int idx = c.l; idx < kCodes * GR; idx += c.G) {
      const int cc = idx / GR, r = idx - cc * GR;
      mine[cc * lds_stride + r] = base + r < c.m ? prof[(uint64_t)cc * ld + base + r] : (L)0;
    }
    __syncthreads();
    const uint4 *rd = myb + (uint64_t)((p + 1) & 1) * passbuf_half + (uint64_t)c.g * c.n_pad;
    uint4 *wr = myb + (uint64_t)(p & 1) * passbuf_half + (uint64_t)c.g * c.n_pad;
    const bool final_pass = p == (int)u.passes - 1;
    if (stop)
      run_pass_sel<P, R, false, true, false, P::kFloor>(false, c, mine, lds_stride, base, p == 0, final_pass, rd, wr,
                                                        e0, vopen, vext, best, top, edge, lane, bj, bi, nullptr,
                                                        &stop_at, &climit);
    else
      run_pass_sel<P, R, false, true>(false, c, mine, lds_stride, base, p == 0, final_pass, rd, wr, e0, vopen, vext,
                                      best, top, edge, lane, bj, bi);
  }
  // the group's optimum: (score desc, column asc, row asc)
  for (int off = 1; off < c.G; off <<= 1) {
    const T ob = (T)__shfl_xor((int)best, off);
    end_merge<P>(ob, best, bj, bi, off);
    best = P::max(best, ob);
  }
  if (c.l == 0 && valid) {
    PairOut po;
    po.swg = (int)best;
    po.flag = 2;  // exact int32
    o.pout[outi] = po;
    o.qend[outi] = (int)best > 0 ? bi[0] : -1;
    o.tend[outi] = (int)best > 0 ? bj[0] : -1;
  }
}

// Persistent DP kernel: every block (one wave) pulls units from *counter.
// R = 8 or 16: rows per lane of every unit; R = 0: each unit's own R (8, 12
// or 16), so one launch holds units of every height.
// END (int32 kernels without the evidence): also the end cell of the
// optimum, first in column-major order, into qend / tend.
template <class P, int R, bool UNG, bool END>
__global__ __launch_bounds__(64, (R == 8 && (!END || P::kPer == 2)) ? 4 : (R == 0 ? 3 : 2)) void dp_kernel(
    const Unit *__restrict__ units, uint32_t nunits, uint32_t *__restrict__ counter,
    const PairDev *__restrict__ pairs, const QueryDev *__restrict__ queries,
    const uint32_t *__restrict__ pair_q, const uint8_t *__restrict__ arena, const uint64_t *__restrict__ toff,
    const uint32_t *__restrict__ tlen, int32_t gap_open, int32_t gap_ext,
    uint4 *__restrict__ passbuf, uint64_t passbuf_half, DpOut out, int lds_stride) {
  using T = typename P::T;
  using L = typename P::L;
  // the pass's profile slice: kCodes rows of lds_stride (>= G*R + 2) entries
  extern __shared__ __align__(16) unsigned char lds_raw[];
  L *const lds = (L *)lds_raw;
  const int lane = threadIdx.x;
  // per block: two boundary halves, then 64 lanes x 16 right-edge words
  uint4 *const myb = passbuf + (uint64_t)blockIdx.x * (2 * passbuf_half + kEdgeQuadsPerBlock);
  T *const edge = (T *)(myb + 2 * passbuf_half);

  const T vopen = P::splat(gap_open);
  const T vext = P::splat(gap_ext);
  // boundary state of the row above the matrix / column left of it
  const T e0 = P::kFloor ? T(0) : T(max(-gap_open, (INT32_MIN / 4) - gap_ext));

  for (;;) {
    uint32_t got = 0;
    if (lane == 0) got = atomicAdd(counter, 1u);
    const uint32_t ui = __builtin_amdgcn_readfirstlane(__shfl(got, 0));
    if (ui >= nunits) break;
    const Unit u = units[ui];
    if constexpr (END && P::kPer == 1 && R == 8) {
      // bit 31 of the query index marks a window unit
      if (u.query >> 31) {
        dp_unit_win<P, R>(u, queries, pair_q, pairs, arena, toff, tlen, lds, lds_stride, myb, edge, passbuf_half, e0,
                          vopen, vext, out, lane);
        continue;
      }
    }
    if constexpr (R == 0) {
      if (u.R == 12)
        dp_unit<P, 12, UNG, END>(u, queries, pairs, arena, toff, tlen, lds, lds_stride, myb, edge, passbuf_half, e0,
                                 vopen, vext, gap_open, gap_ext, out, lane);
      else if (u.R == 16)
        dp_unit<P, 16, UNG, END>(u, queries, pairs, arena, toff, tlen, lds, lds_stride, myb, edge, passbuf_half, e0,
                                 vopen, vext, gap_open, gap_ext, out, lane);
      else
        dp_unit<P, 8, UNG, END>(u, queries, pairs, arena, toff, tlen, lds, lds_stride, myb, edge, passbuf_half, e0,
                                vopen, vext, gap_open, gap_ext, out, lane);
    } else {
      dp_unit<P, R, UNG, END>(u, queries, pairs, arena, toff, tlen, lds, lds_stride, myb, edge, passbuf_half, e0,
                              vopen, vext, gap_open, gap_ext, out, lane);
    }
  }
}

// Clear the unit counter before a launch (a kernel, not a memset, so that a
// captured graph replays it correctly).
__global__ void zero_u32(uint32_t *p, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}


// Resident blocks per CU of a kernel instance with `lds_bytes` of dynamic LDS.
int dp_occupancy(int kind, int R, bool ung, size_t lds_bytes, bool end) {
  int n = 0;
#define ST_OCC(PP, RR, UU, EE) \
  (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, dp_kernel<PP, RR, UU, EE>, 64, lds_bytes)
  if (end) {
    if (kind == 0) ST_OCC(Pk16, 0, false, true);
    else if (kind == 2) ST_OCC(I32N, 8, false, true);
    else if (kind == 6) ST_OCC(I32N8, 8, false, true);
    else if (kind == 5) {
      if (R == 8) ST_OCC(I32F8, 8, false, true); else ST_OCC(I32F8, 16, false, true);
    } else if (kind == 4) ST_OCC(I32N16, 8, false, true);
    else if (kind == 3) {
      if (R == 8) ST_OCC(I32F16, 8, false, true); else ST_OCC(I32F16, 16, false, true);
    } else if (R == 8) ST_OCC(I32F, 8, false, true);
    else ST_OCC(I32F, 16, false, true);
  } else if (kind >= 3) {
    return 0;  // the int16-profile instances track end cells only
  } else if (kind == 0) {
    if (R == 0) {
      if (ung) ST_OCC(Pk16, 0, true, false); else ST_OCC(Pk16, 0, false, false);
    } else if (R == 8) {
      if (ung) ST_OCC(Pk16, 8, true, false); else ST_OCC(Pk16, 8, false, false);
    } else {
      if (ung) ST_OCC(Pk16, 16, true, false); else ST_OCC(Pk16, 16, false, false);
    }
  } else if (kind == 1) {
    if (R == 8) {
      if (ung) ST_OCC(I32F, 8, true, false); else ST_OCC(I32F, 8, false, false);
    } else {
      if (ung) ST_OCC(I32F, 16, true, false); else ST_OCC(I32F, 16, false, false);
    }
  } else {
    if (ung) ST_OCC(I32N, 8, true, false); else ST_OCC(I32N, 8, false, false);
  }
#undef ST_OCC
  return n;
}

hipError_t clear_u32(uint32_t *p, uint32_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(zero_u32, dim3((n + 63) / 64), dim3(64), 0, s, p, n);
  return hipGetLastError();
}

// kind: 0 = Pk16, 1 = I32F, 2 = I32N, 3 = I32F16, 4 = I32N16, 5 = I32F8,
// 6 = I32N8 (the last four: end-cell instances only); R: 8 or 16 (the
// unfloored kinds: 8; Pk16: 0, each unit's own height); end: the score +
// end-cell instances (no evidence)
hipError_t launch_dp(const DpLaunch &a, int kind, int R, bool ung, hipStream_t s, bool end) {
  if ((kind >= 3 && (!end || ung)) || (end && kind == 0 && R != 0)) return hipErrorInvalidValue;
  hipLaunchKernelGGL(zero_u32, dim3(1), dim3(64), 0, s, a.counter, 1u);
#define ST_DP(PP, RR, UU, EE)                                                                       \
  hipLaunchKernelGGL((dp_kernel<PP, RR, UU, EE>), dim3(a.grid), dim3(64),                            \
                     dp_lds_bytes(a.lds_stride, sizeof(typename PP::L), UU), s, a.units, a.nunits,   \
                     a.counter, a.pairs, a.queries, a.pair_q, a.arena, a.toff, a.tlen, a.open, a.ext, a.passbuf, \
                     a.passbuf_half, a.out, a.lds_stride)
  if (end) {
    if (kind == 0) ST_DP(Pk16, 0, false, true);
    else if (kind == 2) ST_DP(I32N, 8, false, true);
    else if (kind == 6) ST_DP(I32N8, 8, false, true);
    else if (kind == 5) {
      if (R == 8) ST_DP(I32F8, 8, false, true); else ST_DP(I32F8, 16, false, true);
    } else if (kind == 4) ST_DP(I32N16, 8, false, true);
    else if (kind == 3) {
      if (R == 8) ST_DP(I32F16, 8, false, true); else ST_DP(I32F16, 16, false, true);
    } else if (R == 8) ST_DP(I32F, 8, false, true);
    else ST_DP(I32F, 16, false, true);
  } else if (kind == 0) {
    if (R == 0) {
      if (ung) ST_DP(Pk16, 0, true, false); else ST_DP(Pk16, 0, false, false);
    } else if (R == 8) {
      if (ung) ST_DP(Pk16, 8, true, false); else ST_DP(Pk16, 8, false, false);
    } else {
      if (ung) ST_DP(Pk16, 16, true, false); else ST_DP(Pk16, 16, false, false);
    }
  } else if (kind == 1) {
    if (R == 8) {
      if (ung) ST_DP(I32F, 8, true, false); else ST_DP(I32F, 8, false, false);
    } else {
      if (ung) ST_DP(I32F, 16, true, false); else ST_DP(I32F, 16, false, false);
    }
  } else {
    if (ung) ST_DP(I32N, 8, true, false); else ST_DP(I32N, 8, false, false);
  }
#undef ST_DP
  return hipGetLastError();
}

}  // namespace st
