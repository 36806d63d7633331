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
//
// Mapping.  A work unit is one wave64 and one query.  Lanes form groups of G
// (1..16, inside a 16-lane DPP row); a group scores two targets at once in
// the int16 halves of 32-bit registers (Pk16) or one target in int32
// (I32F/I32N).  Lane l of a group owns R consecutive query rows of the
// current pass (G*R rows per pass) with H/E/U/B for those rows in VGPRs; the
// group walks the target columns with lane l skewed l steps behind lane l-1,
// so the row boundary (H, F, U, B of the lane's last row) moves to the next
// lane by one DPP row_shr:1 per step.  Between passes the last lane's
// boundary row goes through a per-block global buffer.  The query profile of
// the pass (22 codes x G*R rows) is staged in LDS; targets are padded with the
// zero-score code 21, which never raises any H, E, F or B, so padded columns,
// padded rows and the skew's idle steps need no masking.
//
// Exactness of the int16 pass: E and F are kept floored at 0 (exact for
// ext >= 0, since H >= 0 absorbs every negative E/F), which lets the u16
// saturating subtract do "minus penalty, floor 0" in one op; H + s uses the
// signed saturating add; a pair whose best reaches 32767 is flagged and
// rescored by the exact int32 kernel.  U <= H cell by cell, so an unsaturated
// H also means exact U and B.
#include <hip/hip_runtime.h>

#include "st_device.h"

namespace st {

typedef short s16x2 __attribute__((ext_vector_type(2)));
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));

template <class To, class From>
__device__ __forceinline__ To bc(From v) {
  return __builtin_bit_cast(To, v);
}

// ---- arithmetic policies -----------------------------------------------------

// Two pairs per group in the int16 halves (low = first slot, high = second).
struct Pk16 {
  static constexpr int kPer = 2;
  static constexpr bool kFloor = true;
  using T = uint32_t;
  using L = int16_t;
  __device__ static T splat(int v) { return (uint32_t)(uint16_t)v * 0x10001u; }
  __device__ static T add(T a, T b) {
    return bc<T>(__builtin_elementwise_add_sat(bc<s16x2>(a), bc<s16x2>(b)));
  }
  __device__ static T max(T a, T b) {
    return bc<T>(__builtin_elementwise_max(bc<s16x2>(a), bc<s16x2>(b)));
  }
  // a - c floored at 0 (a, c >= 0)
  __device__ static T subf(T a, T c) {
    return bc<T>(__builtin_elementwise_sub_sat(bc<u16x2>(a), bc<u16x2>(c)));
  }
  __device__ static T score(const L *rowA, const L *rowB, int k) {
    return (uint32_t)(uint16_t)rowA[k] | ((uint32_t)(uint16_t)rowB[k] << 16);
  }
  __device__ static int lo(T v) { return (int)(int16_t)(v & 0xffffu); }
  __device__ static int hi(T v) { return (int)(int16_t)(v >> 16); }
  __device__ static int half(T v, int x) { return x ? hi(v) : lo(v); }
};

// One pair per group in int32, E/F floored at 0 (needs ext >= 0, open >= 0).
struct I32F {
  static constexpr int kPer = 1;
  static constexpr bool kFloor = true;
  using T = int32_t;
  using L = int32_t;
  __device__ static T splat(int v) { return v; }
  __device__ static T add(T a, T b) { return a + b; }
  __device__ static T max(T a, T b) { return a > b ? a : b; }
  __device__ static T subf(T a, T c) {
    return (int32_t)__builtin_elementwise_sub_sat((uint32_t)a, (uint32_t)c);
  }
  __device__ static T score(const L *rowA, const L *, int k) { return rowA[k]; }
  __device__ static int half(T v, int) { return v; }
};

// One pair per group in int32 with the reference's unfloored E/F (any gap
// costs); cells outside the real matrix are masked.
struct I32N {
  static constexpr int kPer = 1;
  static constexpr bool kFloor = false;
  using T = int32_t;
  using L = int32_t;
  __device__ static T splat(int v) { return v; }
  __device__ static T add(T a, T b) { return a + b; }
  __device__ static T max(T a, T b) { return a > b ? a : b; }
  __device__ static T subf(T a, T c) { return a - c; }
  __device__ static T score(const L *rowA, const L *, int k) { return rowA[k]; }
  __device__ static int half(T v, int) { return v; }
};

template <class P>
__device__ __forceinline__ typename P::T shr1(typename P::T v) {
  // lane i <- lane i-1 inside each 16-lane row; row starts keep their value
  // and are overwritten by the caller (they are always group starts)
  return (typename P::T)__builtin_amdgcn_update_dpp((int)v, (int)v, 0x111, 0xf, 0xf, false);
}

__device__ __forceinline__ uint32_t shr1u(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp((int)v, (int)v, 0x111, 0xf, 0xf, false);
}

template <class P>
__device__ __forceinline__ const typename P::L *prof_of(const QueryDev &q);
template <>
__device__ __forceinline__ const int16_t *prof_of<Pk16>(const QueryDev &q) { return q.prof16; }
template <>
__device__ __forceinline__ const int32_t *prof_of<I32F>(const QueryDev &q) { return q.prof32; }
template <>
__device__ __forceinline__ const int32_t *prof_of<I32N>(const QueryDev &q) { return q.prof32; }

// Persistent DP kernel: every block (one wave) pulls units from *counter.
template <class P, int R, bool UNG>
__global__ __launch_bounds__(64) void dp_kernel(
    const Unit *__restrict__ units, uint32_t nunits, uint32_t *__restrict__ counter,
    const PairDev *__restrict__ pairs, const QueryDev *__restrict__ queries,
    const uint8_t *__restrict__ arena, const uint64_t *__restrict__ toff,
    const uint32_t *__restrict__ tlen, int32_t gap_open, int32_t gap_ext,
    uint4 *__restrict__ passbuf, uint64_t passbuf_half, PairOut *__restrict__ pout,
    int32_t *__restrict__ diag_out) {
  using T = typename P::T;
  using L = typename P::L;
  __shared__ L lds[kCodes * kLdsStride];
  const int lane = threadIdx.x;
  uint4 *const myb = passbuf + (uint64_t)blockIdx.x * 2 * passbuf_half;

  const T vopen = P::splat(gap_open);
  const T vext = P::splat(gap_ext);
  // boundary state of the row above the matrix / column left of it
  const T e0 = P::kFloor ? T(0) : T(P::max(-gap_open, (INT32_MIN / 4) - gap_ext));

  for (;;) {
    uint32_t got = 0;
    if (lane == 0) got = atomicAdd(counter, 1u);
    const uint32_t ui = __builtin_amdgcn_readfirstlane(__shfl(got, 0));
    if (ui >= nunits) break;
    const Unit u = units[ui];
    const QueryDev q = queries[u.query];
    const L *prof = prof_of<P>(q);
    const int G = (int)u.G;
    const int l = lane & (G - 1);
    const int g = lane / G;
    const int GR = G * R;
    const int n_pad = (int)u.n_pad;
    const int m = (int)q.m;
    const int m_pad = (int)u.passes * GR;

    // this group's pair slots
    int len[2] = {0, 0};
    const uint8_t *tp[2] = {arena, arena};
    uint64_t d0[2] = {0, 0};
    uint32_t outi[2] = {0, 0};
    bool valid[2] = {false, false};
#pragma unroll
    for (int x = 0; x < P::kPer; x++) {
      const uint32_t slot = (uint32_t)(g * P::kPer + x);
      if (slot < u.npairs) {
        const PairDev pd = pairs[u.pair0 + slot];
        valid[x] = true;
        len[x] = (int)tlen[pd.tid];
        tp[x] = arena + toff[pd.tid];
        d0[x] = pd.diag0;
        outi[x] = pd.out;
      }
    }

    T best = 0;
    for (int p = 0; p < (int)u.passes; p++) {
      const int base = p * GR;
      // stage the profile rows [base, base + GR) of all codes into LDS
      __syncthreads();
      if (sizeof(L) == 2) {
        const int dw = GR / 2;
        for (int idx = lane; idx < kCodes * dw; idx += 64) {
          const int c = idx / dw, r = idx - c * dw;
          const uint32_t v = *(const uint32_t *)(prof + (uint64_t)c * q.ld + base + 2 * r);
          *(uint32_t *)(lds + c * kLdsStride + 2 * r) = v;
        }
      } else {
        for (int idx = lane; idx < kCodes * GR; idx += 64) {
          const int c = idx / GR, r = idx - c * GR;
          lds[c * kLdsStride + r] = prof[(uint64_t)c * q.ld + base + r];
        }
      }
      __syncthreads();

      const uint4 *rd = myb + (uint64_t)((p + 1) & 1) * passbuf_half + (uint64_t)g * n_pad;
      uint4 *wr = myb + (uint64_t)(p & 1) * passbuf_half + (uint64_t)g * n_pad;
      const bool last = p == (int)u.passes - 1;

      T H[R], E[R], U[R], B[R];
#pragma unroll
      for (int k = 0; k < R; k++) {
        H[k] = 0;
        E[k] = e0;
        U[k] = 0;
        B[k] = 0;
      }
      T hdiag = 0, udiag = 0, bdiag = 0;
      T out_h = 0, out_f = e0, out_u = 0, out_b = 0;
      uint32_t code = kSentinel | (kSentinel << 8);
      // lane 0 of the group prefetches the next column's residues
      uint32_t nxt = kSentinel | (kSentinel << 8);
      if (l == 0) {
        const uint32_t a = len[0] > 0 ? tp[0][0] : kSentinel;
        const uint32_t b = len[1] > 0 ? tp[1][0] : kSentinel;
        nxt = a | (b << 8);
      }
      const int steps = n_pad + G - 1;
      for (int t = 0; t < steps; t++) {
        T in_h = shr1<P>(out_h), in_f = shr1<P>(out_f);
        T in_u = 0, in_b = 0;
        if (UNG) {
          in_u = shr1<P>(out_u);
          in_b = shr1<P>(out_b);
        }
        uint32_t in_code = shr1u(code);
        if (l == 0) {
          if (p > 0 && t < n_pad) {
            const uint4 v = rd[t];
            in_h = (T)v.x;
            in_f = (T)v.y;
            in_u = (T)v.z;
            in_b = (T)v.w;
          } else {
            in_h = 0;
            in_f = e0;
            in_u = 0;
            in_b = 0;
          }
          in_code = nxt;
          const int tn = t + 1;
          const uint32_t a = tn < len[0] ? tp[0][tn] : kSentinel;
          const uint32_t b = tn < len[1] ? tp[1][tn] : kSentinel;
          nxt = a | (b << 8);
        }
        code = in_code;
        const int j = t - l;
        const bool live = P::kFloor || (j >= 0 && j < n_pad);
        if (live) {
          const L *rowA = lds + (int)(code & 0xff) * kLdsStride + l * R;
          const L *rowB = lds + (int)((code >> 8) & 0xff) * kLdsStride + l * R;
          T sc[R];
#pragma unroll
          for (int k = 0; k < R; k++) sc[k] = P::score(rowA, rowB, k);
          // diagonal moves, last row first so row k-1 still holds column j-1
#pragma unroll
          for (int k = R - 1; k >= 1; k--) {
            if (UNG) {
              U[k] = P::add(P::max(U[k - 1], 0), sc[k]);
              B[k] = P::max(B[k - 1], U[k]);
            }
            H[k] = P::add(H[k - 1], sc[k]);
          }
          if (UNG) {
            U[0] = P::add(P::max(udiag, 0), sc[0]);
            B[0] = P::max(bdiag, U[0]);
          }
          H[0] = P::add(hdiag, sc[0]);
          // vertical / horizontal gaps, first row first
          T F = in_f;
#pragma unroll
          for (int k = 0; k < R; k++) {
            T h = P::max(P::max(H[k], E[k]), F);
            if (!P::kFloor) h = P::max(h, 0);
            H[k] = h;
            if (P::kFloor) {
              best = P::max(best, h);
            } else if (j < len[0] && base + l * R + k < m) {
              best = P::max(best, h);
            }
            const T ho = P::subf(h, vopen);
            E[k] = P::max(P::subf(E[k], vext), ho);
            F = P::max(P::subf(F, vext), ho);
          }
          out_h = H[R - 1];
          out_f = F;
          if (UNG) {
            out_u = U[R - 1];
            out_b = B[R - 1];
          }
        }
        hdiag = in_h;
        if (UNG) {
          udiag = in_u;
          bdiag = in_b;
        }
        // boundary row for the next pass
        if (!last && l == G - 1 && j >= 0 && j < n_pad) {
          uint4 v;
          v.x = (uint32_t)out_h;
          v.y = (uint32_t)out_f;
          v.z = (uint32_t)out_u;
          v.w = (uint32_t)out_b;
          wr[j] = v;
        }
        if (UNG) {
          // a diagonal ends on the last padded column ...
          if (j == n_pad - 1) {
#pragma unroll
            for (int k = 0; k < R; k++) {
              const int d = (n_pad - 1) - (base + l * R + k);
#pragma unroll
              for (int x = 0; x < P::kPer; x++)
                if (valid[x] && len[x] > 0 && d >= -(m - 1) && d <= len[x] - 1)
                  diag_out[d0[x] + (uint64_t)(d + m - 1)] = P::half(B[k], x);
            }
          }
          // ... or on the last padded row
          if (last && l == G - 1 && j >= 0 && j < n_pad - 1) {
            const int d = j - (m_pad - 1);
#pragma unroll
            for (int x = 0; x < P::kPer; x++)
              if (valid[x] && len[x] > 0 && d >= -(m - 1) && d <= len[x] - 1)
                diag_out[d0[x] + (uint64_t)(d + m - 1)] = P::half(B[R - 1], x);
          }
        }
      }
    }
    // best over the group's lanes
    for (int off = 1; off < G; off <<= 1) {
      const T o = (T)__shfl_xor((int)best, off);
      best = P::max(best, o);
    }
    if (l == 0) {
#pragma unroll
      for (int x = 0; x < P::kPer; x++)
        if (valid[x]) {
          PairOut po;
          po.swg = P::half(best, x);
          // 0: exact int16; 1: int16 saturated (rescore); 2: exact int32
          po.flag = P::kPer == 2 ? (po.swg >= 32767 ? 1 : 0) : 2;
          pout[outi[x]] = po;
        }
    }
  }
}

// Clear the unit counters before a launch (a kernel, not a memset, so that a
// captured graph replays it correctly).
__global__ void zero_u32(uint32_t *p, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

}  // namespace st

// ---- launch entry points used by the host library ------------------------------

namespace st {

struct DpLaunch {
  const Unit *units;
  uint32_t nunits;
  uint32_t *counter;
  const PairDev *pairs;
  const QueryDev *queries;
  const uint8_t *arena;
  const uint64_t *toff;
  const uint32_t *tlen;
  int32_t open, ext;
  uint4 *passbuf;
  uint64_t passbuf_half;
  PairOut *pout;
  int32_t *diag_out;
  uint32_t grid;
};

// kind: 0 = Pk16, 1 = I32F, 2 = I32N; R: 8 or 16
hipError_t launch_dp(const DpLaunch &a, int kind, int R, bool ung, hipStream_t s) {
  hipLaunchKernelGGL(zero_u32, dim3(1), dim3(64), 0, s, a.counter, 1u);
#define ST_DP(PP, RR, UU)                                                                  \
  hipLaunchKernelGGL((dp_kernel<PP, RR, UU>), dim3(a.grid), dim3(64), 0, s, a.units,       \
                     a.nunits, a.counter, a.pairs, a.queries, a.arena, a.toff, a.tlen,      \
                     a.open, a.ext, a.passbuf, a.passbuf_half, a.pout, a.diag_out)
  if (kind == 0) {
    if (R == 8) {
      if (ung) ST_DP(Pk16, 8, true); else ST_DP(Pk16, 8, false);
    } else {
      if (ung) ST_DP(Pk16, 16, true); else ST_DP(Pk16, 16, false);
    }
  } else if (kind == 1) {
    if (R == 8) {
      if (ung) ST_DP(I32F, 8, true); else ST_DP(I32F, 8, false);
    } else {
      if (ung) ST_DP(I32F, 16, true); else ST_DP(I32F, 16, false);
    }
  } else {
    if (ung) ST_DP(I32N, 8, true); else ST_DP(I32N, 8, false);
  }
#undef ST_DP
  return hipGetLastError();
}

}  // namespace st
