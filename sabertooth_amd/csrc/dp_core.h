// Device core of the DP kernels (included by the batch and serving kernels):
// arithmetic policies, lane shuffles, profile gathers and one pass of the
// fused SWG + Kadane recurrence.  The mapping and the exactness argument are
// described in dp_kernels.hip.
#pragma once
#include <hip/hip_runtime.h>

#include "st_device.h"

namespace st {

typedef short s16x2 __attribute__((ext_vector_type(2)));
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef const __attribute__((address_space(1))) uint32_t *gptr_u32;
typedef const __attribute__((address_space(1))) u32x4 *gptr_u32x4;

template <class To, class From>
__device__ __forceinline__ To bc(From v) {
  return __builtin_bit_cast(To, v);
}

// global-address-space loads (plain global_load, never flat)
__device__ __forceinline__ uint32_t ldg_u32(const uint8_t *p) { return *(gptr_u32)(p); }
__device__ __forceinline__ uint4 ldg_u4(const uint4 *p) {
  const u32x4 v = *(gptr_u32x4)(p);
  return make_uint4(v.x, v.y, v.z, v.w);
}

constexpr uint32_t kSent4 = 0x15151515u;  // four zero-score codes
// 1: a step's profile words are gathered one step ahead
#ifndef ST_SCORE_PREFETCH
#define ST_SCORE_PREFETCH 0
#endif
// DP steps per loop iteration: one in the batch evidence kernels, four in
// the end-cell kernels and the serving kernel (measured on MI355X: the
// batched, PSSM and long-pair workloads run 1.0-1.7% faster with one step
// per iteration, the alignment spans 3% and saturated serving 4-8% slower)
template <bool UNG, bool END>
constexpr int step_unroll() {
  return UNG && !END ? 1 : 4;
}

// ---- arithmetic policies -----------------------------------------------------

// Two pairs per group in the int16 halves (low = first slot, high = second).
struct Pk16 {
  static constexpr int kPer = 2;
  static constexpr bool kFloor = true;
  using T = uint32_t;
  using L = int16_t;
  using K = uint32_t;  // score << 17 | (0x1FFFF - x)
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
  __device__ static int half(T v, int x) { return x ? (int)(int16_t)(v >> 16) : (int)(int16_t)(v & 0xffffu); }
  __device__ static K key(int score, uint32_t x) { return ((uint32_t)score << 17) | (0x1FFFFu - x); }
  __device__ static uint64_t key64(K k) {
    return k ? diag_key((int32_t)(k >> 17), 0x1FFFFu - (k & 0x1FFFFu)) : 0ull;
  }
  __device__ static K shfl_xor(K k, int off) { return (K)__shfl_xor((int)k, off); }
  __device__ static K shfl(K k, int src) { return (K)__shfl((int)k, src); }
};

// One pair per group in int32, E/F floored at 0 (needs ext >= 0, open >= 0).
struct I32F {
  static constexpr int kPer = 1;
  static constexpr bool kFloor = true;
  using T = int32_t;
  using L = int32_t;
  using K = uint64_t;
  __device__ static T splat(int v) { return v; }
  __device__ static T add(T a, T b) { return a + b; }
  __device__ static T max(T a, T b) { return a > b ? a : b; }
  __device__ static T subf(T a, T c) {
    return (int32_t)__builtin_elementwise_sub_sat((uint32_t)a, (uint32_t)c);
  }
  __device__ static T score(const L *rowA, const L *, int k) { return rowA[k]; }
  __device__ static int half(T v, int) { return v; }
  __device__ static K key(int score, uint32_t x) { return diag_key(score, x); }
  __device__ static uint64_t key64(K k) { return k; }
  __device__ static K shfl_xor(K k, int off) {
    const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)k, off);
    const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(k >> 32), off);
    return ((uint64_t)hi << 32) | lo;
  }
  __device__ static K shfl(K k, int src) {
    const uint32_t lo = (uint32_t)__shfl((int)(uint32_t)k, src);
    const uint32_t hi = (uint32_t)__shfl((int)(uint32_t)(k >> 32), src);
    return ((uint64_t)hi << 32) | lo;
  }
};

// One pair per group in int32 with the reference's unfloored E/F (any gap
// costs); cells outside the real matrix are masked.
struct I32N : I32F {
  static constexpr bool kFloor = false;
  __device__ static T subf(T a, T c) { return a - c; }
};

// The int32 recurrences over a profile whose values fit int16 (the query's
// int16 profile): the LDS slice is half as large, so a 512-row pass leaves
// room for twice the waves; each value is sign-extended as it is read.
struct I32F16 : I32F {
  using L = int16_t;
};
struct I32N16 : I32N {
  using L = int16_t;
};
// ... and over a profile whose values fit int8 (a quarter of the int32
// slice).
struct I32F8 : I32F {
  using L = int8_t;
};
struct I32N8 : I32N {
  using L = int8_t;
};

template <class K>
__device__ __forceinline__ void insert4(K (&k)[4], K v) {
  // k[0] >= k[1] >= k[2] >= k[3]; keys are distinct
  k[3] = k[3] > (k[2] < v ? k[2] : v) ? k[3] : (k[2] < v ? k[2] : v);
  k[2] = k[2] > (k[1] < v ? k[1] : v) ? k[2] : (k[1] < v ? k[1] : v);
  k[1] = k[1] > (k[0] < v ? k[0] : v) ? k[1] : (k[0] < v ? k[0] : v);
  k[0] = k[0] > v ? k[0] : v;
}

__device__ __forceinline__ uint32_t shr1u(uint32_t v) {
  // lane i <- lane i-1 inside each 16-lane row; row starts read 0 and are
  // overwritten by the caller (they are always group starts)
  return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x111, 0xf, 0xf, true);
}

// lane i <- lane i-1 across the whole wave (groups of 32 or 64 lanes): the
// DPP wave shift (wave_shr:1), a VALU move with no LDS round trip; lane 0
// reads 0 and is overwritten by the caller (a group start)
__device__ __forceinline__ uint32_t shr1w(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x138, 0xf, 0xf, true);
}

template <class P, bool WIDE>
__device__ __forceinline__ typename P::T shr1(typename P::T v) {
  return (typename P::T)(WIDE ? shr1w((uint32_t)v) : shr1u((uint32_t)v));
}

template <class P>
__device__ __forceinline__ const typename P::L *prof_of(const QueryDev &q);
template <>
__device__ __forceinline__ const int16_t *prof_of<Pk16>(const QueryDev &q) { return q.prof16; }
template <>
__device__ __forceinline__ const int32_t *prof_of<I32F>(const QueryDev &q) { return q.prof32; }
template <>
__device__ __forceinline__ const int32_t *prof_of<I32N>(const QueryDev &q) { return q.prof32; }
template <>
__device__ __forceinline__ const int16_t *prof_of<I32F16>(const QueryDev &q) { return q.prof16; }
template <>
__device__ __forceinline__ const int16_t *prof_of<I32N16>(const QueryDev &q) { return q.prof16; }
template <>
__device__ __forceinline__ const int8_t *prof_of<I32F8>(const QueryDev &q) { return q.prof8; }
template <>
__device__ __forceinline__ const int8_t *prof_of<I32N8>(const QueryDev &q) { return q.prof8; }

// One step's profile words: Pk16 reads R/2 dwords (row pairs) per target,
// the int32 kernels R dwords (R/2 over an int16 profile, R/4 over an int8
// one).  Row pointers are dword aligned (the stride and R keep them so).
template <class P, int R>
__device__ __forceinline__ void load_raw(const typename P::L *lds, int stride, uint32_t code, int l,
                                         uint32_t (&raw)[R]) {
  // code * stride < 2^24: full-rate 24-bit multiplies
  if (sizeof(typename P::L) == 1) {
    const uint32_t *ra = (const uint32_t *)(lds + (int)__umul24(code & 0xff, (uint32_t)stride) + l * R);
#pragma unroll
    for (int i = 0; i < R / 4; i++) raw[i] = ra[i];
  } else if (sizeof(typename P::L) == 2 && P::kPer == 1) {
    const uint32_t *ra = (const uint32_t *)(lds + (int)__umul24(code & 0xff, (uint32_t)stride) + l * R);
#pragma unroll
    for (int i = 0; i < R / 2; i++) raw[i] = ra[i];
  } else if (sizeof(typename P::L) == 2) {
    const uint32_t *ra = (const uint32_t *)(lds + (int)__umul24(code & 0xff, (uint32_t)stride) + l * R);
    const uint32_t *rb = (const uint32_t *)(lds + (int)__umul24((code >> 8) & 0xff, (uint32_t)stride) + l * R);
#pragma unroll
    for (int i = 0; i < R / 2; i++) {
      raw[i] = ra[i];
      raw[R / 2 + i] = rb[i];
    }
  } else {
    const uint32_t *ra = (const uint32_t *)(lds + (int)__umul24(code & 0xff, (uint32_t)stride) + l * R);
#pragma unroll
    for (int i = 0; i < R; i++) raw[i] = ra[i];
  }
}

// raw words -> per-row scores (Pk16: first target low half, second high)
template <class P, int R>
__device__ __forceinline__ void unpack_raw(const uint32_t (&raw)[R], typename P::T (&sc)[R]) {
  if (sizeof(typename P::L) == 1) {
    // int32 recurrences over int8 values: row 4i + b in byte b
#pragma unroll
    for (int i = 0; i < R / 4; i++)
#pragma unroll
      for (int b = 0; b < 4; b++) sc[4 * i + b] = (typename P::T)((int32_t)(raw[i] << (24 - 8 * b)) >> 24);
  } else if (sizeof(typename P::L) == 2 && P::kPer == 1) {
    // int32 recurrences over int16 values: row 2i in the low half
#pragma unroll
    for (int i = 0; i < R / 2; i++) {
      sc[2 * i] = (typename P::T)(int32_t)(int16_t)(raw[i] & 0xffffu);
      sc[2 * i + 1] = (typename P::T)((int32_t)raw[i] >> 16);
    }
  } else if (sizeof(typename P::L) == 2) {
#pragma unroll
    for (int i = 0; i < R / 2; i++) {
      sc[2 * i] = (typename P::T)__builtin_amdgcn_perm(raw[R / 2 + i], raw[i], 0x05040100u);
      sc[2 * i + 1] = (typename P::T)__builtin_amdgcn_perm(raw[R / 2 + i], raw[i], 0x07060302u);
    }
  } else {
#pragma unroll
    for (int i = 0; i < R; i++) sc[i] = (typename P::T)raw[i];
  }
}

// Wave priority by request age (serving): every 64 column steps the wave
// sets its priority to the number of thresholds th[k] that the age of its
// request (device wall clock, 100 MHz, low 32 bits) has passed, so older
// requests' units win the SIMD's issue slots over newer ones.
struct AgeClock {
  uint32_t t_pub;
  uint32_t th[3];
};

__device__ __forceinline__ void set_age_priority(const AgeClock &ac) {
  const uint32_t age = (uint32_t)wall_clock64() - ac.t_pub;
  const int lvl = (age >= ac.th[0]) + (age >= ac.th[1]) + (age >= ac.th[2]);
  switch (__builtin_amdgcn_readfirstlane(lvl)) {
    case 0: __builtin_amdgcn_s_setprio(0); break;
    case 1: __builtin_amdgcn_s_setprio(1); break;
    case 2: __builtin_amdgcn_s_setprio(2); break;
    default: __builtin_amdgcn_s_setprio(3); break;
  }
}

// Per-unit, per-group state shared by the passes.
template <class P>
struct GroupCtx {
  int G, l, g, n_pad, m, m_pad;
  int len[2];
  const uint8_t *tp[2];
  bool has[2];
};

// The first column where a group's pair reaches its known score `target`:
// each lane's first column holding it (its running pass maximum rises
// strictly, so pbj is where pbest was first reached), the minimum over the
// group's G lanes, and the limit carried over from earlier passes.
template <class P>
__device__ __forceinline__ int group_stop_column(typename P::T pbest, int pbj, int target, int carried, int G) {
  int col = (int)pbest >= target ? pbj : 0x7fffffff;
  for (int off = 1; off < G; off <<= 1) col = min(col, __shfl_xor(col, off));
  return min(col, carried);
}

// One pass of G*R rows over all n_pad columns.  LAST: the pass holding the
// matrix's last padded row (bottom-row diagonal ends are folded here).
// WIDE: groups of 32 or 64 lanes (boundaries cross 16-lane DPP rows).
// AGE: the serving kernel's wave priority by request age (ac).
// STOP (end-cell kernels, one pair per group): the group's pair is known to
// reach *stop_at; the pass ends once every group has completed the column
// where its pair first reached it, or the column *climit carried over from
// the earlier passes (the first such column so far), whichever comes first:
// a later column cannot hold an earlier optimum in column-major order, and
// the next pass needs the boundary row only up to that column.  *climit is
// updated for the next pass.
template <class P, int R, bool UNG, bool LAST, bool END, bool WIDE, bool AGE = false, bool STOP = false, int UNR = 0>
__device__ __forceinline__ void run_pass(const GroupCtx<P> &c, const typename P::L *lds, int stride, int base,
                                         bool first, bool final_pass, const uint4 *rd, uint4 *wr,
                                         typename P::T e0, typename P::T vopen, typename P::T vext,
                                         typename P::T &best, typename P::K (&top)[P::kPer][4],
                                         typename P::T *edge, int lane, int (&bj)[2], int (&bi)[2],
                                         const AgeClock *ac = nullptr, const int *stop_at = nullptr,
                                         int *climit = nullptr) {
  using T = typename P::T;
  using K = typename P::K;
  constexpr int kPer = P::kPer;
  const int l = c.l, G = c.G, n_pad = c.n_pad, m = c.m;
  constexpr bool wide = WIDE;
  const T kInvalid = (T)0xFFFFFFFFu;  // a negative B: not a diagonal end

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
  // END: this lane's first optimum of the pass (its cells in column-major
  // order: columns ascending, rows ascending in a column), merged into the
  // unit's after the pass; Pk16: one per int16 half (pair slot)
  T pbest = 0;
  int pbj[2] = {0x7fffffff, 0x7fffffff}, pbi[2] = {0, 0};
  // the group's first lane streams residues: cw = columns [4u, 4u + 4),
  // nw = the next four; the boundary row of the previous pass one ahead
  uint32_t cw[2] = {kSent4, kSent4}, nw[2] = {kSent4, kSent4};
  uint4 pb = make_uint4(0, (uint32_t)e0, 0, 0);
  uint32_t code = kSentinel | (kSentinel << 8);
  if (l == 0) {
#pragma unroll
    for (int x = 0; x < kPer; x++) {
      if (c.len[x] > 0) cw[x] = ldg_u32(c.tp[x]);
      if (c.len[x] > 4) nw[x] = ldg_u32(c.tp[x] + 4);
    }
    if (!first && n_pad > 0) pb = ldg_u4(rd);
    code = (cw[0] & 0xffu) | ((cw[1] & 0xffu) << 8);
  }
  // scores of the current step are fetched one step ahead
#if ST_SCORE_PREFETCH
  uint32_t raw[R];
  load_raw<P, R>(lds, stride, code, l, raw);
#endif
  // bottom row (LAST): the value the group's last lane produced at step t is
  // kept by lane t mod G and folded every G steps
  T pending = kInvalid;
  // STOP: the first column holding the known score in the earlier passes
  const int carried = STOP ? *climit : 0;

  const int steps = n_pad + G - 1;
  constexpr int kStepUnroll = UNR ? UNR : step_unroll<UNG, END>();
  for (int t0 = 0; t0 < steps; t0 += kStepUnroll) {
    if constexpr (AGE) {
      if ((t0 & 63) == 0) set_age_priority(*ac);
    }
    if constexpr (STOP) {
      // every 16 steps: this group's first column holding the known score
      // (this pass or an earlier one); the wave leaves once every group
      // has run the steps that complete it (column + G - 1)
      if ((t0 & 15) == 0 && t0 > 0) {
        const int lim = group_stop_column<P>(pbest, pbj[0], *stop_at, *climit, G);
        if (__all(t0 >= lim + G)) break;
      }
    }
#pragma unroll
    for (int s = 0; s < kStepUnroll; s++) {
      const int t = t0 + s;
      if (t >= steps) break;
      // next step's code: the upper lane's current one, or the stream
      uint32_t nc = wide ? shr1w(code) : shr1u(code);
      if (l == 0) {
        if ((t & 3) < 3) {
          const int sh = ((t + 1) & 3) * 8;
          nc = ((cw[0] >> sh) & 0xffu) | (((cw[1] >> sh) & 0xffu) << 8);
        } else {
          nc = (nw[0] & 0xffu) | ((nw[1] & 0xffu) << 8);
          const int nx = t + 5;
#pragma unroll
          for (int x = 0; x < kPer; x++) {
            cw[x] = nw[x];
            nw[x] = nx < c.len[x] ? ldg_u32(c.tp[x] + nx) : kSent4;
          }
        }
      }
#if ST_SCORE_PREFETCH
      uint32_t nraw[R];
      load_raw<P, R>(lds, stride, nc, l, nraw);
#endif

      // boundary from the upper lane (or the previous pass / the matrix edge)
      T in_h = shr1<P, WIDE>(out_h), in_f = shr1<P, WIDE>(out_f);
      T in_u = 0, in_b = 0;
      if (UNG) {
        in_u = shr1<P, WIDE>(out_u);
        in_b = shr1<P, WIDE>(out_b);
      }
      if (l == 0) {
        if (!first && t < n_pad) {
          in_h = (T)pb.x;
          in_f = (T)pb.y;
          in_u = (T)pb.z;
          in_b = (T)pb.w;
        } else {
          in_h = 0;
          in_f = e0;
          in_u = 0;
          in_b = 0;
        }
        if (!first && t + 1 < n_pad) pb = ldg_u4(rd + t + 1);
      }
      const int j = t - l;
      const bool live = P::kFloor || (j >= 0 && j < n_pad);
      if (live) {
        T sc[R];
#if !ST_SCORE_PREFETCH
        uint32_t raw[R];
        load_raw<P, R>(lds, stride, code, l, raw);
#endif
        unpack_raw<P, R>(raw, sc);
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
        // END: the column's maximum over this lane's cells of the real
        // matrix (I32N: cells outside it count as 0, which never improves)
        T cm = 0;
#pragma unroll
        for (int k = 0; k < R; k++) {
          T h = P::max(P::max(H[k], E[k]), F);
          if (!P::kFloor) h = P::max(h, 0);
          H[k] = h;
          if (END) {
            cm = P::max(cm, (P::kFloor || (j < c.len[0] && base + l * R + k < m)) ? h : T(0));
          } else if (P::kFloor) {
            best = P::max(best, h);
          } else if (j < c.len[0] && base + l * R + k < m) {
            best = P::max(best, h);
          }
          const T ho = P::subf(h, vopen);
          E[k] = P::max(P::subf(E[k], vext), ho);
          F = P::max(P::subf(F, vext), ho);
        }
        if constexpr (END && kPer == 1) {
          // STOP: columns past the limit carried from the earlier passes
          // read boundary rows those passes did not write, and cannot hold
          // the optimum: they are not tracked
          if (__builtin_expect(cm > pbest, 0) && (!STOP || j <= carried)) {
            // a strictly larger value in this column: the first optimum of
            // the pass so far is this column's first row holding it (a lane
            // sees its columns in increasing j, a column's rows in
            // increasing i); a rare, divergent branch
            int kk = R - 1;
#pragma unroll
            for (int k = R - 1; k >= 0; k--)
              if (H[k] == cm && (P::kFloor || base + l * R + k < m)) kk = k;
            pbest = cm;
            pbj[0] = j;
            pbi[0] = base + l * R + kk;
          }
        } else if constexpr (END) {
          // the same per int16 half: a half of the column maximum above the
          // half's pass optimum (floored kernel: padded rows and columns
          // hold zero scores, never a strict rise past the real cells)
          const T up = P::max(cm, pbest);
          if (__builtin_expect(up != pbest, 0)) {
#pragma unroll
            for (int x = 0; x < kPer; x++) {
              const int cx = P::half(cm, x);
              int kk = R - 1;
#pragma unroll
              for (int k = R - 1; k >= 0; k--)
                if (P::half(H[k], x) == cx) kk = k;
              const bool rise = cx > P::half(pbest, x);
              pbj[x] = rise ? j : pbj[x];
              pbi[x] = rise ? base + l * R + kk : pbi[x];
            }
            pbest = up;
          }
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
      code = nc;
#if ST_SCORE_PREFETCH
#pragma unroll
      for (int k = 0; k < R; k++) raw[k] = nraw[k];
#endif
      // boundary row for the next pass
      if (!final_pass && l == G - 1 && j >= 0 && j < n_pad) {
        uint4 v;
        v.x = (uint32_t)out_h;
        v.y = (uint32_t)out_f;
        v.z = (uint32_t)out_u;
        v.w = (uint32_t)out_b;
        wr[j] = v;
      }
      if (UNG) {
        // right edge: park this lane's rows; folded after the pass
        if (j == n_pad - 1) {
#pragma unroll
          for (int k = 0; k < R; k++) edge[lane * R + k] = B[k];
        }
        if (LAST) {
          // bottom row: the group's last lane produces one value per step
          const T mine = (j >= 0 && j < n_pad - 1) ? B[R - 1] : kInvalid;
          const int slot = t & (G - 1);
          const T v = G > 1 ? (T)__shfl((int)mine, (lane & ~(G - 1)) + G - 1) : mine;
          if (slot == l) pending = v;
          if (slot == G - 1 || t == steps - 1) {
            // lane l's pending value came from step t - slot + l
            const int d = (t - slot + l - (G - 1)) - (c.m_pad - 1);
#pragma unroll
            for (int x = 0; x < kPer; x++) {
              const int sv = P::half(pending, x);
              if (l <= slot && c.has[x] && sv >= 0 && d >= -(m - 1) && d <= c.len[x] - 1)
                insert4<K>(top[x], P::key(sv, (uint32_t)(d + m - 1)));
            }
            pending = kInvalid;
          }
        }
      }
    }
  }
  if (UNG) {
    // fold this pass's right-edge values (each lane its own rows)
#pragma unroll
    for (int k = 0; k < R; k++) {
      const T v = edge[lane * R + k];
      const int d = (n_pad - 1) - (base + l * R + k);
#pragma unroll
      for (int x = 0; x < kPer; x++)
        if (c.has[x] && d >= -(m - 1) && d <= c.len[x] - 1)
          insert4<K>(top[x], P::key(P::half(v, x), (uint32_t)(d + m - 1)));
    }
  }
  if constexpr (STOP) *climit = group_stop_column<P>(pbest, pbj[0], *stop_at, *climit, G);
  if (END) {
    // the pass's first optimum against the earlier passes' (score desc,
    // column asc, row asc): a later pass holds lower rows, so it wins a tie
    // only at a smaller column
    // (selects, not conditional stores into the arrays: the compiler
    // dropped such stores of the packed kernel's merge)
#pragma unroll
    for (int x = 0; x < kPer; x++) {
      const int pb = P::half(pbest, x), b = P::half(best, x);
      const bool take = pb > b || (pb == b && pb > 0 && (pbj[x] < bj[x] || (pbj[x] == bj[x] && pbi[x] < bi[x])));
      bj[x] = take ? pbj[x] : bj[x];
      bi[x] = take ? pbi[x] : bi[x];
    }
    best = P::max(best, pbest);
  }
}

// One pass with the instance for the group width and the pass position
// chosen once (uniform per unit), not per step.
template <class P, int R, bool UNG, bool END, bool AGE = false, bool STOP = false, int UNR = 0>
__device__ __forceinline__ void run_pass_sel(bool last, const GroupCtx<P> &c, const typename P::L *lds, int stride,
                                             int base, bool first, bool final_pass, const uint4 *rd, uint4 *wr,
                                             typename P::T e0, typename P::T vopen, typename P::T vext,
                                             typename P::T &best, typename P::K (&top)[P::kPer][4],
                                             typename P::T *edge, int lane, int (&bj)[2], int (&bi)[2],
                                             const AgeClock *ac = nullptr, const int *stop_at = nullptr,
                                             int *climit = nullptr) {
  if (c.G > 16) {
    if (last)
      run_pass<P, R, UNG, true, END, true, AGE, STOP, UNR>(c, lds, stride, base, first, true, rd, wr, e0, vopen, vext,
                                                      best, top, edge, lane, bj, bi, ac, stop_at, climit);
    else
      run_pass<P, R, UNG, false, END, true, AGE, STOP, UNR>(c, lds, stride, base, first, final_pass, rd, wr, e0, vopen,
                                                       vext, best, top, edge, lane, bj, bi, ac, stop_at, climit);
  } else {
    if (last)
      run_pass<P, R, UNG, true, END, false, AGE, STOP, UNR>(c, lds, stride, base, first, true, rd, wr, e0, vopen, vext,
                                                       best, top, edge, lane, bj, bi, ac, stop_at, climit);
    else
      run_pass<P, R, UNG, false, END, false, AGE, STOP, UNR>(c, lds, stride, base, first, final_pass, rd, wr, e0,
                                                        vopen, vext, best, top, edge, lane, bj, bi, ac, stop_at,
                                                        climit);
  }
}

// One butterfly step of the group's end-cell merge: the partner lane's
// (score, column, row) per pair slot against this lane's, under (score
// desc, column asc, row asc); the caller then takes the max of the scores.
template <class P>
__device__ __forceinline__ void end_merge(typename P::T o, typename P::T best, int (&bj)[2], int (&bi)[2], int off) {
  int nj[2], ni[2];
#pragma unroll
  for (int x = 0; x < P::kPer; x++) {
    const int oj = __shfl_xor(bj[x], off), oi = __shfl_xor(bi[x], off);
    const int ov = P::half(o, x), bv = P::half(best, x);
    const bool take = ov > bv || (ov == bv && (oj < bj[x] || (oj == bj[x] && oi < bi[x])));
    nj[x] = take ? oj : bj[x];
    ni[x] = take ? oi : bi[x];
  }
#pragma unroll
  for (int x = 0; x < P::kPer; x++) {
    bj[x] = nj[x];
    bi[x] = ni[x];
  }
}

// Outputs of one unit, identical in every lane of a group after the merge:
// the best H (two int16 halves in Pk16), the 4 best diagonal keys per pair,
// the end cell (END) and the group's pair slots.
template <class P>
struct UnitResult {
  typename P::T best;
  typename P::K top[P::kPer][4];
  int bj, bi;
  uint32_t outi[2];
  bool valid[2];
  GroupCtx<P> c;
};

// Score one work unit: a query, npairs pair slots, G lanes per group, n_pad
// columns, `passes` passes of G*R rows.  lds holds kCodes rows of lds_stride
// entries; myb is the block's pass-boundary region (two halves of
// passbuf_half quads, then the right-edge parking area).
template <class P, int R, bool UNG, bool END, bool AGE = false>
__device__ __forceinline__ void score_unit(const QueryDev q, const PairDev *upairs, uint32_t npairs, uint32_t G,
                                           uint32_t n_pad, uint32_t passes, int32_t gap_open, int32_t gap_ext,
                                           const uint8_t *arena, const uint64_t *toff, const uint32_t *tlen,
                                           typename P::L *lds, int lds_stride, uint4 *myb, uint64_t passbuf_half,
                                           int lane, UnitResult<P> &r, const AgeClock *ac = nullptr) {
  using T = typename P::T;
  using L = typename P::L;
  using K = typename P::K;
  constexpr int kPer = P::kPer;
  T *const edge = (T *)(myb + 2 * passbuf_half);
  const T vopen = P::splat(gap_open);
  const T vext = P::splat(gap_ext);
  // boundary state of the row above the matrix / column left of it
  const T e0 = P::kFloor ? T(0) : T(max(-gap_open, (INT32_MIN / 4) - gap_ext));
  const L *prof = prof_of<P>(q);
  GroupCtx<P> &c = r.c;
  c.G = (int)G;
  c.l = lane & (c.G - 1);
  c.g = lane / c.G;
  const int GR = c.G * R;
  c.n_pad = (int)n_pad;
  c.m = (int)q.m;
  c.m_pad = (int)passes * GR;

  // this group's pair slots
  r.outi[0] = r.outi[1] = 0;
  r.valid[0] = r.valid[1] = false;
  c.len[0] = c.len[1] = 0;
  c.tp[0] = c.tp[1] = arena;
#pragma unroll
  for (int x = 0; x < kPer; x++) {
    const uint32_t slot = (uint32_t)(c.g * kPer + x);
    if (slot < npairs) {
      const PairDev pd = upairs[slot];
      r.valid[x] = true;
      c.len[x] = (int)tlen[pd.tid];
      c.tp[x] = arena + toff[pd.tid];
      r.outi[x] = pd.out;
    }
  }
#pragma unroll
  for (int x = 0; x < 2; x++) c.has[x] = r.valid[x] && c.len[x] > 0 && c.m > 0;

#pragma unroll
  for (int x = 0; x < kPer; x++)
#pragma unroll
    for (int k = 0; k < 4; k++) r.top[x][k] = 0;
  T best = 0;
  int bj[2] = {0x7fffffff, 0x7fffffff}, bi[2] = {0, 0};
  for (int p = 0; p < (int)passes; p++) {
    const int base = p * GR;
    // stage the profile rows [base, base + GR) of all codes into LDS
    __syncthreads();
    if (sizeof(L) == 2) {
      const int dw = GR / 2;
      for (int idx = lane; idx < kCodes * dw; idx += 64) {
        const int cc = idx / dw, rr = idx - cc * dw;
        const uint32_t v = *(const uint32_t *)(prof + (uint64_t)cc * q.ld + base + 2 * rr);
        *(uint32_t *)(lds + cc * lds_stride + 2 * rr) = v;
      }
    } else {
      for (int idx = lane; idx < kCodes * GR; idx += 64) {
        const int cc = idx / GR, rr = idx - cc * GR;
        lds[cc * lds_stride + rr] = prof[(uint64_t)cc * q.ld + base + rr];
      }
    }
    __syncthreads();
    const uint4 *rd = myb + (uint64_t)((p + 1) & 1) * passbuf_half + (uint64_t)c.g * c.n_pad;
    uint4 *wr = myb + (uint64_t)(p & 1) * passbuf_half + (uint64_t)c.g * c.n_pad;
    const bool final_pass = p == (int)passes - 1;
    // the serving kernel keeps four steps per loop iteration
    run_pass_sel<P, R, UNG, END, AGE, false, 4>(UNG && final_pass, c, lds, lds_stride, base, p == 0, final_pass, rd,
                                                wr, e0, vopen, vext, best, r.top, edge, lane, bj, bi, ac);
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
        for (int k = 0; k < 4; k++) other[k] = P::shfl_xor(r.top[x][k], off);
#pragma unroll
        for (int k = 0; k < 4; k++) insert4<K>(r.top[x], other[k]);
      }
    }
  }
  r.best = best;
  if (END) {
    r.bj = bj[0];
    r.bi = bi[0];
  }
}

}  // namespace st
