// Device core of the gate (included by the gate and serving kernels): score
// lookups, kadane_segment and the chain/gate rule.
#pragma once
#include <hip/hip_runtime.h>

#include "st_device.h"

namespace st {

// score lookups: code 21 rows of both profiles are zero
struct ProfGlobal {
  const int32_t *p;
  uint32_t ld;
  __device__ int32_t operator()(uint32_t code, int q) const { return p[(uint64_t)code * ld + q]; }
};

// kadane_segment (src/ungapped.rs:69-88) over query rows [lo, hi) of
// diagonal d: the first maximum in query order, restarting when cur <= 0
template <class Acc>
__device__ int32_t kadane_segment_dev(const Acc &acc, const uint8_t *t, int d, int lo, int hi, int *qs,
                                      int *qe) {
  int32_t best = 0, cur = 0;
  int bs = 0, be = 0, cs = lo;
  int q = lo;
  if (q < hi) {
    const int64_t p0 = (int64_t)q + d;
    const int64_t pa = p0 & ~(int64_t)15;
    int skip = (int)(p0 - pa);
    for (int64_t pos = pa; q < hi; pos += 16) {
      const uint4 v = *(const uint4 *)(t + pos);
      const uint32_t w[4] = {v.x, v.y, v.z, v.w};
      int32_t s[16];
#pragma unroll
      for (int b = 0; b < 16; b++) {
        const int qq = q + b - skip;
        s[b] = (b >= skip && qq < hi) ? acc((w[b >> 2] >> ((b & 3) * 8)) & 0xffu, qq) : 0;
      }
#pragma unroll
      for (int b = 0; b < 16; b++) {
        if (b >= skip && q < hi) {
          if (cur <= 0) {
            cur = s[b];
            cs = q;
          } else {
            cur += s[b];
          }
          if (cur > best) {
            best = cur;
            bs = cs;
            be = q;
          }
          q++;
        }
      }
      skip = 0;
    }
  }
  *qs = bs;
  *qe = be;
  return best;
}

// chain + gate of one target over ns <= NMAX scored diagonals
// (src/prefilter.rs:1577-1657, region 1269-1278).  pre (nullable): the
// kadane_segment bounds of every diagonal the chain rule may scan, computed
// beforehand.
template <class Acc, int NMAX = 16>
__device__ GateRes gate_eval(const Acc &acc, const uint8_t *t, const int32_t *sc, const int32_t *dg,
                             const int32_t *lo, const int32_t *hi, int ns, int32_t open, int32_t ext,
                             int32_t min_diag_score, bool regions, const int2 *pre = nullptr) {
  GateRes r;
  r.pass = 0;
  r.chain = 0;
  r.t_lo = 0;
  r.t_hi = 0xffffffffu;
  int32_t best_s = 0;
  int best_k = 0;
  for (int k = 0; k < ns; k++)
    if (sc[k] > best_s) {
      best_s = sc[k];
      best_k = k;
    }
  r.best = best_s;
  r.best_diag = ns > 0 ? dg[best_k] : 0;
  r.score = best_s;
  if (ns == 0) {
    r.pass = 0 >= min_diag_score;
    return r;
  }
  if (ns == 1 && sc[0] < min_diag_score) return r;
  int qs[NMAX], qe[NMAX];
  auto segment = [&](int k) {
    if (pre) {
      qs[k] = pre[k].x;
      qe[k] = pre[k].y;
    } else {
      kadane_segment_dev(acc, t, dg[k], lo[k], hi[k], &qs[k], &qe[k]);
    }
  };
  uint32_t have = 0;
  int32_t chain_s = INT32_MIN;
  int cx = 0, cy = 0;
  if (ns > 1)
    for (int x = 0; x < ns; x++)
      for (int y = 0; y < ns; y++) {
        if (x == y) continue;
        const int32_t sx = sc[x], sy = sc[y];
        const int dx = dg[x], dy = dg[y];
        if (sx <= 0 || sy <= 0 || dx == dy) continue;
        const int ad = dy > dx ? dy - dx : dx - dy;
        const int32_t c = sx + sy - (open + (ad - 1) * ext);
        if (c <= best_s || c <= chain_s) continue;
        if (!(have >> x & 1)) {
          segment(x);
          have |= 1u << x;
        }
        if (!(have >> y & 1)) {
          segment(y);
          have |= 1u << y;
        }
        if (qe[x] < qs[y] && qe[x] + dx < qs[y] + dy) {
          chain_s = c;
          cx = x;
          cy = y;
        }
      }
  int used0 = best_k, used1 = best_k;
  if (chain_s > best_s) {
    r.chain = 1;
    r.score = chain_s;
    used0 = cx;
    used1 = cy;
  }
  if (r.score < min_diag_score) return r;
  r.pass = 1;
  if (regions) {
    uint32_t a = 0xffffffffu, b = 0;
    const int us[2] = {used0, used1};
    for (int u = 0; u < 2; u++) {
      const int k = us[u];
      if (!(have >> k & 1)) {
        segment(k);
        have |= 1u << k;
      }
      const int s0 = qs[k] + dg[k], s1 = qe[k] + dg[k];
      const uint32_t ua = (uint32_t)(s0 > 0 ? s0 : 0), ub = (uint32_t)(s1 > 0 ? s1 : 0);
      a = ua < a ? ua : a;
      b = ub > b ? ub : b;
    }
    r.t_lo = a;
    r.t_hi = b;
  }
  return r;
}

// A pair's top-4 diagonal keys (score desc, diagonal asc, 0 = none; keys
// after the first 0 are ignored) decoded into scores, diagonals and whole
// diagonal ranges (diag_range, src/ungapped.rs:38-46); returns their count.
__device__ __forceinline__ int decode_keys(const uint64_t (&key)[4], int m, int n, int32_t (&ts)[4],
                                           int32_t (&td)[4], int32_t (&lo)[4], int32_t (&hi)[4]) {
  int nt = 0;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    ts[k] = td[k] = lo[k] = hi[k] = 0;
    if (key[k] == 0 || nt < k) continue;
    ts[k] = (int32_t)(key[k] >> 32);
    td[k] = (int32_t)(0xFFFFFFFFu - (uint32_t)key[k]) - (m - 1);
    const int d = td[k];
    const int a0 = -d > 0 ? -d : 0;
    const int b0 = m < n - d ? m : n - d;
    lo[k] = a0;
    hi[k] = b0 > a0 ? b0 : a0;
    nt++;
  }
  return nt;
}

// LDS scratch of gate_wave (one wave64 block).
struct GateLds {
  uint16_t jobs[256];     // lane * 4 + k of each segment scan to run
  int2 seg[256];          // kadane_segment bounds by lane * 4 + k
  int32_t d[256], lo[256], hi[256];
  uint64_t tp[64];        // each lane's target residues
};
constexpr size_t kGateLdsBytes = sizeof(GateLds);

// The gate of up to 64 pairs of one query, one per lane, by a whole wave:
// the ungapped score and best diagonal (the best key) and the chain rule's
// gate score over the keys' whole diagonals.  Every diagonal that can take
// part in a chain beating the best single diagonal (the only chains whose
// segments the rule scans) is listed and the kadane_segment scans are shared
// out over the lanes, instead of each lane scanning its own pair's diagonals
// while the others wait; the chain rule then runs on the precomputed bounds,
// so the results equal gate_eval's.
template <class Acc>
__device__ void gate_wave(const Acc &acc, bool active, const uint8_t *t, int m, int n, const uint64_t (&key)[4],
                          int32_t open, int32_t ext, int lane, GateLds *g, int32_t &ung, int32_t &diag,
                          int32_t &gscore) {
  int32_t ts[4], td[4], lo[4], hi[4];
  const int nt = active ? decode_keys(key, m, n, ts, td, lo, hi) : 0;
  int32_t best_s = 0;
#pragma unroll
  for (int k = 0; k < 4; k++)
    if (k < nt && ts[k] > best_s) best_s = ts[k];
  uint32_t need = 0;
#pragma unroll
  for (int x = 0; x < 4; x++)
#pragma unroll
    for (int y = 0; y < 4; y++) {
      if (x == y || x >= nt || y >= nt) continue;
      if (ts[x] <= 0 || ts[y] <= 0 || td[x] == td[y]) continue;
      const int ad = td[y] > td[x] ? td[y] - td[x] : td[x] - td[y];
      if (ts[x] + ts[y] - (open + (ad - 1) * ext) > best_s) need |= (1u << x) | (1u << y);
    }
#pragma unroll
  for (int k = 0; k < 4; k++) {
    g->d[lane * 4 + k] = td[k];
    g->lo[lane * 4 + k] = lo[k];
    g->hi[lane * 4 + k] = hi[k];
  }
  g->tp[lane] = (uint64_t)(uintptr_t)t;
  int njobs = 0;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const bool b = (need >> k) & 1;
    const uint64_t bal = __ballot(b);
    if (b) g->jobs[njobs + __popcll(bal & ((1ull << lane) - 1))] = (uint16_t)(lane * 4 + k);
    njobs += __popcll(bal);
  }
  __syncthreads();
  for (int j = lane; j < njobs; j += 64) {
    const int e = g->jobs[j];
    int qs, qe;
    kadane_segment_dev(acc, (const uint8_t *)(uintptr_t)g->tp[e >> 2], g->d[e], g->lo[e], g->hi[e], &qs, &qe);
    g->seg[e] = make_int2(qs, qe);
  }
  __syncthreads();
  int32_t gs = 0;
  if (nt) {
    const GateRes r = gate_eval<Acc, 4>(acc, t, ts, td, lo, hi, nt, open, ext, INT32_MIN, false, &g->seg[lane * 4]);
    gs = r.score;
  }
  ung = nt ? ts[0] : 0;
  diag = nt ? td[0] : 0;
  gscore = gs;
  __syncthreads();
}

}  // namespace st
