// CPU baseline — TEST/BENCH INFRASTRUCTURE ONLY.  See avx2_baseline.h.
// A restatement of the reference's AVX2 striped cascade (u8 -> i16 ->
// scalar), src/simd.rs:183-292, 418-591, 715-897, 342-390.
#include "avx2_baseline.h"

#include <immintrin.h>
#include <omp.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cstring>
#include <vector>

namespace {

constexpr int kNprof = 21;

// ---- biased u8 striped profile (StripedU8::build, src/simd.rs:435-472) ----
struct ProfU8 {
  bool ok = false;
  int seg = 0;
  uint8_t bias = 0, limit = 0;
  std::vector<__m256i> data;  // (kNprof + 1) * seg vectors
};

ProfU8 build_u8(const int32_t *rows, int m) {
  ProfU8 p;
  if (m == 0) return p;
  int lo = 0, hi = 0;
  for (int k = 0; k < m * kNprof; k++) {
    lo = std::min(lo, rows[k]);
    hi = std::max(hi, rows[k]);
  }
  const int bias = -lo, limit = 255 - (hi + bias);
  if (bias > 255 || limit < 96) return p;
  p.ok = true;
  p.seg = (m + 31) / 32;
  p.bias = (uint8_t)bias;
  p.limit = (uint8_t)limit;
  p.data.resize((size_t)(kNprof + 1) * p.seg);
  for (int res = 0; res <= kNprof; res++)
    for (int i = 0; i < p.seg; i++) {
      alignas(32) uint8_t b[32];
      for (int l = 0; l < 32; l++) {
        const int col = i + l * p.seg;
        const int v = (col < m && res < kNprof) ? rows[col * kNprof + res] : 0;
        b[l] = (uint8_t)(v + bias);
      }
      p.data[(size_t)res * p.seg + i] = _mm256_load_si256((const __m256i *)b);
    }
  return p;
}

// ---- i16 striped profile (Striped::build, src/simd.rs:730-754) ----
struct ProfI16 {
  bool ok = false;
  int seg = 0;
  std::vector<__m256i> data;
};

ProfI16 build_i16(const int32_t *rows, int m) {
  ProfI16 p;
  if (m == 0) return p;
  p.seg = (m + 15) / 16;
  p.data.resize((size_t)(kNprof + 1) * p.seg);
  for (int res = 0; res <= kNprof; res++)
    for (int i = 0; i < p.seg; i++) {
      alignas(32) int16_t b[16];
      for (int l = 0; l < 16; l++) {
        const int col = i + l * p.seg;
        const int v = (col < m && res < kNprof) ? rows[col * kNprof + res] : 0;
        if (v > 32767 || v < -32768) return ProfI16();
        b[l] = (int16_t)v;
      }
      p.data[(size_t)res * p.seg + i] = _mm256_load_si256((const __m256i *)b);
    }
  p.ok = true;
  return p;
}

struct Scratch {
  std::vector<__m256i> hs, hl, ea;
  std::vector<int32_t> hp, hc, fc;
};

inline __m256i shl1_u8(__m256i v) { return _mm256_alignr_epi8(v, _mm256_permute2x128_si256(v, v, 0x08), 15); }
inline __m256i shl1_i16(__m256i v) { return _mm256_alignr_epi8(v, _mm256_permute2x128_si256(v, v, 0x08), 14); }

inline uint8_t hmax_u8(__m256i v) {
  __m128i m = _mm_max_epu8(_mm256_castsi256_si128(v), _mm256_extracti128_si256(v, 1));
  m = _mm_max_epu8(m, _mm_srli_si128(m, 8));
  m = _mm_max_epu8(m, _mm_srli_si128(m, 4));
  m = _mm_max_epu8(m, _mm_srli_si128(m, 2));
  m = _mm_max_epu8(m, _mm_srli_si128(m, 1));
  return (uint8_t)(_mm_cvtsi128_si32(m) & 0xff);
}

inline int16_t hmax_i16(__m256i v) {
  __m128i m = _mm_max_epi16(_mm256_castsi256_si128(v), _mm256_extracti128_si256(v, 1));
  m = _mm_max_epi16(m, _mm_shuffle_epi32(m, 0x4E));
  m = _mm_max_epi16(m, _mm_shuffle_epi32(m, 0x01));
  m = _mm_max_epi16(m, _mm_shufflelo_epi16(m, 0x01));
  return (int16_t)_mm_extract_epi16(m, 0);
}

// StripedU8::score_avx2 (src/simd.rs:516-591); -1 = headroom exceeded
int score_u8(const ProfU8 &p, const uint8_t *t, int n, int open, int ext, Scratch &s) {
  const int seg = p.seg;
  const __m256i vo = _mm256_set1_epi8((char)(uint8_t)open), ve = _mm256_set1_epi8((char)(uint8_t)ext);
  const __m256i vb = _mm256_set1_epi8((char)p.bias), vz = _mm256_setzero_si256();
  const __m256i vl = _mm256_set1_epi8((char)p.limit);
  s.hs.assign(seg, vz);
  s.hl.assign(seg, vz);
  s.ea.assign(seg, vz);
  __m256i *hs = s.hs.data(), *hl = s.hl.data(), *ea = s.ea.data();
  __m256i vmax = vz;
  for (int j = 0; j < n; j++) {
    const __m256i *prof = p.data.data() + (size_t)std::min<int>(t[j], kNprof) * seg;
    __m256i vh = shl1_u8(hs[seg - 1]);
    __m256i vf = vz;
    std::swap(hs, hl);
    for (int i = 0; i < seg; i++) {
      const __m256i e = ea[i];
      vh = _mm256_subs_epu8(_mm256_adds_epu8(vh, prof[i]), vb);
      vh = _mm256_max_epu8(vh, e);
      vh = _mm256_max_epu8(vh, vf);
      vmax = _mm256_max_epu8(vmax, vh);
      hs[i] = vh;
      const __m256i ho = _mm256_subs_epu8(vh, vo);
      ea[i] = _mm256_max_epu8(_mm256_subs_epu8(e, ve), ho);
      vf = _mm256_max_epu8(_mm256_subs_epu8(vf, ve), ho);
      vh = hl[i];
    }
    for (int k = 0; k < 32; k++) {
      vf = shl1_u8(vf);
      bool done = false;
      for (int i = 0; i < seg; i++) {
        const __m256i vh2 = _mm256_max_epu8(hs[i], vf);
        vmax = _mm256_max_epu8(vmax, vh2);
        hs[i] = vh2;
        const __m256i ho = _mm256_subs_epu8(vh2, vo);
        ea[i] = _mm256_max_epu8(ea[i], ho);
        vf = _mm256_subs_epu8(vf, ve);
        if ((uint32_t)_mm256_movemask_epi8(_mm256_cmpeq_epi8(_mm256_subs_epu8(vf, ho), vz)) == 0xFFFFFFFFu) {
          done = true;
          break;
        }
      }
      if (done) break;
    }
    if ((uint32_t)_mm256_movemask_epi8(_mm256_cmpeq_epi8(_mm256_subs_epu8(vmax, vl), vz)) != 0xFFFFFFFFu)
      return -1;
  }
  return hmax_u8(vmax);
}

// Striped::score_avx2 (src/simd.rs:827-897); -1 = saturated
int score_i16(const ProfI16 &p, const uint8_t *t, int n, int open, int ext, Scratch &s) {
  const int seg = p.seg;
  const __m256i vo = _mm256_set1_epi16((short)open), ve = _mm256_set1_epi16((short)ext);
  const __m256i vz = _mm256_setzero_si256(), vneg = _mm256_set1_epi16(SHRT_MIN);
  const __m256i lane0 = _mm256_set_epi16(0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, SHRT_MIN);
  const __m256i sat = _mm256_set1_epi16(SHRT_MAX);
  s.hs.assign(seg, vz);
  s.hl.assign(seg, vz);
  s.ea.assign(seg, vz);
  __m256i *hs = s.hs.data(), *hl = s.hl.data(), *ea = s.ea.data();
  __m256i vmax = vz;
  for (int j = 0; j < n; j++) {
    const __m256i *prof = p.data.data() + (size_t)std::min<int>(t[j], kNprof) * seg;
    __m256i vh = shl1_i16(hs[seg - 1]);
    __m256i vf = vneg;
    std::swap(hs, hl);
    for (int i = 0; i < seg; i++) {
      const __m256i e = ea[i];
      vh = _mm256_adds_epi16(vh, prof[i]);
      vh = _mm256_max_epi16(vh, e);
      vh = _mm256_max_epi16(vh, vf);
      vh = _mm256_max_epi16(vh, vz);
      vmax = _mm256_max_epi16(vmax, vh);
      hs[i] = vh;
      const __m256i ho = _mm256_subs_epi16(vh, vo);
      ea[i] = _mm256_max_epi16(_mm256_subs_epi16(e, ve), ho);
      vf = _mm256_max_epi16(_mm256_subs_epi16(vf, ve), ho);
      vh = hl[i];
    }
    for (int k = 0; k < 16; k++) {
      vf = _mm256_or_si256(shl1_i16(vf), lane0);
      bool done = false;
      for (int i = 0; i < seg; i++) {
        const __m256i vh2 = _mm256_max_epi16(hs[i], vf);
        vmax = _mm256_max_epi16(vmax, vh2);
        hs[i] = vh2;
        const __m256i ho = _mm256_subs_epi16(vh2, vo);
        ea[i] = _mm256_max_epi16(ea[i], ho);
        vf = _mm256_subs_epi16(vf, ve);
        if (_mm256_movemask_epi8(_mm256_cmpgt_epi16(vf, ho)) == 0) {
          done = true;
          break;
        }
      }
      if (done) break;
    }
  }
  if (_mm256_movemask_epi8(_mm256_cmpeq_epi16(vmax, sat)) != 0) return -1;
  return hmax_i16(vmax);
}

// scalar_gotoh_score (src/simd.rs:342-390)
int score_scalar(const int32_t *rows, int m, const uint8_t *t, int n, int open, int ext, Scratch &s) {
  const int32_t neg = INT32_MIN / 4;
  s.hp.assign(n + 1, 0);
  s.hc.assign(n + 1, 0);
  s.fc.assign(n + 1, neg);
  int32_t best = 0;
  for (int i = 0; i < m; i++) {
    const int32_t *row = rows + (size_t)i * kNprof;
    int32_t e = neg;
    s.hc[0] = 0;
    for (int j = 1; j <= n; j++) {
      const int tr = t[j - 1];
      const int32_t d = s.hp[j - 1] + (tr < kNprof ? row[tr] : 0);
      e = std::max(s.hc[j - 1] - open, e - ext);
      const int32_t f = std::max(s.hp[j] - open, s.fc[j] - ext);
      s.fc[j] = f;
      const int32_t v = std::max(std::max(d, e), std::max(f, 0));
      s.hc[j] = v;
      best = std::max(best, v);
    }
    std::swap(s.hp, s.hc);
  }
  return best;
}

}  // namespace

extern "C" double avx2_score_pairs(uint32_t nq, const int32_t *const *rows, const uint32_t *qlen, uint64_t npairs,
                                   const uint32_t *pair_q, const uint32_t *pair_t, const uint8_t *arena,
                                   const uint64_t *off, const uint32_t *len, int32_t open, int32_t ext, int threads,
                                   int32_t *scores, uint8_t *kernels) {
  // pairs grouped by query; rayon-style parallelism across queries
  std::vector<std::vector<uint64_t>> byq(nq);
  for (uint64_t k = 0; k < npairs; k++) byq[pair_q[k]].push_back(k);
  const auto t0 = std::chrono::steady_clock::now();
#pragma omp parallel num_threads(threads)
  {
    Scratch s;
#pragma omp for schedule(dynamic, 1)
    for (int64_t q = 0; q < (int64_t)nq; q++) {
      if (byq[q].empty()) continue;
      const int m = (int)qlen[q];
      const ProfU8 pu = build_u8(rows[q], m);
      const ProfI16 pi = build_i16(rows[q], m);
      const bool simd_ok = open > ext && open <= 32767 && ext >= 0;
      for (uint64_t k : byq[q]) {
        const uint32_t tid = pair_t[k];
        const uint8_t *t = arena + off[tid];
        const int n = (int)len[tid];
        int sc = -1, kd = 2;
        if (m == 0 || n == 0) {
          sc = 0;
        } else if (simd_ok) {
          if (pu.ok && open <= 255) {
            sc = score_u8(pu, t, n, open, ext, s);
            kd = 0;
          }
          if (sc < 0 && pi.ok) {
            sc = score_i16(pi, t, n, open, ext, s);
            kd = 1;
          }
        }
        if (sc < 0) {
          sc = score_scalar(rows[q], m, t, n, open, ext, s);
          kd = 2;
        }
        scores[k] = sc;
        if (kernels) kernels[k] = (uint8_t)kd;
      }
    }
  }
  const auto t1 = std::chrono::steady_clock::now();
  return std::chrono::duration<double>(t1 - t0).count();
}

// ---- ungapped evidence: score_batch + kadane_avx2 + score_and_gate ----------
// (src/ungapped.rs:188-458, src/prefilter.rs:1498-1660)
namespace {

constexpr uint8_t kSentinelCode = 21;
constexpr size_t kEvChunk = 16384;  // EV_CHUNK, src/prefilter.rs:1381
constexpr int kDiagLanes = 32;      // DiagKernel::Avx2

// Biased u8 lookup tables, 32 bytes per query position (codes 21..31 score
// 0); usable when the bias fits a byte and at least 64 of headroom remain
// (DiagProfile::rebuild, src/ungapped.rs:221-250).
struct DiagTables {
  bool ok = false;
  uint8_t bias = 0, limit = 0;
  std::vector<uint8_t> tab;
};

DiagTables build_diag(const int32_t *rows, int m) {
  DiagTables d;
  int lo = 0, hi = 0;
  for (int k = 0; k < m * kNprof; k++) {
    lo = std::min(lo, rows[k]);
    hi = std::max(hi, rows[k]);
  }
  const int bias = -lo, limit = 255 - (hi + bias);
  if (m == 0 || bias > 255 || limit < 64) return d;
  d.ok = true;
  d.bias = (uint8_t)bias;
  d.limit = (uint8_t)limit;
  d.tab.assign((size_t)m * 32, (uint8_t)bias);
  for (int i = 0; i < m; i++)
    for (int r = 0; r < kNprof; r++) d.tab[(size_t)i * 32 + r] = (uint8_t)(rows[i * kNprof + r] + bias);
  return d;
}

inline int32_t sub_score(const int32_t *rows, int q, uint8_t c) { return c < kNprof ? rows[q * kNprof + c] : 0; }

int32_t kadane_exact(const int32_t *rows, const uint8_t *t, int64_t d, uint32_t lo, uint32_t hi) {
  int32_t best = 0, cur = 0;
  for (uint32_t q = lo; q < hi; q++) {
    cur = std::max(cur, 0) + sub_score(rows, (int)q, t[(int64_t)q + d]);
    best = std::max(best, cur);
  }
  return best;
}

// kadane_segment: the first maximum in query order, restarting at cur <= 0
int32_t kadane_bounds(const int32_t *rows, const uint8_t *t, int64_t d, uint32_t lo, uint32_t hi, uint32_t *qs,
                      uint32_t *qe) {
  int32_t best = 0, cur = 0;
  uint32_t bs = 0, be = 0, cs = lo;
  for (uint32_t q = lo; q < hi; q++) {
    const int32_t s = sub_score(rows, (int)q, t[(int64_t)q + d]);
    if (cur <= 0) {
      cur = s;
      cs = q;
    } else {
      cur += s;
    }
    if (cur > best) {
      best = cur;
      bs = cs;
      be = q;
    }
  }
  *qs = bs;
  *qe = be;
  return best;
}

// 16 rows of 16 bytes (src[r]) -> 16 columns at dst + c * stride, each byte
// clamped to the sentinel code: the unpack network 8 -> 16 -> 32 -> 64 bits
inline void transpose16(const uint8_t *const src[16], uint8_t *dst, size_t stride) {
  const __m128i cap = _mm_set1_epi8((char)kSentinelCode);
  __m128i r[16], u[16];
  for (int k = 0; k < 16; k++) r[k] = _mm_min_epu8(_mm_loadu_si128((const __m128i *)src[k]), cap);
  for (int k = 0; k < 16; k += 2) {
    u[k] = _mm_unpacklo_epi8(r[k], r[k + 1]);
    u[k + 1] = _mm_unpackhi_epi8(r[k], r[k + 1]);
  }
  for (int k = 0; k < 16; k += 4) {
    r[k] = _mm_unpacklo_epi16(u[k], u[k + 2]);
    r[k + 1] = _mm_unpackhi_epi16(u[k], u[k + 2]);
    r[k + 2] = _mm_unpacklo_epi16(u[k + 1], u[k + 3]);
    r[k + 3] = _mm_unpackhi_epi16(u[k + 1], u[k + 3]);
  }
  // r[4j + c]: rows 4j..4j+3, columns 4c..4c+3
  for (int h = 0; h < 2; h++)
    for (int c = 0; c < 4; c++) {
      u[8 * h + 2 * c] = _mm_unpacklo_epi32(r[8 * h + c], r[8 * h + 4 + c]);
      u[8 * h + 2 * c + 1] = _mm_unpackhi_epi32(r[8 * h + c], r[8 * h + 4 + c]);
    }
  // u[8h + k]: rows 8h..8h+7, columns 2k, 2k+1
  for (int k = 0; k < 8; k++) {
    _mm_storeu_si128((__m128i *)(dst + (2 * k) * stride), _mm_unpacklo_epi64(u[k], u[8 + k]));
    _mm_storeu_si128((__m128i *)(dst + (2 * k + 1) * stride), _mm_unpackhi_epi64(u[k], u[8 + k]));
  }
}

struct Lane {
  const uint8_t *t;
  int64_t d;
  uint32_t lo, hi;
};

struct UngScratch {
  std::vector<uint8_t> lm, pm;
  std::vector<uint32_t> cnt, order;
  std::vector<int32_t> sc;
};

// exact kadane_score of up to 32 lanes (score_batch, src/ungapped.rs:285-386)
void score_lanes(const DiagTables &P, const int32_t *rows, int m, const Lane *lanes, int nl, int32_t *out,
                 UngScratch &s) {
  if (!P.ok) {
    for (int l = 0; l < nl; l++) out[l] = kadane_exact(rows, lanes[l].t, lanes[l].d, lanes[l].lo, lanes[l].hi);
    return;
  }
  uint32_t q_lo = (uint32_t)m, q_hi = 0;
  for (int l = 0; l < nl; l++)
    if (lanes[l].lo < lanes[l].hi) {
      q_lo = std::min(q_lo, lanes[l].lo);
      q_hi = std::max(q_hi, lanes[l].hi);
    }
  alignas(32) uint8_t best[kDiagLanes] = {0};
  if (q_lo < q_hi) {
    const uint32_t a = q_lo & ~15u, b = (q_hi + 15) & ~15u, w = b - a;
    if (s.lm.size() < (size_t)kDiagLanes * w) s.lm.resize((size_t)kDiagLanes * w);
    if (s.pm.size() < (size_t)kDiagLanes * w) s.pm.resize((size_t)kDiagLanes * w);
    for (int l = 0; l < kDiagLanes; l++) {
      uint8_t *row = s.lm.data() + (size_t)l * w;
      if (l >= nl || lanes[l].lo >= lanes[l].hi) {
        std::memset(row, kSentinelCode, w);
        continue;
      }
      const uint32_t r0 = lanes[l].lo - a, r1 = lanes[l].hi - a;
      std::memset(row, kSentinelCode, r0);
      std::memcpy(row + r0, lanes[l].t + (int64_t)lanes[l].lo + lanes[l].d, r1 - r0);
      std::memset(row + r1, kSentinelCode, w - r1);
    }
    const uint8_t *src[16];
    for (int g = 0; g < kDiagLanes / 16; g++)
      for (uint32_t pb = 0; pb < w / 16; pb++) {
        for (int r = 0; r < 16; r++) src[r] = s.lm.data() + (size_t)(g * 16 + r) * w + pb * 16;
        transpose16(src, s.pm.data() + (size_t)pb * 16 * kDiagLanes + g * 16, kDiagLanes);
      }
    // kadane_avx2: saturating biased-u8 Kadane, scores by two 16-entry
    // byte shuffles blended on code > 15
    const __m256i vbias = _mm256_set1_epi8((char)P.bias), v15 = _mm256_set1_epi8(15);
    __m256i cur = _mm256_setzero_si256(), bst = _mm256_setzero_si256();
    const uint8_t *tab = P.tab.data();
    for (uint32_t i = q_lo; i < q_hi; i++) {
      const __m256i t = _mm256_loadu_si256((const __m256i *)(s.pm.data() + (size_t)(i - a) * kDiagLanes));
      const __m256i lo = _mm256_broadcastsi128_si256(_mm_loadu_si128((const __m128i *)(tab + (size_t)i * 32)));
      const __m256i hi = _mm256_broadcastsi128_si256(_mm_loadu_si128((const __m128i *)(tab + (size_t)i * 32 + 16)));
      const __m256i sc = _mm256_blendv_epi8(_mm256_shuffle_epi8(lo, t), _mm256_shuffle_epi8(hi, t),
                                            _mm256_cmpgt_epi8(t, v15));
      cur = _mm256_subs_epu8(_mm256_adds_epu8(cur, sc), vbias);
      bst = _mm256_max_epu8(bst, cur);
    }
    _mm256_store_si256((__m256i *)best, bst);
  }
  for (int l = 0; l < nl; l++)
    out[l] = best[l] > P.limit ? kadane_exact(rows, lanes[l].t, lanes[l].d, lanes[l].lo, lanes[l].hi) : best[l];
}

// One EV_CHUNK: jobs [0, nj) (chunk-relative), records rc[0 .. nr)
void gate_chunk(const DiagTables &P, const int32_t *rows, int m, const avx2_job *jb, size_t nj, const avx2_rec *rc,
                size_t nr, uint32_t jbase, const uint8_t *arena, const uint64_t *off, int32_t open, int32_t ext,
                int32_t gate, int regions, avx2_cand *out, uint64_t *nout, uint64_t *chains, int32_t *scores_out,
                UngScratch &s) {
  // counting sort by window start, so a batch's lanes cover nearby rows
  s.cnt.assign((size_t)m + 2, 0);
  for (size_t j = 0; j < nj; j++) s.cnt[std::min<uint32_t>(jb[j].lo, (uint32_t)m) + 1]++;
  for (size_t k = 1; k < s.cnt.size(); k++) s.cnt[k] += s.cnt[k - 1];
  s.order.resize(nj);
  for (size_t j = 0; j < nj; j++) s.order[s.cnt[std::min<uint32_t>(jb[j].lo, (uint32_t)m)]++] = (uint32_t)j;
  s.sc.resize(nj);
  Lane lanes[kDiagLanes];
  int32_t lo_out[kDiagLanes];
  for (size_t c0 = 0; c0 < nj; c0 += kDiagLanes) {
    const int nl = (int)std::min<size_t>(kDiagLanes, nj - c0);
    for (int l = 0; l < nl; l++) {
      const avx2_job &J = jb[s.order[c0 + l]];
      lanes[l] = Lane{arena + off[J.tid], J.diag, J.lo, std::max(J.hi, J.lo)};
    }
    score_lanes(P, rows, m, lanes, nl, lo_out, s);
    for (int l = 0; l < nl; l++) s.sc[s.order[c0 + l]] = lo_out[l];
  }
  if (scores_out) std::memcpy(scores_out, s.sc.data(), nj * 4);
  // chains and the gate per target (src/prefilter.rs:1577-1657)
  for (size_t r = 0; r < nr; r++) {
    const uint32_t j0 = rc[r].j0 - jbase;
    const size_t ns = std::min<size_t>(rc[r].j1 - rc[r].j0, 16);
    if (ns == 0) continue;
    const int32_t *sc = s.sc.data() + j0;
    if (ns == 1 && sc[0] < gate) continue;
    const uint8_t *t = arena + off[rc[r].tid];
    int32_t best = 0;
    size_t bk = 0;
    uint32_t qs[16], qe[16];
    bool have[16] = {false};
    for (size_t k = 0; k < ns; k++)
      if (sc[k] > best) {
        best = sc[k];
        bk = k;
      }
    auto bounds = [&](size_t k) {
      if (!have[k]) {
        const avx2_job &J = jb[j0 + k];
        kadane_bounds(rows, t, J.diag, J.lo, std::max(J.hi, J.lo), &qs[k], &qe[k]);
        have[k] = true;
      }
    };
    int32_t cbest = INT32_MIN;
    size_t cx = 0, cy = 0;
    for (size_t x = 0; ns > 1 && x < ns; x++)
      for (size_t y = 0; y < ns; y++) {
        if (x == y) continue;
        const int64_t dx = jb[j0 + x].diag, dy = jb[j0 + y].diag;
        if (sc[x] <= 0 || sc[y] <= 0 || dx == dy) continue;
        const uint64_t ad = (uint64_t)(dy > dx ? dy - dx : dx - dy);
        const int32_t cv = sc[x] + sc[y] - (open + (int32_t)(ad - 1) * ext);
        if (cv <= best || cv <= cbest) continue;
        bounds(x);
        bounds(y);
        if (qe[x] < qs[y] && (int64_t)qe[x] + dx < (int64_t)qs[y] + dy) {
          cbest = cv;
          cx = x;
          cy = y;
        }
      }
    int32_t score = best;
    size_t used[2] = {bk, bk};
    if (cbest > best) {
      (*chains)++;
      score = cbest;
      used[0] = cx;
      used[1] = cy;
    }
    if (score < gate) continue;
    uint32_t tlo = 0, thi = UINT32_MAX;
    if (regions) {  // region, src/prefilter.rs:1269-1278
      tlo = UINT32_MAX;
      thi = 0;
      for (size_t u : used) {
        bounds(u);
        const int64_t a = (int64_t)qs[u] + jb[j0 + u].diag, b = (int64_t)qe[u] + jb[j0 + u].diag;
        tlo = std::min(tlo, (uint32_t)std::max<int64_t>(a, 0));
        thi = std::max(thi, (uint32_t)std::max<int64_t>(b, 0));
      }
    }
    out[(*nout)++] = avx2_cand{rc[r].tid, rc[r].hits, jb[j0 + bk].diag, score, tlo, thi};
  }
}

}  // namespace

extern "C" double avx2_score_and_gate(uint32_t nq, const int32_t *const *rows, const uint32_t *qlen,
                                      const uint64_t *job_off, const avx2_job *jobs, const uint64_t *rec_off,
                                      const avx2_rec *recs, const uint8_t *arena, const uint64_t *off,
                                      const int32_t *open, const int32_t *ext, const int32_t *gate, int regions,
                                      int threads, uint64_t *n_cand, avx2_cand *cands, uint64_t *chains,
                                      int32_t *job_scores) {
  // each query's records must index its jobs contiguously and in order
  for (uint32_t q = 0; q < nq; q++) {
    uint32_t next = 0;
    for (uint64_t r = rec_off[q]; r < rec_off[q + 1]; r++) {
      if (recs[r].j0 != next || recs[r].j1 < recs[r].j0) return -1.0;
      next = recs[r].j1;
    }
    if (next != job_off[q + 1] - job_off[q]) return -1.0;
  }
  const auto t0 = std::chrono::steady_clock::now();
#pragma omp parallel num_threads(threads)
  {
    UngScratch s;
#pragma omp for schedule(dynamic, 1)
    for (int64_t q = 0; q < (int64_t)nq; q++) {
      const int m = (int)qlen[q];
      const DiagTables P = build_diag(rows[q], m);
      const avx2_job *jb = jobs + job_off[q];
      const avx2_rec *rc = recs + rec_off[q];
      const uint64_t nr = rec_off[q + 1] - rec_off[q];
      uint64_t nc = 0, ch = 0;
      // records accumulate until their jobs reach EV_CHUNK, then the chunk
      // is scored and gated (src/prefilter.rs:1488-1491)
      uint64_t r0 = 0;
      for (uint64_t r = 0; r < nr; r++) {
        const bool last = r + 1 == nr;
        if (!last && rc[r].j1 - rc[r0].j0 < kEvChunk) continue;
        const uint32_t jbase = rc[r0].j0;
        gate_chunk(P, rows[q], m, jb + jbase, rc[r].j1 - jbase, rc + r0, r - r0 + 1, jbase, arena, off, open[q],
                   ext[q], gate[q], regions, cands + rec_off[q], &nc, &ch,
                   job_scores ? job_scores + job_off[q] + jbase : nullptr, s);
        r0 = r + 1;
      }
      n_cand[q] = nc;
      chains[q] = ch;
    }
  }
  const auto t1 = std::chrono::steady_clock::now();
  return std::chrono::duration<double>(t1 - t0).count();
}
