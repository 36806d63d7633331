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
