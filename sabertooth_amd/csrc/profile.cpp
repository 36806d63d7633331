// PSSM construction on the host: MSA text -> numeric MSA -> int8 PSSM, and
// the score rows a profile query hands to the scoring path.
//
// Follows Profile::from_numeric_msa / parse_msa (src/profile.rs:90-195,
// 218-449): Henikoff position weights, weighted match frequencies, Neff from
// column entropy, substitution-matrix pseudocounts and round-half-away
// eighth-bit log-odds.  All profile arithmetic is f32 in the reference's
// order, with the C library's log2f/exp2f (the functions Rust's f32 methods
// call on x86_64 Linux); this file is compiled with -ffp-contract=off.
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "host_internal.h"
#include "../../include/sabertooth_host.h"

namespace st {
namespace {

constexpr uint8_t kGap = 0xFF;
constexpr int kAA = 20;

bool is_ws(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\v' || c == '\f' || c == '\r'; }

std::string trim(const std::string &s, bool left) {
  size_t a = 0, b = s.size();
  if (left)
    while (a < b && is_ws(s[a])) a++;
  while (b > a && is_ws(s[b - 1])) b--;
  return s.substr(a, b - a);
}

// parse_msa (src/profile.rs:372-449)
int parse(const char *text, size_t len, std::vector<std::vector<uint8_t>> &msa, std::string &err) {
  uint8_t a2n[256];
  st_aa2num(a2n);
  std::vector<std::string> raw;
  std::string cur;
  bool started = false;
  size_t pos = 0;
  while (pos <= len) {
    size_t e = pos;
    while (e < len && text[e] != '\n') e++;
    std::string line = trim(std::string(text + pos, e - pos), false);
    if (!line.empty() && line[0] == '>') {
      if (started) raw.push_back(cur);
      cur.clear();
      started = true;
    } else {
      cur += trim(line, true);
    }
    if (e >= len) break;
    pos = e + 1;
  }
  if (started) raw.push_back(cur);
  if (raw.empty()) {
    err = "MSA contains no sequences";
    return ST_ERR_INVALID;
  }
  bool a3m = false;
  for (const std::string &s : raw)
    for (char c : s) a3m |= c >= 'a' && c <= 'z';
  msa.clear();
  if (a3m) {
    // lowercase letters are insertions relative to the query: dropped
    for (const std::string &s : raw) {
      std::vector<uint8_t> row;
      for (char c : s) {
        if (c >= 'a' && c <= 'z') continue;
        if (c == '-' || c == '.')
          row.push_back(kGap);
        else if (c >= 'A' && c <= 'Z')
          row.push_back(a2n[(uint8_t)c]);
      }
      msa.push_back(std::move(row));
    }
  } else {
    // aligned FASTA: match columns are the first row's non-gap positions;
    // shorter rows read as gaps
    size_t width = 0;
    for (const std::string &s : raw) width = std::max(width, s.size());
    const std::string &first = raw[0];
    std::vector<size_t> cols;
    for (size_t i = 0; i < width; i++)
      if (i < first.size() && first[i] != '-' && first[i] != '.') cols.push_back(i);
    for (const std::string &s : raw) {
      std::vector<uint8_t> row;
      row.reserve(cols.size());
      for (size_t i : cols) {
        const char c = i < s.size() ? s[i] : '-';
        if (c == '-' || c == '.')
          row.push_back(kGap);
        else
          row.push_back(a2n[(uint8_t)((c >= 'a' && c <= 'z') ? c - 32 : c)]);
      }
      msa.push_back(std::move(row));
    }
  }
  const size_t q = msa[0].size();
  for (size_t k = 0; k < msa.size(); k++)
    if (msa[k].size() != q) {
      err = "MSA row " + std::to_string(k + 1) + " has " + std::to_string(msa[k].size()) +
            " match columns, expected " + std::to_string(q);
      return ST_ERR_INVALID;
    }
  return ST_OK;
}

}  // namespace
}  // namespace st

using namespace st;

extern "C" {

int st_parse_msa(const char *text, uint64_t len, uint8_t *rows, uint64_t cap, uint64_t *n_rows,
                 uint64_t *query_len) {
  if ((!text && len) || !n_rows || !query_len) return fail(ST_ERR_INVALID, "st_parse_msa: null argument");
  std::vector<std::vector<uint8_t>> msa;
  std::string err;
  const int rc = parse(text ? text : "", len, msa, err);
  if (rc) return fail(rc, "st_parse_msa: " + err);
  *n_rows = msa.size();
  *query_len = msa[0].size();
  const uint64_t need = (uint64_t)msa.size() * msa[0].size();
  if (!rows || cap < need) return ST_OK;  // sizes only
  for (size_t k = 0; k < msa.size(); k++) std::memcpy(rows + k * msa[0].size(), msa[k].data(), msa[k].size());
  return ST_OK;
}

int st_pssm_from_msa(const uint8_t *msa, uint64_t n_rows, uint64_t stride, uint64_t query_len, float pca,
                     float pcb, float score_bias, int8_t *pssm, float *prob_out, float *neff_out,
                     uint8_t *consensus) {
  if (n_rows == 0) return fail(ST_ERR_INVALID, "st_pssm_from_msa: cannot build a profile from an empty MSA");
  if (query_len == 0)
    return fail(ST_ERR_INVALID, "st_pssm_from_msa: cannot build a profile with zero match columns");
  if (!msa || !pssm) return fail(ST_ERR_INVALID, "st_pssm_from_msa: null argument");
  if (stride < query_len)
    return fail(ST_ERR_INVALID, "st_pssm_from_msa: MSA rows are shorter than query_len");
  const size_t N = n_rows, L = query_len;
  auto at = [&](size_t k, size_t p) { return msa[k * stride + p]; };
  double pb[21];
  float R[kAA * kAA];
  st_blosum62(nullptr, pb, R, nullptr);

  // 1. Henikoff position-based weights (src/profile.rs:218-255)
  std::vector<float> w(N, 1e-6f);
  std::vector<uint32_t> nres(N, 0);
  for (size_t k = 0; k < N; k++)
    for (size_t p = 0; p < L; p++) nres[k] += at(k, p) != kGap;
  for (size_t p = 0; p < L; p++) {
    int32_t nl[kAA] = {0};
    for (size_t k = 0; k < N; k++)
      if (at(k, p) < kAA) nl[at(k, p)]++;
    int32_t distinct = 0;
    for (int32_t c : nl) distinct += c > 0;
    if (!distinct) continue;
    for (size_t k = 0; k < N; k++) {
      const uint8_t c = at(k, p);
      if (c < kAA) w[k] += 1.0f / ((float)nl[c] * (float)distinct * ((float)nres[k] + 30.0f));
    }
  }
  // 2. normalise to sum 1 (src/profile.rs:121-127)
  float total = 0.0f;
  for (float x : w) total += x;
  for (float &x : w) x = total > 0.0f ? x / total : 1.0f / (float)N;

  // 3. weighted match frequencies, column-normalised (src/profile.rs:130-139)
  std::vector<float> freq(L * kAA, 0.0f);
  for (size_t p = 0; p < L; p++) {
    float *f = &freq[p * kAA];
    for (size_t k = 0; k < N; k++)
      if (at(k, p) < kAA) f[at(k, p)] += w[k];
    float sum = 0.0f;
    for (int a = 0; a < kAA; a++) sum += f[a];
    if (sum != 0.0f) {
      const float fac = 1.0f / sum;
      for (int a = 0; a < kAA; a++) f[a] *= fac;
    } else {
      for (int a = 0; a < kAA; a++) f[a] = (float)pb[a];
    }
  }

  // 4. Neff per column (src/profile.rs:258-303)
  float hmm = 0.0f;
  for (size_t p = 0; p < L; p++) {
    float ent = 0.0f;
    for (int a = 0; a < kAA; a++) {
      const float f = freq[p * kAA + a];
      if (f > 1e-10f) ent -= f * log2f(f);
    }
    hmm += exp2f(ent);
  }
  hmm /= (float)L;
  const float nlim = (hmm + 1.0f > 10.0f) ? hmm + 1.0f : 10.0f;
  const float scale = log2f((nlim - hmm) / (nlim - 1.0f));
  std::vector<float> neff(L);
  for (size_t p = 0; p < L; p++) {
    float wm = -1.0f / (float)N;
    for (size_t k = 0; k < N; k++)
      if (at(k, p) != kGap) wm += w[k];
    neff[p] = wm < 0.0f ? 1.0f : nlim - (nlim - 1.0f) * exp2f(scale * wm);
  }

  // 5. consensus: largest freq - background above 1e-8 (src/profile.rs:306-323)
  if (consensus)
    for (size_t p = 0; p < L; p++) {
      float best = 1e-8f;
      uint8_t arg = 20;
      for (int a = 0; a < kAA; a++) {
        const float d = freq[p * kAA + a] - (float)pb[a];
        if (d > best) {
          best = d;
          arg = (uint8_t)a;
        }
      }
      consensus[p] = arg;
    }

  // 6. substitution-matrix pseudocounts (src/profile.rs:145-169)
  std::vector<float> prob(freq);
  if (pca > 0.0f)
    for (size_t p = 0; p < L; p++) {
      const float *f = &freq[p * kAA];
      float g[kAA];
      for (int a = 0; a < kAA; a++) {
        float acc = 0.0f;
        for (int b = 0; b < kAA; b++) acc += R[a * kAA + b] * f[b];
        g[a] = acc;
      }
      float tau = pca / (1.0f + neff[p] / pcb);
      if (tau > 1.0f) tau = 1.0f;
      for (int a = 0; a < kAA; a++) prob[p * kAA + a] = (1.0f - tau) * f[a] + tau * g[a];
    }

  // 7. eighth-bit log-odds, round half away, clamp to int8 (src/profile.rs:326-346)
  for (size_t p = 0; p < L; p++)
    for (int a = 0; a < kAA; a++) {
      const float v = 8.0f * log2f(prob[p * kAA + a] / (float)pb[a]) + 8.0f * score_bias;
      float r = v < 0.0f ? v - 0.5f : v + 0.5f;
      r = r < -128.0f ? -128.0f : (r > 127.0f ? 127.0f : r);
      pssm[p * kAA + a] = std::isnan(r) ? 0 : (int8_t)r;
    }
  if (prob_out) std::memcpy(prob_out, prob.data(), L * kAA * sizeof(float));
  if (neff_out) std::memcpy(neff_out, neff.data(), L * sizeof(float));
  return ST_OK;
}

int st_pssm_rows(const int8_t *pssm, uint64_t query_len, int32_t *rows) {
  // ProfileScorer (src/align.rs:81-99, Profile::score src/profile.rs:70-79):
  // target X and codes >= 20 score 0
  if ((!pssm || !rows) && query_len) return fail(ST_ERR_INVALID, "st_pssm_rows: null argument");
  for (uint64_t i = 0; i < query_len; i++) {
    for (int r = 0; r < kAA; r++) rows[i * 21 + r] = pssm[i * kAA + r];
    rows[i * 21 + 20] = 0;
  }
  return ST_OK;
}

}  // extern "C"
