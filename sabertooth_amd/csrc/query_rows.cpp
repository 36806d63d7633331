// Score rows of a sequence query as the default search builds them: local
// composition-bias correction (src/compbias.rs:38-79), the SEG-like
// low-complexity mask (src/mask.rs:52-134) and the row assembly of
// search_query (src/search.rs:432-463): BLOSUM62 rows plus the per-column
// correction, then masked columns replaced by the matrix X row.  f64 in the
// reference's order with the C library's log2; compiled with
// -ffp-contract=off.
#include <cmath>
#include <vector>

#include "host_internal.h"
#include "../../include/sabertooth_host.h"

namespace {

int32_t round_half_away(double x) {
  return x >= 0.0 ? (int32_t)std::floor(x + 0.5) : -(int32_t)std::floor(-x + 0.5);
}

}  // namespace

using namespace st;

extern "C" {

int st_local_bias_correction(const uint8_t *query, uint64_t m, int penalize_only, int32_t *out) {
  if ((!query || !out) && m) return fail(ST_ERR_INVALID, "st_local_bias_correction: null argument");
  int8_t sc[21 * 21];
  double pb[21];
  st_blosum62(sc, pb, nullptr, nullptr);
  for (uint64_t i = 0; i < m; i++)
    if (query[i] > 20) return fail(ST_ERR_INVALID, "st_local_bias_correction: residue code above 20");
  // expected score of residue a against the background
  double expected[21];
  for (int a = 0; a < 21; a++) {
    double e = 0.0;
    for (int b = 0; b < 20; b++) e += pb[b] * (double)sc[a * 21 + b];
    expected[a] = e;
  }
  const uint64_t half = 20;  // window of 40 centred on i
  for (uint64_t i = 0; i < m; i++) {
    const uint64_t lo = i > half ? i - half : 0;
    const uint64_t hi = std::min<uint64_t>(i + half, m);
    const uint8_t a = query[i];
    int64_t sum = 0;
    for (uint64_t j = lo; j < hi; j++)
      if (j != i) sum += sc[a * 21 + query[j]];
    const double delta = expected[a] - (double)sum / (double)(hi - lo);
    const int32_t d = round_half_away(delta);
    out[i] = penalize_only ? std::min(d, 0) : d;
  }
  return ST_OK;
}

int st_low_complexity_mask(const uint8_t *seq, uint64_t n, uint32_t window, double locut, double hicut,
                           uint8_t *mask, uint64_t *n_masked) {
  if ((!seq || !mask) && n) return fail(ST_ERR_INVALID, "st_low_complexity_mask: null argument");
  for (uint64_t i = 0; i < n; i++) mask[i] = 0;
  if (n_masked) *n_masked = 0;
  const uint64_t w = window;
  if (w < 2 || n < w) return ST_OK;
  // Shannon entropy of each window from running c*log2(c) sums
  std::vector<double> clog(w + 1, 0.0);
  for (uint64_t c = 1; c <= w; c++) clog[c] = (double)c * std::log2((double)c);
  const double log_w = std::log2((double)w);
  const uint64_t nwin = n - w + 1;
  std::vector<double> ent(nwin);
  uint64_t counts[32] = {0};
  double sum_clog = 0.0;
  auto sym = [](uint8_t r) { return r < 31 ? r : 31; };
  for (uint64_t k = 0; k < w; k++) {
    const int s = sym(seq[k]);
    sum_clog -= clog[counts[s]];
    counts[s]++;
    sum_clog += clog[counts[s]];
  }
  ent[0] = log_w - sum_clog / (double)w;
  for (uint64_t i = 1; i < nwin; i++) {
    const int o = sym(seq[i - 1]);
    sum_clog -= clog[counts[o]];
    counts[o]--;
    sum_clog += clog[counts[o]];
    const int in = sym(seq[i + w - 1]);
    sum_clog -= clog[counts[in]];
    counts[in]++;
    sum_clog += clog[counts[in]];
    ent[i] = log_w - sum_clog / (double)w;
  }
  // trigger windows (entropy <= locut) extended over neighbours <= hicut
  uint64_t i = 0;
  while (i < nwin) {
    if (ent[i] <= locut + 1e-12) {
      uint64_t lo = i, hi = i;
      while (lo > 0 && ent[lo - 1] <= hicut + 1e-12) lo--;
      while (hi + 1 < nwin && ent[hi + 1] <= hicut + 1e-12) hi++;
      for (uint64_t k = lo; k < hi + w && k < n; k++) mask[k] = 1;
      i = hi + 1;
    } else {
      i++;
    }
  }
  if (n_masked) {
    uint64_t c = 0;
    for (uint64_t k = 0; k < n; k++) c += mask[k];
    *n_masked = c;
  }
  return ST_OK;
}

int st_query_rows(const uint8_t *query, uint64_t m, int comp_bias, int penalize_only, uint32_t mask_window,
                  double mask_locut, double mask_hicut, int32_t *rows, uint64_t *n_masked) {
  if ((!query || !rows) && m) return fail(ST_ERR_INVALID, "st_query_rows: null argument");
  if (n_masked) *n_masked = 0;
  int rc = st_seq_rows(query, m, rows);
  if (rc) return rc;
  if (comp_bias && m) {
    std::vector<int32_t> bias(m);
    rc = st_local_bias_correction(query, m, penalize_only, bias.data());
    if (rc) return rc;
    for (uint64_t i = 0; i < m; i++)
      for (int r = 0; r < 21; r++) rows[i * 21 + r] += bias[i];
  }
  if (mask_window && m) {
    std::vector<uint8_t> mk(m);
    uint64_t cnt = 0;
    rc = st_low_complexity_mask(query, m, mask_window, mask_locut, mask_hicut, mk.data(), &cnt);
    if (rc) return rc;
    int8_t sc[21 * 21];
    st_blosum62(sc, nullptr, nullptr, nullptr);
    for (uint64_t i = 0; i < m; i++)
      if (mk[i])
        for (int r = 0; r < 21; r++) rows[i * 21 + r] = sc[20 * 21 + r];
    if (n_masked) *n_masked = cnt;
  }
  return ST_OK;
}

}  // extern "C"
