// Per-profile statistics (calibrate_profile, src/evalue.rs:714-749): `samples`
// random background targets, lengths drawn from a sample of the database's
// lengths, are scored against the profile on the GPU with the search's gap
// costs; a Gumbel is fitted to the scores by maximum likelihood on the host
// and expressed as Karlin-Altschul (lambda, K) over the search space m x N:
// K = e^(lambda (mu + 1/2)) / (m * mean length).
//
// The random targets follow the reference's streams exactly: sample i uses
// SplitMix64(stream_seed(seed, 0xBEEF, i)) (src/evalue.rs:52-82) to pick a
// length and draw residues from the BLOSUM62 background (BackgroundSampler,
// src/evalue.rs:85-114), so the scores, and the fit, are the reference's.
#include <cmath>
#include <cstring>
#include <vector>

#include "host_internal.h"
#include "../../include/sabertooth_host.h"

namespace {

struct SplitMix64 {
  uint64_t s;
  uint64_t next_u64() {
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double next_f64() { return (double)(next_u64() >> 11) / (double)(1ull << 53); }
  uint64_t below(uint64_t n) { return next_u64() % n; }
};

uint64_t stream_seed(uint64_t seed, uint64_t c, uint64_t i) {
  SplitMix64 g{seed ^ (c * 0xA24BAED4963EE407ull) ^ (i * 0x9FB21C651E98DF25ull)};
  return g.next_u64();
}

struct Background {
  double cum[20];
  Background() {
    double pb[21];
    st_blosum62(nullptr, pb, nullptr, nullptr);
    double total = 0.0;
    for (int i = 0; i < 20; i++) total += pb[i];
    double acc = 0.0;
    for (int i = 0; i < 20; i++) {
      acc += pb[i] / total;
      cum[i] = acc;
    }
    cum[19] = 1.0;
  }
  uint8_t sample(SplitMix64 &g) const {
    const double u = g.next_f64();
    for (int i = 0; i < 20; i++)
      if (u < cum[i]) return (uint8_t)i;
    return 19;
  }
};

}  // namespace

using namespace st;

extern "C" {

int st_calibration_targets(const uint64_t *lengths, uint64_t n_lengths, uint64_t samples, uint64_t seed,
                           uint8_t *arena, uint64_t cap, uint64_t *offsets) {
  if (!offsets || (samples && (!lengths || n_lengths == 0)))
    return fail(ST_ERR_INVALID, "st_calibration_targets: null argument or empty length sample");
  const Background bg;
  offsets[0] = 0;
  for (uint64_t i = 0; i < samples; i++) {
    SplitMix64 g{stream_seed(seed, 0xBEEF, i)};
    const uint64_t n = std::max<uint64_t>(lengths[g.below(n_lengths)], 1);
    const uint64_t o = offsets[i];
    if (arena && o + n <= cap)
      for (uint64_t k = 0; k < n; k++) arena[o + k] = bg.sample(g);
    offsets[i + 1] = o + n;
  }
  return ST_OK;
}

int st_calibrate_profile(st_ctx *ctx, const st_qprof *qp, int32_t gap_open, int32_t gap_extend,
                         const uint64_t *lengths, uint64_t n_lengths, uint64_t samples, uint64_t seed,
                         st_gapped_params *out, st_gumbel *fit) {
  if (!ctx || !qp || !out) return fail(ST_ERR_INVALID, "st_calibrate_profile: null argument");
  if (n_lengths == 0 || !lengths || qp->m == 0)
    return fail(ST_ERR_INVALID, "st_calibrate_profile: empty length sample or empty profile");
  std::vector<uint64_t> off(samples + 1);
  int rc = st_calibration_targets(lengths, n_lengths, samples, seed, nullptr, 0, off.data());
  if (rc) return rc;
  std::vector<uint8_t> arena(off[samples]);
  rc = st_calibration_targets(lengths, n_lengths, samples, seed, arena.data(), arena.size(), off.data());
  if (rc) return rc;
  st_db *db = nullptr;
  rc = st_db_create(ctx, arena.data(), arena.size(), off.data(), samples, &db);
  if (rc) return rc;
  std::vector<uint32_t> tids(samples);
  for (uint64_t i = 0; i < samples; i++) tids[i] = (uint32_t)i;
  std::vector<int32_t> sc(samples);
  rc = st_sw_score_batch(ctx, qp, db, tids.data(), samples, gap_open, gap_extend, sc.data(), nullptr);
  st_db_destroy(db);
  if (rc) return rc;
  std::vector<double> xs(samples);
  for (uint64_t i = 0; i < samples; i++) xs[i] = (double)sc[i];
  st_gumbel g;
  rc = st_fit_gumbel_ml(xs.data(), samples, &g);
  if (rc) return rc;
  uint64_t sum = 0;
  for (uint64_t k = 0; k < n_lengths; k++) sum += lengths[k];
  const double mean_len = (double)sum / (double)n_lengths;
  std::memset(out, 0, sizeof *out);
  out->lambda = g.lambda;
  out->k = std::exp(g.lambda * (g.mu + 0.5)) / ((double)qp->m * mean_len);
  out->has_length = 0;
  if (fit) *fit = g;
  return ST_OK;
}

}  // extern "C"
