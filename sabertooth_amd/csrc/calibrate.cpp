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
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "dp_launch.h"
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

// body(lo, hi) over [0, n) split into contiguous runs, one per thread
template <class F>
void parallel_runs(uint64_t n, F body) {
  const uint64_t hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  const uint64_t nt = std::min<uint64_t>(hw, std::max<uint64_t>(1, n / 256));
  if (nt <= 1) {
    body(0, n);
    return;
  }
  std::vector<std::thread> th;
  for (uint64_t t = 0; t < nt; t++) th.emplace_back(body, n * t / nt, n * (t + 1) / nt);
  for (auto &x : th) x.join();
}

// profile memory of one batch of simulated queries (int16 + int32 entries)
constexpr uint64_t kSimProfileBytes = 1ull << 30;

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

int st_simulate_scores(st_ctx *ctx, int32_t gap_open, int32_t gap_extend, int comp_bias, uint64_t m, uint64_t n,
                       uint64_t samples, uint64_t seed, uint64_t cell_id, double *out) {
  if (!ctx || (samples && !out)) return fail(ST_ERR_INVALID, "st_simulate_scores: null argument");
  if (m > (1u << 24) || n > (1u << 24)) return fail(ST_ERR_INVALID, "st_simulate_scores: length beyond 2^24");
  if (samples == 0) return ST_OK;
  if (m == 0 || n == 0) {
    // an empty query or target scores 0
    for (uint64_t i = 0; i < samples; i++) out[i] = 0.0;
    return ST_OK;
  }
  // the pairs, each from its own stream (query first, then target), and the
  // queries' bias corrections, on host threads
  const Background bg;
  std::vector<uint8_t> qs(samples * m), ts(samples * n);
  std::vector<int32_t> bias(comp_bias ? samples * m : 0);
  std::atomic<int> brc{ST_OK};
  parallel_runs(samples, [&](uint64_t lo, uint64_t hi) {
    for (uint64_t i = lo; i < hi; i++) {
      SplitMix64 g{stream_seed(seed, cell_id, i)};
      for (uint64_t k = 0; k < m; k++) qs[i * m + k] = bg.sample(g);
      for (uint64_t k = 0; k < n; k++) ts[i * n + k] = bg.sample(g);
      if (comp_bias && st_local_bias_correction(&qs[i * m], m, 0, &bias[i * m]) != ST_OK) brc = ST_ERR_INVALID;
    }
  });
  if (brc) return fail(brc, "st_simulate_scores: bias correction failed");
  // the score table indexed by code (code 21 scores 0)
  int8_t b21[21 * 21];
  st_blosum62(b21, nullptr, nullptr, nullptr);
  int8_t mat[kCodes * kCodes] = {};
  for (int a = 0; a < 21; a++)
    for (int c = 0; c < 21; c++) mat[a * kCodes + c] = b21[a * 21 + c];
  // rows in 64-row steps: the planner keeps every pass inside them
  const uint32_t ld = (uint32_t)((m + 63) & ~63ull);
  const uint64_t per = (uint64_t)kCodes * ld;
  const uint64_t chunk = std::max<uint64_t>(1, std::min<uint64_t>(samples, kSimProfileBytes / (per * 6)));
  DevBuf<int8_t> dmat;
  DevBuf<uint8_t> dq;
  DevBuf<int32_t> dbias;
  DevBuf<int16_t> p16;
  DevBuf<int32_t> p32;
  std::vector<int32_t> sc(chunk);
  for (uint64_t s0 = 0; s0 < samples; s0 += chunk) {
    const uint64_t cnt = std::min(chunk, samples - s0);
    std::vector<uint64_t> toff(cnt + 1);
    for (uint64_t i = 0; i <= cnt; i++) toff[i] = i * n;
    st_db *db = nullptr;
    int rc = st_db_create(ctx, ts.data() + s0 * n, cnt * n, toff.data(), cnt, &db);
    if (rc) return rc;
    {
      std::lock_guard<std::mutex> lk(ctx->mu);
      if (hipSetDevice(ctx->device) != hipSuccess || dmat.upload(mat, sizeof mat, ctx->stream) != hipSuccess ||
          dq.upload(qs.data() + s0 * m, cnt * m, ctx->stream) != hipSuccess ||
          (comp_bias && dbias.upload(bias.data() + s0 * m, cnt * m, ctx->stream) != hipSuccess) ||
          p16.reserve(cnt * per) != hipSuccess || p32.reserve(cnt * per) != hipSuccess ||
          launch_sim_prof(dq.p, comp_bias ? dbias.p : nullptr, dmat.p, (uint32_t)m, ld, (uint32_t)cnt, p16.p, p32.p,
                          ctx->stream) != hipSuccess ||
          hipStreamSynchronize(ctx->stream) != hipSuccess) {
        st_db_destroy(db);
        return fail(ST_ERR_DEVICE, "st_simulate_scores: profile build failed");
      }
    }
    // one view per sample into the profiles just built; BLOSUM62 plus a
    // bias of at most a few units fits int16
    std::vector<st_qprof> views(cnt);
    std::vector<const st_qprof *> vq(cnt);
    std::vector<uint64_t> off(cnt + 1);
    std::vector<uint32_t> tids(cnt);
    std::vector<int32_t> op(cnt, gap_open), ex(cnt, gap_extend);
    for (uint64_t i = 0; i < cnt; i++) {
      st_qprof &v = views[i];
      v.ctx = ctx;
      v.m = (uint32_t)m;
      v.ld = ld;
      v.fits16 = true;
      v.fits8 = false;
      v.view.prof16 = p16.p + i * per;
      v.view.prof32 = p32.p + i * per;
      v.view.prof8 = nullptr;
      v.view.m = v.m;
      v.view.ld = ld;
      vq[i] = &v;
      off[i] = i;
      tids[i] = (uint32_t)i;
    }
    off[cnt] = cnt;
    rc = score_multi(ctx, db, vq.data(), cnt, off.data(), tids.data(), op.data(), ex.data(), sc.data());
    st_db_destroy(db);
    if (rc) return rc;
    for (uint64_t i = 0; i < cnt; i++) out[s0 + i] = (double)sc[i];
  }
  return ST_OK;
}

int st_calibrate_sequence(st_ctx *ctx, int32_t gap_open, int32_t gap_extend, int comp_bias, const uint64_t *lengths,
                          uint64_t n_lengths, uint64_t samples, uint64_t seed, st_gapped_params *out,
                          st_calibration_cell *cells, double *ln_k_rms) {
  if (!ctx || !out || (n_lengths && !lengths)) return fail(ST_ERR_INVALID, "st_calibrate_sequence: null argument");
  if (n_lengths == 0) return fail(ST_ERR_INVALID, "st_calibrate_sequence: no lengths");
  std::vector<uint64_t> cm(n_lengths), cn(n_lengths), off(n_lengths + 1, 0);
  std::vector<double> xs(n_lengths * samples);
  for (uint64_t c = 0; c < n_lengths; c++) {
    cm[c] = cn[c] = lengths[c];
    off[c + 1] = off[c] + samples;
    if (int rc = st_simulate_scores(ctx, gap_open, gap_extend, comp_bias, lengths[c], lengths[c], samples, seed, c,
                                    xs.data() + off[c]))
      return rc;
  }
  return st_fit_calibration(cm.data(), cn.data(), off.data(), xs.data(), n_lengths, 200, out, cells, ln_k_rms);
}

}  // extern "C"
