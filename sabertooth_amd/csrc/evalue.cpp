// Statistics of reported hits on the host: the calibrated gapped parameter
// table, BLAST length adjustment, effective search space, bit scores,
// E-values, the minimum raw score for an E-value cut-off and the Gumbel
// maximum-likelihood fit (src/evalue.rs:130-466).  f64 throughout, in the
// reference's operation order with the C library's exp/log (what Rust's f64
// methods call on x86_64 Linux); compiled with -ffp-contract=off.
#include <cmath>
#include <cstring>
#include <limits>
#include <string>
#include <algorithm>
#include <vector>

#include "host_internal.h"
#include "../../include/sabertooth_host.h"

namespace {

// SEQUENCE_TABLE (src/evalue.rs:274-305): BLOSUM62 (bit factor 2, bias
// -0.2), gaps 11/1, without and with composition-bias correction.
struct TableRow {
  int32_t open, ext;
  int comp_bias;
  double lambda, k, adl, beta;
};
const TableRow kTable[] = {
    {11, 1, 0, 0.3666, 0.1177, 2.7, -13.0},
    {11, 1, 1, 0.3649, 0.1076, 2.4, -12.0},
};

}  // namespace

using namespace st;

extern "C" {

int st_evalue_table(int32_t gap_open, int32_t gap_extend, int comp_bias, st_gapped_params *out) {
  if (!out) return fail(ST_ERR_INVALID, "st_evalue_table: null argument");
  for (const TableRow &r : kTable)
    if (r.open == gap_open && r.ext == gap_extend && r.comp_bias == (comp_bias ? 1 : 0)) {
      out->lambda = r.lambda;
      out->k = r.k;
      out->has_length = 1;
      out->reserved = 0;
      out->alpha_d_lambda = r.adl;
      out->beta = r.beta;
      return ST_OK;
    }
  return fail(ST_ERR_INVALID, "st_evalue_table: no calibrated statistics for gap costs " +
                                  std::to_string(gap_open) + "/" + std::to_string(gap_extend) +
                                  " with comp-bias-corr " + std::to_string(comp_bias ? 1 : 0));
}

int st_evalue_params_from(const st_gapped_params *p, uint64_t db_residues, uint64_t db_seqs,
                          st_evalue_params *out) {
  // EValueParams::from_params (src/evalue.rs:333-343)
  if (!p || !out) return fail(ST_ERR_INVALID, "st_evalue_params_from: null argument");
  std::memset(out, 0, sizeof *out);
  out->lambda = p->lambda;
  out->k = p->k;
  out->log_k = std::log(p->k);
  out->db_residues = (double)(db_residues ? db_residues : 1);
  out->db_seqs = (double)(db_seqs ? db_seqs : 1);
  out->has_length = p->has_length ? 1 : 0;
  out->alpha_d_lambda = p->alpha_d_lambda;
  out->beta = p->beta;
  return ST_OK;
}

int st_evalue_params_profile(uint64_t db_residues, st_evalue_params *out) {
  // EValueParams::for_profile (src/evalue.rs:391-403): ungapped ln2/8, K 0.041
  if (!out) return fail(ST_ERR_INVALID, "st_evalue_params_profile: null argument");
  std::memset(out, 0, sizeof *out);
  out->lambda = M_LN2 / 8.0;
  out->k = 0.041;
  out->log_k = std::log(0.041);
  out->db_residues = (double)(db_residues ? db_residues : 1);
  out->db_seqs = 1.0;
  return ST_OK;
}

double st_length_adjustment(double k, double alpha_d_lambda, double beta, double m, double n_res, double n_seqs) {
  // length_adjustment (src/evalue.rs:193-237): integer part of the fixed point
  // of l = (a/l)(ln K + ln((m - l)(N - n l))) + b, bounded so that the
  // effective space stays >= max(m, N) / K
  const double log_k = std::log(k);
  const double a = n_seqs;
  const double mb = m * n_seqs + n_res;
  const double c = n_res * m - std::fmax(m, n_res) / k;
  if (c < 0.0) return 0.0;
  const double ell_max = 2.0 * c / (mb + std::sqrt(std::fmax(mb * mb - 4.0 * a * c, 0.0)));
  double lo = 0.0, hi = ell_max, ell = 0.0;
  bool converged = false;
  for (int i = 1; i <= 20; i++) {
    const double ss = (m - ell) * (n_res - n_seqs * ell);
    const double bar = alpha_d_lambda * (log_k + std::log(std::fmax(ss, 1.0))) + beta;
    if (bar >= ell) {
      lo = ell;
      if (bar - lo <= 1.0) {
        converged = true;
        break;
      }
      if (lo == hi) break;
    } else {
      hi = ell;
    }
    if (lo <= bar && bar <= hi)
      ell = bar;
    else if (i == 1)
      ell = hi;
    else
      ell = 0.5 * (lo + hi);
  }
  double out = std::floor(lo);
  if (converged) {
    const double up = std::ceil(lo);
    if (up <= hi) {
      const double ss = (m - up) * (n_res - n_seqs * up);
      if (alpha_d_lambda * (log_k + std::log(std::fmax(ss, 1.0))) + beta >= up) out = up;
    }
  }
  return std::fmax(out, 0.0);
}

double st_search_space(const st_evalue_params *p, uint64_t query_len) {
  // EValueParams::search_space (src/evalue.rs:427-438)
  if (!p) return 0.0;
  const double m = (double)(query_len ? query_len : 1);
  if (!p->has_length) return m * p->db_residues;
  const double ell = st_length_adjustment(p->k, p->alpha_d_lambda, p->beta, m, p->db_residues, p->db_seqs);
  const double m_eff = std::fmax(std::fmax(m - ell, 1.0 / p->k), 1.0);
  const double n_eff = std::fmax(std::fmax(p->db_residues - p->db_seqs * ell, p->db_residues * 0.01), 1.0);
  return m_eff * n_eff;
}

double st_bit_score(const st_evalue_params *p, int32_t raw) {
  // (lambda S - ln K) / ln 2 (src/evalue.rs:420-423)
  return p ? (p->lambda * (double)raw - p->log_k) / M_LN2 : 0.0;
}

double st_evalue_in_space(const st_evalue_params *p, int32_t raw, double space) {
  // K space e^(-lambda S) (src/evalue.rs:447-450)
  return p ? p->k * space * std::exp(-p->lambda * (double)raw) : 0.0;
}

double st_evalue(const st_evalue_params *p, int32_t raw, uint64_t query_len) {
  return st_evalue_in_space(p, raw, st_search_space(p, query_len));
}

int32_t st_min_score_for(const st_evalue_params *p, double max_evalue, double space) {
  // smallest raw score with E <= max_evalue (src/evalue.rs:453-461)
  if (!p) return 1;
  const double s = std::ceil(std::log(p->k * space / max_evalue) / p->lambda);
  if (!std::isfinite(s)) return 1;
  return (int32_t)std::fmin(std::fmax(s, 1.0), (double)std::numeric_limits<int32_t>::max() / 2.0);
}

double st_gumbel_mu_given_lambda(const double *xs, uint64_t n, double lambda) {
  // ML location for a fixed lambda (src/evalue.rs:131-135)
  if (!xs || n == 0) return std::nan("");
  double xmin = INFINITY;
  for (uint64_t i = 0; i < n; i++) xmin = std::fmin(xmin, xs[i]);
  double s = 0.0;
  for (uint64_t i = 0; i < n; i++) s += std::exp(-lambda * (xs[i] - xmin));
  return xmin - std::log(s / (double)n) / lambda;
}

int st_fit_gumbel_ml(const double *xs, uint64_t n, st_gumbel *out) {
  // fit_gumbel_ml (src/evalue.rs:139-182): Lawless' ML equation for lambda
  // by bisection from a moments bracket; ST_ERR_INVALID when no fit exists
  if (!out || (n && !xs)) return fail(ST_ERR_INVALID, "st_fit_gumbel_ml: null argument");
  if (n < 10) return fail(ST_ERR_INVALID, "st_fit_gumbel_ml: fewer than 10 samples");
  double mean = 0.0;
  for (uint64_t i = 0; i < n; i++) mean += xs[i];
  mean /= (double)n;
  double var = 0.0;
  for (uint64_t i = 0; i < n; i++) var += (xs[i] - mean) * (xs[i] - mean);
  var /= ((double)n - 1.0);
  if (std::isnan(var) || var <= 0.0) return fail(ST_ERR_INVALID, "st_fit_gumbel_ml: degenerate sample");
  double xmin = INFINITY;
  for (uint64_t i = 0; i < n; i++) xmin = std::fmin(xmin, xs[i]);
  auto g = [&](double lam) {
    double s0 = 0.0, s1 = 0.0;
    for (uint64_t i = 0; i < n; i++) {
      const double w = std::exp(-lam * (xs[i] - xmin));
      s0 += w;
      s1 += xs[i] * w;
    }
    return 1.0 / lam - mean + s1 / s0;
  };
  const double lam0 = M_PI / (std::sqrt(var) * std::sqrt(6.0));
  double lo = lam0 / 10.0, hi = lam0 * 10.0;
  if (g(lo) < 0.0 || g(hi) > 0.0) return fail(ST_ERR_INVALID, "st_fit_gumbel_ml: no root in the bracket");
  for (int it = 0; it < 200; it++) {
    const double mid = 0.5 * (lo + hi);
    if (g(mid) > 0.0)
      lo = mid;
    else
      hi = mid;
    if (hi - lo < 1e-12 * hi) break;
  }
  out->lambda = 0.5 * (lo + hi);
  out->mu = st_gumbel_mu_given_lambda(xs, n, out->lambda);
  out->n = n;
  return ST_OK;
}

int st_fit_calibration(const uint64_t *cell_m, const uint64_t *cell_n, const uint64_t *score_off,
                       const double *scores, uint64_t ncells, uint64_t lambda_min_len, st_gapped_params *out,
                       st_calibration_cell *cells, double *ln_k_rms) {
  // fit_calibration (src/evalue.rs:576-668): one lambda shared by the cells
  // whose shorter side is at least lambda_min_len (every cell when none
  // is), each cell's location with it (+1/2 for integer scores), then the
  // length model (alpha/lambda, beta) on the reference's grid that makes
  // ln K the same across cells, K = e^(mean ln K)
  if (!out || (ncells && (!cell_m || !cell_n || !score_off || !scores)))
    return fail(ST_ERR_INVALID, "st_fit_calibration: null argument");
  if (ncells == 0) return fail(ST_ERR_INVALID, "st_fit_calibration: no cells");
  if (score_off[0] != 0) return fail(ST_ERR_INVALID, "st_fit_calibration: score_off[0] must be 0");
  for (uint64_t c = 0; c < ncells; c++)
    if (score_off[c + 1] < score_off[c]) return fail(ST_ERR_INVALID, "st_fit_calibration: score_off decreases");
  // cells by (m, n), ties in input order (a stable sort, as sort_by_key)
  std::vector<uint64_t> ord(ncells);
  for (uint64_t c = 0; c < ncells; c++) ord[c] = c;
  std::stable_sort(ord.begin(), ord.end(), [&](uint64_t a, uint64_t b) {
    return cell_m[a] != cell_m[b] ? cell_m[a] < cell_m[b] : cell_n[a] < cell_n[b];
  });
  auto xs_of = [&](uint64_t c) { return scores + score_off[c]; };
  auto n_of = [&](uint64_t c) { return score_off[c + 1] - score_off[c]; };
  std::vector<uint64_t> pool;
  for (uint64_t c : ord)
    if (std::min(cell_m[c], cell_n[c]) >= lambda_min_len) pool.push_back(c);
  if (pool.empty()) pool = ord;
  for (uint64_t c : ord)
    if (n_of(c) == 0) return fail(ST_ERR_INVALID, "st_fit_calibration: a cell without scores");
  auto loglik = [&](double lam) {
    double ll = 0.0;
    for (uint64_t c : pool) {
      const double *xs = xs_of(c);
      const double mu = st_gumbel_mu_given_lambda(xs, n_of(c), lam);
      for (uint64_t i = 0; i < n_of(c); i++) {
        const double z = lam * (xs[i] - mu);
        ll += std::log(lam) - z - std::exp(-z);
      }
    }
    return ll;
  };
  // golden-section search for the maximum over lambda in [0.05, 1.5]
  double a = 0.05, b = 1.5;
  const double gr = (std::sqrt(5.0) - 1.0) / 2.0;
  double c1 = b - gr * (b - a), c2 = a + gr * (b - a);
  double f1 = loglik(c1), f2 = loglik(c2);
  for (int it = 0; it < 80; it++) {
    if (f1 < f2) {
      a = c1;
      c1 = c2;
      f1 = f2;
      c2 = a + gr * (b - a);
      f2 = loglik(c2);
    } else {
      b = c2;
      c2 = c1;
      f2 = f1;
      c1 = b - gr * (b - a);
      f1 = loglik(c1);
    }
  }
  const double lambda = 0.5 * (a + b);
  std::vector<st_calibration_cell> fitted(ncells);
  for (uint64_t k = 0; k < ncells; k++) {
    const uint64_t c = ord[k];
    const double *xs = xs_of(c);
    const uint64_t n = n_of(c);
    double mean = 0.0;
    for (uint64_t i = 0; i < n; i++) mean += xs[i];
    mean /= (double)n;
    double ss = 0.0;
    for (uint64_t i = 0; i < n; i++) ss += (xs[i] - mean) * (xs[i] - mean);
    st_gumbel free_fit;
    if (int rc = st_fit_gumbel_ml(xs, n, &free_fit)) return rc;
    st_calibration_cell &o = fitted[k];
    o.m = cell_m[c];
    o.n = cell_n[c];
    o.samples = n;
    o.mean = mean;
    o.sd = std::sqrt(ss / ((double)n - 1.0));
    o.lambda_ml = free_fit.lambda;
    o.mu_ml = free_fit.mu;
    o.mu = st_gumbel_mu_given_lambda(xs, n, lambda) + 0.5;
  }
  // the length model: alpha/lambda in 0.1 steps over [0, 20], beta in unit
  // steps over [-80, 20] (accumulated as the reference accumulates them)
  bool found = false;
  double bvar = 0.0, badl = 0.0, bbeta = 0.0, bmean = 0.0;
  std::vector<double> lnks;
  lnks.reserve(ncells);
  for (double adl = 0.0; adl <= 20.0 + 1e-9; adl += 0.1) {
    for (double beta = -80.0; beta <= 20.0 + 1e-9; beta += 1.0) {
      lnks.clear();
      bool ok = true;
      for (const st_calibration_cell &o : fitted) {
        const double ell = std::fmax(adl * lambda * o.mu + beta, 0.0);
        const double me = (double)o.m - ell, ne = (double)o.n - ell;
        if (me < 1.0 || ne < 1.0) {
          ok = false;
          break;
        }
        lnks.push_back(lambda * o.mu - std::log(me * ne));
      }
      if (!ok || lnks.empty()) continue;
      double mean = 0.0;
      for (double v : lnks) mean += v;
      mean /= (double)lnks.size();
      double var = 0.0;
      for (double v : lnks) var += (v - mean) * (v - mean);
      var /= (double)lnks.size();
      if (!found || var < bvar) {
        found = true;
        bvar = var;
        badl = adl;
        bbeta = beta;
        bmean = mean;
      }
    }
  }
  if (!found) return fail(ST_ERR_INVALID, "st_fit_calibration: no length model fits the cells");
  std::memset(out, 0, sizeof *out);
  out->lambda = lambda;
  out->k = std::exp(bmean);
  out->has_length = 1;
  out->alpha_d_lambda = std::round(badl * 10.0) / 10.0;
  out->beta = bbeta;
  if (cells) std::memcpy(cells, fitted.data(), ncells * sizeof *cells);
  if (ln_k_rms) *ln_k_rms = std::sqrt(bvar);
  return ST_OK;
}

double st_ungapped_lambda(void) {
  // solve_lambda (src/evalue.rs:466-498): sum p_i p_j e^(lambda s_ij) = 1 by
  // bisection over the integer BLOSUM62 (diagnostic statistics only)
  int8_t sc[21 * 21];
  double pb[21];
  st_blosum62(sc, pb, nullptr, nullptr);
  auto f = [&](double lam) {
    double acc = 0.0;
    for (int i = 0; i < 20; i++)
      for (int j = 0; j < 20; j++) acc += pb[i] * pb[j] * std::exp(lam * (double)sc[i * 21 + j]);
    return acc - 1.0;
  };
  double lo = 1e-6, hi = 2.0;
  if (f(lo) > 0.0) return 0.3;
  for (int it = 0; it < 60 && f(hi) < 0.0; it++) hi *= 2.0;
  for (int it = 0; it < 100; it++) {
    const double mid = 0.5 * (lo + hi);
    if (f(mid) < 0.0)
      lo = mid;
    else
      hi = mid;
  }
  return 0.5 * (lo + hi);
}

}  // extern "C"
