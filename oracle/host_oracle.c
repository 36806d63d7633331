/*
 * CPU oracle, host-side rows of the scoring path — TEST INFRASTRUCTURE ONLY.
 *
 * Plain-C restatements of the reference's floating-point host functions in
 * the reference's operation order (f32 where the reference computes in f32,
 * glibc log2f/exp2f/exp/log as Rust's f32/f64 methods call on
 * x86_64-linux-gnu; built with -ffp-contract=off so no product is fused):
 *   - the PSSM builder (src/profile.rs:90-195, 218-349),
 *   - local composition-bias correction (src/compbias.rs:38-79),
 *   - the low-complexity mask (src/mask.rs:70-134),
 *   - E-value statistics (src/evalue.rs:130-240, 430-466, 714-749).
 * Only tests/ and the bench's cpu_baseline leg may load it.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define GAP 0xFF
#define NAA 20

/* ---- PSSM (src/profile.rs) ---------------------------------------------- */

/* msa: nrows rows of `stride` codes (0..19 residues, 20 = X, 0xFF = gap);
 * the first qlen columns are match states.  Returns 0, or -1 on the
 * reference's errors (empty MSA, zero columns; src/profile.rs:96-115). */
int or_pssm_from_msa(const uint8_t *msa, size_t nrows, size_t stride, size_t qlen, float pca, float pcb,
                     float score_bias, int8_t *pssm, float *prob_out, float *neff_out, uint8_t *cons_out) {
  if (nrows == 0 || qlen == 0 || stride < qlen) return -1;
  double pb[OR_NPROF];
  float r[NAA * NAA];
  or_blosum62(NULL, pb, r, NULL);
  float pbf[NAA];
  for (int a = 0; a < NAA; a++) pbf[a] = (float)pb[a];

  /* Henikoff position-based weights (src/profile.rs:218-255) */
  float *w = malloc(nrows * sizeof(float));
  uint32_t *nres = malloc(nrows * sizeof(uint32_t));
  for (size_t k = 0; k < nrows; k++) {
    w[k] = 1e-6f;
    uint32_t c = 0;
    for (size_t p = 0; p < qlen; p++)
      if (msa[k * stride + p] != GAP) c++;
    nres[k] = c;
  }
  for (size_t p = 0; p < qlen; p++) {
    int32_t nl[NAA] = {0};
    for (size_t k = 0; k < nrows; k++) {
      uint8_t c = msa[k * stride + p];
      if (c < NAA) nl[c]++;
    }
    int32_t distinct = 0;
    for (int a = 0; a < NAA; a++) distinct += nl[a] > 0;
    if (distinct == 0) continue;
    for (size_t k = 0; k < nrows; k++) {
      uint8_t c = msa[k * stride + p];
      if (c < NAA) {
        float denom = (float)nl[c] * (float)distinct * ((float)nres[k] + 30.0f);
        w[k] += 1.0f / denom;
      }
    }
  }
  /* normalise (src/profile.rs:121-127) */
  float total = 0.0f;
  for (size_t k = 0; k < nrows; k++) total += w[k];
  for (size_t k = 0; k < nrows; k++) w[k] = total > 0.0f ? w[k] / total : 1.0f / (float)nrows;

  /* weighted match frequencies (src/profile.rs:130-139, 372-386) */
  float *freq = calloc(qlen * NAA, sizeof(float));
  for (size_t p = 0; p < qlen; p++) {
    float *f = freq + p * NAA;
    for (size_t k = 0; k < nrows; k++) {
      uint8_t c = msa[k * stride + p];
      if (c < NAA) f[c] += w[k];
    }
    float sum = 0.0f;
    for (int a = 0; a < NAA; a++) sum += f[a];
    if (sum != 0.0f) {
      float fac = 1.0f / sum;
      for (int a = 0; a < NAA; a++) f[a] *= fac;
    } else {
      for (int a = 0; a < NAA; a++) f[a] = (float)pb[a];
    }
  }

  /* Neff (src/profile.rs:258-303) */
  float neff_hmm = 0.0f;
  for (size_t p = 0; p < qlen; p++) {
    float s = 0.0f;
    for (int a = 0; a < NAA; a++) {
      float f = freq[p * NAA + a];
      if (f > 1e-10f) s -= f * log2f(f);
    }
    neff_hmm += exp2f(s);
  }
  neff_hmm /= (float)qlen;
  float nlim = neff_hmm + 1.0f > 10.0f ? neff_hmm + 1.0f : 10.0f;
  float scale = log2f((nlim - neff_hmm) / (nlim - 1.0f));
  float *neff = malloc(qlen * sizeof(float));
  for (size_t p = 0; p < qlen; p++) {
    float wm = -1.0f / (float)nrows;
    for (size_t k = 0; k < nrows; k++)
      if (msa[k * stride + p] != GAP) wm += w[k];
    neff[p] = wm < 0.0f ? 1.0f : nlim - (nlim - 1.0f) * exp2f(scale * wm);
  }

  /* consensus (src/profile.rs:306-323) */
  if (cons_out)
    for (size_t p = 0; p < qlen; p++) {
      float maxw = 1e-8f;
      uint8_t maxa = 20;
      for (int a = 0; a < NAA; a++) {
        float d = freq[p * NAA + a] - pbf[a];
        if (d > maxw) {
          maxw = d;
          maxa = (uint8_t)a;
        }
      }
      cons_out[p] = maxa;
    }

  /* pseudocounts (src/profile.rs:145-169) */
  float *prob = malloc(qlen * NAA * sizeof(float));
  if (pca > 0.0f) {
    for (size_t p = 0; p < qlen; p++) {
      const float *f = freq + p * NAA;
      float g[NAA];
      for (int a = 0; a < NAA; a++) {
        float acc = 0.0f;
        for (int b = 0; b < NAA; b++) acc += r[a * NAA + b] * f[b];
        g[a] = acc;
      }
      float tau = pca / (1.0f + neff[p] / pcb);
      if (tau > 1.0f) tau = 1.0f;
      for (int a = 0; a < NAA; a++) prob[p * NAA + a] = (1.0f - tau) * f[a] + tau * g[a];
    }
  } else {
    memcpy(prob, freq, qlen * NAA * sizeof(float));
  }

  /* log-odds (src/profile.rs:326-346) */
  for (size_t p = 0; p < qlen; p++)
    for (int a = 0; a < NAA; a++) {
      float lp = log2f(prob[p * NAA + a] / pbf[a]);
      float v = 8.0f * lp + 8.0f * score_bias;
      float rd = v < 0.0f ? v - 0.5f : v + 0.5f;
      if (rd < -128.0f) rd = -128.0f;
      if (rd > 127.0f) rd = 127.0f;
      pssm[p * NAA + a] = rd != rd ? 0 : (int8_t)rd;
    }
  if (prob_out) memcpy(prob_out, prob, qlen * NAA * sizeof(float));
  if (neff_out) memcpy(neff_out, neff, qlen * sizeof(float));
  free(w);
  free(nres);
  free(freq);
  free(neff);
  free(prob);
  return 0;
}

/* ---- simulation calibration (src/evalue.rs:130-182, 576-668) ------------ */

/* gumbel_mu_given_lambda (src/evalue.rs:130-134) */
double or_gumbel_mu_given_lambda(const double *xs, size_t n, double lam) {
  double xmin = INFINITY, s = 0.0;
  for (size_t i = 0; i < n; i++) xmin = fmin(xmin, xs[i]);
  for (size_t i = 0; i < n; i++) s += exp(-lam * (xs[i] - xmin));
  return xmin - log(s / (double)n) / lam;
}

/* fit_gumbel_ml (src/evalue.rs:138-182): 0 and (lambda, mu), or -1 */
int or_fit_gumbel_ml(const double *xs, size_t n, double *lam_out, double *mu_out) {
  if (n < 10) return -1;
  double mean = 0.0, var = 0.0, xmin = INFINITY;
  for (size_t i = 0; i < n; i++) mean += xs[i];
  mean /= (double)n;
  for (size_t i = 0; i < n; i++) var += (xs[i] - mean) * (xs[i] - mean);
  var /= ((double)n - 1.0);
  if (var != var || var <= 0.0) return -1;
  for (size_t i = 0; i < n; i++) xmin = fmin(xmin, xs[i]);
  double lam0 = 3.14159265358979323846 / (sqrt(var) * sqrt(6.0));
  double lo = lam0 / 10.0, hi = lam0 * 10.0;
#define OR_G(l, outv)                                  \
  do {                                                 \
    double s0_ = 0.0, s1_ = 0.0;                       \
    for (size_t i_ = 0; i_ < n; i_++) {                \
      double w_ = exp(-(l) * (xs[i_] - xmin));         \
      s0_ += w_;                                       \
      s1_ += xs[i_] * w_;                              \
    }                                                  \
    (outv) = 1.0 / (l) - mean + s1_ / s0_;             \
  } while (0)
  double glo, ghi;
  OR_G(lo, glo);
  OR_G(hi, ghi);
  if (glo < 0.0 || ghi > 0.0) return -1;
  for (int it = 0; it < 200; it++) {
    double mid = 0.5 * (lo + hi), gm;
    OR_G(mid, gm);
    if (gm > 0.0)
      lo = mid;
    else
      hi = mid;
    if (hi - lo < 1e-12 * hi) break;
  }
#undef OR_G
  *lam_out = 0.5 * (lo + hi);
  *mu_out = or_gumbel_mu_given_lambda(xs, n, *lam_out);
  return 0;
}

/* fit_calibration (src/evalue.rs:576-668).  Cells must come sorted by
 * (m, n) (the caller sorts).  out6 = {lambda, K, alpha/lambda, beta,
 * ln K rms, 0}; cell_out (8 per cell) = {m, n, samples, mean, sd,
 * lambda_ml, mu_ml, mu}.  Returns 0, or -1 where the reference returns
 * None. */
static double or_cal_loglik(double lam, const uint64_t *cm, const uint64_t *cn, const uint64_t *off,
                            const double *xs, size_t nc, uint64_t lmin, int any_long) {
  double ll = 0.0;
  for (size_t c = 0; c < nc; c++) {
    if (any_long && (cm[c] < cn[c] ? cm[c] : cn[c]) < lmin) continue;
    const double *x = xs + off[c];
    size_t n = off[c + 1] - off[c];
    double mu = or_gumbel_mu_given_lambda(x, n, lam);
    for (size_t i = 0; i < n; i++) {
      double z = lam * (x[i] - mu);
      ll += log(lam) - z - exp(-z);
    }
  }
  return ll;
}

int or_fit_calibration(const uint64_t *cm, const uint64_t *cn, const uint64_t *off, const double *xs, size_t nc,
                       uint64_t lmin, double *out6, double *cell_out) {
  int any_long = 0;
  for (size_t c = 0; c < nc; c++) any_long |= (cm[c] < cn[c] ? cm[c] : cn[c]) >= lmin;
  /* golden-section search for the maximum over [0.05, 1.5], 80 steps */
  double a = 0.05, b = 1.5, gr = (sqrt(5.0) - 1.0) / 2.0;
  double c1 = b - gr * (b - a), c2 = a + gr * (b - a);
  double f1 = or_cal_loglik(c1, cm, cn, off, xs, nc, lmin, any_long);
  double f2 = or_cal_loglik(c2, cm, cn, off, xs, nc, lmin, any_long);
  for (int it = 0; it < 80; it++) {
    if (f1 < f2) {
      a = c1;
      c1 = c2;
      f1 = f2;
      c2 = a + gr * (b - a);
      f2 = or_cal_loglik(c2, cm, cn, off, xs, nc, lmin, any_long);
    } else {
      b = c2;
      c2 = c1;
      f2 = f1;
      c1 = b - gr * (b - a);
      f1 = or_cal_loglik(c1, cm, cn, off, xs, nc, lmin, any_long);
    }
  }
  double lambda = 0.5 * (a + b);
  for (size_t c = 0; c < nc; c++) {
    const double *x = xs + off[c];
    size_t n = off[c + 1] - off[c];
    double mean = 0.0, ss = 0.0, fl, fm;
    for (size_t i = 0; i < n; i++) mean += x[i];
    mean /= (double)n;
    for (size_t i = 0; i < n; i++) ss += (x[i] - mean) * (x[i] - mean);
    if (or_fit_gumbel_ml(x, n, &fl, &fm)) return -1;
    double *o = cell_out + 8 * c;
    o[0] = (double)cm[c];
    o[1] = (double)cn[c];
    o[2] = (double)n;
    o[3] = mean;
    o[4] = sqrt(ss / ((double)n - 1.0));
    o[5] = fl;
    o[6] = fm;
    o[7] = or_gumbel_mu_given_lambda(x, n, lambda) + 0.5;
  }
  int found = 0;
  double bv = 0.0, ba = 0.0, bb = 0.0, bm = 0.0;
  double *lnk = (double *)malloc((nc ? nc : 1) * sizeof(double));
  for (double adl = 0.0; adl <= 20.0 + 1e-9; adl += 0.1)
    for (double beta = -80.0; beta <= 20.0 + 1e-9; beta += 1.0) {
      size_t k = 0;
      int ok = 1;
      for (size_t c = 0; c < nc; c++) {
        double mu = cell_out[8 * c + 7];
        double ell = fmax(adl * lambda * mu + beta, 0.0);
        double me = (double)cm[c] - ell, ne = (double)cn[c] - ell;
        if (me < 1.0 || ne < 1.0) {
          ok = 0;
          break;
        }
        lnk[k++] = lambda * mu - log(me * ne);
      }
      if (!ok || k == 0) continue;
      double mean = 0.0, var = 0.0;
      for (size_t i = 0; i < k; i++) mean += lnk[i];
      mean /= (double)k;
      for (size_t i = 0; i < k; i++) var += (lnk[i] - mean) * (lnk[i] - mean);
      var /= (double)k;
      if (!found || var < bv) {
        found = 1;
        bv = var;
        ba = adl;
        bb = beta;
        bm = mean;
      }
    }
  free(lnk);
  if (!found) return -1;
  out6[0] = lambda;
  out6[1] = exp(bm);
  out6[2] = round(ba * 10.0) / 10.0;
  out6[3] = bb;
  out6[4] = sqrt(bv);
  out6[5] = 0.0;
  return 0;
}
