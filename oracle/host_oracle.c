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
