/*
 * CPU oracle for the post-prefilter scoring path — TEST INFRASTRUCTURE ONLY.
 * See oracle.h.  Plain C99, scalar, no vectorisation: clarity over speed.
 * Only tests/, the smoke check and the bench's cpu_baseline leg may use it.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define OR_MAX(a, b) ((a) > (b) ? (a) : (b))
#define OR_MIN(a, b) ((a) < (b) ? (a) : (b))

/* ------------------------------------------------------------------------ */
/* alphabet (src/alphabet.rs:36-50)                                          */
/* ------------------------------------------------------------------------ */

static const char OR_NUM2AA[OR_NPROF + 1] = "ACDEFGHIKLMNPQRSTVWYX";

void or_aa2num(uint8_t table[256]) {
  for (int i = 0; i < 256; i++) table[i] = 20; /* everything else -> X */
  for (int k = 0; k < OR_NPROF; k++) {
    unsigned char up = (unsigned char)OR_NUM2AA[k];
    table[up] = (uint8_t)k;
    if (up >= 'A' && up <= 'Z') table[up - 'A' + 'a'] = (uint8_t)k;
  }
}

/* ------------------------------------------------------------------------ */
/* BLOSUM62 (src/matrix.rs:56-175, data src/data/blosum62.out)               */
/* The data file lists rows and columns in the order of the alphabet above,  */
/* so the table below is already in residue-code order.                      */
/* ------------------------------------------------------------------------ */

static const double OR_B62_BACK[OR_NPROF] = {
    0.07422, 0.02469, 0.05363, 0.05431, 0.04742, 0.07415, 0.02621,
    0.06792, 0.05815, 0.09891, 0.02499, 0.04465, 0.03854, 0.03426,
    0.05161, 0.05723, 0.05089, 0.07292, 0.01303, 0.03228, 0.00001};
static const double OR_B62_LAMBDA = 0.34657;
static const double OR_B62_HALFBITS[OR_NPROF][OR_NPROF] = {
    {3.9291, -0.4085, -1.7534, -0.8639, -2.2101, 0.1596, -1.6251, -1.3218, -0.7340, -1.4646, -0.9353, -1.5307, -0.8143, -0.8040, -1.4135, 1.1158, -0.0454, -0.1894, -2.5269, -1.7640, -1.0},
    {-0.4085, 8.5821, -3.4600, -3.6125, -2.3755, -2.5004, -2.9878, -1.2277, -3.0363, -1.2775, -1.4198, -2.6598, -2.7952, -2.9019, -3.3892, -0.8750, -0.8667, -0.8077, -2.3041, -2.4071, -1.0},
    {-1.7534, -3.4600, 5.7742, 1.5103, -3.4839, -1.3135, -1.1189, -3.1212, -0.7018, -3.6057, -3.0585, 1.2717, -1.4801, -0.3134, -1.6058, -0.2610, -1.0507, -3.1426, -4.2143, -3.0650, -1.0},
    {-0.8639, -3.6125, 1.5103, 4.9028, -3.1924, -2.1102, -0.1177, -3.1944, 0.7753, -2.8465, -1.9980, -0.2680, -1.1162, 1.8546, -0.1154, -0.1469, -0.8633, -2.4423, -2.8354, -2.0205, -1.0},
    {-2.2101, -2.3755, -3.4839, -3.1924, 6.0461, -3.1074, -1.2342, -0.1609, -3.0787, 0.4148, 0.0126, -2.9940, -3.5973, -3.1644, -2.7863, -2.3690, -2.1076, -0.8490, 0.9176, 2.9391, -1.0},
    {0.1596, -2.5004, -1.3135, -2.1102, -3.1074, 5.5633, -2.0409, -3.7249, -1.5280, -3.6270, -2.6766, -0.4228, -2.1335, -1.7852, -2.3041, -0.2925, -1.5754, -3.1387, -2.4915, -3.0398, -1.0},
    {-1.6251, -2.9878, -1.1189, -0.1177, -1.2342, -2.0409, 7.5111, -3.2316, -0.7210, -2.7867, -1.5513, 0.5785, -2.1609, 0.4480, -0.2499, -0.8816, -1.6859, -3.1175, -2.3422, 1.6926, -1.0},
    {-1.3218, -1.2277, -3.1212, -3.1944, -0.1609, -3.7249, -3.2316, 3.9985, -2.6701, 1.5216, 1.1268, -3.2170, -2.7567, -2.7696, -2.9902, -2.3482, -0.7176, 2.5470, -2.5805, -1.3314, -1.0},
    {-0.7340, -3.0363, -0.7018, 0.7753, -3.0787, -1.5280, -0.7210, -2.6701, 4.5046, -2.4468, -1.3547, -0.1790, -1.0136, 1.2726, 2.1087, -0.2034, -0.6696, -2.2624, -2.9564, -1.8200, -1.0},
    {-1.4646, -1.2775, -3.6057, -2.8465, 0.4148, -3.6270, -2.7867, 1.5216, -2.4468, 3.8494, 1.9918, -3.3789, -2.8601, -2.1339, -2.1546, -2.4426, -1.1975, 0.7884, -1.6319, -1.0621, -1.0},
    {-0.9353, -1.4198, -3.0585, -1.9980, 0.0126, -2.6766, -1.5513, 1.1268, -1.3547, 1.9918, 5.3926, -2.1509, -2.4764, -0.4210, -1.3671, -1.4809, -0.6663, 0.6872, -1.4248, -0.9949, -1.0},
    {-1.5307, -2.6598, 1.2717, -0.2680, -2.9940, -0.4228, 0.5785, -3.2170, -0.1790, -3.3789, -2.1509, 5.6532, -2.0004, 0.0017, -0.4398, 0.6009, -0.0461, -2.8763, -3.6959, -2.0818, -1.0},
    {-0.8143, -2.7952, -1.4801, -1.1162, -3.5973, -2.1335, -2.1609, -2.7567, -1.0136, -2.8601, -2.4764, -2.0004, 7.3646, -1.2819, -2.1086, -0.8090, -1.0753, -2.3487, -3.6542, -2.9198, -1.0},
    {-0.8040, -2.9019, -0.3134, 1.8546, -3.1644, -1.7852, 0.4480, -2.7696, 1.2726, -2.1339, -0.4210, 0.0017, -1.2819, 5.2851, 0.9828, -0.1011, -0.6753, -2.1984, -1.9465, -1.4211, -1.0},
    {-1.4135, -3.3892, -1.6058, -0.1154, -2.7863, -2.3041, -0.2499, -2.9902, 2.1087, -2.1546, -1.3671, -0.4398, -2.1086, 0.9828, 5.4735, -0.7648, -1.1223, -2.5026, -2.6794, -1.6939, -1.0},
    {1.1158, -0.8750, -0.2610, -0.1469, -2.3690, -0.2925, -0.8816, -2.3482, -0.2034, -2.4426, -1.4809, 0.6009, -0.8090, -0.1011, -0.7648, 3.8844, 1.3811, -1.6462, -2.7519, -1.6858, -1.0},
    {-0.0454, -0.8667, -1.0507, -0.8633, -2.1076, -1.5754, -1.6859, -0.7176, -0.6696, -1.1975, -0.6663, -0.0461, -1.0753, -0.6753, -1.1223, 1.3811, 4.5453, -0.0555, -2.4289, -1.6060, -1.0},
    {-0.1894, -0.8077, -3.1426, -2.4423, -0.8490, -3.1387, -3.1175, 2.5470, -2.2624, 0.7884, 0.6872, -2.8763, -2.3487, -2.1984, -2.5026, -1.6462, -0.0555, 3.7689, -2.8343, -1.2075, -1.0},
    {-2.5269, -2.3041, -4.2143, -2.8354, 0.9176, -2.4915, -2.3422, -2.5805, -2.9564, -1.6319, -1.4248, -3.6959, -3.6542, -1.9465, -2.6794, -2.7519, -2.4289, -2.8343, 10.5040, 2.1542, -1.0},
    {-1.7640, -2.4071, -3.0650, -2.0205, 2.9391, -3.0398, 1.6926, -1.3314, -1.8200, -1.0621, -0.9949, -2.0818, -2.9198, -1.4211, -1.6939, -1.6858, -1.6060, -1.2075, 2.1542, 6.5950, -1.0},
    {-1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0, -1.0}};

void or_blosum62(int8_t score[OR_NPROF * OR_NPROF], double p_back_out[OR_NPROF],
                 float r_out[OR_NAA * OR_NAA], double *lambda_out) {
  double pb[OR_NPROF];
  memcpy(pb, OR_B62_BACK, sizeof pb);
  /* xIsPositive check + background rescale (src/matrix.rs:119-135) */
  int x_pos = 0;
  for (int j = 0; j < OR_NPROF; j++)
    if (OR_B62_HALFBITS[20][j] > 0.0 || OR_B62_HALFBITS[j][20] > 0.0) x_pos = 1;
  if (!x_pos) {
    pb[20] = 1e-5; /* ANY_BACK */
    for (int i = 0; i < OR_NAA; i++) pb[i] *= 1.0 - pb[20];
  }
  /* joint probabilities (src/matrix.rs:137-143) */
  double prob[OR_NPROF][OR_NPROF];
  for (int i = 0; i < OR_NPROF; i++)
    for (int j = 0; j < OR_NPROF; j++)
      prob[i][j] = exp(OR_B62_LAMBDA * OR_B62_HALFBITS[i][j]) * pb[i] * pb[j];
  if (r_out) /* R[a][b] = P(a|b) (src/matrix.rs:145-151) */
    for (int a = 0; a < OR_NAA; a++)
      for (int b = 0; b < OR_NAA; b++)
        r_out[a * OR_NAA + b] = (float)(prob[a][b] / pb[b]);
  if (score) /* round-half-away of 2*log2(P/(pa pb)) - 0.2 (src/matrix.rs:153-162) */
    for (int i = 0; i < OR_NPROF; i++)
      for (int j = 0; j < OR_NPROF; j++) {
        double s = prob[i][j] / (pb[i] * pb[j]);
        double v = 2.0 * log2(s) + (-0.2);
        double rd = v < 0.0 ? v - 0.5 : v + 0.5;
        if (rd < -128.0) rd = -128.0;
        if (rd > 127.0) rd = 127.0;
        score[i * OR_NPROF + j] = (int8_t)rd; /* `as i8` truncates toward 0 */
      }
  if (p_back_out) memcpy(p_back_out, pb, sizeof pb);
  if (lambda_out) *lambda_out = OR_B62_LAMBDA;
}

void or_seq_rows(const uint8_t *q, size_t m, int32_t *rows) {
  int8_t mat[OR_NPROF * OR_NPROF];
  or_blosum62(mat, NULL, NULL, NULL);
  for (size_t i = 0; i < m; i++)
    for (int r = 0; r < OR_NPROF; r++) {
      /* SeqScorer::score indexes the matrix by the query code directly */
      int qa = q[i] < OR_NPROF ? q[i] : 20;
      rows[i * OR_NPROF + r] = mat[qa * OR_NPROF + r];
    }
}

/* ------------------------------------------------------------------------ */
/* SWG (src/simd.rs:347-390)                                                 */
/* ------------------------------------------------------------------------ */

static inline int32_t or_sub(const int32_t *row, uint8_t tr) {
  return tr < OR_NPROF ? row[tr] : 0;
}

int32_t or_gotoh_score_end(const int32_t *rows, size_t m, const uint8_t *t,
                           size_t n, int32_t open, int32_t ext, int64_t *qend,
                           int64_t *tend) {
  const int32_t neg = INT32_MIN / 4;
  int32_t best = 0;
  int64_t bi = 0, bj = INT64_MAX;
  if (m == 0 || n == 0) {
    if (qend) *qend = -1;
    if (tend) *tend = -1;
    return 0;
  }
  int32_t *hp = calloc(n + 1, sizeof *hp), *hc = calloc(n + 1, sizeof *hc);
  int32_t *fc = malloc((n + 1) * sizeof *fc);
  for (size_t j = 0; j <= n; j++) fc[j] = neg;
  for (size_t i = 0; i < m; i++) {
    const int32_t *row = rows + i * OR_NPROF;
    int32_t e = neg;
    hc[0] = 0;
    for (size_t j = 1; j <= n; j++) {
      int32_t diag = hp[j - 1] + or_sub(row, t[j - 1]);
      e = OR_MAX(hc[j - 1] - open, e - ext);
      int32_t f = OR_MAX(hp[j] - open, fc[j] - ext);
      fc[j] = f;
      int32_t v = OR_MAX(OR_MAX(diag, e), OR_MAX(f, 0));
      hc[j] = v;
      /* a tie moves the end only to an earlier column (line 377) */
      if (v > best || (v == best && v > 0 && (int64_t)(j - 1) < bj)) {
        best = v;
        bi = (int64_t)i;
        bj = (int64_t)(j - 1);
      }
    }
    int32_t *tmp = hp;
    hp = hc;
    hc = tmp;
  }
  free(hp);
  free(hc);
  free(fc);
  if (qend) *qend = best > 0 ? bi : -1;
  if (tend) *tend = best > 0 ? bj : -1;
  return best;
}

/* three full matrices, then a column-major scan (src/simd.rs:1124-1195) */
int32_t or_gotoh_textbook(const int32_t *rows, size_t m, const uint8_t *t,
                          size_t n, int32_t open, int32_t ext, int64_t *qend,
                          int64_t *tend) {
  const int32_t ni = INT32_MIN / 4;
  size_t w = n + 1;
  int32_t *h = calloc((m + 1) * w, sizeof *h);
  int32_t *e = malloc((m + 1) * w * sizeof *e);
  int32_t *f = malloc((m + 1) * w * sizeof *f);
  for (size_t k = 0; k < (m + 1) * w; k++) e[k] = f[k] = ni;
  for (size_t i = 1; i <= m; i++)
    for (size_t j = 1; j <= n; j++) {
      e[i * w + j] = OR_MAX(h[i * w + j - 1] - open, e[i * w + j - 1] - ext);
      f[i * w + j] = OR_MAX(h[(i - 1) * w + j] - open, f[(i - 1) * w + j] - ext);
      int32_t d = h[(i - 1) * w + j - 1] + or_sub(rows + (i - 1) * OR_NPROF, t[j - 1]);
      h[i * w + j] = OR_MAX(OR_MAX(d, e[i * w + j]), OR_MAX(f[i * w + j], 0));
    }
  int32_t best = 0;
  int64_t bi = -1, bj = -1;
  for (size_t j = 1; j <= n; j++)
    for (size_t i = 1; i <= m; i++)
      if (h[i * w + j] > best) {
        best = h[i * w + j];
        bi = (int64_t)i - 1;
        bj = (int64_t)j - 1;
      }
  free(h);
  free(e);
  free(f);
  if (qend) *qend = bi;
  if (tend) *tend = bj;
  return best;
}

/* ------------------------------------------------------------------------ */
/* ungapped evidence (src/ungapped.rs:38-88)                                 */
/* ------------------------------------------------------------------------ */

void or_diag_range(size_t m, size_t n, int64_t d, size_t *lo, size_t *hi) {
  int64_t a = -d > 0 ? -d : 0;
  int64_t b = OR_MIN((int64_t)m, (int64_t)n - d);
  if (b <= a) {
    *lo = 0;
    *hi = 0;
  } else {
    *lo = (size_t)a;
    *hi = (size_t)b;
  }
}

int32_t or_kadane_score(const int32_t *rows, const uint8_t *t, int64_t d,
                        size_t lo, size_t hi) {
  int32_t best = 0, cur = 0;
  for (size_t q = lo; q < hi; q++) {
    int32_t s = or_sub(rows + q * OR_NPROF, t[(int64_t)q + d]);
    cur = OR_MAX(cur, 0) + s;
    best = OR_MAX(best, cur);
  }
  return best;
}

int32_t or_kadane_segment(const int32_t *rows, const uint8_t *t, int64_t d,
                          size_t lo, size_t hi, size_t *qs, size_t *qe) {
  int32_t best = 0, cur = 0;
  size_t bs = 0, be = 0, cs = lo;
  for (size_t q = lo; q < hi; q++) {
    int32_t s = or_sub(rows + q * OR_NPROF, t[(int64_t)q + d]);
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

/* ------------------------------------------------------------------------ */
/* gate (src/prefilter.rs:1498-1660, 1269-1278, 246-257)                     */
/* ------------------------------------------------------------------------ */

static int32_t or_gap_cost(int32_t open, int32_t ext, uint64_t len) {
  /* GapCosts::cost (src/align.rs:216-222) */
  return len == 0 ? 0 : open + (int32_t)(len - 1) * ext;
}

typedef struct {
  int32_t score;
  size_t qs, qe;
  int64_t diag;
  int have_pos;
} or_seg;

static void or_pos_of(or_seg *sg, const or_job *jb, const int32_t *rows,
                      const uint8_t *t) {
  if (!sg->have_pos) {
    or_kadane_segment(rows, t, sg->diag, jb->lo, jb->hi, &sg->qs, &sg->qe);
    sg->have_pos = 1;
  }
}

/* chain + gate of one target over its (<= 16) scored jobs; returns 1 and
 * fills *c when it passes (src/prefilter.rs:1577-1657) */
static int or_gate_target(const int32_t *rows, const uint8_t *t,
                          const or_job *jobs, const int32_t *scores, size_t ns,
                          uint32_t tid, uint32_t hits, int32_t open,
                          int32_t ext, int32_t min_diag_score, int regions,
                          or_candidate *c, uint64_t *chains_used) {
  or_seg segs[16];
  if (ns > 16) ns = 16;
  if (ns == 1 && scores[0] < min_diag_score) return 0;
  int32_t best_s = 0;
  size_t best_k = 0;
  for (size_t k = 0; k < ns; k++) {
    segs[k].score = scores[k];
    segs[k].qs = segs[k].qe = 0;
    segs[k].diag = jobs[k].diag;
    segs[k].have_pos = 0;
    if (scores[k] > best_s) {
      best_s = scores[k];
      best_k = k;
    }
  }
  int32_t chain_s = INT32_MIN;
  size_t cx = 0, cy = 0;
  if (ns > 1)
    for (size_t x = 0; x < ns; x++)
      for (size_t y = 0; y < ns; y++) {
        if (x == y) continue;
        int32_t sx = segs[x].score, sy = segs[y].score;
        int64_t dx = segs[x].diag, dy = segs[y].diag;
        if (sx <= 0 || sy <= 0 || dx == dy) continue;
        uint64_t ad = (uint64_t)(dy > dx ? dy - dx : dx - dy);
        int32_t cval = sx + sy - or_gap_cost(open, ext, ad);
        if (cval <= best_s || cval <= chain_s) continue;
        or_pos_of(&segs[x], &jobs[x], rows, t);
        or_pos_of(&segs[y], &jobs[y], rows, t);
        size_t qex = segs[x].qe, qsy = segs[y].qs;
        if (qex < qsy && (int64_t)qex + dx < (int64_t)qsy + dy) {
          chain_s = cval;
          cx = x;
          cy = y;
        }
      }
  int32_t score;
  size_t used[2];
  if (chain_s > best_s) {
    if (chains_used) (*chains_used)++;
    score = chain_s;
    used[0] = cx;
    used[1] = cy;
  } else {
    score = best_s;
    used[0] = used[1] = best_k;
  }
  if (score < min_diag_score) return 0;
  uint32_t lo = 0, hi = UINT32_MAX;
  if (regions) { /* region (src/prefilter.rs:1269-1278) */
    lo = UINT32_MAX;
    hi = 0;
    for (int u = 0; u < 2; u++) {
      or_seg *sg = &segs[used[u]];
      or_pos_of(sg, &jobs[used[u]], rows, t);
      int64_t a = (int64_t)sg->qs + sg->diag, b = (int64_t)sg->qe + sg->diag;
      uint32_t ua = (uint32_t)(a > 0 ? a : 0), ub = (uint32_t)(b > 0 ? b : 0);
      lo = OR_MIN(lo, ua);
      hi = OR_MAX(hi, ub);
    }
  }
  c->target_id = tid;
  c->hits = hits;
  c->best_diag = segs[best_k].diag;
  c->diag_score = score;
  c->t_lo = lo;
  c->t_hi = hi;
  return 1;
}

size_t or_score_and_gate(const int32_t *rows, size_t m, const uint8_t *arena,
                         const uint64_t *offsets, const or_job *jobs,
                         size_t njobs, const or_target_rec *recs, size_t nrecs,
                         int32_t open, int32_t ext, int32_t min_diag_score,
                         int regions, or_candidate *out, uint64_t *chains_used,
                         int32_t *scores_out) {
  (void)m;
  int32_t *scores = malloc((njobs ? njobs : 1) * sizeof *scores);
  for (size_t j = 0; j < njobs; j++) {
    const uint8_t *t = arena + offsets[jobs[j].tid];
    scores[j] = or_kadane_score(rows, t, jobs[j].diag, jobs[j].lo,
                                jobs[j].hi > jobs[j].lo ? jobs[j].hi : jobs[j].lo);
  }
  size_t nc = 0;
  for (size_t r = 0; r < nrecs; r++) {
    const or_target_rec *tr = &recs[r];
    size_t ns = tr->j1 - tr->j0;
    if (ns == 0) continue;
    const uint8_t *t = arena + offsets[tr->tid];
    if (or_gate_target(rows, t, jobs + tr->j0, scores + tr->j0, ns, tr->tid,
                       tr->hits, open, ext, min_diag_score, regions, &out[nc],
                       chains_used))
      nc++;
  }
  if (scores_out) memcpy(scores_out, scores, njobs * sizeof *scores);
  free(scores);
  return nc;
}

int32_t or_size_adjusted_gate(int32_t gate, uint64_t db_residues) {
  double x = log((double)(db_residues > 0 ? db_residues : 1) / 1.0e8);
  double slope = x < 0.0 ? 2.5 : 1.0;
  double v = (double)gate + slope * x;
  /* f64::round: half away from zero */
  double rd = v < 0.0 ? -floor(-v + 0.5) : floor(v + 0.5);
  int32_t g = (int32_t)rd;
  return g < 20 ? 20 : (g > 60 ? 60 : g);
}

/* ------------------------------------------------------------------------ */
/* all-diagonal evidence for a (query, candidate) pair                       */
/* ------------------------------------------------------------------------ */

void or_ung_pair(const int32_t *rows, size_t m, const uint8_t *t, size_t n,
                 int32_t open, int32_t ext, int32_t min_diag_score,
                 int32_t *ung, int32_t *best_diag, int32_t *gate_score,
                 int32_t *pass) {
  int32_t ts[4];
  int64_t td[4];
  int nt = 0;
  if (m > 0 && n > 0)
    for (int64_t d = -(int64_t)(m - 1); d <= (int64_t)n - 1; d++) {
      size_t lo, hi;
      or_diag_range(m, n, d, &lo, &hi);
      int32_t s = or_kadane_score(rows, t, d, lo, hi);
      /* keep the four best, score desc then diagonal asc; diagonals arrive in
       * ascending order, so a tie never displaces an earlier entry */
      int pos = nt;
      while (pos > 0 && s > ts[pos - 1]) pos--;
      if (pos < 4) {
        int last = nt < 4 ? nt : 3;
        for (int k = last; k > pos; k--) {
          ts[k] = ts[k - 1];
          td[k] = td[k - 1];
        }
        ts[pos] = s;
        td[pos] = d;
        if (nt < 4) nt++;
      }
    }
  *ung = nt ? ts[0] : 0;
  *best_diag = nt ? (int32_t)td[0] : 0;
  or_job jobs[4];
  for (int k = 0; k < nt; k++) {
    size_t lo, hi;
    or_diag_range(m, n, td[k], &lo, &hi);
    jobs[k].tid = 0;
    jobs[k].diag = (int32_t)td[k];
    jobs[k].lo = (uint32_t)lo;
    jobs[k].hi = (uint32_t)hi;
  }
  or_candidate c;
  int32_t gs = 0;
  int ok;
  if (nt == 0) {
    ok = 0 >= min_diag_score;
  } else {
    /* run the gate with a threshold that always passes to read the score */
    ok = or_gate_target(rows, t, jobs, ts, (size_t)nt, 0, 0, open, ext,
                        INT32_MIN, 0, &c, NULL);
    gs = c.diag_score;
    ok = ok && gs >= min_diag_score;
  }
  *gate_score = gs;
  *pass = ok;
}

/* ------------------------------------------------------------------------ */
/* SplitMix64 + BackgroundSampler (src/evalue.rs:52-114)                     */
/* ------------------------------------------------------------------------ */

uint64_t or_splitmix_next(uint64_t *state) {
  *state += 0x9E3779B97F4A7C15ull;
  uint64_t z = *state;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

void or_background_sequence(uint64_t *state, size_t len, uint8_t *out) {
  double pb[OR_NPROF], cum[OR_NAA], total = 0.0, acc = 0.0;
  or_blosum62(NULL, pb, NULL, NULL);
  for (int i = 0; i < OR_NAA; i++) total += pb[i];
  for (int i = 0; i < OR_NAA; i++) {
    acc += pb[i] / total;
    cum[i] = acc;
  }
  cum[OR_NAA - 1] = 1.0;
  for (size_t k = 0; k < len; k++) {
    double u = (double)(or_splitmix_next(state) >> 11) / 9007199254740992.0;
    uint8_t r = OR_NAA - 1;
    for (int i = 0; i < OR_NAA; i++)
      if (u < cum[i]) {
        r = (uint8_t)i;
        break;
      }
    out[k] = r;
  }
}
