// Exact local alignment (traceback) of reported hits on the host.
//
// align_canonical (src/align.rs:840-882): the optimal score and its end cell
// (first optimum in column-major order, scalar_gotoh_score_end
// src/simd.rs:347-390; the GPU score+end pass produces the same triple), the
// start from the same rule on the reversed prefixes, then an anchored
// three-state traceback of the start-end rectangle inside a diagonal band
// doubled until the banded optimum reproduces the score
// (anchored_banded_tb src/align.rs:647-786), summarised by finish_alignment
// (src/align.rs:347-404).  Integer arithmetic only.
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "host_internal.h"
#include "../../include/sabertooth_host.h"

namespace {

constexpr int32_t kNeg = INT32_MIN / 4;
constexpr uint8_t kStop = 0, kDiag = 1, kSrcE = 2, kSrcF = 3, kEExt = 4, kFExt = 8;

struct Problem {
  const int32_t *rows;  // [m][21]
  const uint8_t *qres;  // query residues for identity (nullable: none identical)
  uint64_t m;
  const uint8_t *t;
  uint64_t n;
  int32_t open, ext;
  int32_t pair(uint64_t i, uint64_t j) const {
    const uint8_t r = t[j];
    return r < 21 ? rows[i * 21 + r] : 0;
  }
  bool identical(uint64_t i, uint64_t j) const { return qres && qres[i] == t[j] && qres[i] < 20; }
};

// scalar_gotoh_score_end over rows[ri(i)] and t[tj(j)] (index maps allow the
// reversed prefixes without copies); returns score, end (i, j) 0-based
template <class RowAt, class ResAt>
int32_t score_end(uint64_t m, uint64_t n, RowAt row_at, ResAt res_at, int32_t open, int32_t ext, uint64_t *bi_out,
                  uint64_t *bj_out) {
  std::vector<int32_t> hp(n + 1, 0), hc(n + 1, 0), fc(n + 1, kNeg);
  int32_t best = 0;
  uint64_t bi = 0, bj = UINT64_MAX;
  for (uint64_t i = 0; i < m; i++) {
    const int32_t *row = row_at(i);
    int32_t e = kNeg;
    hc[0] = 0;
    for (uint64_t j = 1; j <= n; j++) {
      const uint8_t tr = res_at(j - 1);
      const int32_t sc = tr < 21 ? row[tr] : 0;
      const int32_t diag = hp[j - 1] + sc;
      e = std::max(hc[j - 1] - open, e - ext);
      const int32_t f = std::max(hp[j] - open, fc[j] - ext);
      fc[j] = f;
      const int32_t v = std::max(std::max(diag, e), std::max(f, 0));
      hc[j] = v;
      if (v > best || (v == best && v > 0 && j - 1 < bj)) {
        best = v;
        bi = i;
        bj = j - 1;
      }
    }
    std::swap(hp, hc);
  }
  *bi_out = bi;
  *bj_out = bj;
  return best;
}

void finish(std::vector<uint8_t> &ops_rev, int32_t score, uint64_t qs, uint64_t ts, const Problem &p,
            st_alignment *out, std::string &cigar) {
  std::reverse(ops_rev.begin(), ops_rev.end());
  uint64_t i = qs, j = ts, ident = 0, mism = 0, gaps = 0, opens = 0;
  uint8_t prev = 0;
  for (uint8_t op : ops_rev) {
    if (op == 'M') {
      if (p.identical(i, j))
        ident++;
      else
        mism++;
      i++;
      j++;
    } else if (op == 'I') {
      gaps++;
      if (prev != 'I') opens++;
      j++;
    } else {
      gaps++;
      if (prev != 'D') opens++;
      i++;
    }
    prev = op;
  }
  std::memset(out, 0, sizeof *out);
  out->raw_score = score;
  out->query_start = qs;
  out->query_end = i - 1;
  out->target_start = ts;
  out->target_end = j - 1;
  out->aln_len = ops_rev.size();
  out->identities = ident;
  out->mismatches = mism;
  out->gaps = gaps;
  out->gap_opens = opens;
  cigar.clear();
  for (size_t k = 0; k < ops_rev.size();) {
    size_t r = 1;
    while (k + r < ops_rev.size() && ops_rev[k + r] == ops_rev[k]) r++;
    cigar += std::to_string(r);
    cigar += (char)ops_rev[k];
    k += r;
  }
}

// anchored_banded_tb: rectangle [qs, qe] x [ts, te], path from the aligned
// pair (qs, ts) to (qe, te); true when the banded optimum equals `score`
bool banded_tb(const Problem &p, uint64_t qs, uint64_t qe, uint64_t ts, uint64_t te, uint64_t radius,
               int32_t score, st_alignment *out, std::string &cigar) {
  const int64_t m = (int64_t)(qe + 1 - qs), n = (int64_t)(te + 1 - ts);
  const int64_t dmin = std::min<int64_t>(0, n - m) - (int64_t)radius;
  const int64_t dmax = std::max<int64_t>(0, n - m) + (int64_t)radius;
  const int64_t bw = dmax - dmin + 1;
  std::vector<uint8_t> tb((size_t)((m + 1) * bw), kStop);
  std::vector<int32_t> hp(n + 1, kNeg), hc(n + 1, kNeg), fc(n + 1, kNeg);
  hp[0] = 0;
  int64_t prev_hi = 0;
  for (int64_t a = 1; a <= m; a++) {
    const int64_t lo = std::max<int64_t>(a + dmin, 1), hi = std::min<int64_t>(a + dmax, n);
    for (int64_t b = prev_hi + 1; b <= hi; b++) {
      hp[b] = kNeg;
      fc[b] = kNeg;
    }
    hc[lo - 1] = kNeg;
    int32_t e = kNeg;
    const uint64_t qi = qs + a - 1;
    for (int64_t b = lo; b <= hi; b++) {
      const int32_t d = hp[b - 1];
      const int32_t diag = d > kNeg / 2 ? d + p.pair(qi, ts + b - 1) : kNeg;
      const int32_t e_open = hc[b - 1] - p.open, e_ext = e - p.ext;
      uint8_t bits = 0;
      if (e_ext > e_open) {
        e = e_ext;
        bits |= kEExt;
      } else {
        e = e_open;
      }
      const int32_t f_open = hp[b] - p.open, f_ext = fc[b] - p.ext;
      int32_t f;
      if (f_ext > f_open) {
        bits |= kFExt;
        f = f_ext;
      } else {
        f = f_open;
      }
      fc[b] = f;
      int32_t v = diag;
      uint8_t src = kDiag;
      if (e > v) {
        v = e;
        src = kSrcE;
      }
      if (f > v) {
        v = f;
        src = kSrcF;
      }
      if (v <= kNeg / 2) {
        v = kNeg;
        src = kStop;
      }
      hc[b] = v;
      tb[(size_t)(a * bw + (b - a - dmin))] = bits | src;
    }
    prev_hi = hi;
    std::swap(hp, hc);
  }
  if (hp[n] != score) return false;
  std::vector<uint8_t> ops;
  int64_t a = m, b = n;
  int state = 0;
  while (!(a == 0 && b == 0)) {
    if (a == 0 || b == 0) return false;
    const int64_t off = b - a - dmin;
    if (off < 0 || off >= bw) return false;
    const uint8_t cell = tb[(size_t)(a * bw + off)];
    if (state == 0) {
      const uint8_t src = cell & 3;
      if (src == kDiag) {
        ops.push_back('M');
        a--;
        b--;
      } else if (src == kSrcE) {
        state = 1;
      } else if (src == kSrcF) {
        state = 2;
      } else {
        return false;
      }
    } else if (state == 1) {
      ops.push_back('I');
      if (!(cell & kEExt)) state = 0;
      b--;
    } else {
      ops.push_back('D');
      if (!(cell & kFExt)) state = 0;
      a--;
    }
  }
  if (state != 0) return false;
  finish(ops, score, qs, ts, p, out, cigar);
  return true;
}

// anchored_start (src/align.rs:565-613): start of the first row-major optimum
// ending at (ie, je); the canonical path reaches it only if the reverse pass
// disagreed with the forward one
bool anchored_start(const Problem &p, uint64_t ie, uint64_t je, int32_t score, uint64_t *is, uint64_t *js) {
  const uint64_t m = ie + 1, n = je + 1;
  std::vector<int32_t> hp(n + 1, kNeg), hc(n + 1, kNeg), fc(n + 1, kNeg);
  hp[0] = 0;
  for (uint64_t a = 1; a <= m; a++) {
    int32_t e = kNeg;
    hc[0] = kNeg;
    const uint64_t qi = ie + 1 - a;
    bool alive = false;
    for (uint64_t b = 1; b <= n; b++) {
      const uint64_t tj = je + 1 - b;
      const int32_t d = hp[b - 1];
      const int32_t diag = d > kNeg / 2 ? d + p.pair(qi, tj) : kNeg;
      e = std::max(hc[b - 1] - p.open, e - p.ext);
      const int32_t f = std::max(hp[b] - p.open, fc[b] - p.ext);
      if (diag == score) {
        *is = qi;
        *js = tj;
        return true;
      }
      int32_t v = std::max(std::max(diag, e), f);
      if (v <= 0)
        v = kNeg;
      else
        alive = true;
      hc[b] = v;
      fc[b] = f <= 0 ? kNeg : f;
      if (e <= 0) e = kNeg;
    }
    if (!alive) break;
    std::swap(hp, hc);
  }
  return false;
}

int emit_cigar(const std::string &c, char *buf, uint64_t cap, uint64_t *len) {
  if (len) *len = c.size();
  if (buf && cap > 0) {
    const uint64_t k = std::min<uint64_t>(cap - 1, c.size());
    std::memcpy(buf, c.data(), k);
    buf[k] = 0;
  }
  return ST_OK;
}

}  // namespace

using namespace st;

extern "C" {

int st_sw_score_end_host(const int32_t *rows, uint64_t m, const uint8_t *t, uint64_t n, int32_t gap_open,
                         int32_t gap_extend, int32_t *score, int64_t *qend, int64_t *tend) {
  if (!score || (m && !rows) || (n && !t)) return fail(ST_ERR_INVALID, "st_sw_score_end_host: null argument");
  uint64_t bi = 0, bj = 0;
  int32_t s = 0;
  if (m && n)
    s = score_end(
        m, n, [&](uint64_t i) { return rows + i * 21; }, [&](uint64_t j) { return t[j]; }, gap_open, gap_extend,
        &bi, &bj);
  *score = s > 0 ? s : 0;
  if (qend) *qend = s > 0 ? (int64_t)bi : -1;
  if (tend) *tend = s > 0 ? (int64_t)bj : -1;
  return ST_OK;
}

// Step 3 of align_canonical (src/align.rs:869-881): the anchored banded
// traceback between a known start and end, the band doubled until the score
// is reproduced.  When even the full width does not reproduce it there is no
// alignment (*found stays 0), as the reference's release build returns None
// (src/align.rs:875-878).
static int traceback_span(const Problem &p, uint64_t is, uint64_t ie, uint64_t js, uint64_t je, int32_t score,
                          uint32_t band_radius, st_alignment *out, char *cigar, uint64_t cigar_cap,
                          uint64_t *cigar_len, int *found) {
  const uint64_t rows_n = ie + 1 - is, cols_n = je + 1 - js;
  uint64_t r = std::max<uint32_t>(band_radius, 1);
  std::string cg;
  for (;;) {
    if (banded_tb(p, is, ie, js, je, r, score, out, cg)) break;
    if (r >= rows_n + cols_n) {
      std::memset(out, 0, sizeof *out);
      return ST_OK;
    }
    r = std::min(r * 2, rows_n + cols_n);
  }
  *found = 1;
  return emit_cigar(cg, cigar, cigar_cap, cigar_len);
}

int st_align_canonical(const int32_t *rows, const uint8_t *qres, uint64_t m, const uint8_t *t, uint64_t n,
                       int32_t gap_open, int32_t gap_extend, uint32_t band_radius, const st_score_end *known,
                       st_alignment *out, char *cigar, uint64_t cigar_cap, uint64_t *cigar_len, int *found) {
  if (!out || !found || (m && !rows) || (n && !t)) return fail(ST_ERR_INVALID, "st_align_canonical: null argument");
  *found = 0;
  if (cigar_len) *cigar_len = 0;
  if (cigar && cigar_cap) cigar[0] = 0;
  std::memset(out, 0, sizeof *out);
  if (m == 0 || n == 0) return ST_OK;
  const Problem p{rows, qres, m, t, n, gap_open, gap_extend};
  // 1. forward score and column-major-first end cell
  int32_t score;
  uint64_t ie, je;
  if (known) {
    score = known->score;
    if (score <= 0) return ST_OK;
    if (known->query_end < 0 || known->target_end < 0 || (uint64_t)known->query_end >= m ||
        (uint64_t)known->target_end >= n)
      return fail(ST_ERR_INVALID, "st_align_canonical: end cell outside the problem");
    ie = (uint64_t)known->query_end;
    je = (uint64_t)known->target_end;
  } else {
    score = score_end(
        m, n, [&](uint64_t i) { return rows + i * 21; }, [&](uint64_t j) { return t[j]; }, gap_open, gap_extend,
        &ie, &je);
    if (score <= 0) return ST_OK;
  }
  // 2. the same rule on the reversed prefixes gives the start
  uint64_t ra = 0, rb = 0, is, js;
  const int32_t rs = score_end(
      ie + 1, je + 1, [&](uint64_t i) { return rows + (ie - i) * 21; }, [&](uint64_t j) { return t[je - j]; },
      gap_open, gap_extend, &ra, &rb);
  if (rs == score) {
    is = ie - ra;
    js = je - rb;
  } else if (!anchored_start(p, ie, je, score, &is, &js)) {
    return ST_OK;  // anchored_start found no start: None (src/align.rs:861-862)
  }
  return traceback_span(p, is, ie, js, je, score, band_radius, out, cigar, cigar_cap, cigar_len, found);
}

int st_align_span(const int32_t *rows, const uint8_t *qres, uint64_t m, const uint8_t *t, uint64_t n,
                  int32_t gap_open, int32_t gap_extend, uint32_t band_radius, const st_span *span,
                  st_alignment *out, char *cigar, uint64_t cigar_cap, uint64_t *cigar_len, int *found) {
  if (!out || !found || !span || (m && !rows) || (n && !t)) return fail(ST_ERR_INVALID, "st_align_span: null argument");
  *found = 0;
  if (cigar_len) *cigar_len = 0;
  if (cigar && cigar_cap) cigar[0] = 0;
  std::memset(out, 0, sizeof *out);
  if (m == 0 || n == 0 || span->score <= 0) return ST_OK;
  const bool no_start = span->query_start == -1 && span->target_start == -1;
  if (span->query_end < 0 || span->target_end < 0 || (uint64_t)span->query_end >= m ||
      (uint64_t)span->target_end >= n ||
      (!no_start && (span->query_start < 0 || span->target_start < 0 || span->query_start > span->query_end ||
                     span->target_start > span->target_end)))
    return fail(ST_ERR_INVALID, "st_align_span: span outside the problem");
  const Problem p{rows, qres, m, t, n, gap_open, gap_extend};
  const uint64_t ie = (uint64_t)span->query_end, je = (uint64_t)span->target_end;
  uint64_t is, js;
  if (no_start) {
    // the reversed pass did not reproduce the score: the anchored search
    // from the end, as align_canonical falls back to
    if (!anchored_start(p, ie, je, span->score, &is, &js))
      return ST_OK;  // anchored_start found no start: None (src/align.rs:861-862)
  } else {
    is = (uint64_t)span->query_start;
    js = (uint64_t)span->target_start;
  }
  return traceback_span(p, is, ie, js, je, span->score, band_radius, out, cigar, cigar_cap, cigar_len, found);
}

int st_rescore_alignment(const int32_t *rows, const uint8_t *qres, uint64_t m, const uint8_t *t, uint64_t n,
                         int32_t gap_open, int32_t gap_extend, const st_alignment *aln, const char *cigar,
                         int32_t *score) {
  // rescore_alignment (src/align.rs:1064-1135): walk the CIGAR, recompute the
  // score and every count; ST_ERR_INVALID on any disagreement
  if (!aln || !cigar || !score || (m && !rows) || (n && !t))
    return fail(ST_ERR_INVALID, "st_rescore_alignment: null argument");
  const Problem p{rows, qres, m, t, n, gap_open, gap_extend};
  uint64_t i = aln->query_start, j = aln->target_start, cols = 0, ident = 0, mism = 0, gapc = 0, opens = 0;
  int64_t sc = 0;
  char prev = 0, first = 0, last = 0;
  uint64_t num = 0;
  bool have = false;
  for (const char *c = cigar; *c; c++) {
    if (*c >= '0' && *c <= '9') {
      num = num * 10 + (uint64_t)(*c - '0');
      have = true;
      continue;
    }
    if (!have || num == 0) return fail(ST_ERR_INVALID, "st_rescore_alignment: CIGAR op without a positive count");
    const char op = *c;
    if (op != 'M' && op != 'I' && op != 'D') return fail(ST_ERR_INVALID, "st_rescore_alignment: unknown CIGAR op");
    if (op == prev) return fail(ST_ERR_INVALID, "st_rescore_alignment: adjacent runs with the same op");
    if (!first) first = op;
    last = op;
    if (op == 'M') {
      for (uint64_t k = 0; k < num; k++) {
        if (i >= m || j >= n) return fail(ST_ERR_INVALID, "st_rescore_alignment: M out of bounds");
        sc += p.pair(i, j);
        if (p.identical(i, j))
          ident++;
        else
          mism++;
        i++;
        j++;
      }
    } else {
      const int64_t cost = (int64_t)gap_open + (int64_t)(num - 1) * gap_extend;
      sc -= cost;
      if (op == 'I') {
        if (j + num > n) return fail(ST_ERR_INVALID, "st_rescore_alignment: I out of bounds");
        j += num;
      } else {
        if (i + num > m) return fail(ST_ERR_INVALID, "st_rescore_alignment: D out of bounds");
        i += num;
      }
      gapc += num;
      opens++;
    }
    cols += num;
    prev = op;
    num = 0;
    have = false;
  }
  if (have) return fail(ST_ERR_INVALID, "st_rescore_alignment: trailing count without op");
  if (first != 'M' || last != 'M')
    return fail(ST_ERR_INVALID, "st_rescore_alignment: a local alignment begins and ends with an aligned pair");
  if (i != aln->query_end + 1 || j != aln->target_end + 1 || cols != aln->aln_len || ident != aln->identities ||
      mism != aln->mismatches || gapc != aln->gaps || opens != aln->gap_opens)
    return fail(ST_ERR_INVALID, "st_rescore_alignment: counts or coordinates disagree with the CIGAR");
  if (sc != aln->raw_score) return fail(ST_ERR_INVALID, "st_rescore_alignment: the CIGAR rescores differently");
  *score = (int32_t)sc;
  return ST_OK;
}

}  // extern "C"
