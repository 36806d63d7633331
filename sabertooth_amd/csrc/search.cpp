// The host side of one query's search after scoring (search_query's gapped
// stage and deferred traceback, src/search.rs:532-666) and the .m8 output row
// (Hit::to_m8, src/search.rs:201-216).
//
// Scores come from the GPU (st_sw_score_batch / st_batch) or any exact
// source; the gapped stage keeps candidates with score >= min_score_for(
// max_evalue, search space) and score > 0, orders them by (score desc,
// target asc), aligns them with align_canonical in that order until max_hits
// hits are accepted (coverage filters skip a hit and continue), and orders
// the hits by E-value, bit score and target ordinal (hit_order).
#include <algorithm>
#include <cerrno>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "host_internal.h"
#include "../../include/sabertooth_host.h"

using namespace st;

namespace {

// Rust's {:.2e}: the C library's correctly rounded mantissa, then the
// exponent without '+' or zero padding (2.92e-1, 8.38e-55, 0.00e0)
std::string rust_e2(double x) {
  if (std::isnan(x)) return "NaN";
  if (std::isinf(x)) return x > 0 ? "inf" : "-inf";
  char b[64];
  std::snprintf(b, sizeof b, "%.2e", x);
  std::string s(b);
  const size_t e = s.find('e');
  const int ex = std::atoi(s.c_str() + e + 1);
  return s.substr(0, e) + "e" + std::to_string(ex);
}

// Rust's {:.Nf} for finite values is the correctly rounded fixed form
std::string rust_f(double x, int prec) {
  if (std::isnan(x)) return "NaN";
  if (std::isinf(x)) return x > 0 ? "inf" : "-inf";
  char b[512];
  std::snprintf(b, sizeof b, "%.*f", prec, x);
  return b;
}

bool hit_less(const st_hit &a, const st_hit &b) {
  // hit_order (src/search.rs:191-197): E-value asc, bits desc, target asc
  if (a.evalue != b.evalue) return a.evalue < b.evalue;
  if (a.bit_score != b.bit_score) return a.bit_score > b.bit_score;
  return a.target < b.target;
}

}  // namespace

extern "C" {

int st_hit_order(const st_hit *a, const st_hit *b) {
  if (!a || !b) return 0;
  if (hit_less(*a, *b)) return -1;
  if (hit_less(*b, *a)) return 1;
  return 0;
}

int st_gapped_stage(const int32_t *scores, const uint32_t *tids, uint64_t n, int32_t min_score, uint64_t *order,
                    uint64_t *n_pass) {
  // Pass filter and order (src/search.rs:537-580)
  if (!n_pass || (n && (!scores || !tids || !order))) return fail(ST_ERR_INVALID, "st_gapped_stage: null argument");
  uint64_t k = 0;
  for (uint64_t i = 0; i < n; i++)
    if (scores[i] >= min_score && scores[i] > 0) order[k++] = i;
  std::sort(order, order + k, [&](uint64_t a, uint64_t b) {
    if (scores[a] != scores[b]) return scores[a] > scores[b];
    return tids[a] < tids[b];
  });
  *n_pass = k;
  return ST_OK;
}

static int search_hits_impl(const int32_t *rows, const uint8_t *qres, uint64_t m, const uint8_t *arena,
                            const uint64_t *offsets, uint64_t n_targets, const uint32_t *cand_tids,
                            const int32_t *scores, const st_span *spans, uint64_t n_cands, int32_t gap_open,
                            int32_t gap_extend, const st_evalue_params *ev, double max_evalue, uint64_t max_hits,
                            double min_query_cov, double min_target_cov, uint32_t band_radius, st_hit *hits,
                            char *cigars, uint64_t cigar_stride, uint64_t cap, uint64_t *n_hits,
                            st_search_counters *counters) {
  if (!ev || !n_hits || (m && !rows) || (n_cands && (!cand_tids || !scores || !arena || !offsets)) ||
      (cap && !hits))
    return fail(ST_ERR_INVALID, "st_search_hits: null argument");
  *n_hits = 0;
  st_search_counters cnt{};
  const double space = st_search_space(ev, m);
  const int32_t min_score = st_min_score_for(ev, max_evalue, space);
  std::vector<uint64_t> order(n_cands);
  uint64_t np = 0;
  int rc = st_gapped_stage(scores, cand_tids, n_cands, min_score, order.data(), &np);
  if (rc) return rc;
  cnt.scored = n_cands;
  cnt.passed_score = np;
  std::vector<st_hit> out;
  std::vector<std::string> cig;
  for (uint64_t k = 0; k < np; k++) {
    if (max_hits > 0 && out.size() >= max_hits) break;
    const uint64_t c = order[k];
    const uint32_t tid = cand_tids[c];
    if (tid >= n_targets) return fail(ST_ERR_INVALID, "st_search_hits: target id out of range");
    const uint8_t *t = arena + offsets[tid];
    const uint64_t n = offsets[tid + 1] - offsets[tid];
    st_alignment a;
    std::vector<char> cg(2 * (m + n) + 64);
    uint64_t cl = 0;
    int found = 0;
    cnt.tracebacks++;
    if (spans) {
      if (spans[c].score != scores[c]) return fail(ST_ERR_INVALID, "st_search_hits_spans: span score != score");
      rc = st_align_span(rows, qres, m, t, n, gap_open, gap_extend, band_radius, &spans[c], &a, cg.data(), cg.size(),
                         &cl, &found);
    } else {
      rc = st_align_canonical(rows, qres, m, t, n, gap_open, gap_extend, band_radius, nullptr, &a, cg.data(),
                              cg.size(), &cl, &found);
    }
    if (rc) return rc;
    if (!found) continue;
    if (a.raw_score != scores[c]) cnt.score_mismatches++;
    const double qcov = (double)(a.query_end + 1 - a.query_start) / (double)(m ? m : 1);
    const double tcov = (double)(a.target_end + 1 - a.target_start) / (double)(n ? n : 1);
    if ((min_query_cov > 0.0 && qcov < min_query_cov) || (min_target_cov > 0.0 && tcov < min_target_cov)) {
      cnt.rejected_coverage++;
      continue;
    }
    const double e = st_evalue_in_space(ev, a.raw_score, space);
    if (e > max_evalue) continue;
    st_hit h{};
    h.target = tid;
    h.raw_score = a.raw_score;
    h.pident = a.aln_len == 0 ? 0.0 : 100.0 * (double)a.identities / (double)a.aln_len;
    h.aln_len = a.aln_len;
    h.mismatches = a.mismatches;
    h.gap_opens = a.gap_opens;
    h.q_start = a.query_start;
    h.q_end = a.query_end;
    h.t_start = a.target_start;
    h.t_end = a.target_end;
    h.evalue = e;
    h.bit_score = st_bit_score(ev, a.raw_score);
    h.qlen = m;
    h.tlen = n;
    out.push_back(h);
    cig.emplace_back(cg.data(), cl);
  }
  std::vector<size_t> idx(out.size());
  for (size_t i = 0; i < idx.size(); i++) idx[i] = i;
  std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) { return hit_less(out[a], out[b]); });
  const uint64_t nh = std::min<uint64_t>(cap, out.size());
  for (uint64_t i = 0; i < nh; i++) {
    hits[i] = out[idx[i]];
    if (cigars && cigar_stride) {
      const std::string &c = cig[idx[i]];
      const size_t k = std::min<size_t>(cigar_stride - 1, c.size());
      std::memcpy(cigars + i * cigar_stride, c.data(), k);
      cigars[i * cigar_stride + k] = 0;
    }
  }
  *n_hits = out.size();
  cnt.hits = out.size();
  if (counters) {
    counters->scored += cnt.scored;
    counters->passed_score += cnt.passed_score;
    counters->tracebacks += cnt.tracebacks;
    counters->score_mismatches += cnt.score_mismatches;
    counters->rejected_coverage += cnt.rejected_coverage;
    counters->hits += cnt.hits;
  }
  return ST_OK;
}

int st_search_hits(const int32_t *rows, const uint8_t *qres, uint64_t m, const uint8_t *arena,
                   const uint64_t *offsets, uint64_t n_targets, const uint32_t *cand_tids, const int32_t *scores,
                   uint64_t n_cands, int32_t gap_open, int32_t gap_extend, const st_evalue_params *ev,
                   double max_evalue, uint64_t max_hits, double min_query_cov, double min_target_cov,
                   uint32_t band_radius, st_hit *hits, char *cigars, uint64_t cigar_stride, uint64_t cap,
                   uint64_t *n_hits, st_search_counters *counters) {
  return search_hits_impl(rows, qres, m, arena, offsets, n_targets, cand_tids, scores, nullptr, n_cands, gap_open,
                          gap_extend, ev, max_evalue, max_hits, min_query_cov, min_target_cov, band_radius, hits,
                          cigars, cigar_stride, cap, n_hits, counters);
}

int st_search_hits_spans(const int32_t *rows, const uint8_t *qres, uint64_t m, const uint8_t *arena,
                         const uint64_t *offsets, uint64_t n_targets, const uint32_t *cand_tids,
                         const int32_t *scores, const st_span *spans, uint64_t n_cands, int32_t gap_open,
                         int32_t gap_extend, const st_evalue_params *ev, double max_evalue, uint64_t max_hits,
                         double min_query_cov, double min_target_cov, uint32_t band_radius, st_hit *hits,
                         char *cigars, uint64_t cigar_stride, uint64_t cap, uint64_t *n_hits,
                         st_search_counters *counters) {
  if (n_cands && !spans) return fail(ST_ERR_INVALID, "st_search_hits_spans: null spans");
  return search_hits_impl(rows, qres, m, arena, offsets, n_targets, cand_tids, scores, spans, n_cands, gap_open,
                          gap_extend, ev, max_evalue, max_hits, min_query_cov, min_target_cov, band_radius, hits,
                          cigars, cigar_stride, cap, n_hits, counters);
}

int st_format_m8(const char *query_id, const char *target_id, const st_hit *h, char *buf, uint64_t cap,
                 uint64_t *len) {
  // Hit::to_m8 (src/search.rs:201-216): 1-based coordinates, pident {:.3},
  // E-value {:.2e}, bits {:.1}, tab separated, no newline
  if (!query_id || !target_id || !h) return fail(ST_ERR_INVALID, "st_format_m8: null argument");
  std::string s = std::string(query_id) + "\t" + target_id + "\t" + rust_f(h->pident, 3) + "\t" +
                  std::to_string(h->aln_len) + "\t" + std::to_string(h->mismatches) + "\t" +
                  std::to_string(h->gap_opens) + "\t" + std::to_string(h->q_start + 1) + "\t" +
                  std::to_string(h->q_end + 1) + "\t" + std::to_string(h->t_start + 1) + "\t" +
                  std::to_string(h->t_end + 1) + "\t" + rust_e2(h->evalue) + "\t" + rust_f(h->bit_score, 1);
  if (len) *len = s.size();
  if (buf && cap) {
    const uint64_t k = std::min<uint64_t>(cap - 1, s.size());
    std::memcpy(buf, s.data(), k);
    buf[k] = 0;
  }
  return ST_OK;
}

int st_parse_candidate_trace(const char *text, uint64_t len, const uint8_t *id_bytes, const uint64_t *id_off,
                             uint64_t n_targets, uint64_t *n_queries, uint64_t *n_cands, char *qid_bytes,
                             uint64_t qid_cap, uint64_t *qid_off, uint64_t *cand_off, uint32_t *cand_tids,
                             int32_t *diag_scores, uint64_t *cells) {
  // lines "query \t target \t prefilter score \t DP cells" (candidate_trace,
  // src/search.rs:358-385, written at src/search.rs:522-530)
  if (!n_queries || !n_cands || (len && !text) || (n_targets && (!id_bytes || !id_off)))
    return fail(ST_ERR_INVALID, "st_parse_candidate_trace: null argument");
  std::unordered_map<std::string, uint32_t> tid_of;
  tid_of.reserve(n_targets * 2);
  for (uint64_t t = 0; t < n_targets; t++)
    tid_of.emplace(std::string((const char *)id_bytes + id_off[t], id_off[t + 1] - id_off[t]), (uint32_t)t);
  struct Row {
    uint32_t q, tid;
    int32_t score;
    uint64_t cells;
  };
  std::vector<Row> rows;
  std::vector<std::string> qids;
  std::unordered_map<std::string, uint32_t> q_of;
  uint64_t pos = 0, line = 0;
  while (pos < len) {
    uint64_t e = pos;
    while (e < len && text[e] != '\n') e++;
    line++;
    std::string l(text + pos, e - pos);
    pos = e + 1;
    if (!l.empty() && l.back() == '\r') l.pop_back();
    if (l.empty()) continue;
    std::string f[4];
    size_t a = 0;
    int nf = 0;
    for (; nf < 4; nf++) {
      const size_t b = l.find('\t', a);
      if (nf == 3) {
        if (b != std::string::npos) break;
        f[nf] = l.substr(a);
      } else {
        if (b == std::string::npos) break;
        f[nf] = l.substr(a, b - a);
        a = b + 1;
      }
    }
    if (nf != 4) return fail(ST_ERR_INVALID, "st_parse_candidate_trace: line " + std::to_string(line) + " has not 4 fields");
    auto it = tid_of.find(f[1]);
    if (it == tid_of.end())
      return fail(ST_ERR_INVALID, "st_parse_candidate_trace: line " + std::to_string(line) + ": unknown target " + f[1]);
    char *end = nullptr;
    errno = 0;
    const long long sc = std::strtoll(f[2].c_str(), &end, 10);
    if (f[2].empty() || *end || errno || sc < INT32_MIN || sc > INT32_MAX)
      return fail(ST_ERR_INVALID, "st_parse_candidate_trace: line " + std::to_string(line) + ": bad score");
    const unsigned long long ce = std::strtoull(f[3].c_str(), &end, 10);
    if (f[3].empty() || *end || errno || f[3][0] == '-')
      return fail(ST_ERR_INVALID, "st_parse_candidate_trace: line " + std::to_string(line) + ": bad cell count");
    auto qi = q_of.find(f[0]);
    uint32_t q;
    if (qi == q_of.end()) {
      q = (uint32_t)qids.size();
      q_of.emplace(f[0], q);
      qids.push_back(f[0]);
    } else {
      q = qi->second;
    }
    rows.push_back(Row{q, it->second, (int32_t)sc, (uint64_t)ce});
  }
  // candidates grouped per query (first appearance), in trace order
  std::stable_sort(rows.begin(), rows.end(), [](const Row &x, const Row &y) { return x.q < y.q; });
  *n_queries = qids.size();
  *n_cands = rows.size();
  uint64_t qb = 0;
  for (const std::string &q : qids) qb += q.size();
  if (qid_off) {
    qid_off[0] = 0;
    for (size_t k = 0; k < qids.size(); k++) qid_off[k + 1] = qid_off[k] + qids[k].size();
  }
  if (qid_bytes && qid_cap >= qb) {
    uint64_t o = 0;
    for (const std::string &q : qids) {
      std::memcpy(qid_bytes + o, q.data(), q.size());
      o += q.size();
    }
  }
  if (cand_off) {
    std::fill(cand_off, cand_off + qids.size() + 1, 0);
    for (const Row &r : rows) cand_off[r.q + 1]++;
    for (size_t k = 0; k < qids.size(); k++) cand_off[k + 1] += cand_off[k];
  }
  for (size_t k = 0; k < rows.size(); k++) {
    if (cand_tids) cand_tids[k] = rows[k].tid;
    if (diag_scores) diag_scores[k] = rows[k].score;
    if (cells) cells[k] = rows[k].cells;
  }
  return ST_OK;
}

int st_format_candidate_trace(const char *query_id, const char *const *target_ids, const int32_t *diag_scores,
                              const uint64_t *cells, uint64_t n, char *buf, uint64_t cap, uint64_t *len) {
  if (!len || (n && (!query_id || !target_ids || !diag_scores || !cells)))
    return fail(ST_ERR_INVALID, "st_format_candidate_trace: null argument");
  std::string out;
  for (uint64_t k = 0; k < n; k++) {
    if (!target_ids[k]) return fail(ST_ERR_INVALID, "st_format_candidate_trace: null target id");
    out += std::string(query_id) + "\t" + target_ids[k] + "\t" + std::to_string(diag_scores[k]) + "\t" +
           std::to_string(cells[k]) + "\n";
  }
  *len = out.size();
  if (buf && cap) {
    const uint64_t k = std::min<uint64_t>(cap - 1, out.size());
    std::memcpy(buf, out.data(), k);
    buf[k] = 0;
  }
  return ST_OK;
}

int st_format_m8_extra(const st_hit *h, const char *cigar, char *buf, uint64_t cap, uint64_t *len) {
  // Hit::extra_cols (src/search.rs:219-230): raw score, query and target
  // lengths, query and target coverage {:.3} (span / max(len, 1)), CIGAR
  if (!h) return fail(ST_ERR_INVALID, "st_format_m8_extra: null hit");
  const double qcov = (double)(h->q_end + 1 - h->q_start) / (double)std::max<uint64_t>(h->qlen, 1);
  const double tcov = (double)(h->t_end + 1 - h->t_start) / (double)std::max<uint64_t>(h->tlen, 1);
  const std::string s = std::to_string(h->raw_score) + "\t" + std::to_string(h->qlen) + "\t" +
                        std::to_string(h->tlen) + "\t" + rust_f(qcov, 3) + "\t" + rust_f(tcov, 3) + "\t" +
                        (cigar ? cigar : "");
  if (len) *len = s.size();
  if (buf && cap) {
    const uint64_t k = std::min<uint64_t>(cap - 1, s.size());
    std::memcpy(buf, s.data(), k);
    buf[k] = 0;
  }
  return ST_OK;
}

int st_format_m8_block(const char *query_id, const char *const *target_ids, const st_hit *hits,
                       const char *const *cigars, uint64_t n, int extra, int header, char *buf, uint64_t cap,
                       uint64_t *len) {
  // write_header / write_hits (src/main.rs:667-687): one '\n'-terminated
  // line per hit, the extra columns after a tab when asked for
  if (!len || (n && (!query_id || !target_ids || !hits)) || (extra && n && !cigars))
    return fail(ST_ERR_INVALID, "st_format_m8_block: null argument");
  std::string out;
  if (header) {
    out += "query\ttarget\tpident\talnlen\tmismatch\tgapopen\tqstart\tqend\ttstart\ttend\tevalue\tbits";
    if (extra) out += "\traw\tqlen\ttlen\tqcov\ttcov\tcigar";
    out += "\n";
  }
  std::vector<char> line;
  for (uint64_t i = 0; i < n; i++) {
    if (!target_ids[i]) return fail(ST_ERR_INVALID, "st_format_m8_block: null target id");
    uint64_t l = 0;
    int rc = st_format_m8(query_id, target_ids[i], &hits[i], nullptr, 0, &l);
    if (rc) return rc;
    line.resize(l + 1);
    st_format_m8(query_id, target_ids[i], &hits[i], line.data(), line.size(), &l);
    out.append(line.data(), l);
    if (extra) {
      rc = st_format_m8_extra(&hits[i], cigars[i], nullptr, 0, &l);
      if (rc) return rc;
      line.resize(l + 1);
      st_format_m8_extra(&hits[i], cigars[i], line.data(), line.size(), &l);
      out += "\t";
      out.append(line.data(), l);
    }
    out += "\n";
  }
  *len = out.size();
  if (buf && cap) {
    const uint64_t k = std::min<uint64_t>(cap - 1, out.size());
    std::memcpy(buf, out.data(), k);
    buf[k] = 0;
  }
  return ST_OK;
}

}  // extern "C"
