// Seeded synthetic protein workloads (host only, no GPU).
//
// Residues are drawn i.i.d. from the BLOSUM62 background with SplitMix64 as
// the reference's BackgroundSampler does (src/evalue.rs:52-114); homologs
// substitute residues from the conditional R[a][b] = P(a|b)
// (src/matrix.rs:145-151) and add indel runs.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

extern "C" int st_blosum62(int8_t *score, double *p_back, float *r, double *lambda);

namespace {

struct Rng {
  uint64_t s;
  uint64_t next() {
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double f64() { return (double)(next() >> 11) / 9007199254740992.0; }
  uint64_t below(uint64_t n) { return next() % n; }
};

struct Tables {
  double cum[20];
  // per source residue b: cumulative P(a|b) over a != b, renormalised
  double sub[20][20];
  Tables() {
    double pb[21];
    float r[400];
    st_blosum62(nullptr, pb, r, nullptr);
    double tot = 0, acc = 0;
    for (int i = 0; i < 20; i++) tot += pb[i];
    for (int i = 0; i < 20; i++) {
      acc += pb[i] / tot;
      cum[i] = acc;
    }
    cum[19] = 1.0;
    for (int b = 0; b < 20; b++) {
      double t = 0;
      for (int a = 0; a < 20; a++)
        if (a != b) t += r[a * 20 + b];
      double c = 0;
      for (int a = 0; a < 20; a++) {
        if (a != b) c += r[a * 20 + b] / t;
        sub[b][a] = c;
      }
      sub[b][19] = 1.0;
      if (b == 19) sub[b][18] = 1.0;
    }
  }
};

const Tables &tables() {
  static Tables t;
  return t;
}

uint8_t background(Rng &g) {
  const double u = g.f64();
  const Tables &t = tables();
  for (int i = 0; i < 20; i++)
    if (u < t.cum[i]) return (uint8_t)i;
  return 19;
}

uint8_t substitute(Rng &g, uint8_t b) {
  const Tables &t = tables();
  if (b >= 20) return background(g);
  const double u = g.f64();
  for (int a = 0; a < 20; a++)
    if (a != b && u < t.sub[b][a]) return (uint8_t)a;
  return b == 19 ? 18 : 19;
}

}  // namespace

extern "C" {

// BackgroundSampler::sequence: n residues; *state advances.
void sy_background(uint64_t *state, uint64_t n, uint8_t *out) {
  Rng g{*state};
  for (uint64_t k = 0; k < n; k++) out[k] = background(g);
  *state = g.s;
}

// count lengths exp(ln(median) + sigma * z), z standard normal (Box-Muller),
// rounded and clipped to [lo, hi].
void sy_lognormal(uint64_t *state, uint64_t count, double median, double sigma, uint32_t lo, uint32_t hi,
                  uint32_t *out) {
  Rng g{*state};
  for (uint64_t k = 0; k < count; k++) {
    double u1 = g.f64();
    const double u2 = g.f64();
    if (u1 < 1e-300) u1 = 1e-300;
    const double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    double v = std::exp(std::log(median) + sigma * z);
    v = std::floor(v + 0.5);
    if (v < lo) v = lo;
    if (v > hi) v = hi;
    out[k] = (uint32_t)v;
  }
  *state = g.s;
}

// uniform integers in [lo, hi]
void sy_uniform(uint64_t *state, uint64_t count, uint32_t lo, uint32_t hi, uint32_t *out) {
  Rng g{*state};
  for (uint64_t k = 0; k < count; k++) out[k] = lo + (uint32_t)g.below((uint64_t)hi - lo + 1);
  *state = g.s;
}

// A homolog of src[0..m): each position is kept with probability `identity`
// or substituted from P(a|b) (a != b); at each position an indel starts with
// probability `indel_rate` (insertion of background residues or deletion,
// even odds) of length U(1, indel_max).  The result is cut or extended with
// background residues to exactly out_len.  Returns out_len.
uint64_t sy_homolog(uint64_t *state, const uint8_t *src, uint64_t m, double identity, double indel_rate,
                    uint32_t indel_max, uint64_t out_len, uint8_t *out) {
  Rng g{*state};
  std::vector<uint8_t> h;
  h.reserve(m + 64);
  for (uint64_t i = 0; i < m;) {
    if (indel_rate > 0 && g.f64() < indel_rate) {
      const uint32_t L = 1 + (uint32_t)g.below(indel_max);
      if (g.below(2) == 0) {
        for (uint32_t k = 0; k < L; k++) h.push_back(background(g));
      } else {
        i += L;
        continue;
      }
    }
    h.push_back(g.f64() < identity ? src[i] : substitute(g, src[i]));
    i++;
  }
  // place the homologous stretch at a random offset inside out_len
  if (h.size() > out_len) {
    const uint64_t cut = g.below(h.size() - out_len + 1);
    std::memcpy(out, h.data() + cut, out_len);
  } else {
    const uint64_t pad = out_len - h.size();
    const uint64_t left = pad ? g.below(pad + 1) : 0;
    for (uint64_t k = 0; k < left; k++) out[k] = background(g);
    std::memcpy(out + left, h.data(), h.size());
    for (uint64_t k = left + h.size(); k < out_len; k++) out[k] = background(g);
  }
  *state = g.s;
  return out_len;
}

// n distinct-ish random ids in [0, range) (with replacement)
void sy_pick(uint64_t *state, uint64_t n, uint64_t range, uint32_t *out) {
  Rng g{*state};
  for (uint64_t k = 0; k < n; k++) out[k] = (uint32_t)g.below(range);
  *state = g.s;
}

// One a3m member row of an MSA built around `root` (m codes): each root
// column keeps its residue with probability `identity` (else a P(a|b)
// substitute) as an uppercase letter; a gap run of 1..10 columns ('-')
// starts with probability gap_rate; after a column, an insertion of 1..5
// lowercase background residues follows with probability insert_rate.
// Writes at most cap bytes of text (no newline) and returns the length.
uint64_t sy_msa_member(uint64_t *state, const uint8_t *root, uint64_t m, double identity, double gap_rate,
                       double insert_rate, char *out, uint64_t cap) {
  static const char kAA[] = "ACDEFGHIKLMNPQRSTVWYX";
  Rng g{*state};
  uint64_t n = 0;
  auto put = [&](char c) {
    if (n < cap) out[n] = c;
    n++;
  };
  for (uint64_t i = 0; i < m;) {
    if (g.f64() < gap_rate) {
      const uint64_t L = std::min<uint64_t>(1 + g.below(10), m - i);
      for (uint64_t k = 0; k < L; k++) put('-');
      i += L;
      continue;
    }
    const uint8_t r = g.f64() < identity ? root[i] : substitute(g, root[i]);
    put(kAA[r < 21 ? r : 20]);
    i++;
    if (i < m && g.f64() < insert_rate) {
      const uint64_t L = 1 + g.below(5);
      for (uint64_t k = 0; k < L; k++) put((char)(kAA[background(g)] + 32));
    }
  }
  *state = g.s;
  return n;
}

}  // extern "C"
