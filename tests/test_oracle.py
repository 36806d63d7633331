"""Pins of the CPU oracle (no GPU).

The oracle restates the reference's scalar kernels; these tests pin it to the
reference's own known answers and test generators, to brute force on tiny
inputs and to closed forms, so that a dropped term, a swapped gap state or a
wrong tie rule fails here before the oracle is trusted as the GPU checker.
"""
import itertools
import math

import numpy as np
import pytest

import oracle as O


class XorShift:
    """The reference's test RNG (src/simd.rs:1038-1051)."""

    def __init__(self, seed):
        self.s = seed & 0xFFFFFFFFFFFFFFFF

    def next(self):
        x = self.s
        x ^= (x << 13) & 0xFFFFFFFFFFFFFFFF
        x ^= x >> 7
        x ^= (x << 17) & 0xFFFFFFFFFFFFFFFF
        self.s = x
        return x

    def range(self, n):
        return self.next() % n

    def seq(self, n, alpha):
        return np.array([self.range(alpha) for _ in range(n)], dtype=np.uint8)


# ---- alphabet / matrix --------------------------------------------------------

def test_alphabet_roundtrip_and_ambiguity():
    t = O.aa2num()
    for i, c in enumerate("ACDEFGHIKLMNPQRSTVWYX"):
        assert t[ord(c)] == i and t[ord(c.lower())] == i
    for c in "BZJUO*?.":
        assert t[ord(c)] == 20


def test_blosum62_known_cells():
    s, pb, r, lam = O.blosum62()
    # the MMseqs2 reconstruction at bit factor 2 / bias -0.2 (README 400/400):
    # W-W is 10 here (not the NCBI 11), C-C 8, X row -1 everywhere
    aa = {c: i for i, c in enumerate("ACDEFGHIKLMNPQRSTVWYX")}
    assert s[aa["W"], aa["W"]] == 10
    assert s[aa["C"], aa["C"]] == 8
    assert s[aa["A"], aa["A"]] == 4
    assert (s[20] == -1).all() and (s[:, 20] == -1).all()
    assert (s == s.T).all()
    assert all(s[i, i] > 0 for i in range(20))
    assert s[aa["W"], aa["W"]] == s.max()
    # background after the X rescale sums to ~1; R columns are P(.|b)
    assert abs(pb[:20].sum() - 1.0) < 0.01 and pb[20] == 1e-5
    assert np.all(np.abs(r.sum(axis=0) - 1.0) < 0.02)
    assert abs(lam - 0.34657) < 1e-12


def test_blosum62_matches_closed_form_of_the_data_file():
    # half-bit value v -> round-half-away(2*log2(exp(lambda*v)) - 0.2): the
    # pb factors cancel exactly, so each cell depends on one datum only
    s, _, _, lam = O.blosum62()
    for v, want in [(10.5040, 10), (3.9291, 4), (-1.0, -1), (-4.2143, -4), (0.0017, 0), (-0.0454, 0)]:
        x = 2 * math.log2(math.exp(lam * v)) - 0.2
        assert int(x - 0.5 if x < 0 else x + 0.5) == want


# ---- SWG ------------------------------------------------------------------------

def brute_local(rows, t, open_, ext):
    """Best local alignment score by enumerating every alignment path of
    tiny inputs (independent of any DP formulation)."""
    m, n = len(rows), len(t)
    best = 0

    def sc(i, j):
        return int(rows[i][t[j]]) if t[j] < 21 else 0

    # walk: state at (i, j) having just aligned i-1/j-1 ... enumerate paths
    def walk(i, j, total, last):
        nonlocal best
        if total > best:
            best = total
        if i < m and j < n:
            walk(i + 1, j + 1, total + sc(i, j), "M")
        # gaps may not start a local alignment or follow nothing (they never help)
        if last is not None:
            if j < n:
                walk(i, j + 1, total - (ext if last == "I" else open_), "I")
            if i < m:
                walk(i + 1, j, total - (ext if last == "D" else open_), "D")

    for i0 in range(m):
        for j0 in range(n):
            walk(i0 + 1, j0 + 1, sc(i0, j0), "M")
    return best


@pytest.mark.parametrize("open_,ext", [(11, 1), (2, 1), (3, 3), (6, 2), (1, 1), (5, 0)])
def test_gotoh_equals_brute_force(open_, ext):
    rng = XorShift(2024 + open_ * 7 + ext)
    for _ in range(60):
        m, n = 1 + rng.range(4), 1 + rng.range(5)
        alpha = 2 + rng.range(19)
        q, t = rng.seq(m, alpha), rng.seq(n, alpha)
        rows = O.seq_rows(q)
        want = brute_local(rows, t, open_, ext)
        assert O.gotoh_score(rows, t, open_, ext) == want
        assert O.gotoh_textbook(rows, t, open_, ext)[0] == want


def test_gotoh_scalar_matches_textbook_reference_generator():
    # the reference's gap list and generator (src/simd.rs:1144-1164)
    rng = XorShift(12345)
    for open_, ext in [(1, 1), (2, 1), (3, 1), (2, 0), (4, 1), (11, 1), (6, 2), (5, 5)]:
        for _ in range(60):
            ql, tl = 1 + rng.range(90), 1 + rng.range(90)
            alpha = 2 + rng.range(19)
            q, t = rng.seq(ql, alpha), rng.seq(tl, alpha)
            rows = O.seq_rows(q)
            a = O.gotoh_score_end(rows, t, open_, ext)
            b = O.gotoh_textbook(rows, t, open_, ext)
            assert a == b, (open_, ext, q, t)


def test_end_cell_column_major_first():
    # tiny alphabets give many tied optima (src/simd.rs:1197-1234)
    rng = XorShift(424242)
    for open_, ext in [(11, 1), (2, 1), (6, 2), (3, 3), (1, 1)]:
        for rnd in range(80):
            ql, tl = 1 + rng.range(30), 1 + rng.range(30)
            alpha = 2 + rng.range(3) if rnd % 3 == 0 else 2 + rng.range(19)
            q, t = rng.seq(ql, alpha), rng.seq(tl, alpha)
            if rnd % 7 == 0:
                t[0] = 20
            rows = O.seq_rows(q)
            # independent column-major scan of the full matrix
            sc, qe, te = O.gotoh_textbook(rows, t, open_, ext)
            assert O.gotoh_score_end(rows, t, open_, ext) == (sc, qe, te)


def test_known_answers_from_reference_tests():
    W = O.aa2num()[ord("W")]
    q = np.full(3000, W, np.uint8)
    assert O.gotoh_score_end(O.seq_rows(q), q, 11, 1) == (30000, 2999, 2999)  # src/simd.rs:1233
    q4 = np.full(4000, W, np.uint8)
    assert O.gotoh_score(O.seq_rows(q4), q4, 11, 1) == 40000  # src/simd.rs:1327-1348
    assert O.gotoh_score(np.full((3, 21), 40000), [0, 1, 2], 11, 1) == 120000  # src/simd.rs:1350-1355
    # empty query / target score 0 (src/simd.rs:253-255)
    assert O.gotoh_score(np.zeros((0, 21)), [1, 2], 11, 1) == 0
    assert O.gotoh_score(O.seq_rows([1, 2]), [], 11, 1) == 0


def test_out_of_alphabet_codes_score_zero():
    rng = XorShift(7)
    q = rng.seq(30, 20)
    t = XorShift(9).seq(40, 20)
    rows = O.seq_rows(q)
    t2 = t.copy()
    t2[7] = 200
    t3 = t.copy()
    t3[7] = 21
    assert O.gotoh_score(rows, t2, 11, 1) == O.gotoh_score(rows, t3, 11, 1)
    # a code >= 21 is a zero column: equal to a row of zeros at that position
    rows_x = rows.copy()
    assert O.gotoh_textbook(rows_x, t2, 11, 1)[0] == O.gotoh_score(rows, t2, 11, 1)


def test_identical_sequences_align_fully():
    rng = XorShift(5)
    for _ in range(20):
        q = rng.seq(1 + rng.range(200), 20)
        rows = O.seq_rows(q)
        self_sum = int(sum(rows[i][q[i]] for i in range(len(q))))
        assert O.gotoh_score(rows, q, 11, 1) >= self_sum


# ---- ungapped ---------------------------------------------------------------------

def test_diag_range_closed_form():
    for m, n in [(5, 7), (7, 5), (1, 1), (3, 0)]:
        for d in range(-m - 2, n + 3):
            lo, hi = O.diag_range(m, n, d)
            cells = [i for i in range(m) if 0 <= i + d < n]
            if cells:
                assert (lo, hi) == (cells[0], cells[-1] + 1)
            else:
                assert (lo, hi) == (0, 0)


def test_kadane_equals_brute_force_max_subarray():
    sm = O.SplitMix64(3)
    for _ in range(200):
        m = 1 + sm.next_u64() % 40
        n = 1 + sm.next_u64() % 60
        q = np.array([sm.next_u64() % 21 for _ in range(m)], np.uint8)
        t = np.array([sm.next_u64() % 22 for _ in range(n)], np.uint8)
        rows = O.seq_rows(q)
        for d in range(-m - 1, n + 1):
            lo, hi = O.diag_range(m, n, d)
            s = [int(rows[i][t[i + d]]) if t[i + d] < 21 else 0 for i in range(lo, hi)]
            brute = max([0] + [sum(s[a:b]) for a in range(len(s)) for b in range(a + 1, len(s) + 1)])
            assert O.kadane_score(rows, t, d, lo, hi) == brute
            sc, qs, qe = O.kadane_segment(rows, t, d, lo, hi)
            assert sc == brute
            if brute > 0:
                # first maximum in query order: earliest end, then latest start
                ends = [b for b in range(1, len(s) + 1)
                        if max(sum(s[a:b]) for a in range(b)) == brute]
                e = ends[0]
                starts = [a for a in range(e) if sum(s[a:e]) == brute]
                assert (qs, qe) == (lo + starts[-1], lo + e - 1)


# ---- gate ---------------------------------------------------------------------------

def _db(seqs):
    offs = np.zeros(len(seqs) + 1, np.uint64)
    offs[1:] = np.cumsum([len(s) for s in seqs])
    return np.concatenate(seqs).astype(np.uint8), offs


def test_chain_rescues_indel_split_homolog():
    # the construction of src/prefilter.rs:1938-1958: a repeat on a wrong
    # diagonal plus the real homology split by a 4-residue insertion
    q = O.encode("MKVLLACDEFGHIKLMNPQRSTVWYHHHHWWCCMM")
    t = O.encode("QRST" * 7 + "MKVLLACDEFGHIKLMNP" + "GGGG" + "QRSTVWYHHHHWWCCMM")
    rows = O.seq_rows(q)
    arena, offs = _db([t])
    m, n = len(q), len(t)
    diags = [28, 32, 0, 4]
    jobs = np.zeros(len(diags), dtype=O.JOB_DT)
    for k, d in enumerate(diags):
        lo, hi = O.diag_range(m, n, d)
        jobs[k] = (0, d, lo, hi)
    recs = np.array([(0, 9, 0, len(diags))], dtype=O.REC_DT)
    cands, chains, scores = O.score_and_gate(rows, arena, offs, jobs, recs, 11, 1, 0, regions=True)
    s28, s32 = int(scores[0]), int(scores[1])
    block1 = sum(int(rows[i][q[i]]) for i in range(18))
    block2 = sum(int(rows[i][q[i]]) for i in range(18, 35))
    assert s28 >= block1 and s32 >= block2
    assert chains == 1 and len(cands) == 1
    c = cands[0]
    assert c["diag_score"] == s28 + s32 - (11 + 3 * 1)  # gap of |32-28| = 4
    assert c["diag_score"] > max(scores)
    assert c["best_diag"] == diags[int(np.argmax(scores))]
    # region spans both segments on the target
    assert c["t_lo"] <= 28 and c["t_hi"] >= 28 + 34
    # the all-diagonal evidence of the same pair finds the same chain
    ung, bd, gs, ok = O.ung_pair(rows, t, 11, 1, 0)
    assert ung == max(scores) and gs == c["diag_score"] and ok


def test_gate_threshold_and_single_job_shortcut():
    q = O.encode("MKVLLACDEFGHIK")
    t = O.encode("MKVLLACDEFGHIKMKVLLAC")
    rows = O.seq_rows(q)
    arena, offs = _db([t])
    lo, hi = O.diag_range(len(q), len(t), 0)
    jobs = np.array([(0, 0, lo, hi)], dtype=O.JOB_DT)
    recs = np.array([(0, 1, 0, 1)], dtype=O.REC_DT)
    _, _, sc = O.score_and_gate(rows, arena, offs, jobs, recs, 11, 1, 0)
    s = int(sc[0])
    assert s == sum(int(rows[i][q[i]]) for i in range(len(q)))
    assert len(O.score_and_gate(rows, arena, offs, jobs, recs, 11, 1, s)[0]) == 1
    assert len(O.score_and_gate(rows, arena, offs, jobs, recs, 11, 1, s + 1)[0]) == 0


def test_size_adjusted_gate_operating_points():
    # src/prefilter.rs:2089-2112
    assert O.size_adjusted_gate(42, 100_000_000) == 42
    assert O.size_adjusted_gate(42, 2_611_087) == 33
    assert O.size_adjusted_gate(42, 68_310_060) == 41
    assert O.size_adjusted_gate(42, 265_123_580) == 43
    assert O.size_adjusted_gate(42, 0) == 20 and O.size_adjusted_gate(42, 10) == 20
    e2 = math.e ** 2
    assert O.size_adjusted_gate(42, int(1e8 / e2)) == round(42 - 5.0)
    assert O.size_adjusted_gate(42, int(1e8 * e2)) == 44


def test_ung_pair_definition_by_enumeration():
    sm = O.SplitMix64(11)
    for _ in range(40):
        m = 1 + sm.next_u64() % 25
        n = 1 + sm.next_u64() % 25
        q = np.array([sm.next_u64() % 20 for _ in range(m)], np.uint8)
        t = np.array([sm.next_u64() % 20 for _ in range(n)], np.uint8)
        rows = O.seq_rows(q)
        per = []
        for d in range(-(m - 1), n):
            lo, hi = O.diag_range(m, n, d)
            per.append((O.kadane_score(rows, t, d, lo, hi), d))
        best = max(s for s, _ in per)
        bd = min(d for s, d in per if s == best)
        ung, bdiag, gs, ok = O.ung_pair(rows, t, 11, 1, 30)
        assert (ung, bdiag) == (best, bd)
        top = sorted(per, key=lambda x: (-x[0], x[1]))[:4]
        # the gate score is the best single or the best co-linear chain of two
        chain = best
        for (sx, dx), (sy, dy) in itertools.permutations(top, 2):
            if sx <= 0 or sy <= 0 or dx == dy:
                continue
            _, qsx, qex = O.kadane_segment(rows, t, dx, *O.diag_range(m, n, dx))
            _, qsy, qey = O.kadane_segment(rows, t, dy, *O.diag_range(m, n, dy))
            if qex < qsy and qex + dx < qsy + dy:
                chain = max(chain, sx + sy - (11 + (abs(dy - dx) - 1)))
        assert gs == chain and ok == (gs >= 30)


# ---- RNG --------------------------------------------------------------------------------

def test_splitmix64_published_vector():
    # SplitMix64 from seed 0: the published first outputs of Vigna's generator
    s = O.SplitMix64(0)
    assert s.next_u64() == 0xE220A8397B1DCDAF
    assert s.next_u64() == 0x6E789E6AA1B965F4
    assert s.next_u64() == 0x06C45D188009454F


def test_background_sampler_frequencies():
    x = O.SplitMix64(1).background(200_000)
    _, pb, _, _ = O.blosum62()
    f = np.bincount(x, minlength=20)[:20] / len(x)
    assert np.all(np.abs(f - pb[:20] / pb[:20].sum()) < 0.004)


# ---- the seed-hit stage (hits -> top diagonals -> jobs) -----------------------

def kmer_index(targets, k):
    """Exact k-mer posting lists of targets: postings (target << 32 | pos)
    sorted by (k-mer, target, position), and k-mer -> [start, end)."""
    ents = []
    for t, s in enumerate(targets):
        for p in range(len(s) - k + 1):
            ents.append((tuple(int(x) for x in s[p:p + k]), t, p))
    ents.sort()
    post = np.array([(t << 32) | p for _, t, p in ents], dtype=np.uint64)
    rng = {}
    for i, (km, _, _) in enumerate(ents):
        a, b = rng.get(km, (i, i))
        rng[km] = (a, i + 1)
    return post, rng


def query_seeds(q, k, rng_of, extra=0, n_post=0, rng=None):
    seeds = []
    for i in range(len(q) - k + 1):
        km = tuple(int(x) for x in q[i:i + k])
        if km in rng_of:
            a, b = rng_of[km]
            seeds.append((i, a, b))
    for _ in range(extra):
        a = int(rng.integers(0, n_post))
        seeds.append((int(rng.integers(0, len(q))), a, int(min(n_post, a + rng.integers(0, 6)))))
    return seeds


def _brute_jobs(m, seeds, post, tlen, kmax, min_hits, diag_hits, window, span):
    from collections import defaultdict
    per = defaultdict(list)
    for q, a, b in seeds:
        for v in post[a:b]:
            per[int(v) >> 32].append(((int(v) & 0xFFFFFFFF) - q, q))
    jobs, recs = [], []
    for t in sorted(per):
        hits = per[t]
        if len(hits) < min_hits:
            continue
        runs = {}
        for d, q in hits:
            c, lo, hi = runs.get(d, (0, q, q))
            runs[d] = (c + 1, min(lo, q), max(hi, q))
        # the kmax best; jobs for at most 16 of them (src/prefilter.rs:1466)
        ranked = sorted(runs.items(), key=lambda kv: (-kv[1][0], kv[0]))[:min(kmax, 16)]
        if ranked[0][1][0] < diag_hits:
            continue
        j0 = len(jobs)
        for d, (c, qmin, qmax) in ranked:
            lo, hi = max(-d, 0), min(m, int(tlen[t]) - d)
            if hi <= lo:
                lo = hi = 0
            if window > 0:
                lo, hi = max(lo, qmin - window, 0), min(hi, qmax + span + window)
            jobs.append((t, d, lo, max(hi, lo)))
        recs.append((t, len(hits), j0, len(jobs)))
    return jobs, recs


@pytest.mark.parametrize("kmax,min_hits,diag_hits,window", [(4, 1, 1, 48), (1, 1, 1, 0), (16, 2, 2, 5),
                                                            (3, 1, 2, 0), (2, 3, 1, 1), (20, 1, 1, 0),
                                                            (40, 1, 2, 3)])
def test_seed_hit_jobs_against_brute_force(kmax, min_hits, diag_hits, window):
    rng = np.random.default_rng(kmax * 100 + min_hits * 10 + diag_hits)
    targets = [rng.integers(0, 4, size=int(rng.integers(3, 90)), dtype=np.uint8) for _ in range(60)]
    post, of = kmer_index(targets, 3)
    tlen = np.array([len(t) for t in targets])
    for _ in range(6):
        q = rng.integers(0, 4, size=int(rng.integers(3, 70)), dtype=np.uint8)
        seeds = query_seeds(q, 3, of, extra=10, n_post=len(post), rng=rng)
        jobs, recs = O.seed_hits_to_jobs(len(q), seeds, post, tlen, kmax, min_hits, diag_hits, window, 3)
        bj, br = _brute_jobs(len(q), seeds, post, tlen, kmax, min_hits, diag_hits, window, 3)
        assert [tuple(int(x) for x in j) for j in jobs] == bj
        assert [tuple(int(x) for x in r) for r in recs] == br
        # the multiset of hits alone decides: seed order does not matter
        perm = [seeds[i] for i in rng.permutation(len(seeds))]
        j2, r2 = O.seed_hits_to_jobs(len(q), perm, post, tlen, kmax, min_hits, diag_hits, window, 3)
        assert (j2 == jobs).all() and (r2 == recs).all()
