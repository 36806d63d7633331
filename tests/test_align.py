"""Exact traceback of reported hits (host C++) vs the oracle's textbook
Gotoh and the reference's alignment tests (src/align.rs:1222-1382)."""
import numpy as np
import pytest

import oracle as O
from sabertooth_amd import GapCosts
from sabertooth_amd import host as H
from test_oracle import XorShift


def _enc(s):
    return O.encode(s)


def _mutate(rng, s, indel_len):
    # the reference's mutate() (src/align.rs:1186-1204)
    out, i = [], 0
    while i < len(s):
        r = rng.range(100)
        if r < 4:
            i += 1 + rng.range(indel_len)
            continue
        if r < 8:
            for _ in range(1 + rng.range(indel_len)):
                out.append(rng.range(20))
        out.append(rng.range(20) if r < 30 else int(s[i]))
        i += 1
    return np.array(out, dtype=np.uint8)


def _check(q, t, gaps, radii=(1, 3, 16)):
    rows = O.seq_rows(q)
    want = O.gotoh_score(rows, t, gaps.open, gaps.extend)
    se = H.sw_score_end_host(rows, t, gaps)
    ref = O.gotoh_textbook(rows, t, gaps.open, gaps.extend)
    assert se[0] == want and (se if want > 0 else (0, -1, -1)) == (ref if want > 0 else (0, -1, -1))
    for r in radii:
        a = H.align_canonical(rows, q, t, gaps, band_radius=r)
        assert (a.raw_score if a else 0) == want, (r, q, t)
        if a:
            assert H.rescore_alignment(rows, q, t, gaps, a) == a.raw_score
            assert (a.query_end, a.target_end) == (se[1], se[2])
            b = H.align_canonical(rows, q, t, gaps, band_radius=r, known=se)
            assert b == a


def test_traceback_rescores_exactly_on_random_gapped_pairs():
    rng = XorShift(0x9E3779B97F4A7C15)
    for o, e in [(11, 1), (5, 1), (3, 1), (8, 2), (2, 2), (1, 1)]:
        for _ in range(120):
            ql = 1 + rng.range(70)
            q = rng.seq(ql, 20)
            t = _mutate(rng, q, 6) if rng.range(3) > 0 else rng.seq(1 + rng.range(70), 20)
            if len(t) == 0:
                continue
            _check(q, t, GapCosts(o, e))


@pytest.mark.parametrize("q,t", [
    ("MKVLLACDEFGHIKLMNPQRSTVWYACDEFGHIKMKV", "MKVLLACDEFGHIKLMNPQR" + "G" * 30 + "STVWYACDEFGHIKMKV"),
    ("MKVLLACDEFGHIKLMNPQR" + "W" * 25 + "STVWYACDEFGHIK", "MKVLLACDEFGHIKLMNPQRSTVWYACDEFGHIK"),
    ("ACDEFACDEFACDEF", "ACDEFACDEFACDEFACDEFACDEF"), ("WWWW", "WWWWWWWWWW"), ("W", "W"), ("W", "AAAAW"),
    ("A", "W"), ("MKVL", "MKVLXXXXMKVL"), ("MKXXLLAXCDE", "MKVVLLAXCDE"), ("XXXXXX", "XXXXXX"),
])
def test_long_gaps_repeats_ties_and_boundaries(q, t):
    # src/align.rs:1233-1265
    _check(_enc(q), _enc(t), GapCosts(11, 1))


def test_empty_inputs():
    q = _enc("MKV")
    assert H.align_canonical(O.seq_rows(q), q, np.zeros(0, np.uint8), GapCosts(11, 1)) is None
    assert H.align_canonical(np.zeros((0, 21), np.int32), None, q, GapCosts(11, 1)) is None


def test_v02_traceback_regression_case():
    # src/align.rs:1267-1281
    q, t = _enc("WFMMHGWW"), _enc("WFMWW")
    a = H.align_canonical(O.seq_rows(q), q, t, GapCosts(11, 1))
    assert a.raw_score == O.gotoh_score(O.seq_rows(q), t, 11, 1)
    assert H.rescore_alignment(O.seq_rows(q), q, t, GapCosts(11, 1), a) == a.raw_score
    assert a.gap_opens == 1 and a.gaps == 3


def test_long_gap_is_a_single_run():
    # src/align.rs:1283-1296
    q = _enc("MKVLLACDEFGHIKLMNPQRSTVWYMKVLLACDEFGHIK")
    t = _enc("MKVLLACDEFGHIKLMNPQRGGGGSTVWYMKVLLACDEFGHIK")
    a = H.align_canonical(O.seq_rows(q), q, t, GapCosts(11, 1))
    assert a.cigar == "20M4I19M" and a.gap_opens == 1 and a.gaps == 4
    assert a.mismatches == 0 and a.identities == 39


def test_identical_sequences_align_fully():
    # src/align.rs:1298-1311
    q = _enc("MKVLLACDE")
    a = H.align_canonical(O.seq_rows(q), q, q, GapCosts(11, 1))
    assert (a.query_start, a.query_end, a.identities, a.cigar) == (0, 8, 9, "9M")
    assert abs(a.pct_identity() - 100.0) < 1e-9 and abs(a.query_cov(9) - 1.0) < 1e-12


def test_large_problem_matches_score():
    # src/align.rs:1350-1365 (m*n above the reference's full-traceback limit)
    rng = XorShift(42)
    q = rng.seq(4500, 20)
    t = np.concatenate([rng.seq(2000, 20), _mutate(rng, q[1000:3500], 8)])
    rows = O.seq_rows(q)
    a = H.align_canonical(rows, q, t, GapCosts(11, 1))
    assert H.rescore_alignment(rows, q, t, GapCosts(11, 1), a) == a.raw_score
    assert a.raw_score == O.gotoh_score(rows, t, 11, 1)


def test_profile_rows_align():
    # src/align.rs:1336-1348
    from sabertooth_amd.host import Profile
    p = Profile.from_msa_str(">q\nMKVLLACDE\n>h\nMKILLACDE\n>i\nMRVLLSCDE\n")
    t = _enc("MKVLLACDE")
    a = H.align_canonical(p.rows(), p.query_num, t, GapCosts(44, 4))
    assert a.raw_score > 0 and a.query_end - a.query_start + 1 == 9
    H.rescore_alignment(p.rows(), p.query_num, t, GapCosts(44, 4), a)


def test_rescore_rejects_a_wrong_cigar():
    q = _enc("MKVLLACDEFGHIKLMNPQRSTVWYMKVLLACDEFGHIK")
    t = _enc("MKVLLACDEFGHIKLMNPQRGGGGSTVWYMKVLLACDEFGHIK")
    a = H.align_canonical(O.seq_rows(q), q, t, GapCosts(11, 1))
    a.cigar = "19M4I20M"
    with pytest.raises(Exception):
        H.rescore_alignment(O.seq_rows(q), q, t, GapCosts(11, 1), a)


def _best_from(rows, t, o, e, i0, j0, ie, je):
    """Best score of an alignment that begins by pairing (i0, j0) and ends
    with a pair anywhere in [i0, ie] x [j0, je] (full Gotoh, no floor)."""
    neg = -(1 << 40)
    m, n = ie - i0 + 1, je - j0 + 1
    H = [[neg] * n for _ in range(m)]
    E = [[neg] * n for _ in range(m)]
    F = [[neg] * n for _ in range(m)]
    best = neg
    for i in range(m):
        for j in range(n):
            s = int(rows[i0 + i][t[j0 + j]])
            diag = 0 if i == 0 and j == 0 else (H[i - 1][j - 1] if i > 0 and j > 0 else neg)
            M = diag + s if diag > neg else neg
            E[i][j] = max(H[i][j - 1] - o, E[i][j - 1] - e) if j > 0 else neg
            F[i][j] = max(H[i - 1][j] - o, F[i - 1][j] - e) if i > 0 else neg
            H[i][j] = max(M, E[i][j], F[i][j])
            best = max(best, M)
    return best


def test_span_start_is_the_last_optimal_start_by_brute_force():
    # the reversed pass's column-major-first end is, in forward coordinates,
    # the optimal start with the largest target position, then the largest
    # query position, among alignments inside the end cell's prefix rectangle
    rng = XorShift(0xA5A5F00D)
    checked = 0
    for o, e in [(11, 1), (3, 1), (2, 2), (1, 1), (4, 0)]:
        for _ in range(40):
            q = rng.seq(1 + rng.range(7), 3)
            t = rng.seq(1 + rng.range(7), 3)
            rows = O.seq_rows(q)
            sc, qs, qe, ts, te = O.span(rows, t, o, e)
            if sc <= 0:
                assert (qs, qe, ts, te) == (-1, -1, -1, -1)
                continue
            starts = [(j0, i0) for i0 in range(qe + 1) for j0 in range(te + 1)
                      if _best_from(rows, t, o, e, i0, j0, qe, te) == sc]
            assert max(starts) == (ts, qs), (q, t, o, e)
            checked += 1
    assert checked > 100


def test_align_span_equals_align_canonical():
    rng = XorShift(0x51A2)
    for o, e in [(11, 1), (6, 2), (2, 1), (-1, 1), (-2, 3)]:
        for _ in range(60):
            q = rng.seq(1 + rng.range(90), 20)
            t = _mutate(rng, q, 5) if rng.range(3) > 0 else rng.seq(1 + rng.range(90), 20)
            if len(t) == 0:
                continue
            rows = O.seq_rows(q)
            sp = np.zeros(1, H.SPAN_DT)[0]
            (sp["score"], sp["query_start"], sp["query_end"], sp["target_start"],
             sp["target_end"]) = O.span(rows, t, o, e)
            a = H.align_canonical(rows, q, t, GapCosts(o, e))
            b = H.align_span(rows, q, t, GapCosts(o, e), sp)
            assert a == b
            if a and sp["query_start"] >= 0:
                assert (a.query_start, a.query_end, a.target_start, a.target_end) == \
                    (sp["query_start"], sp["query_end"], sp["target_start"], sp["target_end"])


def test_align_span_rejects_bad_spans():
    q, t = _enc("MKVLLACDEFGHIK"), _enc("MKVLLACDEFGHIK")
    rows = O.seq_rows(q)
    sp = np.zeros(1, H.SPAN_DT)[0]
    (sp["score"], sp["query_start"], sp["query_end"], sp["target_start"],
     sp["target_end"]) = O.span(rows, t, 11, 1)
    assert H.align_span(rows, q, t, GapCosts(11, 1), sp).cigar == "14M"
    for field, value in [("query_end", 14), ("target_end", 20), ("target_start", 14), ("query_start", -1)]:
        bad = sp.copy()
        bad[field] = value
        with pytest.raises(Exception):
            H.align_span(rows, q, t, GapCosts(11, 1), bad)
    # a start that no alignment of the score reaches: no alignment, as the
    # reference returns None when no band reproduces the score
    bad = sp.copy()
    bad["query_start"], bad["target_start"] = 3, 3
    assert H.align_span(rows, q, t, GapCosts(11, 1), bad) is None
    zero = np.zeros(1, H.SPAN_DT)[0]
    assert H.align_span(rows, q, t, GapCosts(11, 1), zero) is None


def test_negative_gap_costs_fall_back_like_the_reference():
    # with a bonus for opening a gap the reversed pass can score differently:
    # the start is then searched from the end (src/align.rs:859-863), and
    # where nothing reproduces the score there is no alignment (None)
    rng = XorShift(0x77)
    n_nostart, n_none, n_found = 0, 0, 0
    for _ in range(150):
        q = rng.seq(1 + rng.range(30), 3)
        t = rng.seq(1 + rng.range(30), 3)
        rows = O.seq_rows(q)
        sc, qs, qe, ts, te = O.span(rows, t, -1, 1)
        sp = np.zeros(1, H.SPAN_DT)[0]
        sp["score"], sp["query_start"], sp["query_end"], sp["target_start"], sp["target_end"] = sc, qs, qe, ts, te
        a = H.align_canonical(rows, q, t, GapCosts(-1, 1))
        assert a == H.align_span(rows, q, t, GapCosts(-1, 1), sp)
        n_nostart += sc > 0 and qs == -1
        n_none += a is None
        if a:
            n_found += 1
            assert a.raw_score == sc
    assert n_nostart > 20 and n_none > 20 and n_found > 20
