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
