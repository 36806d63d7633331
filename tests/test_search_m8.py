"""The host search pipeline after scoring (rows, gapped stage, traceback,
statistics, .m8 formatting) reproduces the reference's .m8 fixture: the 24
fixture sequences searched against themselves with the default sequence
settings (composition-bias correction, hard masking in scoring, E <= 10)."""
import os

import numpy as np
import pytest

import oracle as O
from oracle import host as OH
from sabertooth_amd import GapCosts
from sabertooth_amd import host as H

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _fasta(path):
    ids, seqs = [], []
    for line in open(path):
        line = line.strip()
        if line.startswith(">"):
            ids.append(line[1:].split()[0])
            seqs.append("")
        elif ids:
            seqs[-1] += line
    return ids, [O.encode(s) for s in seqs]


def _m8_lines(ids, seqs, scorer, spanner=None):
    offs = np.zeros(len(seqs) + 1, np.uint64)
    offs[1:] = np.cumsum([len(s) for s in seqs])
    arena = np.concatenate(seqs)
    ev = H.EValueParams.sequence(11, 1, True, int(offs[-1]), len(seqs))
    params = H.SearchParams(max_evalue=10.0)
    out, cnt = {}, np.zeros(1, H.COUNTERS_DT)
    for qi, q in enumerate(seqs):
        rows = H.query_rows(q)
        tids = np.arange(len(seqs), dtype=np.uint32)
        scores = scorer(rows, seqs)
        spans = spanner(rows, seqs) if spanner else None
        hits, cig = H.search_hits(rows, q, arena, offs, tids, scores, GapCosts(11, 1), ev, params, cnt, spans=spans)
        out[ids[qi]] = [H.format_m8(ids[qi], ids[int(h["target"])], h) + "\t" + c for h, c in zip(hits, cig)]
    return out, cnt


def _oracle_scores(rows, seqs):
    return np.array([O.gotoh_score(rows, t, 11, 1) for t in seqs], dtype=np.int32)


def test_fixture_rows_are_reproduced():
    ids, seqs = _fasta(os.path.join(GOLD, "v1_fixture.fa"))
    got, cnt = _m8_lines(ids, seqs, _oracle_scores)
    lines = open(os.path.join(GOLD, "v1_fixture_k2.v03.m8")).read().strip().split("\n")
    assert lines[0] == H.M8_HEADER
    want = {}
    for l in lines[1:]:
        want.setdefault(l.split("\t")[0], []).append(l)
    assert sum(len(v) for v in want.values()) == 88
    for q, rows in want.items():
        # the fixture's prefilter can only drop candidates: its rows are ours,
        # in the same order
        mine = [r.rsplit("\t", 1)[0] for r in got[q]]
        pos = [mine.index(r) if r in mine else -1 for r in rows]
        assert -1 not in pos, (q, [r for r, p in zip(rows, pos) if p < 0], mine)
        assert pos == sorted(pos)
    assert cnt["score_mismatches"][0] == 0


def _oracle_spans(rows, seqs):
    sp = np.zeros(len(seqs), H.SPAN_DT)
    for k, t in enumerate(seqs):
        (sp[k]["score"], sp[k]["query_start"], sp[k]["query_end"], sp[k]["target_start"],
         sp[k]["target_end"]) = O.span(rows, t, 11, 1)
    return sp


def test_hits_from_known_spans_equal_the_full_traceback():
    # st_search_hits_spans (tracebacks from spans, no host score pass) gives
    # the same hits, CIGARs and counters as st_search_hits
    ids, seqs = _fasta(os.path.join(GOLD, "v1_fixture.fa"))
    a, ca = _m8_lines(ids, seqs, _oracle_scores)
    b, cb = _m8_lines(ids, seqs, _oracle_scores, _oracle_spans)
    assert a == b and sum(len(v) for v in a.values()) >= 88
    assert (ca == cb).all()


def test_rust_number_formatting():
    h = np.zeros(1, H.HIT_DT)[0]
    h["pident"], h["evalue"], h["bit_score"], h["aln_len"] = 1.5625, 0.0, 0.05, 64
    line = H.format_m8("q", "t", h).split("\t")
    # exact ties round half to even; zero prints 0.00e0
    assert line[2] == "1.562" and line[10] == "0.00e0" and line[11] == "0.1"
    for e, s in [(0.292, "2.92e-1"), (8.38e-55, "8.38e-55"), (123456.0, "1.23e5"), (1.0, "1.00e0")]:
        h["evalue"] = e
        assert H.format_m8("q", "t", h).split("\t")[10] == s == OH.rust_e2(e)
    h["pident"] = 2.0 / 3.0 * 100
    assert H.format_m8("q", "t", h).split("\t")[2] == "66.667"


def test_gapped_stage_filter_and_order():
    scores = np.array([5, 40, 40, 0, 39, 41, -3, 40], np.int32)
    tids = np.array([7, 3, 1, 9, 4, 8, 2, 0], np.uint32)
    order = H.gapped_stage(scores, tids, 40)
    assert [int(tids[i]) for i in order] == [8, 0, 1, 3]
    assert len(H.gapped_stage(scores, tids, -5)) == 6  # s > 0 also required


def test_extra_columns_and_output_blocks():
    h = np.zeros(1, H.HIT_DT)[0]
    h["raw_score"], h["qlen"], h["tlen"] = 57, 3, 16
    h["q_start"], h["q_end"], h["t_start"], h["t_end"] = 0, 1, 3, 3
    # 2/3 -> 0.667; 1/16 = 0.0625 exactly: Rust rounds the tie to even, 0.062
    assert H.format_m8_extra(h, "2M") == "57\t3\t16\t0.667\t0.062\t2M"
    h["qlen"], h["q_end"] = 0, 0
    assert H.format_m8_extra(h, "") == "57\t0\t16\t1.000\t0.062\t"
    # the writer's blocks: header, then one line per hit
    ids, seqs = _fasta(os.path.join(GOLD, "v1_fixture.fa"))
    got, _ = _m8_lines(ids, seqs, _oracle_scores)
    q = ids[0]
    offs = np.zeros(len(seqs) + 1, np.uint64)
    offs[1:] = np.cumsum([len(x) for x in seqs])
    rows = H.query_rows(seqs[0])
    ev = H.EValueParams.sequence(11, 1, True, int(offs[-1]), len(seqs))
    hits, cig = H.search_hits(rows, seqs[0], np.concatenate(seqs), offs, np.arange(len(seqs), dtype=np.uint32),
                              _oracle_scores(rows, seqs), GapCosts(11, 1), ev, H.SearchParams(max_evalue=10.0))
    tids = [ids[int(x["target"])] for x in hits]
    plain = H.format_m8_block(q, tids, hits, header=True)
    lines = plain.split("\n")
    assert lines[0] == H.M8_HEADER and lines[-1] == "" and len(lines) == len(hits) + 2
    assert lines[1:-1] == [r.rsplit("\t", 1)[0] for r in got[q]]
    ext = H.format_m8_block(q, tids, hits, cig, extra=True, header=True).split("\n")
    assert ext[0] == H.M8_HEADER + "\traw\tqlen\ttlen\tqcov\ttcov\tcigar"
    for a, b, hh, c in zip(ext[1:-1], lines[1:-1], hits, cig):
        assert a == b + "\t" + H.format_m8_extra(hh, c) and a.endswith("\t" + c)
    assert H.format_m8_block(q, [], hits[:0]) == ""


def test_candidate_trace_round_trip_and_errors():
    ids = ["seq0", "seq1", "tgt|2", "x y"]
    # the writer's line format (src/search.rs:526): query, target, score, cells
    text = (H.format_candidate_trace("qA", ["seq1", "x y"], [41, -3], [1200, 7])
            + H.format_candidate_trace("qB", ["tgt|2"], [0], [0])
            + "qA\tseq0\t55\t18446744073709551615\n")
    assert text.split("\n")[0] == "qA\tseq1\t41\t1200"
    qids, per = H.parse_candidate_trace(text, ids)
    assert qids == ["qA", "qB"]
    assert per[0][0].tolist() == [1, 3, 0] and per[0][1].tolist() == [41, -3, 55]
    assert per[0][2].tolist() == [1200, 7, 2 ** 64 - 1]
    assert per[1][0].tolist() == [2] and per[1][1].tolist() == [0]
    # blank lines and CRLF are tolerated; the empty trace has no queries
    assert H.parse_candidate_trace("\r\nqA\tseq0\t1\t2\r\n\n", ids)[0] == ["qA"]
    assert H.parse_candidate_trace("", ids) == ([], [])
    for bad in ["qA\tseq0\t1\n", "qA\tseq9\t1\t2\n", "qA\tseq0\tx\t2\n", "qA\tseq0\t1\t-2\n",
                "qA\tseq0\t1\t2\t3\n", "qA\tseq0\t99999999999\t2\n"]:
        with pytest.raises(Exception):
            H.parse_candidate_trace(bad, ids)
