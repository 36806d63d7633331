"""Spans of local alignments on the device (the forward and reversed-prefix
score passes of align_canonical) vs the oracle, bit-exact, and the hits and
CIGARs built from them vs the host's full traceback."""
import numpy as np
import pytest

import oracle as O
import sabertooth_amd as st
from sabertooth_amd import GapCosts
from sabertooth_amd import host as H

pytestmark = pytest.mark.gpu

FIELDS = ("score", "query_start", "query_end", "target_start", "target_end")


def _mutate(rng, q, ident, indel):
    out = []
    for c in q:
        r = rng.random()
        if r < indel:
            out.extend(rng.integers(0, 20, size=int(rng.integers(1, 6))).tolist())
        elif r < 2 * indel:
            continue
        out.append(int(c) if rng.random() < ident else int(rng.integers(0, 20)))
    return np.array(out, np.uint8)


def _check(ctx, rows_list, targets, hit_lists, gaps):
    db = st.TargetDb.from_sequences(ctx, targets)
    qps = [st.QueryProfile.from_rows(ctx, r) for r in rows_list]
    got = st.align_spans(ctx, db, qps, hit_lists, gaps)
    k = 0
    for qi, hl in enumerate(hit_lists):
        for tid in hl:
            want = O.span(rows_list[qi], targets[int(tid)], gaps[qi].open, gaps[qi].extend)
            assert tuple(int(got[k][f]) for f in FIELDS) == want, (qi, int(tid))
            k += 1
    assert k == len(got)
    return got, db


def test_spans_match_the_oracle(ctx):
    rng = np.random.default_rng(21)
    queries = [rng.integers(0, 20, size=L, dtype=np.uint8) for L in (1, 7, 60, 233, 700, 1500)]
    targets = [rng.integers(0, 20, size=int(rng.integers(1, 900)), dtype=np.uint8) for _ in range(80)]
    for q in queries:
        for _ in range(4):
            targets.append(np.concatenate([rng.integers(0, 20, size=int(rng.integers(0, 50)), dtype=np.uint8),
                                           _mutate(rng, q, 0.6, 0.03)]))
    targets.append(np.zeros(0, np.uint8))  # empty target: score 0
    hit_lists = [rng.choice(len(targets), size=int(rng.integers(5, 40)), replace=False).astype(np.uint32)
                 for _ in queries]
    # the query's own homologs, one twice
    for i in range(len(queries)):
        h = np.arange(80 + 4 * i, 84 + 4 * i, dtype=np.uint32)
        hit_lists[i] = np.concatenate([hit_lists[i], h, h[:1], [len(targets) - 1]]).astype(np.uint32)
    gaps = [GapCosts(11, 1), GapCosts(11, 1), GapCosts(6, 2), GapCosts(2, 1), GapCosts(11, 1), GapCosts(44, 4)]
    got, _ = _check(ctx, [O.seq_rows(q) for q in queries], targets, hit_lists, gaps)
    assert (got["score"] > 0).sum() > 100 and (got["score"] == 0).any()


def test_spans_with_ties_pssm_rows_and_negative_gaps(ctx):
    rng = np.random.default_rng(22)
    # two- and three-letter sequences: many equal optima, so both passes'
    # column-major tie rule decides the cells
    queries = [rng.integers(0, 2, size=L, dtype=np.uint8) for L in (5, 40, 300)]
    queries += [rng.integers(0, 3, size=L, dtype=np.uint8) for L in (17, 120)]
    targets = [rng.integers(0, 2, size=int(rng.integers(1, 400)), dtype=np.uint8) for _ in range(30)]
    targets += [rng.integers(0, 3, size=int(rng.integers(1, 400)), dtype=np.uint8) for _ in range(30)]
    rows = [O.seq_rows(q) for q in queries]
    pssm = rng.integers(-30, 40, size=(150, 21)).astype(np.int32)
    pssm[:, 20] = 0
    rows.append(pssm)
    rows.append(O.seq_rows(queries[3]))
    hit_lists = [np.arange(60, dtype=np.uint32)] * len(rows)
    gaps = [GapCosts(11, 1), GapCosts(1, 1), GapCosts(3, 1), GapCosts(11, 1), GapCosts(2, 0), GapCosts(44, 4),
            GapCosts(-1, 1)]
    _check(ctx, rows, targets, hit_lists, gaps)


def test_spans_profiles_in_and_outside_int16(ctx):
    # one call over queries whose profiles fit int8 (their row-reversed
    # copies are int8), int16 but not int8 (int16 copies) and neither (int32
    # copies), with floored and unfloored gap costs on each
    rng = np.random.default_rng(24)
    queries = [rng.integers(0, 20, size=L, dtype=np.uint8) for L in (90, 400)]
    targets = [rng.integers(0, 20, size=int(rng.integers(1, 500)), dtype=np.uint8) for _ in range(40)]
    for q in queries:
        targets += [_mutate(rng, q, 0.7, 0.02) for _ in range(3)]
    wide = rng.integers(-40000, 45000, size=(230, 21)).astype(np.int32)
    wide[:, 20] = 0
    mid = rng.integers(-1500, 2500, size=(170, 21)).astype(np.int32)
    mid[:, 20] = 0
    rows = [O.seq_rows(queries[0]), wide, O.seq_rows(queries[1]), wide[:57], mid, mid[:40]]
    hits = np.arange(len(targets), dtype=np.uint32)
    gaps = [GapCosts(11, 1), GapCosts(30000, 2000), GapCosts(-1, 1), GapCosts(-5000, 300), GapCosts(1100, 100),
            GapCosts(-200, 50)]
    _check(ctx, rows, targets, [hits] * len(rows), gaps)


@pytest.mark.parametrize("G", [1, 4, 16, 64])
def test_reverse_windows_every_group_width(ctx, monkeypatch, G):
    # the reverse pass puts up to 64 / G reversed prefixes of one query in a
    # unit, each group with its own profile window; many hits per query with
    # different end rows, int16 and int32 profiles, floored and unfloored gaps
    monkeypatch.setenv("ST_FORCE_G", str(G))
    rng = np.random.default_rng(40 + G)
    queries = [rng.integers(0, 20, size=L, dtype=np.uint8) for L in (35, 180, 620)]
    targets = [rng.integers(0, 20, size=int(rng.integers(1, 700)), dtype=np.uint8) for _ in range(50)]
    for q in queries:
        for _ in range(6):
            a = int(rng.integers(0, len(q)))
            b = int(rng.integers(a + 1, len(q) + 1))
            targets.append(np.concatenate([rng.integers(0, 20, size=int(rng.integers(0, 40)), dtype=np.uint8),
                                           _mutate(rng, q[a:b], 0.75, 0.02)]))
    wide = rng.integers(-40000, 45000, size=(140, 21)).astype(np.int32)
    wide[:, 20] = 0
    rows = [O.seq_rows(q) for q in queries] + [wide, O.seq_rows(queries[1])]
    hits = np.arange(len(targets), dtype=np.uint32)
    gaps = [GapCosts(11, 1), GapCosts(3, 1), GapCosts(11, 1), GapCosts(30000, 2000), GapCosts(-1, 1)]
    _check(ctx, rows, targets, [hits] * len(rows), gaps)


def test_spans_launches_count_both_passes(ctx):
    # the diagnostic of the last call: forward launches, then reverse ones,
    # whose useful cells are the two passes' cells exactly
    rng = np.random.default_rng(25)
    queries = [rng.integers(0, 20, size=L, dtype=np.uint8) for L in (50, 300)]
    targets = [rng.integers(0, 20, size=int(rng.integers(1, 400)), dtype=np.uint8) for _ in range(30)]
    for q in queries:
        targets += [_mutate(rng, q, 0.7, 0.02) for _ in range(3)]
    rows = [O.seq_rows(q) for q in queries]
    hits = [np.arange(len(targets), dtype=np.uint32)] * 2
    got, db = _check(ctx, rows, targets, hits, [GapCosts(11, 1)] * 2)
    li, nf = st.align_spans_launches(ctx)
    assert 0 < nf < len(li)
    fwd = sum(len(r) * len(targets[int(t)]) for r, h in zip(rows, hits) for t in h)
    pos = got["score"] > 0
    rev = int(((got["query_end"][pos] + 1) * (got["target_end"][pos] + 1)).sum())
    assert int(li["useful_cells"][:nf].sum()) == fwd and int(li["useful_cells"][nf:].sum()) == rev
    assert (li["exec_cells"] >= li["useful_cells"]).all() and (li["ms"] > 0).all()


def test_long_spans_above_int16(ctx):
    W = O.aa2num()[ord("W")]
    rng = np.random.default_rng(23)
    q = np.concatenate([rng.integers(0, 20, size=500, dtype=np.uint8), np.full(3600, W, np.uint8),
                        rng.integers(0, 20, size=300, dtype=np.uint8)])
    t = np.concatenate([rng.integers(0, 20, size=200, dtype=np.uint8), np.full(3500, W, np.uint8)])
    got, _ = _check(ctx, [O.seq_rows(q)], [t, q], [np.array([0, 1], np.uint32)], [GapCosts(11, 1)])
    assert got[0]["score"] >= 35000 and got[1]["score"] > 36000


def test_hits_from_device_spans_equal_the_host_traceback(ctx):
    rng = np.random.default_rng(24)
    q = rng.integers(0, 20, size=350, dtype=np.uint8)
    targets = [rng.integers(0, 20, size=int(rng.integers(50, 600)), dtype=np.uint8) for _ in range(200)]
    for k in range(0, 200, 7):
        targets[k] = np.concatenate([targets[k][:30], _mutate(rng, q, 0.5, 0.02)])
    db = st.TargetDb.from_sequences(ctx, targets)
    qp = st.QueryProfile.from_rows(ctx, H.query_rows(q))
    rows = H.query_rows(q)
    tids = np.arange(len(targets), dtype=np.uint32)
    scores = qp.sw_score_batch(db, tids, GapCosts(11, 1))
    spans = st.align_spans(ctx, db, [qp], [tids], [GapCosts(11, 1)])
    assert (spans["score"] == scores).all()
    offs = np.zeros(len(targets) + 1, np.uint64)
    offs[1:] = np.cumsum([len(t) for t in targets])
    arena = np.concatenate(targets)
    ev = H.EValueParams.sequence(11, 1, True, int(offs[-1]), len(targets))
    params = H.SearchParams(max_evalue=10.0)
    a = H.search_hits(rows, q, arena, offs, tids, scores, GapCosts(11, 1), ev, params)
    b = H.search_hits(rows, q, arena, offs, tids, scores, GapCosts(11, 1), ev, params, spans=spans)
    assert len(a[0]) >= 25 and (a[0] == b[0]).all() and a[1] == b[1]


def test_no_hits_and_zero_scores(ctx):
    db = st.TargetDb.from_sequences(ctx, [np.array([1, 2, 3], np.uint8)])
    qp = st.QueryProfile.from_rows(ctx, np.full((4, 21), -5, np.int32))
    assert len(st.align_spans(ctx, db, [qp], [np.zeros(0, np.uint32)], [GapCosts(11, 1)])) == 0
    got = st.align_spans(ctx, db, [qp], [np.array([0, 0], np.uint32)], [GapCosts(11, 1)])
    assert [tuple(int(r[f]) for f in FIELDS) for r in got] == [(0, -1, -1, -1, -1)] * 2


@pytest.mark.parametrize("G", ["1", "8", "64"])
def test_reverse_early_stop_equals_full_rectangles(ctx, monkeypatch, G):
    # short local matches inside long targets: the reverse pass stops at the
    # first column holding the forward score, and with ST_SPANS_STOP=0 it
    # runs every reversed prefix to its end; both equal the oracle (several
    # passes per problem at the narrow group widths)
    monkeypatch.setenv("ST_FORCE_G", G)
    rng = np.random.default_rng(77)
    queries = [rng.integers(0, 20, size=L, dtype=np.uint8) for L in (45, 300, 610)]
    targets = []
    for q in queries:
        for _ in range(5):
            a = int(rng.integers(0, len(q) // 2))
            seg = _mutate(rng, q[a:a + int(rng.integers(20, 60))], 0.8, 0.02)
            targets.append(np.concatenate([rng.integers(0, 20, size=int(rng.integers(100, 700)), dtype=np.uint8),
                                           seg, rng.integers(0, 20, size=int(rng.integers(0, 300)),
                                                             dtype=np.uint8)]))
    targets += [rng.integers(0, 20, size=int(rng.integers(1, 800)), dtype=np.uint8) for _ in range(20)]
    hits = [np.arange(len(targets), dtype=np.uint32)] * len(queries)
    rows = [O.seq_rows(q) for q in queries]
    gaps = [GapCosts(11, 1), GapCosts(6, 2), GapCosts(11, 1)]
    stopped, db = _check(ctx, rows, targets, hits, gaps)
    monkeypatch.setenv("ST_SPANS_STOP", "0")
    qps = [st.QueryProfile.from_rows(ctx, r) for r in rows]
    full = st.align_spans(ctx, db, qps, hits, gaps)
    for f in FIELDS:
        assert (stopped[f] == full[f]).all(), f
