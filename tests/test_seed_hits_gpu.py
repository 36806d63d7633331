"""The prefilter's seed-hit stage on the device (hits -> top diagonals ->
ungapped jobs and target records) vs the oracle, bit-exact, and chained into
score_and_gate on the device."""
import numpy as np
import pytest

import oracle as O
import sabertooth_amd as st
from sabertooth_amd import GapCosts
from sabertooth_amd.core import Postings, seed_hits_to_jobs
from test_oracle import kmer_index, query_seeds

pytestmark = pytest.mark.gpu


def _db(ctx, targets):
    return st.TargetDb.from_sequences(ctx, targets)


@pytest.mark.parametrize("kmax,min_hits,diag_hits,window", [(4, 1, 1, 48), (1, 1, 1, 0), (16, 2, 2, 5),
                                                            (3, 1, 2, 0), (2, 3, 1, 1), (20, 1, 1, 0),
                                                            (40, 1, 2, 3)])
def test_jobs_match_the_oracle(ctx, kmax, min_hits, diag_hits, window):
    rng = np.random.default_rng(40 + kmax + 3 * min_hits + 7 * diag_hits)
    targets = [rng.integers(0, 4, size=int(rng.integers(3, 400)), dtype=np.uint8) for _ in range(300)]
    post, of = kmer_index(targets, 3)
    tlen = np.array([len(t) for t in targets])
    db = _db(ctx, targets)
    dp = Postings(ctx, post)
    assert len(dp) == len(post)
    for _ in range(4):
        q = rng.integers(0, 4, size=int(rng.integers(3, 300)), dtype=np.uint8)
        seeds = query_seeds(q, 3, of, extra=50, n_post=len(post), rng=rng)
        gj, gr = seed_hits_to_jobs(ctx, db, dp, len(q), seeds, kmax, min_hits, diag_hits, window, 3)
        oj, orr = O.seed_hits_to_jobs(len(q), seeds, post, tlen, kmax, min_hits, diag_hits, window, 3)
        assert len(gr) == len(orr) and len(gj) == len(oj)
        assert (gj == oj).all() and (gr == orr).all()
    dp.close()


def test_realistic_alphabet_chained_into_the_gate(ctx):
    # 20-letter targets with planted homologs; seeds are the query's exact
    # 4-mers; the jobs go through score_and_gate on the device and on the
    # oracle, and the homologs come out as candidates
    rng = np.random.default_rng(9)
    q = rng.integers(0, 20, size=250, dtype=np.uint8)
    targets = [rng.integers(0, 20, size=int(rng.integers(30, 600)), dtype=np.uint8) for _ in range(400)]
    for k in range(0, 400, 40):
        h = q.copy()
        h[rng.random(len(h)) < 0.3] = rng.integers(0, 20)
        targets[k] = np.concatenate([targets[k][:50], h])
    post, of = kmer_index(targets, 4)
    tlen = np.array([len(t) for t in targets])
    db = _db(ctx, targets)
    dp = Postings(ctx, post)
    seeds = query_seeds(q, 4, of)
    gj, gr = seed_hits_to_jobs(ctx, db, dp, len(q), seeds, 4, 1, 1, 48, 4)
    oj, orr = O.seed_hits_to_jobs(len(q), seeds, post, tlen, 4, 1, 1, 48, 4)
    assert (gj == oj).all() and (gr == orr).all() and len(gr) > 10
    rows = O.seq_rows(q)
    qp = st.QueryProfile.from_rows(ctx, rows)
    cands, chains, js = st.score_and_gate(qp, db, gj, gr, GapCosts(11, 1), 35)
    arena = np.concatenate(targets)
    offs = np.zeros(len(targets) + 1, np.uint64)
    offs[1:] = np.cumsum(tlen)
    oc, ochains, ojs = O.score_and_gate(rows, arena, offs, oj, orr, 11, 1, 35)
    assert (cands == oc).all() and chains == ochains and (js == ojs).all()
    assert set(range(0, 400, 40)) <= set(int(c["target_id"]) for c in cands)
    dp.close()


def test_edges(ctx):
    targets = [np.array([0, 1, 2, 3, 0, 1, 2], np.uint8), np.array([3, 3, 3], np.uint8)]
    post, of = kmer_index(targets, 2)
    db = _db(ctx, targets)
    dp = Postings(ctx, post)
    # no seeds, empty ranges
    assert [len(x) for x in seed_hits_to_jobs(ctx, db, dp, 5, [])] == [0, 0]
    assert [len(x) for x in seed_hits_to_jobs(ctx, db, dp, 5, [(0, 3, 3), (4, 0, 0)])] == [0, 0]
    # one hit; a one-residue query
    gj, gr = seed_hits_to_jobs(ctx, db, dp, 1, [(0, 0, 1)], 4, 1, 1, 0, 2)
    oj, orr = O.seed_hits_to_jobs(1, [(0, 0, 1)], post, [7, 3], 4, 1, 1, 0, 2)
    assert (gj == oj).all() and (gr == orr).all() and len(gr) == 1
    # a seed outside the postings or the query is refused
    with pytest.raises(st.StError):
        seed_hits_to_jobs(ctx, db, dp, 5, [(0, 0, len(post) + 1)])
    with pytest.raises(st.StError):
        seed_hits_to_jobs(ctx, db, dp, 5, [(5, 0, 1)])
    # a posting naming a target outside the database is refused
    bad = Postings(ctx, np.array([(7 << 32) | 1], np.uint64))
    with pytest.raises(st.StError):
        seed_hits_to_jobs(ctx, db, bad, 5, [(0, 0, 1)])
    with pytest.raises(st.StError):
        seed_hits_to_jobs(ctx, db, dp, 5, [(0, 0, 1)], max_diagonals=0)
    bad.close()
    dp.close()


def test_the_reference_index_drives_the_stage(ctx):
    # the fixture index's own postings on the device; each fixture query
    # seeded with its exact 2-mers (the postings of its 2-mer groups)
    import os
    from sabertooth_amd import host as H
    idx = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "v1_fixture_k2.stidx")
    res, off = H.stidx_targets(idx)
    post = H.stidx_postings(idx)
    db = st.TargetDb.from_stidx(ctx, idx)
    dp = Postings.from_stidx(ctx, idx)
    assert len(dp) == len(post)
    t = (post >> np.uint64(32)).astype(np.int64)
    p = (post & np.uint64(0xFFFFFFFF)).astype(np.int64)
    km = [int(res[int(off[a]) + b]) * 21 + int(res[int(off[a]) + b + 1]) for a, b in zip(t, p)]
    ranges = {}
    for i, k in enumerate(km):
        a, _ = ranges.get(k, (i, i))
        ranges[k] = (a, i + 1)
    tlen = np.diff(off.astype(np.int64))
    for qi in range(len(tlen)):
        q = res[int(off[qi]):int(off[qi + 1])]
        seeds = [(i, *ranges[int(q[i]) * 21 + int(q[i + 1])]) for i in range(len(q) - 1)
                 if int(q[i]) * 21 + int(q[i + 1]) in ranges]
        gj, gr = seed_hits_to_jobs(ctx, db, dp, len(q), seeds, 4, 1, 1, 48, 2)
        oj, orr = O.seed_hits_to_jobs(len(q), seeds, post, tlen, 4, 1, 1, 48, 2)
        assert (gj == oj).all() and (gr == orr).all() and qi in set(int(x) for x in gr["tid"])
    dp.close()


def test_many_queries_in_one_call(ctx):
    from sabertooth_amd.core import seed_hits_to_jobs_batch
    rng = np.random.default_rng(77)
    targets = [rng.integers(0, 4, size=int(rng.integers(3, 300)), dtype=np.uint8) for _ in range(200)]
    post, of = kmer_index(targets, 3)
    tlen = np.array([len(t) for t in targets])
    db = _db(ctx, targets)
    dp = Postings(ctx, post)
    queries = [rng.integers(0, 4, size=int(L), dtype=np.uint8) for L in (5, 1, 200, 3, 77, 150, 2)]
    seeds = [query_seeds(q, 3, of, extra=5, n_post=len(post), rng=rng) if len(q) >= 3 else [] for q in queries]
    seeds[3] = []  # a query without seeds
    for kmax, mh, dh, win in [(4, 1, 1, 48), (16, 2, 2, 0)]:
        out = seed_hits_to_jobs_batch(ctx, db, dp, [len(q) for q in queries], seeds, kmax, mh, dh, win, 3)
        assert len(out) == len(queries)
        for qi, (gj, gr) in enumerate(out):
            oj, orr = O.seed_hits_to_jobs(len(queries[qi]), seeds[qi], post, tlen, kmax, mh, dh, win, 3)
            assert (gj == oj).all() and (gr == orr).all(), qi
            sj, sr = seed_hits_to_jobs(ctx, db, dp, len(queries[qi]), seeds[qi], kmax, mh, dh, win, 3)
            assert (sj == gj).all() and (sr == gr).all(), qi
    assert seed_hits_to_jobs_batch(ctx, db, dp, [], []) == []
    dp.close()


def test_seeds_to_candidates_on_the_device(ctx):
    # several queries (planted homologs, per-query gaps and gates, one query
    # without seeds): the device chain equals the stage's jobs pushed
    # through score_and_gate per query, on the device and on the oracle
    from sabertooth_amd.core import seed_hits_to_candidates_batch
    rng = np.random.default_rng(21)
    qs = [rng.integers(0, 20, size=int(L), dtype=np.uint8) for L in (250, 40, 600, 120, 9)]
    targets = [rng.integers(0, 20, size=int(rng.integers(30, 700)), dtype=np.uint8) for _ in range(500)]
    for k in range(0, 500, 25):
        h = qs[k % len(qs)].copy()
        h[rng.random(len(h)) < 0.3] = rng.integers(0, 20)
        targets[k] = np.concatenate([targets[k][:40], h])
    post, of = kmer_index(targets, 4)
    tlen = np.array([len(t) for t in targets])
    arena = np.concatenate(targets)
    offs = np.zeros(len(targets) + 1, np.uint64)
    offs[1:] = np.cumsum(tlen)
    db = _db(ctx, targets)
    dp = Postings(ctx, post)
    seeds = [query_seeds(q, 4, of) for q in qs]
    seeds[3] = []
    rows = [O.seq_rows(q) for q in qs]
    qps = [st.QueryProfile.from_rows(ctx, r) for r in rows]
    gaps = [GapCosts(11, 1), GapCosts(8, 2), GapCosts(11, 1), GapCosts(5, 1), GapCosts(11, 1)]
    gates = [35, 30, 45, 20, 10]
    for regions, params in [(False, (4, 1, 1, 48)), (True, (16, 2, 1, 0))]:
        cands, chains = seed_hits_to_candidates_batch(db, dp, qps, seeds, gaps, gates, regions, *params, span=4)
        total = 0
        for qi in range(len(qs)):
            gj, gr = seed_hits_to_jobs(ctx, db, dp, len(qs[qi]), seeds[qi], *params, 4)
            c1, ch1, _ = st.score_and_gate(qps[qi], db, gj, gr, gaps[qi], gates[qi], regions)
            oc, och, _ = O.score_and_gate(rows[qi], arena, offs, gj, gr, gaps[qi].open, gaps[qi].extend, gates[qi],
                                          regions)
            assert len(cands[qi]) == len(c1) == len(oc), qi
            assert (cands[qi] == c1).all() and (c1 == oc).all(), qi
            assert ch1 == och
            total += ch1
        assert chains == total
        assert sum(len(c) for c in cands) > 15
    # a query whose seeds find nothing, alone
    cands, chains = seed_hits_to_candidates_batch(db, dp, [qps[3]], [[]], [gaps[3]], [20])
    assert len(cands[0]) == 0 and chains == 0
    for q in qps:
        q.close()
    dp.close()


def test_candidates_profile_in_lds_and_in_global_memory(ctx):
    # dense 4-letter hits give thousands of jobs per query, so blocks hold
    # one query's run: queries of 300 and 200 rows read their profile from
    # LDS; one of 1,300 rows (longer than a block stages) and one whose rows
    # exceed int16 read the int32 profile in global memory
    from sabertooth_amd.core import seed_hits_to_candidates_batch
    rng = np.random.default_rng(5)
    targets = [rng.integers(0, 4, size=int(rng.integers(50, 400)), dtype=np.uint8) for _ in range(1500)]
    post, of = kmer_index(targets, 4)
    tlen = np.array([len(t) for t in targets])
    arena = np.concatenate(targets)
    offs = np.zeros(len(targets) + 1, np.uint64)
    offs[1:] = np.cumsum(tlen)
    db = _db(ctx, targets)
    dp = Postings(ctx, post)
    qs = [rng.integers(0, 4, size=L, dtype=np.uint8) for L in (300, 1300, 250, 200)]
    rows = [O.seq_rows(q) for q in qs]
    rows[2] = np.array(rows[2], dtype=np.int32)
    rows[2][::7, 1] = 40000  # column R: beyond int16
    rows[2][3::11, 2] = -40000
    qps = [st.QueryProfile.from_rows(ctx, r) for r in rows]
    seeds = [query_seeds(q, 4, of) for q in qs]
    gaps = [GapCosts(11, 1), GapCosts(11, 1), GapCosts(9, 2), GapCosts(11, 1)]
    gates = [110, 140, 400000, 100]
    for regions in (False, True):
        cands, chains = seed_hits_to_candidates_batch(db, dp, qps, seeds, gaps, gates, regions, span=4)
        total = 0
        for qi in range(len(qs)):
            gj, gr = seed_hits_to_jobs(ctx, db, dp, len(qs[qi]), seeds[qi], 4, 1, 1, 48, 4)
            assert len(gj) > 2048
            c1, ch1, js = st.score_and_gate(qps[qi], db, gj, gr, gaps[qi], gates[qi], regions)
            oc, och, ojs = O.score_and_gate(rows[qi], arena, offs, gj, gr, gaps[qi].open, gaps[qi].extend,
                                            gates[qi], regions)
            assert (js == ojs).all(), qi
            assert len(cands[qi]) == len(c1) == len(oc) and 0 < len(oc) < len(gr), qi
            assert (cands[qi] == c1).all() and (c1 == oc).all(), qi
            assert ch1 == och
            total += ch1
        assert chains == total
    for q in qps:
        q.close()
    dp.close()


def test_candidates_capacity_error_reports_the_count(ctx):
    # a too-small output buffer: ST_ERR_CAPACITY, *n_out = the candidate
    # count and the chain count untouched; a buffer of that size succeeds
    # and accumulates it
    import ctypes as C
    from sabertooth_amd import _lib
    from sabertooth_amd.core import _seed_array, seed_hits_to_candidates_batch
    rng = np.random.default_rng(33)
    targets = [rng.integers(0, 4, size=int(rng.integers(50, 300)), dtype=np.uint8) for _ in range(300)]
    post, of = kmer_index(targets, 4)
    db = _db(ctx, targets)
    dp = Postings(ctx, post)
    q = rng.integers(0, 4, size=200, dtype=np.uint8)
    qp = st.QueryProfile.from_rows(ctx, O.seq_rows(q))
    seeds = [query_seeds(q, 4, of)]
    want, want_chains = seed_hits_to_candidates_batch(db, dp, [qp], seeds, [GapCosts(11, 1)], [60], span=4)
    assert 1 < len(want[0])
    so, sd = _seed_array(seeds)
    hp = np.zeros(1, dtype=_lib.HIT_PARAMS_DT)
    hp["max_diagonals"], hp["min_hits"], hp["diag_hits"], hp["diag_window"], hp["span"] = 4, 1, 1, 48, 4
    qh = (C.c_void_p * 1)(qp.handle)
    op, ex, gt = (np.array([v], np.int32) for v in (11, 1, 60))
    co = np.zeros(2, np.uint64)
    for cap in (1, len(want[0])):
        out = np.zeros(cap, dtype=_lib.CAND_DT)
        n, chains = C.c_uint64(0), C.c_uint64(5)
        rc = _lib.lib().st_seed_hits_to_candidates_batch(
            ctx.handle, db.handle, dp.handle, 1, qh, _lib.ptr(so), _lib.ptr(sd), _lib.ptr(hp), _lib.ptr(op),
            _lib.ptr(ex), _lib.ptr(gt), 0, _lib.ptr(out), cap, C.byref(n), _lib.ptr(co), C.byref(chains))
        assert n.value == len(want[0])
        if cap < len(want[0]):
            assert rc == _lib.ST_ERR_CAPACITY and chains.value == 5
        else:
            assert rc == 0 and (out == want[0]).all() and list(co) == [0, len(want[0])]
            assert chains.value == 5 + want_chains
    qp.close()
    dp.close()
