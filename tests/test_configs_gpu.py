"""The benchmark configurations other than the bench line, as parity cases,
at their full sizes: plain all-pairs (32 queries x 20,000 targets), PSSM
queries from the PSSM builder (16 profiles x 200,000 targets, 44/4, gate 30)
and long pairs (64 queries of 2,000-35,000 x 5,000 targets of 1,000-30,000,
9.2e13 cells), plus long pairs mixed with short ones in one launch.  Every
SWG score of the first two is checked against the AVX2 cascade restated from
the reference, and every output of a seeded sample against the oracle
(bit-exact).  The long configuration, too large for a CPU check of every
pair, is checked on sampled pairs, on the closed form of a query aligned
with itself, and on each checked query scored alone against its rows of the
full batch.  The two configurations the bench also times run here at full
size as well: the batched uneven configuration (1,000 queries x 3,000
candidates over 68 M residues: every SWG score against the AVX2 cascade,
1,500 sampled pairs' full outputs against the oracle, the HIP-graph replay
equal to the eager run) and the serving configuration (a 265 M-residue
resident database, 2,000 requests of which about 10% are profiles: every
served row equal to the plain batch, sampled pairs of both kinds against the
oracle).  Execution-mode parity follows the reference's own check that every
mode gives the same hits (tests/cli_db_equivalence.rs:495-528)."""
import os

import numpy as np
import pytest

import oracle as O
import sabertooth_amd as st
from sabertooth_amd import synth

pytestmark = pytest.mark.gpu
THREADS = min(16, len(os.sched_getaffinity(0)))


def _run(ctx, rows, arena, offs, cands, gaps, gates):
    db = st.TargetDb(ctx, arena, offs)
    qps = [st.QueryProfile.from_rows(ctx, r) for r in rows]
    b = st.Batch(ctx, db, qps, cands, [st.GapCosts(*g) for g in gaps], gates)
    b.run()
    return b, b.results()


def _check(rows, arena, offs, cands, gaps, gates, b, res, sample, seed):
    off = b.offsets.astype(np.int64)
    # every SWG score against the AVX2 cascade (per gap setting)
    for g in sorted(set(gaps)):
        qs = [q for q in range(len(rows)) if gaps[q] == g]
        pq = np.concatenate([np.full(len(cands[q]), q, np.uint32) for q in qs])
        pt = np.concatenate([np.asarray(cands[q], np.uint32) for q in qs])
        sc, _, _ = O.avx2_score_pairs(rows, pq, pt, arena, offs, g[0], g[1], threads=THREADS)
        got = np.concatenate([res["swg"][off[q]:off[q + 1]] for q in qs])
        assert (sc == got).all(), (g, int((sc != got).sum()))
    # all outputs of a sample against the oracle
    rng = np.random.default_rng(seed)
    for k in rng.choice(len(res), size=min(sample, len(res)), replace=False):
        q = int(np.searchsorted(off, k, side="right") - 1)
        t = arena[int(offs[cands[q][k - off[q]]]):int(offs[cands[q][k - off[q]] + 1])]
        o, e = gaps[q]
        want = (O.gotoh_score(rows[q], t, o, e),) + O.ung_pair(rows[q], t, o, e, gates[q])
        r = res[k]
        got = (int(r["swg"]), int(r["ung"]), int(r["best_diag"]), int(r["gate_score"]), bool(r["pass"]))
        assert got == want, (q, k)


def test_plain_allpairs_full(ctx):
    w = synth.plain_allpairs(seed=1)
    assert len(w.queries) == 32 and len(w.offsets) - 1 == 20000
    rows = [O.seq_rows(q) for q in w.queries]
    gaps = [(11, 1)] * len(rows)
    gates = [w.gate] * len(rows)
    b, res = _run(ctx, rows, w.arena, w.offsets, w.cand_lists, gaps, gates)
    _check(rows, w.arena, w.offsets, w.cand_lists, gaps, gates, b, res, 300, 1)


def test_pssm_queries_from_the_builder_full(ctx):
    w = synth.pssm_search(seed=1)
    assert len(w.cand_lists) == 16 and len(w.offsets) - 1 == 200000
    rows = w.meta["rows"]
    assert all(r.shape[1] == 21 and (r[:, 20] == 0).all() for r in rows)
    gaps = [(44, 4)] * len(rows)
    gates = [30] * len(rows)
    b, res = _run(ctx, rows, w.arena, w.offsets, w.cand_lists, gaps, gates)
    _check(rows, w.arena, w.offsets, w.cand_lists, gaps, gates, b, res, 300, 2)
    # planted family members score far above background
    assert res["swg"].max() > 500


def test_long_pairs_mixed_with_short(ctx):
    w = synth.long_pairs(seed=1, n_queries=4, n_targets=24, qmax=9000, tmax=8000)
    short = synth.plain_allpairs(seed=3, n_queries=4, n_targets=200)
    # one database: the long targets then the short ones
    arena = np.concatenate([w.arena, short.arena])
    offs = np.concatenate([w.offsets, w.offsets[-1] + short.offsets[1:]])
    nl = len(w.offsets) - 1
    rows = [O.seq_rows(q) for q in w.queries] + [O.seq_rows(q) for q in short.queries]
    cands = [np.arange(nl, dtype=np.uint32) for _ in w.queries] + \
        [np.arange(nl, nl + 200, dtype=np.uint32) for _ in short.queries]
    # a long query against itself: a score far above 32767
    arena = np.concatenate([arena, w.queries[0]])
    offs = np.concatenate([offs, [offs[-1] + len(w.queries[0])]]).astype(np.uint64)
    cands[0] = np.concatenate([cands[0], [len(offs) - 2]]).astype(np.uint32)
    gaps = [(11, 1)] * len(rows)
    gates = [w.gate] * len(rows)
    b, res = _run(ctx, rows, arena, offs, cands, gaps, gates)
    assert res["swg"].max() > 32767 and (res["kernel"] == st.KERNEL_I32).any()
    _check(rows, arena, offs, cands, gaps, gates, b, res, 12, 3)


def test_long_pairs_full_size(ctx):
    w = synth.long_pairs(seed=1)
    nq, nt = len(w.queries), len(w.offsets) - 1
    assert (nq, nt) == (64, 5000) and w.cells() > 9e13
    # the first query is also a target: its self-alignment has a closed form
    arena = np.concatenate([w.arena, w.queries[0]])
    offs = np.concatenate([w.offsets, [w.offsets[-1] + len(w.queries[0])]]).astype(np.uint64)
    rows = [O.seq_rows(q) for q in w.queries]
    cands = [np.arange(nt, dtype=np.uint32) for _ in range(nq)]
    cands[0] = np.arange(nt + 1, dtype=np.uint32)
    gaps = [(11, 1)] * nq
    gates = [w.gate] * nq
    b, res = _run(ctx, rows, arena, offs, cands, gaps, gates)
    off = b.offsets.astype(np.int64)
    # every diagonal score of BLOSUM62 is positive: the main diagonal's
    # Kadane segment is the whole query, the largest possible ungapped score
    self_row = res[off[0] + nt]
    diag_sum = int(rows[0][np.arange(len(w.queries[0])), w.queries[0]].sum())
    assert int(self_row["ung"]) == diag_sum and int(self_row["best_diag"]) == 0
    assert int(self_row["swg"]) >= diag_sum > 32767 and self_row["kernel"] == st.KERNEL_I32
    assert bool(self_row["pass"])
    # a random pair and the best-scoring planted pair against the oracle
    rng = np.random.default_rng(4)
    picks = [int(rng.integers(len(res))), int(np.argmax(np.where(np.arange(len(res)) == off[0] + nt, -1,
                                                                 res["swg"])))]
    for k in picks:
        q = int(np.searchsorted(off, k, side="right") - 1)
        tid = int(cands[q][k - off[q]])
        t = arena[int(offs[tid]):int(offs[tid + 1])]
        want = (O.gotoh_score(rows[q], t, 11, 1),) + O.ung_pair(rows[q], t, 11, 1, w.gate)
        r = res[k]
        got = (int(r["swg"]), int(r["ung"]), int(r["best_diag"]), int(r["gate_score"]), bool(r["pass"]))
        assert got == want, (q, tid, len(rows[q]), len(t))
    # two queries alone give the rows they have in the full batch
    for q in (1, 37):
        b1, r1 = _run(ctx, [rows[q]], arena, offs, [cands[q]], [(11, 1)], [w.gate])
        assert (r1 == res[off[q]:off[q + 1]]).all(), q
        b1.close()
    # the two shortest queries: every one of their 5,000 SWG scores against
    # the AVX2 cascade (about 2.5e11 cells on the host)
    short = sorted(range(nq), key=lambda q: len(w.queries[q]))[:2]
    for q in short:
        pq = np.full(nt, q, np.uint32)
        pt = np.arange(nt, dtype=np.uint32)
        sc, _, _ = O.avx2_score_pairs(rows, pq, pt, arena, offs, 11, 1, threads=THREADS)
        got = res["swg"][off[q]:off[q] + nt]
        assert (sc == got).all(), (q, int((sc != got).sum()))
    b.close()


def test_batched_uneven_full(ctx):
    # the bench's configuration: 3 M pairs, 1.88e11 cells, gate 41
    w = synth.batched_uneven(seed=1)
    assert (len(w.queries), len(w.offsets) - 1, w.pairs()) == (1000, 370000, 3000000)
    assert 6.5e7 < int(w.offsets[-1]) < 7.1e7 and w.gate == 41
    rows = [st.seq_rows(q) for q in w.queries]
    gaps = [(11, 1)] * len(rows)
    gates = [w.gate] * len(rows)
    b, res = _run(ctx, rows, w.arena, w.offsets, w.cand_lists, gaps, gates)
    # planted homologs pass the gate, and some pairs were rescored in int32
    # or not, as the scores require: every score is checked below
    assert res["pass"].sum() > 100000
    _check(rows, w.arena, w.offsets, w.cand_lists, gaps, gates, b, res, 1500, 5)
    # the DP phase replayed from a HIP graph: the same results, twice, with
    # an eager run between the replays
    b.run(graph=True)
    assert (b.results() == res).all()
    b.run()
    b.run(graph=True)
    assert (b.results() == res).all()
    b.close()


def test_serving_full(ctx):
    # the serving configuration: 1.4 M targets (265 M residues) resident,
    # 2,000 requests with 3,000 candidates each, 10% profile queries (44/4,
    # gate 30), the rest plain (11/1, size-adjusted gate 43)
    w = synth.serving(seed=1)
    n = len(w.rows)
    assert n == 2000 and len(w.offsets) - 1 == 1400000 and 2.5e8 < int(w.offsets[-1]) < 2.8e8
    prof = np.array(w.profile)
    assert 150 < prof.sum() < 250 and set(w.gates) == {30, 43}
    db = st.TargetDb(ctx, w.arena, w.offsets)
    qps = [st.QueryProfile.from_rows(ctx, r) for r in w.rows]
    gaps = [st.GapCosts(*g) for g in w.gaps]
    b = st.Batch(ctx, db, qps, w.cands, gaps, w.gates)
    b.run()
    ref = b.results()
    off = b.offsets.astype(np.int64)
    b.close()
    reqs = [(qps[k], w.cands[k], gaps[k], w.gates[k]) for k in range(n)]
    # every request through the persistent serving kernel, arriving as a
    # Poisson stream near saturation so that requests overlap on the device
    rng = np.random.default_rng(8)
    arr = np.cumsum(rng.exponential(1e3 / 12000.0, size=n))
    srv = st.Server(ctx, db, workers=3, persistent=True)
    outs, lat = srv.run_trace(reqs, arr)
    stats = srv.stats()
    srv.close()
    assert stats["published"] > 0.9 * n
    bad = [k for k in range(n) if not (outs[k] == ref[off[k]:off[k + 1]]).all()]
    assert not bad, bad[:10]
    # a quarter of the requests through the micro-batch slots
    sub = list(range(0, n, 4))
    srv = st.Server(ctx, db, workers=3, max_batch_queries=64, persistent=False)
    outs2, _ = srv.run_trace([reqs[k] for k in sub], np.zeros(len(sub)))
    srv.close()
    bad = [k for k, o in zip(sub, outs2) if not (o == ref[off[k]:off[k + 1]]).all()]
    assert not bad, bad[:10]
    # sampled pairs of profile and plain requests against the oracle
    picks = list(rng.choice(np.nonzero(prof)[0], 25)) + list(rng.choice(np.nonzero(~prof)[0], 25))
    for k in picks:
        j = int(rng.integers(len(w.cands[k])))
        t = w.target(int(w.cands[k][j]))
        o, e = w.gaps[k]
        want = (O.gotoh_score(w.rows[k], t, o, e),) + O.ung_pair(w.rows[k], t, o, e, w.gates[k])
        r = outs[k][j]
        got = (int(r["swg"]), int(r["ung"]), int(r["best_diag"]), int(r["gate_score"]), bool(r["pass"]))
        assert got == want, (k, j)
    for q in qps:
        q.close()
    db.close()
