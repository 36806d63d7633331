"""The benchmark configurations other than the bench line, as parity cases:
plain all-pairs (reduced), PSSM queries from the PSSM builder (44/4, gate 30)
and long pairs mixed with short ones (int32 scores above 32767).  Every SWG
score is checked against the AVX2 cascade restated from the reference, and
every output of a seeded sample against the oracle (bit-exact)."""
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


def test_plain_allpairs_reduced(ctx):
    w = synth.plain_allpairs(seed=1, n_queries=8, n_targets=3000)
    rows = [O.seq_rows(q) for q in w.queries]
    gaps = [(11, 1)] * len(rows)
    gates = [w.gate] * len(rows)
    b, res = _run(ctx, rows, w.arena, w.offsets, w.cand_lists, gaps, gates)
    _check(rows, w.arena, w.offsets, w.cand_lists, gaps, gates, b, res, 300, 1)


def test_pssm_queries_from_the_builder(ctx):
    w = synth.pssm_search(seed=1, n_profiles=16, n_targets=4000)
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
