"""Ungapped jobs, gate and the all-diagonal batch on the GPU vs the oracle
(bit-exact)."""
import numpy as np
import pytest

import oracle as O
import sabertooth_amd as st
from sabertooth_amd import GapCosts

pytestmark = pytest.mark.gpu


def _db(seqs):
    offs = np.zeros(len(seqs) + 1, np.uint64)
    offs[1:] = np.cumsum([len(s) for s in seqs])
    return np.concatenate(seqs).astype(np.uint8), offs


def _jobs_for(rng, q, targets, per_target_max=6, homolog_every=5):
    m = len(q)
    jobs, recs = [], []
    for tid, t in enumerate(targets):
        n = len(t)
        j0 = len(jobs)
        k = int(rng.integers(1, per_target_max + 1))
        for _ in range(k):
            d = int(rng.integers(-m - 2, n + 2)) if tid % homolog_every else 0
            lo, hi = O.diag_range(m, n, d)
            if rng.random() < 0.3 and hi > lo + 4:
                a = int(rng.integers(lo, hi))
                b = a + int(rng.integers(0, hi - a)) + 1
                lo, hi = a, b
            jobs.append((tid, d, lo, hi))
        recs.append((tid, k + 3, j0, len(jobs)))
    return np.array(jobs, dtype=O.JOB_DT), np.array(recs, dtype=O.REC_DT)


def _workload(rng, m, ntargets):
    q = rng.integers(0, 21, size=m, dtype=np.uint8)
    targets = []
    for k in range(ntargets):
        if k % 5 == 0:
            t = q.copy()
            for _ in range(m // 10):
                t[int(rng.integers(0, m))] = rng.integers(0, 20)
        else:
            t = rng.integers(0, 22, size=int(rng.integers(1, 300)), dtype=np.uint8)
        targets.append(t)
    return q, targets


@pytest.mark.parametrize("case", range(6))
def test_jobs_match_kadane(ctx, case):
    rng = np.random.default_rng(17 + case)
    q, targets = _workload(rng, 1 + int(rng.integers(0, 400 if case % 3 == 0 else 90)), 70)
    rows = O.seq_rows(q)
    jobs, _ = _jobs_for(rng, q, targets)
    arena, offs = _db(targets)
    db = st.TargetDb(ctx, arena, offs)
    qp = st.QueryProfile.from_rows(ctx, rows)
    got = st.score_batch(qp, db, jobs)
    for k, j in enumerate(jobs):
        t = targets[j["tid"]]
        want = O.kadane_score(rows, t, int(j["diag"]), int(j["lo"]), max(int(j["hi"]), int(j["lo"])))
        assert got[k] == want, (case, k)


@pytest.mark.parametrize("regions", [False, True])
@pytest.mark.parametrize("case", range(4))
def test_score_and_gate_matches_oracle(ctx, case, regions):
    rng = np.random.default_rng(100 + case)
    q, targets = _workload(rng, 30 + int(rng.integers(0, 200)), 120)
    rows = O.seq_rows(q)
    jobs, recs = _jobs_for(rng, q, targets, per_target_max=16)
    arena, offs = _db(targets)
    db = st.TargetDb(ctx, arena, offs)
    qp = st.QueryProfile.from_rows(ctx, rows)
    for gate in [0, 20, 35, 60]:
        want, wchains, wscores = O.score_and_gate(rows, arena, offs, jobs, recs, 11, 1, gate, regions)
        got, gchains, gscores = st.score_and_gate(qp, db, jobs, recs, GapCosts(11, 1), gate, regions)
        assert (gscores == wscores).all()
        assert gchains == wchains
        assert len(got) == len(want)
        for a, b in zip(got, want):
            assert tuple(a) == tuple(b), (gate, a, b)


def test_chain_rescue_on_gpu(ctx):
    q = O.encode("MKVLLACDEFGHIKLMNPQRSTVWYHHHHWWCCMM")
    t = O.encode("QRST" * 7 + "MKVLLACDEFGHIKLMNP" + "GGGG" + "QRSTVWYHHHHWWCCMM")
    rows = O.seq_rows(q)
    db = st.TargetDb.from_sequences(ctx, [t])
    qp = st.QueryProfile.from_rows(ctx, rows)
    b = st.Batch(ctx, db, [qp], [np.array([0], np.uint32)], [GapCosts(11, 1)], [0])
    b.run()
    r = b.results()[0]
    ung, bd, gs, ok = O.ung_pair(rows, t, 11, 1, 0)
    assert (r["ung"], r["best_diag"], r["gate_score"], bool(r["pass"])) == (ung, bd, gs, ok)
    assert gs > ung  # the chain wins


def _batch_check(ctx, queries, targets, cand_lists, gaps, gates, rows_list=None):
    arena, offs = _db(targets)
    db = st.TargetDb(ctx, arena, offs)
    rows_list = rows_list or [O.seq_rows(q) for q in queries]
    qps = [st.QueryProfile.from_rows(ctx, r) for r in rows_list]
    b = st.Batch(ctx, db, qps, cand_lists, gaps, gates)
    b.run()
    res = b.results()
    k = 0
    for qi, cl in enumerate(cand_lists):
        rows = rows_list[qi]
        g = gaps[qi]
        for tid in cl:
            t = targets[int(tid)]
            swg = O.gotoh_score(rows, t, g.open, g.extend)
            ung, bd, gs, ok = O.ung_pair(rows, t, g.open, g.extend, gates[qi])
            r = res[k]
            assert (int(r["swg"]), int(r["ung"]), int(r["best_diag"]), int(r["gate_score"]), bool(r["pass"])) == \
                (swg, ung, bd, gs, ok), (qi, int(tid), len(rows), len(t))
            k += 1
    return res


def test_batch_all_outputs_match_oracle(ctx):
    rng = np.random.default_rng(5)
    queries = [rng.integers(0, 20, size=L, dtype=np.uint8) for L in [1, 9, 50, 130, 257, 400, 700]]
    targets = [rng.integers(0, 20, size=int(rng.integers(1, 600)), dtype=np.uint8) for _ in range(240)]
    # homologs of each query
    for i, q in enumerate(queries):
        for k in range(6):
            t = q.copy()
            for _ in range(len(q) // 4):
                t[int(rng.integers(0, len(q)))] = rng.integers(0, 20)
            targets.append(np.concatenate([rng.integers(0, 20, size=int(rng.integers(0, 40)), dtype=np.uint8), t]))
    cands = [rng.choice(len(targets), size=int(rng.integers(1, 120)), replace=False).astype(np.uint32)
             for _ in queries]
    gaps = [GapCosts(11, 1)] * len(queries)
    _batch_check(ctx, queries, targets, cands, gaps, [35] * len(queries))


def test_batch_mixed_gaps_pssm_and_saturation(ctx):
    rng = np.random.default_rng(6)
    W = O.aa2num()[ord("W")]
    qw = np.full(3500, W, np.uint8)
    q1 = rng.integers(0, 20, size=120, dtype=np.uint8)
    pssm = rng.integers(-40, 60, size=(90, 21)).astype(np.int32)
    pssm[:, 20] = 0
    targets = [qw, qw[:2000], rng.integers(0, 20, size=300, dtype=np.uint8)]
    targets += [rng.integers(0, 20, size=int(rng.integers(1, 250)), dtype=np.uint8) for _ in range(60)]
    rows = [O.seq_rows(qw), O.seq_rows(q1), pssm]
    cands = [np.array([0, 1, 2, 5], np.uint32), np.arange(len(targets), dtype=np.uint32),
             np.arange(2, len(targets), dtype=np.uint32)]
    gaps = [GapCosts(11, 1), GapCosts(6, 2), GapCosts(44, 4)]
    res = _batch_check(ctx, [None] * 3, targets, cands, gaps, [41, 35, 30], rows_list=rows)
    assert res[0]["kernel"] == st.KERNEL_I32 and res[0]["swg"] == 35000
