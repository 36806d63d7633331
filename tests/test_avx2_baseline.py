"""The CPU baselines restated from the reference's AVX2 paths give the
oracle's results: the ungapped stage (32-lane biased-u8 Kadane batches,
exact recompute above the u8 limit, chain and gate) equals the scalar
score_and_gate on seed-hit jobs, over several queries, EV_CHUNK chunking,
regions, saturating diagonals and profile rows that leave no u8 headroom."""
import numpy as np
import pytest

import oracle as O
from test_oracle import kmer_index, query_seeds


def _case(rng, nq, ntargets, k=3, alphabet=20):
    targets = [rng.integers(0, alphabet, size=int(rng.integers(20, 700)), dtype=np.uint8) for _ in range(ntargets)]
    queries = [rng.integers(0, alphabet, size=int(rng.integers(30, 600)), dtype=np.uint8) for _ in range(nq)]
    # planted homologs: long diagonals whose Kadane score exceeds the u8 limit
    for i, q in enumerate(queries):
        for j in range(3):
            t = q.copy()
            t[rng.random(len(t)) < 0.1] = rng.integers(0, alphabet)
            targets[(i * 7 + j * 3) % ntargets] = np.concatenate([targets[(i * 7 + j * 3) % ntargets][:40], t])
    post, of = kmer_index(targets, k)
    tlen = np.array([len(t) for t in targets])
    arena = np.concatenate(targets)
    offs = np.zeros(len(targets) + 1, np.uint64)
    offs[1:] = np.cumsum(tlen)
    jl, rl = [], []
    for q in queries:
        seeds = query_seeds(q, k, of)
        j, r = O.seed_hits_to_jobs(len(q), seeds, post, tlen, 4, 1, 1, 48, k)
        jl.append(j)
        rl.append(r)
    return queries, arena, offs, jl, rl


@pytest.mark.parametrize("regions", [False, True])
def test_ungapped_baseline_equals_the_oracle(regions):
    rng = np.random.default_rng(5 + regions)
    queries, arena, offs, jl, rl = _case(rng, 6, 400)
    rows = [O.seq_rows(q) for q in queries]
    # a profile-like query whose score range leaves no u8 headroom (scalar)
    prow = rng.integers(-90, 120, size=(len(queries[0]), 21)).astype(np.int32)
    prow[:, 20] = 0
    rows.append(prow)
    jl.append(jl[0])
    rl.append(rl[0])
    gates = [35, 20, 35, 60, 0, 35, 300]
    opens, exts = [11] * 6 + [44], [1] * 6 + [4]
    got, chains, js, _ = O.avx2_score_and_gate(rows, jl, rl, arena, offs, opens, exts, gates, regions, threads=3,
                                               want_scores=True)
    k = 0
    sat = 0
    for q in range(len(rows)):
        want, wch, wjs = O.score_and_gate(rows[q], arena, offs, jl[q], rl[q], opens[q], exts[q], gates[q], regions)
        assert (got[q] == want).all() and chains[q] == wch, q
        assert (js[k:k + len(jl[q])] == wjs).all(), q
        sat += int((wjs > 240).sum())
        k += len(jl[q])
    assert sat > 0 and sum(len(g) for g in got) > 20 and sum(chains) > 0


def _direct_jobs(rng, m, targets, per_target=(1, 9)):
    jobs, recs = [], []
    for tid, t in enumerate(targets):
        j0 = len(jobs)
        for _ in range(int(rng.integers(*per_target))):
            d = int(rng.integers(-m + 1, len(t)))
            lo, hi = O.diag_range(m, len(t), d)
            if rng.random() < 0.4 and hi > lo + 4:
                a = int(rng.integers(lo, hi))
                lo, hi = a, a + int(rng.integers(0, hi - a)) + 1
            jobs.append((tid, d, lo, hi))
        recs.append((tid, len(jobs) - j0 + 1, j0, len(jobs)))
    return np.array(jobs, dtype=O.JOB_DT), np.array(recs, dtype=O.REC_DT)


def test_ungapped_baseline_chunks_like_the_reference():
    # queries with more than one EV_CHUNK (16,384) of jobs
    rng = np.random.default_rng(9)
    targets = [rng.integers(0, 20, size=int(rng.integers(20, 400)), dtype=np.uint8) for _ in range(5000)]
    queries = [rng.integers(0, 20, size=int(rng.integers(100, 300)), dtype=np.uint8) for _ in range(2)]
    for q in queries:
        for k in range(0, 5000, 97):
            targets[k] = np.concatenate([targets[k][:30], q])
    arena = np.concatenate(targets)
    offs = np.zeros(len(targets) + 1, np.uint64)
    offs[1:] = np.cumsum([len(t) for t in targets])
    jl, rl = zip(*[_direct_jobs(rng, len(q), targets) for q in queries])
    assert min(len(j) for j in jl) > 16384
    rows = [O.seq_rows(q) for q in queries]
    got, chains, js, _ = O.avx2_score_and_gate(rows, jl, rl, arena, offs, [11, 11], [1, 1], [20, 20], False,
                                               threads=2, want_scores=True)
    k = 0
    for q in range(2):
        want, wch, wjs = O.score_and_gate(rows[q], arena, offs, jl[q], rl[q], 11, 1, 20, False)
        assert (got[q] == want).all() and chains[q] == wch
        assert (js[k:k + len(jl[q])] == wjs).all()
        k += len(jl[q])
