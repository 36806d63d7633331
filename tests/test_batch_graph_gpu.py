"""A batch's DP phase captured into a HIP graph and replayed: every replay
matches the eager run and the oracle, bit-exact, with eager work between."""
import numpy as np
import pytest

import oracle as O
import sabertooth_amd as st
from sabertooth_amd import GapCosts

pytestmark = pytest.mark.gpu

FIELDS = ["swg", "ung", "best_diag", "gate_score", "pass", "kernel"]


def _db(ctx, targets):
    lens = np.array([len(t) for t in targets], np.uint64)
    offs = np.zeros(len(targets) + 1, np.uint64)
    offs[1:] = np.cumsum(lens)
    arena = np.concatenate(targets).astype(np.uint8) if targets else np.zeros(0, np.uint8)
    return st.TargetDb(ctx, arena, offs)


def _workload(seed):
    rng = np.random.default_rng(seed)
    W = O.aa2num()[ord("W")]
    qw = np.full(3400, W, np.uint8)  # reaches the int16 clamp: rescored in int32
    queries = [qw] + [rng.integers(0, 20, size=L, dtype=np.uint8) for L in (40, 190, 520)]
    targets = [qw, qw[:3300], qw[:2500]]
    targets += [rng.integers(0, 20, size=int(rng.integers(1, 900)), dtype=np.uint8) for _ in range(300)]
    for q in queries[1:]:
        for _ in range(4):
            t = q.copy()
            for _ in range(len(q) // 5):
                t[int(rng.integers(0, len(q)))] = rng.integers(0, 20)
            targets.append(t)
    cands = [np.array([0, 1, 2, 7, 8], np.uint32)]
    cands += [rng.choice(len(targets), size=int(rng.integers(60, 200)), replace=False).astype(np.uint32)
              for _ in queries[1:]]
    gaps = [GapCosts(11, 1), GapCosts(11, 1), GapCosts(6, 2), GapCosts(44, 4)]
    gates = [41, 35, 30, 38]
    return queries, targets, cands, gaps, gates


def _make(ctx, seed):
    queries, targets, cands, gaps, gates = _workload(seed)
    db = _db(ctx, targets)
    rows = [O.seq_rows(q) for q in queries]
    qps = [st.QueryProfile.from_rows(ctx, r) for r in rows]
    b = st.Batch(ctx, db, qps, cands, gaps, gates)
    return b, (db, qps), (rows, targets, cands, gaps, gates)


def _oracle_rows(rows, targets, cands, gaps, gates):
    out = []
    for qi, cl in enumerate(cands):
        for tid in cl:
            t = targets[int(tid)]
            swg = O.gotoh_score(rows[qi], t, gaps[qi].open, gaps[qi].extend)
            ung, bd, gs, ok = O.ung_pair(rows[qi], t, gaps[qi].open, gaps[qi].extend, gates[qi])
            out.append((swg, ung, bd, gs, ok))
    return out


def _rows(res):
    return [tuple(int(r[f]) for f in FIELDS) for r in res]


def test_graph_replays_match_eager_and_oracle(ctx):
    g, keep_g, inp = _make(ctx, 11)
    other, keep_o, _ = _make(ctx, 12)
    eager, keep_e, _ = _make(ctx, 11)
    # the graph batch's first results come from the captured graph alone
    g.run(graph=True)
    got = g.results()
    dp_ms, _, _, resc = g.timing()
    want = _oracle_rows(*inp)
    assert [(r[0], r[1], r[2], r[3], bool(r[4])) for r in _rows(got)] == want
    assert resc > 0 and got[0]["kernel"] == st.KERNEL_I32 and int(got[0]["swg"]) == 34000
    eager.run()
    ref = _rows(eager.results())
    ref_resc = eager.timing()[3]
    assert _rows(got) == ref and resc == ref_resc
    ref_dp = eager.timing()[0]
    for k in range(4):
        # eager work on the same context and streams between replays
        other.run()
        if k % 2:
            g.run()
            assert _rows(g.results()) == ref
        g.run(graph=True)
        dp_ms, _, _, resc = g.timing()
        # a replay that skipped the clamp-list clear would list every
        # clamped pair twice; one that skipped the counter reset would do
        # no DP work at all
        assert resc == ref_resc, (k, resc, ref_resc)
        assert dp_ms > 0.2 * ref_dp, (k, dp_ms, ref_dp)
        assert _rows(g.results()) == ref, k
        assert (g.launches()["ms"] == 0).all()
    for b in (g, other, eager):
        b.close()


def test_graph_of_an_empty_batch(ctx):
    db = _db(ctx, [np.array([1, 2, 3], np.uint8)])
    qp = st.QueryProfile.from_rows(ctx, O.seq_rows(np.array([4, 5], np.uint8)))
    b = st.Batch(ctx, db, [qp], [np.zeros(0, np.uint32)], [GapCosts(11, 1)], [0])
    b.run(graph=True)
    b.run(graph=True)
    assert len(b.results()) == 0
    b.close()
