"""st_score_and_gate_noalloc (no device allocation, fixed device scratch)
through the C ABI: the same candidates, chain count and job scores as
st_score_and_gate, on two calls in a row (the second smaller, so stale
scratch contents from the first would show), checked against the oracle as
well; sizes beyond the scratch are refused before anything runs."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
import sabertooth_amd as st
from sabertooth_amd import GapCosts, _lib

pytestmark = pytest.mark.gpu


def _workload(rng, m, ntargets, per_target_max=16):
    q = rng.integers(0, 20, size=m, dtype=np.uint8)
    targets = []
    for k in range(ntargets):
        if k % 6 == 0:
            t = q.copy()
            t[rng.integers(0, m, size=m // 4)] = rng.integers(0, 20)
        else:
            t = rng.integers(0, 21, size=int(rng.integers(1, 400)), dtype=np.uint8)
        targets.append(t)
    jobs, recs = [], []
    for tid, t in enumerate(targets):
        j0 = len(jobs)
        k = int(rng.integers(0 if tid % 11 == 5 else 1, per_target_max + 1))
        for _ in range(k):
            d = 0 if tid % 6 == 0 and not jobs[j0:] else int(rng.integers(-m + 1, len(t)))
            lo, hi = O.diag_range(m, len(t), d)
            if rng.random() < 0.3 and hi > lo + 4:
                a = int(rng.integers(lo, hi))
                lo, hi = a, a + int(rng.integers(0, hi - a)) + 1
            jobs.append((tid, d, lo, hi))
        recs.append((tid, k + 2, j0, len(jobs)))
    return q, targets, np.array(jobs, dtype=O.JOB_DT), np.array(recs, dtype=O.REC_DT)


def _noalloc(qp, db, jobs, recs, gaps, gate, regions, chains0=0):
    L = _lib.lib()
    out = np.zeros(max(len(recs), 1), dtype=_lib.CAND_DT)
    n, chains = C.c_uint64(0), C.c_uint64(chains0)
    js = np.zeros(max(len(jobs), 1), np.int32)
    rc = L.st_score_and_gate_noalloc(qp.ctx.handle, qp.handle, db.handle, _lib.ptr(jobs), len(jobs), _lib.ptr(recs),
                                     len(recs), gaps.open, gaps.extend, gate, 1 if regions else 0, _lib.ptr(out),
                                     C.byref(n), C.byref(chains), _lib.ptr(js))
    _lib.check(rc)
    return out[: n.value].copy(), chains.value - chains0, js[: len(jobs)].copy()


def test_noalloc_equals_score_and_gate_twice(ctx):
    rng = np.random.default_rng(77)
    calls = [(_workload(rng, 260, 2500), 35, False), (_workload(rng, 90, 300), 20, True)]
    for (q, targets, jobs, recs), gate, regions in calls:
        assert len(jobs) <= 131072 and len(recs) <= 131072
        rows = O.seq_rows(q)
        db = st.TargetDb.from_sequences(ctx, targets)
        qp = st.QueryProfile.from_rows(ctx, rows)
        want, wchains, wscores = st.score_and_gate(qp, db, jobs, recs, GapCosts(11, 1), gate, regions)
        got, gchains, gscores = _noalloc(qp, db, jobs, recs, GapCosts(11, 1), gate, regions, chains0=7)
        assert len(want) > 10 and (got == want).all() and gchains == wchains and (gscores == wscores).all()
        arena = np.concatenate(targets)
        offs = np.zeros(len(targets) + 1, np.uint64)
        offs[1:] = np.cumsum([len(t) for t in targets])
        oc, ochains, ojs = O.score_and_gate(rows, arena, offs, jobs, recs, 11, 1, gate, regions)
        assert (got == oc).all() and gchains == ochains and (gscores == ojs).all()
        qp.close()
        db.close()
    print(f"no-alloc score_and_gate: two calls ({len(calls[0][0][2])} then {len(calls[1][0][2])} jobs) equal",
          flush=True)


def test_noalloc_refuses_sizes_beyond_its_scratch(ctx):
    q = np.arange(20, dtype=np.uint8)
    db = st.TargetDb.from_sequences(ctx, [q])
    qp = st.QueryProfile.from_rows(ctx, O.seq_rows(q))
    jobs = np.zeros(131073, dtype=O.JOB_DT)
    recs = np.zeros(1, dtype=O.REC_DT)
    with pytest.raises(st.StError) as e:
        _noalloc(qp, db, jobs, recs, GapCosts(11, 1), 0, False)
    assert e.value.code == _lib.ST_ERR_INVALID and "131072" in str(e.value)
    # empty input: nothing to do, nothing written
    got, ch, _ = _noalloc(qp, db, jobs[:0], recs[:0], GapCosts(11, 1), 0, False)
    assert len(got) == 0 and ch == 0
    qp.close()
    db.close()
