"""Serving loop: every request's results are the same alone and in the mix
(bit-exact: integer outputs), equal to the plain batch and to the oracle."""
import numpy as np
import pytest

import oracle as O
import sabertooth_amd as st
from sabertooth_amd import GapCosts

pytestmark = pytest.mark.gpu


def _fields(r):
    return np.stack([r["swg"], r["ung"], r["best_diag"], r["gate_score"], r["pass"].astype(np.int32),
                     r["kernel"].astype(np.int32)])


@pytest.fixture(scope="module")
def setup(ctx):
    rng = np.random.default_rng(2024)
    s = O.SplitMix64(11)
    targets = [s.background(int(L)) for L in rng.integers(1, 900, size=2500)]
    queries = [s.background(int(L)) for L in rng.integers(5, 700, size=20)]
    # planted homologs so gates pass and scores are large
    for k, q in enumerate(queries):
        t = q.copy()
        t[rng.integers(0, len(q), size=len(q) // 4)] = rng.integers(0, 20)
        targets[k * 7] = np.concatenate([targets[k * 7][:30], t])
    rows = [O.seq_rows(q) for q in queries]
    # a PSSM-like query with profile gap costs
    prow = rng.integers(-60, 60, size=(230, 21)).astype(np.int32)
    prow[:, 20] = 0
    rows.append(prow)
    # a query whose self-score saturates int16 (rescored in int32)
    W = O.aa2num()[ord("W")]
    wq = np.full(3400, W, np.uint8)
    rows.append(O.seq_rows(wq))
    targets.append(wq.copy())
    # negative gap costs: the exact unfloored int32 kernel
    rows.append(O.seq_rows(queries[1]))
    db = st.TargetDb.from_sequences(ctx, targets)
    qps = [st.QueryProfile.from_rows(ctx, r) for r in rows]
    nt = len(targets)
    cands = [rng.choice(nt - 1, size=int(rng.integers(0, 400)), replace=False).astype(np.uint32) for _ in rows]
    cands[-2] = np.array([nt - 1, 3, 4], np.uint32)
    cands[-1] = cands[-1][:60]
    cands[3] = np.zeros(0, np.uint32)  # an empty candidate list
    gaps = [GapCosts(11, 1)] * len(queries) + [GapCosts(44, 4), GapCosts(11, 1), GapCosts(3, -1)]
    gates = [35] * len(queries) + [30, 41, 20]
    b = st.Batch(ctx, db, qps, cands, gaps, gates)
    b.run()
    res = b.results()
    ref = [res[int(b.offsets[q]):int(b.offsets[q + 1])] for q in range(len(qps))]
    return dict(db=db, qps=qps, rows=rows, targets=targets, cands=cands, gaps=gaps, gates=gates, ref=ref)


def test_batch_reference_matches_oracle(setup):
    rng = np.random.default_rng(5)
    for q in range(len(setup["qps"])):
        c = setup["cands"][q]
        for k in rng.choice(len(c), size=min(len(c), 12), replace=False) if len(c) else []:
            t = setup["targets"][int(c[k])]
            g = setup["gaps"][q]
            want = (O.gotoh_score(setup["rows"][q], t, g.open, g.extend),) + \
                O.ung_pair(setup["rows"][q], t, g.open, g.extend, setup["gates"][q])
            r = setup["ref"][q][k]
            assert (int(r["swg"]), int(r["ung"]), int(r["best_diag"]), int(r["gate_score"]),
                    bool(r["pass"])) == want, (q, k)
    assert setup["ref"][-2]["swg"][0] == 34000 and setup["ref"][-2]["kernel"][0] == st.KERNEL_I32


@pytest.mark.parametrize("persistent", [False, True])
def test_alone_equals_batch(ctx, setup, persistent):
    srv = st.Server(ctx, setup["db"], workers=2, persistent=persistent)
    try:
        for q, qp in enumerate(setup["qps"]):
            t = srv.submit(qp, setup["cands"][q], setup["gaps"][q], setup["gates"][q])
            out, lat = srv.wait(t, timeout_ms=60000)
            assert lat > 0
            assert (_fields(out) == _fields(setup["ref"][q])).all(), q
        s = srv.stats()
        assert s["queries"] == len(setup["qps"]) and s["rescored"] >= 1
    finally:
        srv.close()


@pytest.mark.parametrize("persistent", [False, True])
def test_in_the_mix_equals_alone(ctx, setup, persistent):
    srv = st.Server(ctx, setup["db"], workers=3, max_batch_queries=5, persistent=persistent)
    try:
        order = list(range(len(setup["qps"]))) * 3
        tickets = [(q, srv.submit(setup["qps"][q], setup["cands"][q], setup["gaps"][q], setup["gates"][q]))
                   for q in order]
        for q, t in tickets:
            out, _ = srv.wait(t, timeout_ms=60000)
            assert (_fields(out) == _fields(setup["ref"][q])).all(), q
        s = srv.stats()
        assert s["queries"] == len(order)
        assert s["max_batch_queries"] <= 5 and (persistent or s["max_batch_queries"] > 1)
    finally:
        srv.close()


@pytest.mark.parametrize("persistent", [False, True])
def test_open_loop_trace_equals_alone(ctx, setup, persistent):
    srv = st.Server(ctx, setup["db"], workers=4, persistent=persistent)
    try:
        rng = np.random.default_rng(9)
        qi = rng.integers(0, len(setup["qps"]), size=60)
        arrivals = np.cumsum(rng.exponential(0.4, size=len(qi)))
        reqs = [(setup["qps"][q], setup["cands"][q], setup["gaps"][q], setup["gates"][q]) for q in qi]
        outs, lat = srv.run_trace(reqs, arrivals)
        assert (lat > 0).all()
        for q, out in zip(qi, outs):
            assert (_fields(out) == _fields(setup["ref"][q])).all(), q
    finally:
        srv.close()


@pytest.mark.parametrize("persistent", [False, True])
def test_tiny_budget_admits_one_at_a_time(ctx, setup, persistent):
    srv = st.Server(ctx, setup["db"], workers=4, state_budget_bytes=1, persistent=persistent)
    try:
        tickets = [(q, srv.submit(setup["qps"][q], setup["cands"][q], setup["gaps"][q], setup["gates"][q]))
                   for q in range(len(setup["qps"]))]
        for q, t in tickets:
            out, _ = srv.wait(t, timeout_ms=60000)
            assert (_fields(out) == _fields(setup["ref"][q])).all(), q
        assert srv.stats()["max_batch_queries"] <= 1
    finally:
        srv.close()


def test_small_arena_restarts_the_serving_kernel(ctx, setup):
    # a 1 MiB arena holds a few dozen requests: publishing drains and
    # relaunches the serving kernel many times; results do not change
    srv = st.Server(ctx, setup["db"], workers=2, persistent=True, arena_mib=1)
    try:
        order = list(range(len(setup["qps"]))) * 4
        tickets = [(q, srv.submit(setup["qps"][q], setup["cands"][q], setup["gaps"][q], setup["gates"][q]))
                   for q in order]
        for q, t in tickets:
            out, _ = srv.wait(t, timeout_ms=60000)
            assert (_fields(out) == _fields(setup["ref"][q])).all(), q
        s = srv.stats()
        assert s["kernel_launches"] >= 2 and s["published"] > 0
    finally:
        srv.close()


def test_serving_kernel_relaunches_after_idling(ctx, setup):
    import time
    srv = st.Server(ctx, setup["db"], workers=1, persistent=True)
    try:
        for q in (0, 1):
            out, _ = srv.wait(srv.submit(setup["qps"][q], setup["cands"][q], setup["gaps"][q], setup["gates"][q]),
                              timeout_ms=60000)
            assert (_fields(out) == _fields(setup["ref"][q])).all(), q
            time.sleep(0.05)  # longer than the kernel's 5 ms idle limit
        assert srv.stats()["kernel_launches"] >= 2
    finally:
        srv.close()


def test_bad_requests_fail_cleanly(ctx, setup):
    srv = st.Server(ctx, setup["db"], workers=1)
    try:
        with pytest.raises(st.StError):
            srv.submit(setup["qps"][0], np.array([len(setup["targets"])], np.uint32), GapCosts(11, 1), 35)
        with pytest.raises(st.StError):
            srv.wait(123456, timeout_ms=10)
    finally:
        srv.close()


@pytest.mark.parametrize("prio", ["0", "1,2,3"])
def test_age_priorities_leave_results_unchanged(ctx, setup, monkeypatch, prio):
    # the serving kernel raises a unit's wave priority with its request's
    # age; off, and with thresholds so small that every unit reaches
    # priority 3, every row still equals the plain batch's
    monkeypatch.setenv("ST_SERVE_PRIO_US", prio)
    srv = st.Server(ctx, setup["db"], workers=2, persistent=True)
    try:
        rng = np.random.default_rng(31)
        qi = rng.integers(0, len(setup["qps"]), size=40)
        arrivals = np.cumsum(rng.exponential(0.2, size=len(qi)))
        reqs = [(setup["qps"][q], setup["cands"][q], setup["gaps"][q], setup["gates"][q]) for q in qi]
        outs, _ = srv.run_trace(reqs, arrivals)
        for q, out in zip(qi, outs):
            assert (_fields(out) == _fields(setup["ref"][q])).all(), q
    finally:
        srv.close()
