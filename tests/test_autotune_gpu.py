"""A shape table read from ST_SHAPE_TABLE steers the planner; results stay
bit-exact whatever shapes it picks."""
import numpy as np
import pytest

import oracle as O
import sabertooth_amd as st
from sabertooth_amd import GapCosts, autotune

pytestmark = pytest.mark.gpu


def _batch(ctx, rng):
    targets = [rng.integers(0, 20, size=int(rng.integers(20, 400)), dtype=np.uint8) for _ in range(200)]
    queries = [rng.integers(0, 20, size=L, dtype=np.uint8) for L in (50, 180, 420)]
    db = st.TargetDb.from_sequences(ctx, targets)
    qps = [st.QueryProfile.from_rows(ctx, O.seq_rows(q)) for q in queries]
    cands = [np.arange(200, dtype=np.uint32)] * 3
    return db, qps, cands, targets, queries


def test_table_steers_the_plan_not_the_results(ctx, tmp_path, monkeypatch):
    rng = np.random.default_rng(3)
    db, qps, cands, targets, queries = _batch(ctx, rng)
    monkeypatch.delenv("ST_SHAPE_TABLE", raising=False)
    b0 = st.Batch(ctx, db, qps, cands, [GapCosts(11, 1)] * 3, [35] * 3)
    b0.run()
    ref = b0.results()
    # a table under which one-lane groups look cheapest
    costs = {(R, G): (0.01 if (R, G) == (8, 1) else 5.0) for R in autotune.ROWS for G in autotune.GROUPS}
    p = tmp_path / "t.txt"
    p.write_text(autotune.table_text(costs))
    monkeypatch.setenv("ST_SHAPE_TABLE", str(p))
    b1 = st.Batch(ctx, db, qps, cands, [GapCosts(11, 1)] * 3, [35] * 3)
    b1.run()
    # one-lane groups: 64 groups x 2 targets = 128 pairs per unit
    assert b1.stats()[1] == 3 * 2 and b0.stats()[1] > 6
    assert all((b1.results()[f] == ref[f]).all() for f in ref.dtype.names)
    k = 0
    for qi, cl in enumerate(cands):
        for tid in cl[:20]:
            assert int(b1.results()[k + int(np.where(cl == tid)[0][0])]["swg"]) == \
                O.gotoh_score(O.seq_rows(queries[qi]), targets[int(tid)], 11, 1)
        k += len(cl)
    # a table that cannot be read fails batch creation loudly
    monkeypatch.setenv("ST_SHAPE_TABLE", str(p.with_name("missing.txt")))
    with pytest.raises(st.StError):
        st.Batch(ctx, db, qps, cands, [GapCosts(11, 1)] * 3, [35] * 3)
    for b in (b0, b1):
        b.close()
