"""From the reference's index file to its .m8 rows on the device: the
targets of tests/data/v1_fixture_k2.stidx loaded straight into a device
database, every query of the fixture scored on the GPU, its hits' spans from
the device score passes, the host traceback from those spans, and the rows
of tests/data/v1_fixture_k2.v03.m8 reproduced in order."""
import os

import numpy as np
import pytest

import sabertooth_amd as st
from sabertooth_amd import GapCosts
from sabertooth_amd import host as H

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
IDX = os.path.join(GOLD, "v1_fixture_k2.stidx")


def test_index_to_m8_rows_on_the_device(ctx):
    db = st.TargetDb.from_stidx(ctx, IDX)
    res, offs = H.stidx_targets(IDX)
    ids = H.stidx_ids(IDX)
    n = len(ids)
    assert len(db) == n and db.residues() == len(res)
    plain = st.TargetDb(ctx, res, offs)
    ev = H.EValueParams.sequence(11, 1, True, int(offs[-1]), n)
    params = H.SearchParams(max_evalue=10.0)
    tids = np.arange(n, dtype=np.uint32)
    got = {}
    for qi in range(n):
        q = res[int(offs[qi]):int(offs[qi + 1])]
        rows = H.query_rows(q)
        qp = st.QueryProfile.from_rows(ctx, rows)
        scores = qp.sw_score_batch(db, tids, GapCosts(11, 1))
        assert (scores == qp.sw_score_batch(plain, tids, GapCosts(11, 1))).all()
        spans = st.align_spans(ctx, db, [qp], [tids], [GapCosts(11, 1)])
        hits, cig = H.search_hits(rows, q, res, offs, tids, scores, GapCosts(11, 1), ev, params, spans=spans)
        ref_hits, ref_cig = H.search_hits(rows, q, res, offs, tids, scores, GapCosts(11, 1), ev, params)
        assert (hits == ref_hits).all() and cig == ref_cig
        got[ids[qi]] = [H.format_m8(ids[qi], ids[int(h["target"])], h) for h in hits]
    lines = open(os.path.join(GOLD, "v1_fixture_k2.v03.m8")).read().strip().split("\n")[1:]
    want = {}
    for l in lines:
        want.setdefault(l.split("\t")[0], []).append(l)
    assert sum(len(v) for v in want.values()) == 88
    for q, rows in want.items():
        pos = [got[q].index(r) if r in got[q] else -1 for r in rows]
        assert -1 not in pos and pos == sorted(pos), q
