"""GPU score + end cell (int32 end-tracking DP) vs the oracle's full-matrix
column-major-first rule (src/simd.rs:1167-1195), bit-exact."""
import numpy as np
import pytest

import oracle as O
import sabertooth_amd as st
from sabertooth_amd import GapCosts
from test_oracle import XorShift

pytestmark = pytest.mark.gpu


def _want(rows, t, g):
    s, qe, te = O.gotoh_textbook(rows, t, g.open, g.extend)
    return (s, qe, te) if s > 0 else (0, -1, -1)


def _check(ctx, rows, targets, g):
    db = st.TargetDb.from_sequences(ctx, targets)
    qp = st.QueryProfile.from_rows(ctx, rows)
    out = qp.sw_score_end_with(db, np.arange(len(targets)), g)
    for k, t in enumerate(targets):
        got = (int(out[k]["score"]), int(out[k]["query_end"]), int(out[k]["target_end"]))
        assert got == _want(rows, t, g), (k, len(rows), len(t), g)


@pytest.mark.parametrize("o,e", [(11, 1), (2, 1), (6, 2), (3, 3), (1, 1), (3, -1), (-2, 1)])
def test_end_rule_with_many_ties(ctx, o, e):
    # the reference's generator: tiny alphabets make many tied optima
    rng = XorShift(424242 + o * 7 + e)
    for rnd in range(6):
        alpha = 2 + rng.range(3) if rnd % 3 == 0 else 2 + rng.range(19)
        q = rng.seq(1 + rng.range(70), alpha)
        targets = [rng.seq(1 + rng.range(70), alpha) for _ in range(40)]
        for k in range(0, 40, 7):
            targets[k][0] = 20
        _check(ctx, O.seq_rows(q), targets, GapCosts(o, e))


def test_long_and_saturating(ctx):
    W = O.aa2num()[ord("W")]
    q = np.full(3000, W, np.uint8)
    rows = O.seq_rows(q)
    db = st.TargetDb.from_sequences(ctx, [q, np.zeros(0, np.uint8), q[:10]])
    out = st.QueryProfile.from_rows(ctx, rows).sw_score_end_with(db, [0, 1, 2], GapCosts(11, 1))
    assert tuple(int(x) for x in (out[0]["score"], out[0]["query_end"], out[0]["target_end"])) == (30000, 2999, 2999)
    assert (int(out[1]["score"]), int(out[1]["query_end"])) == (0, -1)
    assert (int(out[2]["score"]), int(out[2]["query_end"]), int(out[2]["target_end"])) == (100, 9, 9)


@pytest.mark.parametrize("G", [1, 4, 16, 64])
def test_every_group_width(ctx, monkeypatch, G):
    monkeypatch.setenv("ST_FORCE_G", str(G))
    rng = np.random.default_rng(G)
    for m in (5, 130, 700):
        q = rng.integers(0, 4, size=m).astype(np.uint8)
        targets = [rng.integers(0, 4, size=int(L)).astype(np.uint8) for L in rng.integers(1, 600, size=20)]
        targets.append(q.copy())
        _check(ctx, O.seq_rows(q), targets, GapCosts(11, 1))


@pytest.mark.parametrize("o,e", [(11, 1), (0, 0), (-3, 1)])
def test_end_rule_profile_int16_not_int8(ctx, o, e):
    # values beyond int8: the int16 profile is staged (the int8-profile
    # instances take the BLOSUM queries of the other tests)
    rng = np.random.default_rng(51 + abs(o) + e)
    rows = rng.integers(-2000, 3000, size=(130, 21)).astype(np.int32)
    rows[:, 20] = 0
    targets = [rng.integers(0, 4, size=int(L)).astype(np.uint8) for L in rng.integers(1, 300, size=25)]
    _check(ctx, rows, targets, GapCosts(o * 100, e * 50))


@pytest.mark.parametrize("o,e", [(30000, 2000), (0, 0), (-3000, 100)])
def test_end_rule_profile_outside_int16(ctx, o, e):
    # values beyond int16: the int32 profile is staged (the int16-profile
    # instances take the other tests' queries)
    rng = np.random.default_rng(31 + abs(o) + e)
    rows = rng.integers(-20000, 40000, size=(150, 21)).astype(np.int32)
    rows[:, 20] = 0
    targets = [rng.integers(0, 4, size=int(L)).astype(np.uint8) for L in rng.integers(1, 300, size=25)]
    _check(ctx, rows, targets, GapCosts(o, e))
