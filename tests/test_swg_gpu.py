"""SWG score on the GPU vs the oracle, bit-exact (integer path: no tolerance)."""
import numpy as np
import pytest

import oracle as O
import sabertooth_amd as st
from sabertooth_amd import GapCosts

pytestmark = pytest.mark.gpu


def _rand_seqs(rng, count, lo, hi, alpha):
    return [rng.integers(0, alpha, size=int(rng.integers(lo, hi + 1)), dtype=np.uint8) for _ in range(count)]


def _check(ctx, rows, targets, gaps, want_kernel=None):
    db = st.TargetDb.from_sequences(ctx, targets)
    qp = st.QueryProfile.from_rows(ctx, rows)
    sc, kn = qp.sw_score_with(db, np.arange(len(targets)), gaps)
    for k, t in enumerate(targets):
        want = O.gotoh_score(rows, t, gaps.open, gaps.extend)
        assert sc[k] == want, (k, len(rows), len(t), gaps, int(sc[k]), want)
    if want_kernel is not None:
        assert (kn == want_kernel).all()
    return sc, kn


@pytest.mark.parametrize("open_,ext", [(11, 1), (1, 1), (2, 1), (3, 1), (2, 0), (4, 1), (6, 2), (5, 5), (44, 4),
                                       (1, 3), (0, 0)])
def test_random_pairs_all_gap_costs(ctx, open_, ext):
    rng = np.random.default_rng(1000 + open_ * 31 + ext)
    for rnd in range(4):
        alpha = int(rng.integers(2, 21))
        q = rng.integers(0, alpha, size=int(rng.integers(1, 200)), dtype=np.uint8)
        targets = _rand_seqs(rng, 150, 1, 260, alpha)
        # some targets embed the query: high scores
        for k in range(0, len(targets), 9):
            p = int(rng.integers(0, len(targets[k]) + 1))
            targets[k] = np.concatenate([targets[k][:p], q, targets[k][p:]])
        _check(ctx, O.seq_rows(q), targets, GapCosts(open_, ext))


@pytest.mark.parametrize("m", [1, 2, 7, 8, 9, 16, 17, 31, 63, 64, 65, 127, 128, 129, 255, 256, 257, 300, 513, 1111])
def test_query_lengths_cover_every_launch_shape(ctx, m):
    rng = np.random.default_rng(m)
    q = rng.integers(0, 20, size=m, dtype=np.uint8)
    lens = [1, 2, 3, 15, 16, 17, 40, 100, 150, 151, 400, 1000]
    targets = [rng.integers(0, 20, size=L, dtype=np.uint8) for L in lens]
    targets += [q.copy(), np.concatenate([targets[5], q[: m // 2], targets[3]])]
    targets += _rand_seqs(rng, 70, 1, 500, 20)
    _check(ctx, O.seq_rows(q), targets, GapCosts(11, 1))


def test_empty_targets_and_single_residue(ctx):
    rng = np.random.default_rng(3)
    q = rng.integers(0, 20, size=50, dtype=np.uint8)
    targets = [np.zeros(0, np.uint8), rng.integers(0, 20, size=1, dtype=np.uint8), np.zeros(0, np.uint8),
               rng.integers(0, 20, size=80, dtype=np.uint8)]
    sc, _ = _check(ctx, O.seq_rows(q), targets, GapCosts(11, 1))
    assert sc[0] == 0 and sc[2] == 0


def test_x_and_out_of_alphabet_codes(ctx):
    rng = np.random.default_rng(7)
    q = rng.integers(0, 20, size=30, dtype=np.uint8)
    t = rng.integers(0, 20, size=40, dtype=np.uint8)
    t[[5, 15, 20]] = 20
    t2 = t.copy()
    t2[7] = 200
    t3 = t.copy()
    t3[[1, 9, 33]] = [21, 255, 99]
    _check(ctx, O.seq_rows(q), [t, t2, t3], GapCosts(11, 1))


def test_saturation_rescored_in_int32(ctx):
    W = O.aa2num()[ord("W")]
    q = np.full(4000, W, np.uint8)
    rows = O.seq_rows(q)
    n_ok = 32767 // 10 - 2
    sc, kn = _check(ctx, rows, [q, q[:n_ok], q[:3000]], GapCosts(11, 1))
    assert sc[0] == 40000 and kn[0] == st.KERNEL_I32
    assert sc[1] == n_ok * 10 and kn[1] == st.KERNEL_I16
    assert sc[2] == 30000 and kn[2] == st.KERNEL_I16


@pytest.mark.parametrize("force", [("64", "16"), ("32", "16"), ("64", "8")])
def test_forced_shape_the_int32_rescoring_cannot_take(ctx, monkeypatch, force):
    # forced tuning shapes apply to every plan; the int32 rescoring of a
    # saturated pair cannot take 64 x 16 (its profile slice outgrows LDS),
    # so that plan falls back to a free shape: every score stays exact
    monkeypatch.setenv("ST_FORCE_G", force[0])
    monkeypatch.setenv("ST_FORCE_R", force[1])
    W = O.aa2num()[ord("W")]
    q = np.full(3400, W, np.uint8)
    rows = O.seq_rows(q)
    sc, kn = _check(ctx, rows, [q, q[:1000]], GapCosts(11, 1))
    assert sc[0] == 34000 and kn[0] == st.KERNEL_I32
    assert sc[1] == 10000 and kn[1] == st.KERNEL_I16


def test_huge_rows_use_int32(ctx):
    rows = np.full((3, 21), 40000, np.int32)
    sc, kn = _check(ctx, rows, [np.array([0, 1, 2], np.uint8)], GapCosts(11, 1))
    assert sc[0] == 120000 and kn[0] == st.KERNEL_I32


def test_negative_gap_costs_exact(ctx):
    rng = np.random.default_rng(17)
    for open_, ext in [(3, -1), (-2, 1), (0, -1)]:
        q = rng.integers(0, 20, size=int(rng.integers(1, 40)), dtype=np.uint8)
        targets = _rand_seqs(rng, 30, 1, 50, 20)
        _check(ctx, O.seq_rows(q), targets, GapCosts(open_, ext))


def test_adjacent_gap_runs_across_rows(ctx):
    # strongly negative mismatches make adjacent gap runs optimal
    # (src/simd.rs:1236-1261)
    rows = np.full((40, 21), -40, np.int32)
    for c in range(40):
        rows[c, c % 20] = 30
    t0 = np.array([x % 20 for x in list(range(20)) + [5, 5, 5] + list(range(26, 40))], np.uint8)
    rng = np.random.default_rng(77)
    targets = [t0] + _rand_seqs(rng, 200, 10, 70, 20)
    for open_, ext in [(44, 4), (20, 1), (12, 2)]:
        _check(ctx, rows, targets, GapCosts(open_, ext))


def test_pssm_like_wide_rows(ctx):
    rng = np.random.default_rng(99)
    for m in [6, 35, 130, 600]:
        rows = rng.integers(-128, 128, size=(m, 21)).astype(np.int32)
        rows[:, 20] = 0
        targets = _rand_seqs(rng, 60, 1, 400, 20)
        _check(ctx, rows, targets, GapCosts(44, 4))


@pytest.mark.parametrize("G", [1, 2, 4, 8, 16, 32, 64])
@pytest.mark.parametrize("R", [8, 12, 16])
def test_every_group_width_and_rows(ctx, monkeypatch, G, R):
    # every launch shape on multi-pass queries, with all outputs checked
    monkeypatch.setenv("ST_FORCE_G", str(G))
    monkeypatch.setenv("ST_FORCE_R", str(R))
    rng = np.random.default_rng(G * 100 + R)
    queries = [rng.integers(0, 20, size=L, dtype=np.uint8) for L in (5, 70, 333, 1100)]
    targets = [rng.integers(0, 20, size=int(L), dtype=np.uint8) for L in rng.integers(1, 700, size=40)]
    for q in queries:
        t = q.copy()
        t[rng.integers(0, len(q), size=len(q) // 5)] = rng.integers(0, 20)
        targets.append(t)
    rows = [O.seq_rows(q) for q in queries]
    db = st.TargetDb.from_sequences(ctx, targets)
    qps = [st.QueryProfile.from_rows(ctx, r) for r in rows]
    cands = [np.arange(len(targets), dtype=np.uint32) for _ in queries]
    b = st.Batch(ctx, db, qps, cands, [GapCosts(11, 1)] * len(queries), [35] * len(queries))
    b.run()
    res = b.results()
    k = 0
    for qi in range(len(queries)):
        for t in targets:
            want = (O.gotoh_score(rows[qi], t, 11, 1),) + O.ung_pair(rows[qi], t, 11, 1, 35)
            r = res[k]
            got = (int(r["swg"]), int(r["ung"]), int(r["best_diag"]), int(r["gate_score"]), bool(r["pass"]))
            assert got == want, (G, R, qi, len(t))
            k += 1


def test_long_query_long_targets_use_wide_groups(ctx):
    rng = np.random.default_rng(4242)
    q = rng.integers(0, 20, size=3000, dtype=np.uint8)
    targets = [rng.integers(0, 20, size=4000, dtype=np.uint8) for _ in range(5)]
    hom = q.copy()
    hom[rng.integers(0, 3000, size=900)] = rng.integers(0, 20)
    targets.append(np.concatenate([targets[0][:500], hom]))
    _check(ctx, O.seq_rows(q), targets, GapCosts(11, 1))
