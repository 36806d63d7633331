"""Sequence statistics by simulation (calibrate_sequence,
src/evalue.rs:537-700).

The reference's SEQUENCE_TABLE (src/evalue.rs:275-301) is its own output of
this procedure: `sabertooth calibrate --table --samples 20000 --seed 24301`
over the default lengths, printed to 4 decimals (lambda, K) and 1 decimal
(alpha/lambda, beta) (src/main.rs:2279).  Reproducing those printed values
pins the random pairs (SplitMix64 streams, background sampler), the exact
scores, the shared-lambda likelihood, the per-cell locations and the
length-model grid against reference output: the oracle does it on the CPU
(AVX2 cascade scores), the product with GPU scores."""
import json
import math
import os

import numpy as np
import pytest

import oracle as O
from oracle import host as OH
from sabertooth_amd import GapCosts
from sabertooth_amd import host as H

# src/evalue.rs:277-299: (comp_bias, lambda, K, alpha/lambda, beta)
TABLE = [(False, 0.3666, 0.1177, 2.7, -13.0), (True, 0.3649, 0.1076, 2.4, -12.0)]
# the same calibration run's JSON report (bench/results/calibration_table.json
# of the reference: every float printed {:.6}, src/metrics.rs:85-90), with
# each cell's own maximum-likelihood Gumbel (fit_gumbel_ml) and its
# location under the shared lambda
with open(os.path.join(os.path.dirname(__file__), "golden", "calibration_table.json")) as _f:
    REPORT = json.load(_f)


def _six(x):
    return float("%.6f" % x)


def _check_report(comp_bias, p, cells, rms):
    """p = (lambda, K, alpha/lambda, beta); cells in (m, n) order, each
    (m, n, samples, mean, sd, lambda_ml, mu_ml, mu)."""
    want = [c for c in REPORT["configs"] if c["comp_bias_corr"] == comp_bias][0]
    assert REPORT["samples_per_cell"] == H.TABLE_SAMPLES and REPORT["seed"] == H.TABLE_SEED
    assert (_six(p[0]), _six(p[1]), _six(p[2]), _six(p[3]), _six(rms)) == \
        (want["lambda"], want["k"], want["alpha_d_lambda"], want["beta"], want["ln_k_rms"])
    assert len(cells) == len(want["cells"])
    for c, w in zip(cells, want["cells"]):
        assert (int(c[0]), int(c[1]), int(c[2])) == (w["m"], w["n"], w["samples"])
        got = tuple(_six(float(v)) for v in c[3:8])
        assert got == (w["mean"], w["sd"], w["lambda_ml"], w["mu_ml"], w["mu_pooled"]), (w["m"], got)


def _printed(p):
    """The table's printed form ({:.4} / {:.1}, src/main.rs:2279)."""
    lam, k, adl, beta = p
    return (float("%.4f" % lam), float("%.4f" % k), float("%.1f" % adl), float("%.1f" % beta))


def _oracle_cells(comp_bias, lengths, samples, seed, threads=8):
    sc21 = np.asarray(O.blosum62()[0], dtype=np.int32).reshape(21, 21)
    cells = []
    for ci, l in enumerate(lengths):
        qs, ts = OH.simulate_pairs(seed, ci, l, l, samples)
        # BLOSUM62 rows (seq_rows) of every query at once, plus its bias
        allrows = sc21[qs]
        if comp_bias:
            allrows = allrows + OH.local_bias_correction_rows(qs)[:, :, None]
        rows = list(allrows)
        arena = ts.reshape(-1)
        offs = np.arange(samples + 1, dtype=np.uint64) * l
        sc, _, _ = O.avx2_score_pairs(rows, np.arange(samples), np.arange(samples), arena, offs, 11, 1,
                                      threads=threads)
        cells.append((l, l, sc.astype(np.float64)))
    return cells


def _gumbel_cells(seed):
    rng = np.random.default_rng(seed)
    cells = []
    for m, n in [(480, 480), (60, 60), (220, 220), (100, 130), (1000, 1000), (320, 320)]:
        lam = 0.37
        mu = (math.log(0.12 * m * n)) / lam
        u = rng.random(int(rng.integers(300, 900)))
        cells.append((m, n, np.floor(mu - np.log(-np.log(u)) / lam)))
    return cells


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_fit_calibration_product_equals_oracle(seed):
    cells = _gumbel_cells(seed)
    got = H.fit_calibration(cells)
    want = OH.fit_calibration(cells)
    p, gc, grms = got
    assert (p["lambda"], p["k"], p["length"][0], p["length"][1]) == want[0]
    assert grms == want[2]
    for k in range(len(cells)):
        c = gc[k]
        assert (c["m"], c["n"], c["samples"], c["mean"], c["sd"], c["lambda_ml"], c["mu_ml"], c["mu"]) == \
            tuple(want[1][k]), k


def test_fit_calibration_pools_lambda_over_long_cells_only():
    # the shared lambda comes from cells with min(m, n) >= 200; with none
    # that long every cell pools (src/evalue.rs:579-582)
    cells = _gumbel_cells(4)
    short = [c for c in cells if min(c[0], c[1]) < 200]
    longc = [c for c in cells if min(c[0], c[1]) >= 200]
    full = H.fit_calibration(cells)[0]["lambda"]
    only_long = H.fit_calibration(longc)[0]["lambda"]
    assert full == only_long
    assert H.fit_calibration(short)[0]["lambda"] == OH.fit_calibration(short)[0][0]
    # the per-cell locations under the shared lambda are ML locations + 1/2
    _, gc, _ = H.fit_calibration(cells)
    for k, c in enumerate(sorted(cells, key=lambda c: (c[0], c[1]))):
        assert gc[k]["mu"] == OH.gumbel_mu_given_lambda([float(x) for x in c[2]], full) + 0.5


def test_fit_calibration_refuses_like_the_reference():
    # a cell whose free Gumbel fit fails (constant scores) makes the whole
    # calibration None
    cells = _gumbel_cells(5)
    cells[2] = (cells[2][0], cells[2][1], np.full(50, 7.0))
    assert H.fit_calibration(cells) is None
    assert OH.fit_calibration(cells) is None


def test_oracle_rows_are_seq_rows():
    qs, _ = OH.simulate_pairs(7, 0, 90, 1, 20)
    sc21 = np.asarray(O.blosum62()[0], dtype=np.int32).reshape(21, 21)
    for q in qs:
        assert (sc21[q] == O.seq_rows(q)).all()
    b = OH.local_bias_correction_rows(qs)
    for i, q in enumerate(qs):
        assert (b[i] == OH.local_bias_correction(q)).all()


@pytest.mark.parametrize("entry", TABLE, ids=["plain", "comp_bias"])
def test_sequence_table_reproduced_by_the_oracle(entry):
    # gaps 11/1, the table's exact settings, scored by the oracle's AVX2
    # cascade on the CPU
    cells = _oracle_cells(entry[0], H.DEFAULT_CALIBRATION_LENGTHS, H.TABLE_SAMPLES, H.TABLE_SEED)
    p, oc, rms = OH.fit_calibration(cells)
    assert _printed(p) == entry[1:]
    _check_report(entry[0], p, [tuple(c) for c in oc], rms)


# ---- the product path: GPU scores, host fit ------------------------------------

@pytest.mark.gpu
@pytest.mark.parametrize("comp_bias", [False, True])
def test_simulate_scores_equal_oracle(ctx, comp_bias):
    # the reference's reproducibility test's cell (src/evalue.rs:858-867):
    # m = n = 120, 300 samples, seed 9, cell 3; every score against the
    # oracle's scalar Gotoh, and two calls agree
    got = H.simulate_scores(ctx, GapCosts(11, 1), comp_bias, 120, 120, 300, 9, 3)
    qs, ts = OH.simulate_pairs(9, 3, 120, 120, 300)
    bias = OH.local_bias_correction_rows(qs) if comp_bias else None
    for i in range(300):
        rows = O.seq_rows(qs[i])
        if comp_bias:
            rows = rows + bias[i][:, None]
        assert got[i] == O.gotoh_score(rows, ts[i], 11, 1), i
    assert (H.simulate_scores(ctx, GapCosts(11, 1), comp_bias, 120, 120, 300, 9, 3) == got).all()


@pytest.mark.gpu
def test_simulate_scores_uneven_lengths_and_gaps(ctx):
    # m != n, other gap costs (negative ones go to the unfloored int32
    # kernel), an empty side
    for (m, n, o, e) in [(37, 410, 11, 1), (300, 64, 5, 2), (90, 90, -1, 1), (0, 50, 11, 1), (50, 0, 11, 1)]:
        got = H.simulate_scores(ctx, GapCosts(o, e), False, m, n, 64, 77, 5)
        qs, ts = OH.simulate_pairs(77, 5, m, n, 64)
        for i in range(64):
            want = O.gotoh_score(O.seq_rows(qs[i]), ts[i], o, e) if m and n else 0
            assert got[i] == want, (m, n, o, e, i)


@pytest.mark.gpu
def test_simulate_scores_longest_cell_equals_avx2(ctx):
    # the table's longest cell (1000 x 1000), 3000 samples with bias, every
    # score against the AVX2 cascade
    got = H.simulate_scores(ctx, GapCosts(11, 1), True, 1000, 1000, 3000, H.TABLE_SEED, 7)
    qs, ts = OH.simulate_pairs(H.TABLE_SEED, 7, 1000, 1000, 3000)
    sc21 = np.asarray(O.blosum62()[0], dtype=np.int32).reshape(21, 21)
    rows = list(sc21[qs] + OH.local_bias_correction_rows(qs)[:, :, None])
    offs = np.arange(3001, dtype=np.uint64) * 1000
    sc, _, _ = O.avx2_score_pairs(rows, np.arange(3000), np.arange(3000), ts.reshape(-1), offs, 11, 1, threads=8)
    assert (got == sc).all()


@pytest.mark.gpu
@pytest.mark.parametrize("entry", TABLE, ids=["plain", "comp_bias"])
def test_sequence_table_reproduced_on_device(ctx, entry):
    # the product's calibrate_sequence at the table's provenance settings
    p, cells, rms = H.calibrate_sequence(ctx, GapCosts(11, 1), entry[0], H.DEFAULT_CALIBRATION_LENGTHS,
                                         H.TABLE_SAMPLES, H.TABLE_SEED)
    pt = (p["lambda"], p["k"], p["length"][0], p["length"][1])
    assert _printed(pt) == entry[1:]
    _check_report(entry[0], pt, [tuple(c) for c in cells.tolist()], rms)


@pytest.mark.gpu
@pytest.mark.parametrize("comp_bias", [False, True])
def test_table_matches_fresh_calibration(ctx, comp_bias):
    # src/evalue.rs:870-891: an independent seed and fewer samples; lambda
    # within 6 %, E at the 1e-3 threshold within a factor 2
    p, _, _ = H.calibrate_sequence(ctx, GapCosts(11, 1), comp_bias, [100, 220, 480], 1500, 0xD1FF5EED)
    t = H.EValueParams.sequence(11, 1, comp_bias, 1, 1)
    assert abs(p["lambda"] - t.lam) / t.lam < 0.06
    for qlen, n_res, n_seqs in [(150, 3_500_000, 14_000), (400, 50_000_000, 150_000)]:
        a = H.EValueParams.sequence(11, 1, comp_bias, n_res, n_seqs)
        b = H.EValueParams.gapped(p["lambda"], p["k"], n_res, n_seqs, length=p["length"])
        s = a.min_score_for(1e-3, a.search_space(qlen))
        r = a.evalue(s, qlen) / b.evalue(s, qlen)
        assert 0.5 < r < 2.0, (comp_bias, qlen, r)
