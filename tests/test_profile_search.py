"""Profile (PSSM) search after scoring, against the reference's profile-search
fixture and the mathematics of the per-profile Gumbel fit.

Fixture: `demo/hits.m8` of the reference, the profile of `demo/family.a3m`
searched against the five targets of `demo/targets.fasta`
(`sabertooth profilesearch`, src/main.rs:1523-1569; profile defaults: gaps
44/4, E <= 10, src/search.rs:129-137).  Its E-value and bits columns follow
the ungapped diagnostic statistics `EValueParams::for_profile`
(lambda = ln2/8, K = 0.041, space m*N; src/evalue.rs:386-398): the three
decoy rows are reproduced column for column with them.  The two homolog rows
carry raw scores 960 and 363 (the only integers that print their bits and
E-value under those statistics), where the current PSSM builder and SWG give
991 and 496 (the PSSM itself reproduces the reference's `demo/family.pssm`
cell for cell): the file was written by an earlier scorer, before
per-profile calibration, so it cannot pin calibrate_profile.

calibrate_profile (src/evalue.rs:714-749) is pinned piece by piece instead:
its random targets follow SplitMix64 (published vectors, test_oracle) and the
background sampler; its scores are exact SWG scores (the GPU path against the
oracle); its fit is checked here against Lawless' maximum-likelihood
equations themselves (the score equation changes sign across the returned
lambda within the bisection's tolerance, the location equation holds, and
scipy's Brent root of the same equation agrees), and K against its closed
form e^(lambda (mu + 1/2)) / (m * mean length).
"""
import math
import os

import numpy as np
import pytest

import oracle as O
from oracle import host as OH
from sabertooth_amd import GapCosts
from sabertooth_amd import host as H

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SEED = 0x5EED  # SearchParams::calibration_seed default, src/search.rs:125


def _fasta(path):
    ids, seqs = [], []
    for line in open(path):
        line = line.strip()
        if line.startswith(">"):
            ids.append(line[1:].split()[0])
            seqs.append("")
        elif ids:
            seqs[-1] += line
    return ids, [O.encode(s) for s in seqs]


def _demo():
    ids, seqs = _fasta(os.path.join(GOLD, "demo_targets.fasta"))
    prof = H.Profile.from_msa_str(open(os.path.join(GOLD, "family.a3m")).read())
    offs = np.zeros(len(seqs) + 1, np.uint64)
    offs[1:] = np.cumsum([len(s) for s in seqs])
    return ids, seqs, prof, offs


def _fixture_rows():
    lines = open(os.path.join(GOLD, "demo_hits.m8")).read().strip().split("\n")
    assert lines[0] == H.M8_HEADER
    return [l.split("\t") for l in lines[1:]]


def _m8(prof, ids, seqs, offs, scores, ev):
    hits, cig = H.search_hits(prof.rows(), prof.query_num, np.concatenate(seqs), offs,
                              np.arange(len(seqs), dtype=np.uint32), scores, GapCosts(44, 4), ev,
                              H.SearchParams(max_evalue=10.0))
    return [H.format_m8("family", ids[int(h["target"])], h).split("\t") for h in hits], hits


def test_demo_fixture_decoy_rows_and_statistics():
    ids, seqs, prof, offs = _demo()
    rows = prof.rows()
    scores = np.array([O.gotoh_score(rows, t, 44, 4) for t in seqs], np.int32)
    N = int(offs[-1])
    ev = H.EValueParams.for_profile(N)
    assert ev.lam == math.log(2) / 8 and ev.k == 0.041 and ev.search_space(len(rows)) == len(rows) * N
    got, _ = _m8(prof, ids, seqs, offs, scores, ev)
    want = _fixture_rows()
    assert len(want) == 5 and sorted(r[1] for r in got) == sorted(r[1] for r in want)
    by = {r[1]: r for r in got}
    decoys = ["membrane_decoy", "random_decoy_2", "random_decoy_1"]
    for r in want:
        if r[1] in decoys:
            assert by[r[1]] == r, (by[r[1]], r)
    # the decoys come out in the fixture's order (E ascending), after the homologs
    assert [r[1] for r in got][2:] == decoys == [r[1] for r in want][2:]
    # the homolog rows: their raw scores under the same statistics
    sp = ev.search_space(len(rows))
    raw = {}
    for r in want:
        raw[r[1]] = [s for s in range(1, 5000)
                     if f"{ev.bit_score(s):.1f}" == r[11] and OH.rust_e2(ev.evalue_in_space(s, sp)) == r[10]]
    assert raw["true_homolog_diverged"] == [960] and raw["another_kinase_homolog"] == [363]
    assert raw["membrane_decoy"] == [111] and raw["random_decoy_2"] == [101] and raw["random_decoy_1"] == [98]
    # the current scorer: the decoys' scores, and higher homolog scores
    sc = dict(zip(ids, scores.tolist()))
    assert [sc[d] for d in decoys] == [111, 101, 98]
    assert sc["true_homolog_diverged"] == 991 and sc["another_kinase_homolog"] == 496


def _lawless_g(xs, lam):
    # Lawless' ML score equation for the Gumbel scale: 1/lam - mean + E_w[x],
    # weights e^(-lam x) (shifted by min x for range; the ratio is unchanged)
    x = np.asarray(xs, np.float64)
    w = np.exp(-lam * (x - x.min()))
    return 1.0 / lam - x.mean() + float((x * w).sum() / w.sum())


def _check_fit(xs, lam, mu):
    x = np.asarray(xs, np.float64)
    # the score equation changes sign across lam within the bisection's
    # relative tolerance (1e-12 of hi; a margin of a few ulps of g)
    eps = 4e-12
    assert _lawless_g(x, lam * (1 - eps)) > 0 > _lawless_g(x, lam * (1 + eps))
    # an independent root finder on the same equation
    from scipy.optimize import brentq
    lam0 = math.pi / (x.std(ddof=1) * math.sqrt(6.0))
    root = brentq(lambda l: _lawless_g(x, l), lam0 / 10, lam0 * 10, xtol=1e-15, rtol=1e-14)
    assert abs(root - lam) <= 2e-12 * lam
    # the location equation: e^(-lam mu) = mean(e^(-lam x))
    xm = x.min()
    mu_eq = xm - math.log(float(np.exp(-lam * (x - xm)).mean())) / lam
    assert abs(mu - mu_eq) <= 1e-12 * max(1.0, abs(mu))


def test_gumbel_fit_solves_the_ml_equations_on_profile_scores():
    # the demo profile's 2000 calibration scores (oracle SWG, the reference's
    # random targets and lengths = the database's length sample)
    ids, seqs, prof, offs = _demo()
    rows = prof.rows()
    lengths = [len(s) for s in seqs]
    ts = OH.calibration_targets(lengths, 2000, SEED)
    xs = [float(O.gotoh_score(rows, t, 44, 4)) for t in ts]
    f = H.fit_gumbel_ml(xs)
    assert f is not None
    lam, mu = f
    _check_fit(xs, lam, mu)
    assert 0.05 < lam < 0.12 and 40 < mu < 120
    # K's closed form over the search space m x N (src/evalue.rs:740-745)
    cal = OH.calibrate_profile(rows, 44, 4, lengths, 2000, SEED)
    assert cal[2] == (lam, mu)
    assert cal[1] == math.exp(lam * (mu + 0.5)) / (len(rows) * (sum(lengths) / len(lengths)))


@pytest.mark.parametrize("lam,mu,n", [(0.27, 31.0, 5000), (0.08, 70.0, 2000), (1.3, 5.0, 300)])
def test_gumbel_fit_solves_the_ml_equations_on_gumbel_draws(lam, mu, n):
    sm = O.SplitMix64(int(lam * 1000) + n)
    xs = []
    for _ in range(n):
        u = max((sm.next_u64() >> 11) / float(1 << 53), 1e-300)
        xs.append(round(mu - math.log(-math.log(u)) / lam))  # integer scores, as SWG gives
    f = H.fit_gumbel_ml(xs)
    assert f is not None
    _check_fit(xs, *f)


def test_gumbel_fit_refuses_like_the_reference():
    # fewer than 10 samples, zero variance, NaN: no fit (src/evalue.rs:139-150)
    assert H.fit_gumbel_ml([1.0] * 9) is None
    assert H.fit_gumbel_ml([7.0] * 40) is None
    assert H.fit_gumbel_ml([1.0, 2.0] * 10 + [float("nan")]) is None


@pytest.mark.gpu
def test_demo_profile_search_on_device(ctx):
    # the device scores, the GPU-scored calibration and the resulting .m8
    # rows equal the oracle's; with calibrated statistics the decoy rows
    # keep every column but the E-value and bits
    import sabertooth_amd as st
    ids, seqs, prof, offs = _demo()
    rows = prof.rows()
    db = st.TargetDb(ctx, np.concatenate(seqs), offs)
    qp = st.QueryProfile.from_rows(ctx, rows)
    sc, _ = qp.sw_score_with(db, np.arange(len(seqs), dtype=np.uint32), GapCosts(44, 4))
    want = np.array([O.gotoh_score(rows, t, 44, 4) for t in seqs], np.int32)
    assert (sc == want).all()
    lengths = [len(s) for s in seqs]
    lam, k, fit = H.calibrate_profile(qp, GapCosts(44, 4), lengths, 2000, SEED)
    assert (lam, k, fit) == OH.calibrate_profile(rows, 44, 4, lengths, 2000, SEED)
    ev = H.EValueParams.gapped(lam, k, int(offs[-1]), len(seqs))
    got, hits = _m8(prof, ids, seqs, offs, sc, ev)
    ref, _ = _m8(prof, ids, seqs, offs, want, ev)
    assert got == ref and len(got) == 5
    fx = {r[1]: r for r in _fixture_rows()}
    for r in got:
        if r[1] in ("membrane_decoy", "random_decoy_2", "random_decoy_1"):
            assert r[:10] == fx[r[1]][:10]
    for h in hits:
        assert h["evalue"] == k * len(rows) * int(offs[-1]) * math.exp(-lam * int(h["raw_score"]))
    qp.close()
    db.close()
