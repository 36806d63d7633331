"""Hit statistics (host C++) vs the oracle, the reference's unit tests and
the E-value / bit-score columns of the reference's .m8 fixture."""
import math
import os

import numpy as np
import pytest

import oracle as O
from oracle import host as OH
from sabertooth_amd import host as H

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _fasta(path):
    seqs, name = {}, None
    for line in open(path):
        line = line.strip()
        if line.startswith(">"):
            name = line[1:].split()[0]
            seqs[name] = ""
        elif name:
            seqs[name] += line
    return seqs


def test_length_adjustment_product_equals_oracle():
    rng = np.random.default_rng(1)
    for _ in range(3000):
        k = float(rng.uniform(0.01, 0.5))
        adl = float(rng.uniform(0.0, 10.0))
        beta = float(rng.uniform(-40, 5))
        m = float(rng.integers(1, 40000))
        n_res = float(10 ** rng.uniform(2, 10))
        n_seqs = float(max(1, int(n_res / rng.uniform(20, 5000))))
        a = H.length_adjustment(k, adl, beta, m, n_res, n_seqs)
        b = OH.length_adjustment(k, adl, beta, m, n_res, n_seqs)
        assert a == b, (k, adl, beta, m, n_res, n_seqs, a, b)


def test_length_adjustment_is_a_fixed_point():
    # src/evalue.rs:806-821
    k, adl, beta = 0.041, 7.1, -30.0
    for m, n, ns in [(150.0, 3.5e6, 14000.0), (300.0, 1e9, 3e6), (1000.0, 1e7, 3e4)]:
        ell = H.length_adjustment(k, adl, beta, m, n, ns)
        rhs = adl * (math.log(k) + math.log((m - ell) * (n - ns * ell))) + beta
        assert 0 <= ell < m and rhs >= ell - 1e-9
        up = ell + 1.0
        assert adl * (math.log(k) + math.log((m - up) * (n - ns * up))) + beta < up
    assert H.length_adjustment(k, adl, beta, 5.0, 5.0, 1.0) == 0.0


@pytest.mark.parametrize("cb", [False, True])
def test_statistics_product_equals_oracle(cb):
    rng = np.random.default_rng(7 + cb)
    for _ in range(300):
        N = int(10 ** rng.uniform(3, 10))
        ns = max(1, int(N / rng.uniform(50, 500)))
        p = H.EValueParams.sequence(11, 1, cb, N, ns)
        q = OH.EValueParams.sequence(11, 1, cb, N, ns)
        for qlen in rng.integers(1, 6000, size=5):
            qlen = int(qlen)
            sp = p.search_space(qlen)
            assert sp == q.search_space(qlen)
            for raw in rng.integers(-5, 3000, size=5):
                raw = int(raw)
                assert p.bit_score(raw) == q.bit_score(raw)
                assert p.evalue(raw, qlen) == q.evalue(raw, qlen)
            for e in (1e-10, 1e-3, 1.0, 10.0, 1e6):
                assert p.min_score_for(e, sp) == q.min_score_for(e, sp)
    with pytest.raises(Exception):
        H.EValueParams.sequence(10, 1, False, 1000, 10)


def test_evalues_are_monotone_and_bits_consistent():
    # src/evalue.rs:824-839
    p = H.EValueParams.sequence(11, 1, False, 1_000_000, 4000)
    assert p.evalue(200, 300) < p.evalue(40, 300)
    assert p.bit_score(200) > p.bit_score(40)
    space = p.search_space(300)
    e = p.evalue(55, 300)
    assert abs(e / (space * 2.0 ** (-p.bit_score(55))) - 1.0) < 1e-9
    s = p.min_score_for(1e-3, space)
    assert p.evalue_in_space(s, space) <= 1e-3 < p.evalue_in_space(s - 1, space)


def test_ungapped_lambda_and_v02_overconfidence():
    # src/evalue.rs:842-860
    lam = H.ungapped_lambda()
    assert abs(lam - 0.389) < 0.002
    legacy = H.EValueParams.gapped(lam, 0.041, 1_000_000, 1)
    gapped = H.EValueParams.sequence(11, 1, False, 1_000_000, 4000)
    assert gapped.lam < legacy.lam
    s = gapped.min_score_for(1e-3, gapped.search_space(300))
    ratio = gapped.evalue(s, 300) / legacy.evalue(s, 300)
    assert 8.0 < ratio < 12.0


def test_gumbel_ml_recovers_known_parameters():
    # src/evalue.rs:793-804: SplitMix64(1) draws, mu - ln(-ln u)/lambda
    sm = O.SplitMix64(1)
    lam, mu = 0.27, 31.0
    xs = []
    for _ in range(20000):
        u = max((sm.next_u64() >> 11) / float(1 << 53), 1e-300)
        xs.append(mu - math.log(-math.log(u)) / lam)
    f = H.fit_gumbel_ml(xs)
    assert abs(f[0] - lam) / lam < 0.03 and abs(f[1] - mu) < 0.2
    g = OH.fit_gumbel_ml(xs)
    assert f == g
    assert H.fit_gumbel_ml(xs[:9]) is None and H.fit_gumbel_ml([3.0] * 50) is None


def test_fixture_bits_and_evalues():
    # every row of the reference's .m8 fixture (24 sequences searched against
    # themselves, default sequence search: comp-bias table entry) has an
    # integer raw score whose bit score and E-value print as the row does
    seqs = _fasta(os.path.join(GOLD, "v1_fixture.fa"))
    N, ns = sum(len(s) for s in seqs.values()), len(seqs)
    p = H.EValueParams.sequence(11, 1, True, N, ns)
    q = OH.EValueParams.sequence(11, 1, True, N, ns)
    rows = [l.split("\t") for l in open(os.path.join(GOLD, "v1_fixture_k2.v03.m8")).read().strip().split("\n")[1:]]
    assert len(rows) == 88
    for r in rows:
        m = len(seqs[r[0]])
        space = p.search_space(m)
        raw = [S for S in range(1, 3000) if f"{p.bit_score(S):.1f}" == r[11]
               and OH.rust_e2(p.evalue_in_space(S, space)) == r[10]]
        assert raw, r
        assert q.search_space(m) == space
    # the no-comp-bias entry does not reproduce them
    p0 = H.EValueParams.sequence(11, 1, False, N, ns)
    r = rows[0]
    sp0 = p0.search_space(len(seqs[r[0]]))
    assert not [S for S in range(1, 3000) if f"{p0.bit_score(S):.1f}" == r[11]
                and OH.rust_e2(p0.evalue_in_space(S, sp0)) == r[10]]
