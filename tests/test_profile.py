"""PSSM builder (host C++) vs the oracle and the reference's demo fixture."""
import os

import numpy as np
import pytest

from oracle import host as OH
from sabertooth_amd import StError
from sabertooth_amd.host import Profile, PseudoCountParams, parse_msa

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NUM2AA = "ACDEFGHIKLMNPQRSTVWYX"


def _gold():
    a3m = open(os.path.join(GOLD, "family.a3m")).read()
    pssm, neff, res = OH.read_pssm_table(open(os.path.join(GOLD, "family.pssm")).read())
    return a3m, pssm, neff, res


def test_oracle_reproduces_demo_pssm():
    # demo/family.pssm is the reference's PSSM of demo/family.a3m
    a3m, want, wneff, res = _gold()
    msa, L = OH.parse_msa(a3m)
    pssm, _, neff, _ = OH.pssm_from_msa(msa, L)
    assert L == 131 and msa.shape == (4, 131)
    assert (pssm.astype(np.int32) == want).all()
    assert [f"{float(x):.2f}" for x in neff] == wneff
    assert "".join(NUM2AA[c] for c in msa[0]) == res


def test_oracle_fixture_catches_a_wrong_pseudocount():
    # a plausible slip (pcb and pca swapped) must fail the fixture
    a3m, want, _, _ = _gold()
    msa, L = OH.parse_msa(a3m)
    pssm, _, _, _ = OH.pssm_from_msa(msa, L, pca=1.5, pcb=1.0)
    assert (pssm.astype(np.int32) != want).any()


def test_product_reproduces_demo_pssm():
    a3m, want, wneff, res = _gold()
    p = Profile.from_msa_str(a3m)
    assert p.query_len == 131 and p.set_size == 4
    assert (p.pssm.astype(np.int32) == want).all()
    assert [f"{float(x):.2f}" for x in p.neff] == wneff
    rows = p.rows()
    assert rows.shape == (131, 21) and (rows[:, :20] == want).all() and (rows[:, 20] == 0).all()


def _random_msa_text(rng, nseq, L, a3m):
    aa = "ACDEFGHIKLMNPQRSTVWY"
    root = "".join(aa[i] for i in rng.integers(0, 20, size=L))
    out = []
    for k in range(nseq):
        s = list(root)
        for i in range(L):
            u = rng.random()
            if k and u < 0.3:
                s[i] = aa[rng.integers(0, 20)]
            elif k and u < 0.4:
                s[i] = "-"
            elif k and u < 0.42:
                s[i] = "X"
        if a3m and k:
            for _ in range(int(rng.integers(0, 4))):
                p = int(rng.integers(0, len(s) + 1))
                s.insert(p, aa[rng.integers(0, 20)].lower())
        out.append(f">s{k}\n" + "".join(s))
    return "\n".join(out) + "\n"


@pytest.mark.parametrize("seed", range(12))
def test_product_equals_oracle_on_random_msas(seed):
    rng = np.random.default_rng(seed)
    text = _random_msa_text(rng, int(rng.integers(1, 60)), int(rng.integers(1, 400)), a3m=seed % 2 == 1)
    m1, L1 = parse_msa(text)
    m2, L2 = OH.parse_msa(text)
    assert L1 == L2 and (m1 == m2).all()
    params = PseudoCountParams(*[(1.0, 1.5, 0.0), (0.0, 1.5, 0.0), (2.5, 0.7, 0.3)][seed % 3])
    p = Profile.from_numeric_msa(m1, L1, params)
    pssm, prob, neff, cons = OH.pssm_from_msa(m2, L2, params.pca, params.pcb, params.score_bias)
    assert (p.pssm == pssm).all()
    assert (p.prob.view(np.uint32) == prob.view(np.uint32)).all()
    assert (p.neff.view(np.uint32) == neff.view(np.uint32)).all()
    assert (p.consensus == cons).all()


def test_single_sequence_profile_tracks_matrix():
    # src/profile.rs:455-468
    p = Profile.from_msa_str(">q\nMKVLA\n")
    assert p.query_len == 5 and p.set_size == 1
    M, P = NUM2AA.index("M"), NUM2AA.index("P")
    assert p.score(0, M) > p.score(0, P)


def test_conserved_column_scores_higher():
    # src/profile.rs:470-487
    p = Profile.from_msa_str(">a\nCA\n>b\nCK\n>c\nCD\n>d\nCE\n")
    assert p.score(0, NUM2AA.index("C")) > p.score(1, NUM2AA.index("A"))
    assert abs(float(p.neff[0]) - float(p.neff[1])) < 1e-4


def test_a3m_inserts_are_dropped():
    # src/profile.rs:489-502
    rows, L = parse_msa(">q\nMKVLA\n>h\nMKaVLA\n")
    assert L == 5 and rows.shape == (2, 5)


def test_errors_not_panics():
    # src/profile.rs:504-531
    with pytest.raises(StError):
        parse_msa("")
    with pytest.raises(StError):
        Profile.from_numeric_msa(np.zeros((0, 4), np.uint8))
    with pytest.raises(StError):
        parse_msa(">q\nMKVLA\n>bad\nMKVLAaG\n")
    with pytest.raises(StError):
        Profile.from_numeric_msa(np.zeros((2, 0), np.uint8))


def test_aligned_fasta_tolerates_short_rows_and_lowercase_free_text():
    # any lowercase selects a3m mode, where ragged match columns are an error
    text = ">q\nMK-VL.A\n>h\nMKIV\n>i\nmkivlaa"
    with pytest.raises(StError):
        parse_msa(text)
    with pytest.raises(ValueError):
        OH.parse_msa(text)
    text = ">q\nMK-VL.A*\n>h\nMKIV\n>i\nMKIVLAAW"
    r1, L1 = parse_msa(text)
    r2, L2 = OH.parse_msa(text)
    assert L1 == L2 == 6 and (r1 == r2).all()
