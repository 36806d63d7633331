"""Sequence-query rows (host C++): composition-bias correction, the
low-complexity mask and the default row assembly, vs the oracle and the
reference's unit tests."""
import numpy as np
import pytest

import oracle as O
from oracle import host as OH
from sabertooth_amd import host as H


def _enc(s):
    return O.encode(s)


@pytest.mark.parametrize("seed", range(8))
def test_bias_and_mask_equal_oracle(seed):
    rng = np.random.default_rng(seed)
    sm = O.SplitMix64(seed + 100)
    for _ in range(6):
        n = int(rng.integers(0, 400))
        q = sm.background(n)
        # low-complexity stretches
        if n > 60:
            a = int(rng.integers(0, n - 40))
            q[a:a + 30] = rng.integers(0, 3, size=30)
        for po in (False, True):
            assert (H.local_bias_correction(q, po) == OH.local_bias_correction(q, po)).all()
        for mp in (H.MaskParams(), H.MaskParams.scoring_default(), H.MaskParams(2, 0.9, 1.0)):
            assert (H.low_complexity_mask(q, mp) == OH.low_complexity_mask(q, mp.window, mp.locut, mp.hicut)).all()
        assert (H.query_rows(q) == OH.query_rows(q)).all()
        assert (H.query_rows(q, comp_bias=False, mask=None) == O.seq_rows(q)).all()


def test_biased_regions_are_discounted():
    # src/compbias.rs:87-99
    c = H.local_bias_correction(_enc("Q" * 60))
    assert (c <= -5).all()
    d = H.local_bias_correction(_enc("MKVLLACDEFGHIKLMNPQRSTVWYACDEFGHIKLMNPQRSTVWYMKVLLAC"))
    assert abs(d.mean()) < 1.5
    assert len(H.local_bias_correction(_enc(""))) == 0 and len(H.local_bias_correction(_enc("W"))) == 1


def test_homopolymers_and_repeats_are_masked():
    # src/mask.rs:141-152
    s = _enc("MKVLLACDEFGHIK" + "Q" * 30 + "MKVLLACDEFGHIK")
    m = H.low_complexity_mask(s)
    assert m[14:44].all() and not m[:8].any()
    r = _enc("MKVLLACDEFGHIKLMNPQRSTVWY" + "GPP" * 10 + "ACDEFGHIKLMNPQRSTVWY")
    assert H.low_complexity_mask(r).sum() >= 30


def test_ordinary_protein_is_mostly_unmasked():
    # src/mask.rs:155-161
    s = _enc("MADENKLKLGSGSFGEVFLVKHKESGNIYAMKKIRLEVEEEGVPSTAIREISLLKELKHENIVRLYDVVHSDKKLTLVFEYLDKDLKKYMDSCPEF"
             "AKDPRLIKMFLYQILRGLAYCHRRKVLHRDLKPQNLLIDRRTNSLKLADFGLARAFGVPLRTYTHEVVTLWYRAPEILLG")
    assert H.low_complexity_mask(s).mean() < 0.1
    assert not H.low_complexity_mask(_enc("QQQQQ")).any()


def test_masked_columns_take_the_x_row():
    q = _enc("MKVLLACDEFGHIK" + "Q" * 30 + "MKVLLACDEFGHIK")
    rows = H.query_rows(q)
    mk = H.low_complexity_mask(q, H.MaskParams.scoring_default())
    x = np.asarray(O.blosum62()[0], dtype=np.int32).reshape(21, 21)[20]
    assert mk.any() and (rows[mk] == x).all()
    bias = H.local_bias_correction(q)
    assert (rows[~mk] == O.seq_rows(q)[~mk] + bias[~mk][:, None]).all()
