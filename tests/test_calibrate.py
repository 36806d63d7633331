"""Per-profile statistics: the random targets follow the reference's
streams exactly (CPU), and the GPU-scored calibration equals the oracle's
(GPU)."""
import math

import numpy as np
import pytest

import oracle as O
from oracle import host as OH
from sabertooth_amd import GapCosts
from sabertooth_amd import host as H


def test_stream_seed_matches_splitmix():
    # stream_seed(seed, c, i) = SplitMix64(seed ^ c*A ^ i*B).next_u64()
    assert OH.stream_seed(0, 0, 0) == O.SplitMix64(0).next_u64()


@pytest.mark.parametrize("seed", [0x5EED, 5, 0xD1FF5EED])
def test_calibration_targets_equal_oracle(seed):
    rng = np.random.default_rng(seed & 0xFFFF)
    lengths = rng.integers(0, 700, size=int(rng.integers(1, 60)))
    arena, off = H.calibration_targets(lengths, 300, seed)
    want = OH.calibration_targets(lengths, 300, seed)
    assert len(off) == 301
    for i, t in enumerate(want):
        assert (arena[int(off[i]):int(off[i + 1])] == t).all(), i


def _profile(msa):
    from sabertooth_amd.host import Profile
    return Profile.from_msa_str(msa)


@pytest.mark.gpu
def test_calibrate_profile_equals_oracle(ctx):
    import sabertooth_amd as st
    p = _profile(">a\nMKVLLACDEFGHIKLMNPQRSTVWYACDEFGHIK\n>b\nMKILLSCDEFGHLKLMNPQRSTVWFACDEYGHIK\n")
    lengths = [100 + 7 * i for i in range(50)]
    qp = st.QueryProfile.from_rows(ctx, p.rows())
    got = H.calibrate_profile(qp, GapCosts(44, 4), lengths, samples=600, seed=5)
    want = OH.calibrate_profile(p.rows(), 44, 4, lengths, 600, 5)
    assert got == want


@pytest.mark.gpu
def test_profile_calibration_is_per_profile(ctx):
    # src/evalue.rs:893-912
    import sabertooth_amd as st
    pa = _profile(">a\nMKVLLACDEFGHIKLMNPQRSTVWYACDEFGHIK\n>b\nMKILLSCDEFGHLKLMNPQRSTVWFACDEYGHIK\n")
    pb = _profile(">a\nWWWWCCCCWWWWCCCCWWWWCCCCWWWWCCCCWW\n")
    lens = [100 + 7 * i for i in range(50)]
    ga = H.calibrate_profile(st.QueryProfile.from_rows(ctx, pa.rows()), GapCosts(44, 4), lens, 1500, 5)
    gb = H.calibrate_profile(st.QueryProfile.from_rows(ctx, pb.rows()), GapCosts(44, 4), lens, 1500, 5)
    assert ga[0] < math.log(2) / 8.0 * 1.05
    ea = H.EValueParams.gapped(ga[0], ga[1], 1_000_000, 4000).evalue(300, pa.query_len)
    eb = H.EValueParams.gapped(gb[0], gb[1], 1_000_000, 4000).evalue(300, pb.query_len)
    assert abs(math.log(ea / eb)) > 0.5
