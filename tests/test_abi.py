"""The C ABI library: loads, exports every declared symbol, host-only calls
agree with the oracle, and device calls fail cleanly without a GPU."""
import os
import re

import numpy as np
import pytest

import oracle as O
from sabertooth_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b((?:st|sy)_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    names = declared_symbols()
    assert len(names) >= 25
    for n in names:
        assert hasattr(L, n), n
    # and the ctypes signatures cover the scoring ABI
    for n in _lib.EXPORTED:
        assert n in names


def test_host_inputs_match_oracle():
    import sabertooth_amd as st

    s, pb, r, lam = st.blosum62()
    s2, pb2, r2, lam2 = O.blosum62()
    assert (s == s2).all() and (pb == pb2).all() and (r == r2).all() and lam == lam2
    assert (st.aa2num() == O.aa2num()).all()
    q = O.SplitMix64(5).background(300)
    assert (st.seq_rows(q) == O.seq_rows(q)).all()
    for n in [0, 10, 2_611_087, 68_310_060, 100_000_000, 265_123_580, 10**12]:
        assert st.size_adjusted_gate(42, n) == O.size_adjusted_gate(42, n)


def test_no_gpu_fails_loudly():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    import sabertooth_amd as st

    with pytest.raises(st.StError) as e:
        st.Context(0)
    assert e.value.code == _lib.ST_ERR_DEVICE


def test_synthetic_workloads_are_deterministic():
    from sabertooth_amd import synth

    a = synth.plain_allpairs(seed=1, n_queries=4, n_targets=300)
    b = synth.plain_allpairs(seed=1, n_queries=4, n_targets=300)
    assert all((x == y).all() for x, y in zip(a.queries, b.queries))
    assert (a.arena == b.arena).all() and (a.offsets == b.offsets).all()
    tl = a.target_lengths
    assert tl.min() >= 30 and tl.max() <= 2000
    assert all(50 <= len(q) <= 500 for q in a.queries)
    # residue frequencies follow the background
    _, pb, _, _ = O.blosum62()
    f = np.bincount(a.arena, minlength=20)[:20] / len(a.arena)
    assert np.abs(f - pb[:20] / pb[:20].sum()).max() < 0.01


def test_candidates_wrapper_checks_per_query_lengths():
    # the wrapper refuses per-query lists of the wrong length before the C
    # entry point could read past them (no device is touched)
    from sabertooth_amd import GapCosts
    from sabertooth_amd.core import _seed_array, seed_hits_to_candidates_batch
    qs = [object(), object()]
    seeds = [[(0, 0, 1)], [(1, 0, 2)]]
    for gaps, gates, sl in [([GapCosts(11, 1)], [30, 30], seeds), ([GapCosts(11, 1)] * 2, [30], seeds),
                            ([GapCosts(11, 1)] * 2, [30, 30], seeds[:1])]:
        with pytest.raises(ValueError):
            seed_hits_to_candidates_batch(None, None, qs, sl, gaps, gates)
    so, sd = _seed_array(seeds)
    with pytest.raises(ValueError):
        seed_hits_to_candidates_batch(None, None, qs, None, [GapCosts(11, 1)] * 2, [30, 30],
                                      seed_arrays=(so, sd[:1]))
