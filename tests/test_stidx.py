"""Reading a Sabertooth index file's target sections (src/index.rs) against
the reference's own index fixture: tests/data/v1_fixture_k2.stidx, built by
the reference from tests/data/v1_fixture.fa.  The header and the four
section checksums in the file pin the restated Hasher64; corrupted copies
must be refused as SeedIndex::open refuses them."""
import os
import shutil

import numpy as np
import pytest

import oracle as O
from sabertooth_amd import host as H

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
IDX = os.path.join(GOLD, "v1_fixture_k2.stidx")


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


def test_fixture_header_and_checksums():
    info = H.stidx_info(IDX, verify=True)
    assert (int(info["version"]), int(info["alphabet"]), int(info["num_seqs"]), int(info["total_residues"])) == \
        (1, 1, 24, 1748)


def test_fixture_targets_and_ids_equal_the_fasta():
    ids, seqs = _fasta(os.path.join(GOLD, "v1_fixture.fa"))
    res, off = H.stidx_targets(IDX)
    assert len(off) == len(seqs) + 1 and off[0] == 0
    for k, s in enumerate(seqs):
        assert (res[int(off[k]):int(off[k + 1])] == s).all(), k
    assert H.stidx_ids(IDX) == ids
    assert int(H.stidx_info(IDX)["max_seq_len"]) == max(len(s) for s in seqs)


def _corrupt(tmp_path, name, edit):
    p = str(tmp_path / name)
    shutil.copyfile(IDX, p)
    os.chmod(p, 0o644)
    b = bytearray(open(p, "rb").read())
    b = edit(b)
    open(p, "wb").write(bytes(b))
    return p


def _flip(off):
    def f(b):
        b[off] ^= 0x01
        return b
    return f


def test_corrupt_copies_are_refused(tmp_path):
    res_off = 4096  # the residue section follows the header
    cases = {
        "magic": _flip(0),
        "header": _flip(200),              # a header field: checksum mismatch
        "checksum": _flip(4095),           # the stored checksum itself
        "footer": _flip(29976 - 12),       # footer magic
        "length": _flip(29976 - 8),        # footer's file length
    }
    for name, edit in cases.items():
        p = _corrupt(tmp_path, name, edit)
        with pytest.raises(Exception):
            H.stidx_info(p)
    # truncated: no footer, or smaller than a header
    p = _corrupt(tmp_path, "trunc", lambda b: b[:-100])
    with pytest.raises(Exception):
        H.stidx_info(p)
    p = _corrupt(tmp_path, "tiny", lambda b: b[:1000])
    with pytest.raises(Exception):
        H.stidx_info(p)
    # a residue byte: opens (as SeedIndex::open does), fails verification
    p = _corrupt(tmp_path, "residue", _flip(res_off + 5))
    H.stidx_info(p)
    with pytest.raises(Exception):
        H.stidx_info(p, verify=True)
    with pytest.raises(Exception):
        H.stidx_info(os.path.join(GOLD, "v1_fixture.fa"))
    with pytest.raises(Exception):
        H.stidx_info(str(tmp_path / "missing.stidx"))



def test_fixture_postings_decode_to_the_two_mer_lists():
    # the index was built with k = 2: its postings are every 2-mer start of
    # the targets whose residues are both in the 20-letter alphabet, grouped
    # by 2-mer, each group sorted by (target, position)
    res, off = H.stidx_targets(IDX)
    post = H.stidx_postings(IDX)
    t = (post >> np.uint64(32)).astype(np.int64)
    p = (post & np.uint64(0xFFFFFFFF)).astype(np.int64)
    L = np.diff(off.astype(np.int64))
    assert (p + 2 <= L[t]).all()
    want = {(a, b) for a in range(len(L)) for b in range(int(L[a]) - 1)
            if res[int(off[a]) + b] < 20 and res[int(off[a]) + b + 1] < 20}
    assert {(int(a), int(b)) for a, b in zip(t, p)} == want and len(post) == len(want)
    km = [int(res[int(off[a]) + b]) * 21 + int(res[int(off[a]) + b + 1]) for a, b in zip(t, p)]
    groups = [km[0]] + [km[i] for i in range(1, len(km)) if km[i] != km[i - 1]]
    assert len(groups) == len(set(groups))
    assert all(km[i] != km[i - 1] or (t[i], p[i]) > (t[i - 1], p[i - 1]) for i in range(1, len(km)))
