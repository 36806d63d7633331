"""CPU oracle of the scoring path's host-side rows — TEST INFRASTRUCTURE ONLY.

The MSA parser is a pure-Python restatement (text handling only); the
floating-point functions are C (``host_oracle.c``, built into liboracle.so
with -ffp-contract=off) so their f32/f64 operations and libm calls match the
reference's.  Only ``tests/`` and ``bench.py``'s cpu_baseline leg may import
this module.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _p, aa2num, lib

GAP = 0xFF
_WS = " \t\n\v\f\r"
_SIGS = False


def _L():
    global _SIGS
    L = lib()
    if not _SIGS:
        P = C.c_void_p
        L.or_pssm_from_msa.restype = C.c_int
        L.or_pssm_from_msa.argtypes = [P, C.c_size_t, C.c_size_t, C.c_size_t, C.c_float, C.c_float, C.c_float,
                                       P, P, P, P]
        _SIGS = True
    return L


def parse_msa(data: str):
    """parse_msa (src/profile.rs:372-449): (rows as a uint8 matrix, query_len).
    Raises ValueError on the reference's errors."""
    a2n = aa2num()
    raw, cur, started = [], [], False
    for line in data.split("\n"):
        line = line.rstrip(_WS)
        if line.startswith(">"):
            if started:
                raw.append("".join(cur))
                cur = []
            started = True
        else:
            cur.append(line.strip(_WS))
    if started:
        raw.append("".join(cur))
    if not raw:
        raise ValueError("MSA contains no sequences")
    is_a3m = any(any("a" <= ch <= "z" for ch in s) for s in raw)
    msa = []
    if is_a3m:
        for s in raw:
            row = []
            for ch in s:
                if "a" <= ch <= "z":
                    continue
                if ch in "-.":
                    row.append(GAP)
                elif "A" <= ch <= "Z":
                    row.append(int(a2n[ord(ch)]))
            msa.append(row)
    else:
        width = max(len(s) for s in raw)
        first = raw[0]
        cols = [i for i in range(width) if i < len(first) and first[i] not in "-."]
        for s in raw:
            row = []
            for i in cols:
                ch = s[i] if i < len(s) else "-"
                row.append(GAP if ch in "-." else int(a2n[ord(ch.upper()) & 0xFF]))
            msa.append(row)
    qlen = len(msa[0])
    for k, row in enumerate(msa):
        if len(row) != qlen:
            raise ValueError(f"MSA row {k + 1} has {len(row)} match columns, expected {qlen}")
    return np.array(msa, dtype=np.uint8).reshape(len(msa), qlen), qlen


def pssm_from_msa(msa: np.ndarray, qlen: int | None = None, pca=1.0, pcb=1.5, score_bias=0.0):
    """Profile::from_numeric_msa (src/profile.rs:90-182):
    (pssm int8 [L][20], prob f32 [L][20], neff f32 [L], consensus u8 [L])."""
    msa = np.ascontiguousarray(msa, dtype=np.uint8)
    n, stride = msa.shape
    L = stride if qlen is None else qlen
    pssm = np.zeros((max(L, 1), 20), np.int8)
    prob = np.zeros((max(L, 1), 20), np.float32)
    neff = np.zeros(max(L, 1), np.float32)
    cons = np.zeros(max(L, 1), np.uint8)
    rc = _L().or_pssm_from_msa(_p(msa), n, stride, L, pca, pcb, score_bias, _p(pssm), _p(prob), _p(neff), _p(cons))
    if rc != 0:
        raise ValueError("cannot build a profile from an empty MSA or with zero match columns")
    return pssm[:L], prob[:L], neff[:L], cons[:L]


def read_pssm_table(text: str):
    """Parse a PSSM text table (Profile::pssm_table layout,
    src/profile.rs:197-215): (int pssm [L][20], neff strings [L], query letters)."""
    lines = [l for l in text.split("\n") if l.strip()]
    header = lines[0].split()
    assert header[:2] == ["Pos", "Res"] and header[-1] == "Neff"
    rows, neff, res = [], [], []
    for l in lines[1:]:
        tok = l.split()
        res.append(tok[1])
        rows.append([int(x) for x in tok[2:22]])
        neff.append(tok[22])
    return np.array(rows, dtype=np.int32), neff, "".join(res)
