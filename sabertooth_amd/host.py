"""Host-side rows of the scoring path (no GPU): the PSSM builder.

Thin wrappers over ``include/sabertooth_host.h``; names and argument
meanings follow the reference (src/profile.rs).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from ._lib import check, lib, ptr

GAP = 0xFF
_P = C.c_void_p
_SIGS = [
    ("st_parse_msa", C.c_int, [_P, C.c_uint64, _P, C.c_uint64, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    ("st_pssm_from_msa", C.c_int, [_P, C.c_uint64, C.c_uint64, C.c_uint64, C.c_float, C.c_float, C.c_float,
                                   _P, _P, _P, _P]),
    ("st_pssm_rows", C.c_int, [_P, C.c_uint64, _P]),
]
EXPORTED = [s[0] for s in _SIGS]
_DONE = False


def _L():
    global _DONE
    L = lib()
    if not _DONE:
        for name, res, args in _SIGS:
            f = getattr(L, name)
            f.restype, f.argtypes = res, args
        _DONE = True
    return L


@dataclass
class PseudoCountParams:
    """src/profile.rs:47-66 (MMseqs2 defaults)."""
    pca: float = 1.0
    pcb: float = 1.5
    score_bias: float = 0.0


@dataclass
class Profile:
    """Profile (src/profile.rs:31-45): int8 PSSM [L][20], f32 probabilities,
    Neff per column, consensus and the query row of the MSA."""
    query_len: int
    set_size: int
    pssm: np.ndarray
    prob: np.ndarray
    neff: np.ndarray
    consensus: np.ndarray
    query_num: np.ndarray

    @staticmethod
    def from_numeric_msa(msa: np.ndarray, query_len: int | None = None,
                         params: PseudoCountParams = PseudoCountParams()) -> "Profile":
        msa = np.ascontiguousarray(msa, dtype=np.uint8)
        if msa.ndim != 2:
            raise ValueError("msa must be a 2-D array of residue codes")
        n, stride = msa.shape
        L = stride if query_len is None else int(query_len)
        pssm = np.zeros((max(L, 1), 20), np.int8)
        prob = np.zeros((max(L, 1), 20), np.float32)
        neff = np.zeros(max(L, 1), np.float32)
        cons = np.zeros(max(L, 1), np.uint8)
        check(_L().st_pssm_from_msa(ptr(msa), n, stride, L, params.pca, params.pcb, params.score_bias,
                                    ptr(pssm), ptr(prob), ptr(neff), ptr(cons)))
        return Profile(L, n, pssm[:L], prob[:L], neff[:L], cons[:L], msa[0, :L].copy())

    @staticmethod
    def from_msa_str(text: str, params: PseudoCountParams = PseudoCountParams()) -> "Profile":
        msa, L = parse_msa(text)
        return Profile.from_numeric_msa(msa, L, params)

    def score(self, pos: int, target_residue: int) -> int:
        return int(self.pssm[pos, target_residue]) if target_residue < 20 else 0

    def rows(self) -> np.ndarray:
        """Score rows int32 [L][21] of the profile query (X column 0)."""
        out = np.zeros((max(self.query_len, 1), 21), np.int32)
        p = np.ascontiguousarray(self.pssm, dtype=np.int8)
        check(_L().st_pssm_rows(ptr(p), self.query_len, ptr(out)))
        return out[: self.query_len]


def parse_msa(text: str):
    """parse_msa (src/profile.rs:372-449): (codes uint8 [rows][L], L)."""
    b = text.encode()
    buf = C.create_string_buffer(b, len(b) + 1)
    n, L = C.c_uint64(0), C.c_uint64(0)
    check(_L().st_parse_msa(buf, len(b), None, 0, C.byref(n), C.byref(L)))
    out = np.zeros(max(n.value * L.value, 1), np.uint8)
    check(_L().st_parse_msa(buf, len(b), ptr(out), out.size, C.byref(n), C.byref(L)))
    return out[: n.value * L.value].reshape(n.value, L.value), int(L.value)
