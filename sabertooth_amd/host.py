"""Host-side rows of the scoring path (no GPU): the PSSM builder.

Thin wrappers over ``include/sabertooth_host.h``; names and argument
meanings follow the reference (src/profile.rs).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

from ._lib import SPAN_DT, check, lib, ptr

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


# ---- hit statistics (src/evalue.rs) ------------------------------------------

GAPPED_PARAMS_DT = np.dtype([("lambda", "<f8"), ("k", "<f8"), ("has_length", "<i4"), ("reserved", "<i4"),
                             ("alpha_d_lambda", "<f8"), ("beta", "<f8")])
EVALUE_PARAMS_DT = np.dtype([("lambda", "<f8"), ("k", "<f8"), ("log_k", "<f8"), ("db_residues", "<f8"),
                             ("db_seqs", "<f8"), ("has_length", "<i4"), ("reserved", "<i4"),
                             ("alpha_d_lambda", "<f8"), ("beta", "<f8")])
GUMBEL_DT = np.dtype([("lambda", "<f8"), ("mu", "<f8"), ("n", "<u8")])
_D = C.c_double
_SIGS += [
    ("st_evalue_table", C.c_int, [C.c_int32, C.c_int32, C.c_int, _P]),
    ("st_evalue_params_from", C.c_int, [_P, C.c_uint64, C.c_uint64, _P]),
    ("st_evalue_params_profile", C.c_int, [C.c_uint64, _P]),
    ("st_length_adjustment", _D, [_D, _D, _D, _D, _D, _D]),
    ("st_search_space", _D, [_P, C.c_uint64]),
    ("st_bit_score", _D, [_P, C.c_int32]),
    ("st_evalue", _D, [_P, C.c_int32, C.c_uint64]),
    ("st_evalue_in_space", _D, [_P, C.c_int32, _D]),
    ("st_min_score_for", C.c_int32, [_P, _D, _D]),
    ("st_gumbel_mu_given_lambda", _D, [_P, C.c_uint64, _D]),
    ("st_fit_gumbel_ml", C.c_int, [_P, C.c_uint64, _P]),
    ("st_fit_calibration", C.c_int, [_P, _P, _P, _P, C.c_uint64, C.c_uint64, _P, _P, _P]),
    ("st_ungapped_lambda", _D, []),
]
CALIBRATION_CELL_DT = np.dtype([("m", "<u8"), ("n", "<u8"), ("samples", "<u8"), ("mean", "<f8"), ("sd", "<f8"),
                                ("lambda_ml", "<f8"), ("mu_ml", "<f8"), ("mu", "<f8")])
EXPORTED = [s[0] for s in _SIGS]


def length_adjustment(k, alpha_d_lambda, beta, m, n_res, n_seqs) -> float:
    return float(_L().st_length_adjustment(k, alpha_d_lambda, beta, m, n_res, n_seqs))


class EValueParams:
    """EValueParams (src/evalue.rs:311-461) bound to a database size."""

    def __init__(self, rec: np.ndarray):
        self._r = rec

    @staticmethod
    def sequence(open_: int, ext: int, comp_bias: bool, db_residues: int, db_seqs: int) -> "EValueParams":
        g = np.zeros(1, GAPPED_PARAMS_DT)
        check(_L().st_evalue_table(open_, ext, 1 if comp_bias else 0, ptr(g)))
        return EValueParams.from_params(g, db_residues, db_seqs)

    @staticmethod
    def from_params(g: np.ndarray, db_residues: int, db_seqs: int) -> "EValueParams":
        r = np.zeros(1, EVALUE_PARAMS_DT)
        check(_L().st_evalue_params_from(ptr(g), db_residues, db_seqs, ptr(r)))
        return EValueParams(r)

    @staticmethod
    def gapped(lam: float, k: float, db_residues: int, db_seqs: int, length=None) -> "EValueParams":
        g = np.zeros(1, GAPPED_PARAMS_DT)
        g["lambda"], g["k"] = lam, k
        if length is not None:
            g["has_length"], g["alpha_d_lambda"], g["beta"] = 1, length[0], length[1]
        return EValueParams.from_params(g, db_residues, db_seqs)

    @staticmethod
    def for_profile(db_residues: int) -> "EValueParams":
        r = np.zeros(1, EVALUE_PARAMS_DT)
        check(_L().st_evalue_params_profile(db_residues, ptr(r)))
        return EValueParams(r)

    @property
    def lam(self) -> float:
        return float(self._r["lambda"][0])

    @property
    def k(self) -> float:
        return float(self._r["k"][0])

    def search_space(self, qlen: int) -> float:
        return float(_L().st_search_space(ptr(self._r), qlen))

    def bit_score(self, raw: int) -> float:
        return float(_L().st_bit_score(ptr(self._r), raw))

    def evalue(self, raw: int, qlen: int) -> float:
        return float(_L().st_evalue(ptr(self._r), raw, qlen))

    def evalue_in_space(self, raw: int, space: float) -> float:
        return float(_L().st_evalue_in_space(ptr(self._r), raw, space))

    def min_score_for(self, max_evalue: float, space: float) -> int:
        return int(_L().st_min_score_for(ptr(self._r), max_evalue, space))


def fit_gumbel_ml(xs):
    """(lambda, mu) of the ML Gumbel fit, or None (src/evalue.rs:139-182)."""
    x = np.ascontiguousarray(xs, dtype=np.float64)
    out = np.zeros(1, GUMBEL_DT)
    rc = _L().st_fit_gumbel_ml(ptr(x), len(x), ptr(out))
    if rc != 0:
        return None
    return float(out["lambda"][0]), float(out["mu"][0])


def _params_tuple(g):
    return {"lambda": float(g["lambda"][0]), "k": float(g["k"][0]),
            "length": (float(g["alpha_d_lambda"][0]), float(g["beta"][0])) if g["has_length"][0] else None}


def fit_calibration(cells, lambda_min_len: int = 200):
    """fit_calibration (src/evalue.rs:576-668) over [(m, n, scores)]:
    ({"lambda", "k", "length": (alpha/lambda, beta)}, cells record array in
    (m, n) order, ln K rms), or None where the reference returns None."""
    cm = np.array([c[0] for c in cells], np.uint64)
    cn = np.array([c[1] for c in cells], np.uint64)
    off = np.zeros(len(cells) + 1, np.uint64)
    off[1:] = np.cumsum([len(c[2]) for c in cells])
    xs = np.ascontiguousarray(np.concatenate([np.asarray(c[2], np.float64) for c in cells]) if cells
                              else np.zeros(0), dtype=np.float64)
    g = np.zeros(1, GAPPED_PARAMS_DT)
    out = np.zeros(max(len(cells), 1), CALIBRATION_CELL_DT)
    rms = C.c_double(0.0)
    rc = _L().st_fit_calibration(ptr(cm), ptr(cn), ptr(off), ptr(xs), len(cells), lambda_min_len, ptr(g), ptr(out),
                                 C.byref(rms))
    if rc == 1:  # ST_ERR_INVALID: no fit
        return None
    check(rc)
    return _params_tuple(g), out[: len(cells)], float(rms.value)


def ungapped_lambda() -> float:
    return float(_L().st_ungapped_lambda())


# ---- sequence-query rows (src/compbias.rs, src/mask.rs, src/search.rs) -------

@dataclass(frozen=True)
class MaskParams:
    """MaskParams (src/mask.rs:27-66)."""
    window: int = 12
    locut: float = 2.2
    hicut: float = 2.5

    @staticmethod
    def scoring_default() -> "MaskParams":
        return MaskParams(15, 2.3, 2.3)


_SIGS += [
    ("st_local_bias_correction", C.c_int, [_P, C.c_uint64, C.c_int, _P]),
    ("st_low_complexity_mask", C.c_int, [_P, C.c_uint64, C.c_uint32, _D, _D, _P, C.POINTER(C.c_uint64)]),
    ("st_query_rows", C.c_int, [_P, C.c_uint64, C.c_int, C.c_int, C.c_uint32, _D, _D, _P,
                                C.POINTER(C.c_uint64)]),
]
EXPORTED = [s[0] for s in _SIGS]


def _codes(q) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(q, dtype=np.uint8).reshape(-1))


def local_bias_correction(query, penalize_only: bool = False) -> np.ndarray:
    q = _codes(query)
    out = np.zeros(max(len(q), 1), np.int32)
    check(_L().st_local_bias_correction(ptr(q), len(q), 1 if penalize_only else 0, ptr(out)))
    return out[: len(q)]


def low_complexity_mask(seq, p: MaskParams = MaskParams()) -> np.ndarray:
    q = _codes(seq)
    out = np.zeros(max(len(q), 1), np.uint8)
    n = C.c_uint64(0)
    check(_L().st_low_complexity_mask(ptr(q), len(q), p.window, p.locut, p.hicut, ptr(out), C.byref(n)))
    return out[: len(q)].astype(bool)


def query_rows(query, comp_bias: bool = True, penalize_only: bool = False,
               mask: MaskParams | None = MaskParams.scoring_default()) -> np.ndarray:
    """Rows int32 [m][21] of a sequence query as the default search builds
    them (src/search.rs:432-463)."""
    q = _codes(query)
    rows = np.zeros((max(len(q), 1), 21), np.int32)
    n = C.c_uint64(0)
    w, lo, hi = (mask.window, mask.locut, mask.hicut) if mask is not None else (0, 0.0, 0.0)
    check(_L().st_query_rows(ptr(q), len(q), 1 if comp_bias else 0, 1 if penalize_only else 0, w, lo, hi,
                             ptr(rows), C.byref(n)))
    return rows[: len(q)]


# ---- Sabertooth index files (<db>.stidx, src/index.rs) ------------------------

STIDX_INFO_DT = np.dtype([("version", "<u4"), ("alphabet", "<u4"), ("num_seqs", "<u8"),
                          ("total_residues", "<u8"), ("max_seq_len", "<u8"), ("id_bytes", "<u8")])
_SIGS += [
    ("st_stidx_info_read", C.c_int, [C.c_char_p, C.c_int, _P]),
    ("st_stidx_read_targets", C.c_int, [C.c_char_p, _P, _P]),
    ("st_stidx_read_ids", C.c_int, [C.c_char_p, _P, _P]),
    ("st_stidx_read_postings", C.c_int, [C.c_char_p, _P, C.c_uint64, C.POINTER(C.c_uint64)]),
]
EXPORTED = [s[0] for s in _SIGS]


def stidx_info(path: str, verify: bool = False):
    """Header figures of an index file after SeedIndex::open's checks
    (and, with verify, the target sections' checksums)."""
    out = np.zeros(1, STIDX_INFO_DT)
    check(_L().st_stidx_info_read(os.fsencode(path), 1 if verify else 0, ptr(out)))
    return out[0]


def stidx_targets(path: str):
    """(residues u8, offsets u64) of an index file's target sections."""
    info = stidx_info(path)
    res = np.zeros(max(int(info["total_residues"]), 1), np.uint8)
    off = np.zeros(int(info["num_seqs"]) + 1, np.uint64)
    check(_L().st_stidx_read_targets(os.fsencode(path), ptr(res), ptr(off)))
    return res[: int(info["total_residues"])], off


def stidx_postings(path: str) -> np.ndarray:
    """The index's posting section decoded to (target << 32) | position."""
    n = C.c_uint64(0)
    check(_L().st_stidx_read_postings(os.fsencode(path), None, 0, C.byref(n)))
    out = np.zeros(max(n.value, 1), np.uint64)
    check(_L().st_stidx_read_postings(os.fsencode(path), ptr(out), len(out), C.byref(n)))
    return out[: n.value]


def stidx_ids(path: str):
    """Target identifiers of an index file (SeedIndex::id)."""
    info = stidx_info(path)
    b = np.zeros(max(int(info["id_bytes"]), 1), np.uint8)
    off = np.zeros(int(info["num_seqs"]) + 1, np.uint64)
    check(_L().st_stidx_read_ids(os.fsencode(path), ptr(b), ptr(off)))
    raw = b.tobytes()
    return [raw[int(off[i]):int(off[i + 1])].decode("utf-8", "replace") for i in range(len(off) - 1)]


# ---- exact traceback of reported hits (src/align.rs) -------------------------

ALIGNMENT_DT = np.dtype([("raw_score", "<i4"), ("reserved", "<i4"), ("query_start", "<u8"), ("query_end", "<u8"),
                         ("target_start", "<u8"), ("target_end", "<u8"), ("aln_len", "<u8"),
                         ("identities", "<u8"), ("mismatches", "<u8"), ("gaps", "<u8"), ("gap_opens", "<u8")])
SCORE_END_DT = np.dtype([("score", "<i4"), ("reserved", "<i4"), ("query_end", "<i8"), ("target_end", "<i8")])
_SIGS += [
    ("st_sw_score_end_host", C.c_int, [_P, C.c_uint64, _P, C.c_uint64, C.c_int32, C.c_int32, _P,
                                       C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    ("st_align_canonical", C.c_int, [_P, _P, C.c_uint64, _P, C.c_uint64, C.c_int32, C.c_int32, C.c_uint32, _P, _P,
                                     _P, C.c_uint64, C.POINTER(C.c_uint64), C.POINTER(C.c_int)]),
    ("st_align_span", C.c_int, [_P, _P, C.c_uint64, _P, C.c_uint64, C.c_int32, C.c_int32, C.c_uint32, _P, _P,
                                _P, C.c_uint64, C.POINTER(C.c_uint64), C.POINTER(C.c_int)]),
    ("st_rescore_alignment", C.c_int, [_P, _P, C.c_uint64, _P, C.c_uint64, C.c_int32, C.c_int32, _P, _P, _P]),
]
EXPORTED = [s[0] for s in _SIGS]
DEFAULT_BAND_RADIUS = 16


@dataclass
class Alignment:
    """Alignment (src/align.rs:224-270)."""
    raw_score: int
    query_start: int
    query_end: int
    target_start: int
    target_end: int
    aln_len: int
    identities: int
    mismatches: int
    gaps: int
    gap_opens: int
    cigar: str

    def pct_identity(self) -> float:
        return 0.0 if self.aln_len == 0 else 100.0 * self.identities / self.aln_len

    def query_cov(self, qlen: int) -> float:
        return (self.query_end + 1 - self.query_start) / max(qlen, 1)

    def target_cov(self, tlen: int) -> float:
        return (self.target_end + 1 - self.target_start) / max(tlen, 1)

    def _rec(self) -> np.ndarray:
        r = np.zeros(1, ALIGNMENT_DT)
        for f in ALIGNMENT_DT.names:
            if f != "reserved":
                r[f] = getattr(self, f)
        return r


def _rows(rows) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(rows, dtype=np.int32).reshape(-1, 21))


def sw_score_end_host(rows, target, gaps):
    """(score, query_end, target_end) with the column-major-first end, or
    (0, -1, -1) (src/simd.rs:347-390)."""
    r, t = _rows(rows), _codes(target)
    s, qe, te = C.c_int32(0), C.c_int64(0), C.c_int64(0)
    check(_L().st_sw_score_end_host(ptr(r), len(r), ptr(t), len(t), gaps.open, gaps.extend, C.byref(s),
                                    C.byref(qe), C.byref(te)))
    return int(s.value), int(qe.value), int(te.value)


def align_canonical(rows, qres, target, gaps, band_radius: int = DEFAULT_BAND_RADIUS, known=None):
    """align_canonical (src/align.rs:840-882): an Alignment or None.
    known = (score, query_end, target_end) of the forward pass, if at hand."""
    r, t = _rows(rows), _codes(target)
    q = _codes(qres) if qres is not None else None
    kn = None
    if known is not None:
        kn = np.zeros(1, SCORE_END_DT)
        kn["score"], kn["query_end"], kn["target_end"] = known
    out = np.zeros(1, ALIGNMENT_DT)
    cap = 2 * (len(r) + len(t)) + 64
    buf = C.create_string_buffer(cap)
    n, found = C.c_uint64(0), C.c_int(0)
    check(_L().st_align_canonical(ptr(r), ptr(q) if q is not None else None, len(r), ptr(t), len(t), gaps.open,
                                  gaps.extend, band_radius, ptr(kn) if kn is not None else None, ptr(out), buf, cap,
                                  C.byref(n), C.byref(found)))
    if not found.value:
        return None
    o = out[0]
    return Alignment(int(o["raw_score"]), int(o["query_start"]), int(o["query_end"]), int(o["target_start"]),
                     int(o["target_end"]), int(o["aln_len"]), int(o["identities"]), int(o["mismatches"]),
                     int(o["gaps"]), int(o["gap_opens"]), buf.value.decode())


def align_span(rows, qres, target, gaps, span, band_radius: int = DEFAULT_BAND_RADIUS):
    """align_canonical's traceback from a known span (a SPAN_DT record, e.g.
    from align_spans on the device): an Alignment or None."""
    r, t = _rows(rows), _codes(target)
    q = _codes(qres) if qres is not None else None
    sp = np.zeros(1, SPAN_DT)
    sp[0] = span
    out = np.zeros(1, ALIGNMENT_DT)
    cap = 2 * (len(r) + len(t)) + 64
    buf = C.create_string_buffer(cap)
    n, found = C.c_uint64(0), C.c_int(0)
    check(_L().st_align_span(ptr(r), ptr(q) if q is not None else None, len(r), ptr(t), len(t), gaps.open,
                             gaps.extend, band_radius, ptr(sp), ptr(out), buf, cap, C.byref(n), C.byref(found)))
    if not found.value:
        return None
    o = out[0]
    return Alignment(int(o["raw_score"]), int(o["query_start"]), int(o["query_end"]), int(o["target_start"]),
                     int(o["target_end"]), int(o["aln_len"]), int(o["identities"]), int(o["mismatches"]),
                     int(o["gaps"]), int(o["gap_opens"]), buf.value.decode())


def rescore_alignment(rows, qres, target, gaps, aln: Alignment) -> int:
    """rescore_alignment (src/align.rs:1064-1135): the recomputed score;
    raises StError on any inconsistency."""
    r, t = _rows(rows), _codes(target)
    q = _codes(qres) if qres is not None else None
    s = C.c_int32(0)
    check(_L().st_rescore_alignment(ptr(r), ptr(q) if q is not None else None, len(r), ptr(t), len(t), gaps.open,
                                    gaps.extend, ptr(aln._rec()), aln.cigar.encode(), C.byref(s)))
    return int(s.value)


# ---- gapped stage, hits and .m8 rows (src/search.rs) -------------------------

HIT_DT = np.dtype([("target", "<u4"), ("raw_score", "<i4"), ("pident", "<f8"), ("aln_len", "<u8"),
                   ("mismatches", "<u8"), ("gap_opens", "<u8"), ("q_start", "<u8"), ("q_end", "<u8"),
                   ("t_start", "<u8"), ("t_end", "<u8"), ("evalue", "<f8"), ("bit_score", "<f8"), ("qlen", "<u8"),
                   ("tlen", "<u8")])
COUNTERS_DT = np.dtype([(f, "<u8") for f in ("scored", "passed_score", "tracebacks", "score_mismatches",
                                             "rejected_coverage", "hits")])
M8_HEADER = "query\ttarget\tpident\talnlen\tmismatch\tgapopen\tqstart\tqend\ttstart\ttend\tevalue\tbits"
_SIGS += [
    ("st_hit_order", C.c_int, [_P, _P]),
    ("st_gapped_stage", C.c_int, [_P, _P, C.c_uint64, C.c_int32, _P, C.POINTER(C.c_uint64)]),
    ("st_search_hits", C.c_int, [_P, _P, C.c_uint64, _P, _P, C.c_uint64, _P, _P, C.c_uint64, C.c_int32, C.c_int32,
                                 _P, _D, C.c_uint64, _D, _D, C.c_uint32, _P, _P, C.c_uint64, C.c_uint64,
                                 C.POINTER(C.c_uint64), _P]),
    ("st_search_hits_spans", C.c_int, [_P, _P, C.c_uint64, _P, _P, C.c_uint64, _P, _P, _P, C.c_uint64, C.c_int32,
                                       C.c_int32, _P, _D, C.c_uint64, _D, _D, C.c_uint32, _P, _P, C.c_uint64,
                                       C.c_uint64, C.POINTER(C.c_uint64), _P]),
    ("st_format_m8", C.c_int, [C.c_char_p, C.c_char_p, _P, _P, C.c_uint64, C.POINTER(C.c_uint64)]),
    ("st_format_m8_extra", C.c_int, [_P, C.c_char_p, _P, C.c_uint64, C.POINTER(C.c_uint64)]),
    ("st_parse_candidate_trace", C.c_int, [C.c_char_p, C.c_uint64, _P, _P, C.c_uint64, C.POINTER(C.c_uint64),
                                           C.POINTER(C.c_uint64), _P, C.c_uint64, _P, _P, _P, _P, _P]),
    ("st_format_candidate_trace", C.c_int, [C.c_char_p, _P, _P, _P, C.c_uint64, _P, C.c_uint64,
                                            C.POINTER(C.c_uint64)]),
    ("st_format_m8_block", C.c_int, [C.c_char_p, _P, _P, _P, C.c_uint64, C.c_int, C.c_int, _P, C.c_uint64,
                                     C.POINTER(C.c_uint64)]),
]
EXPORTED = [s[0] for s in _SIGS]


@dataclass
class SearchParams:
    """The gapped-stage settings of SearchParams (src/search.rs:59-141)."""
    max_evalue: float = 1e-3
    max_hits: int = 300
    min_query_cov: float = 0.0
    min_target_cov: float = 0.0
    band_radius: int = DEFAULT_BAND_RADIUS


def gapped_stage(scores, tids, min_score: int) -> np.ndarray:
    s = np.ascontiguousarray(scores, dtype=np.int32)
    t = np.ascontiguousarray(tids, dtype=np.uint32)
    order = np.zeros(max(len(s), 1), np.uint64)
    n = C.c_uint64(0)
    check(_L().st_gapped_stage(ptr(s), ptr(t), len(s), min_score, ptr(order), C.byref(n)))
    return order[: n.value].astype(np.int64)


def search_hits(rows, qres, arena, offsets, cand_tids, scores, gaps, ev: EValueParams,
                params: SearchParams = SearchParams(), counters: np.ndarray | None = None, spans=None):
    """Hits (HIT_DT array, hit_order) and their CIGARs for one query.  With
    spans (SPAN_DT per candidate, from align_spans) every traceback starts
    from its span and the host runs no score pass."""
    r = _rows(rows)
    q = _codes(qres) if qres is not None else None
    ar = np.ascontiguousarray(arena, dtype=np.uint8)
    of = np.ascontiguousarray(offsets, dtype=np.uint64)
    ct = np.ascontiguousarray(cand_tids, dtype=np.uint32)
    sc = np.ascontiguousarray(scores, dtype=np.int32)
    cap = len(ct) if params.max_hits == 0 else min(len(ct), params.max_hits)
    hits = np.zeros(max(cap, 1), HIT_DT)
    maxlen = int((of[1:] - of[:-1]).max()) if len(of) > 1 else 0
    stride = 2 * (len(r) + maxlen) + 64
    cig = C.create_string_buffer(max(cap, 1) * stride)
    n = C.c_uint64(0)
    cnt = counters if counters is not None else np.zeros(1, COUNTERS_DT)
    common = (params.max_evalue, params.max_hits, params.min_query_cov, params.min_target_cov, params.band_radius,
              ptr(hits), cig, stride, cap, C.byref(n), ptr(cnt))
    qp = ptr(q) if q is not None else None
    if spans is None:
        check(_L().st_search_hits(ptr(r), qp, len(r), ptr(ar), ptr(of), len(of) - 1, ptr(ct), ptr(sc), len(ct),
                                  gaps.open, gaps.extend, ptr(ev._r), *common))
    else:
        sp = np.ascontiguousarray(spans, dtype=SPAN_DT)
        if len(sp) != len(ct):
            raise ValueError("one span per candidate")
        check(_L().st_search_hits_spans(ptr(r), qp, len(r), ptr(ar), ptr(of), len(of) - 1, ptr(ct), ptr(sc),
                                        ptr(sp), len(ct), gaps.open, gaps.extend, ptr(ev._r), *common))
    k = min(cap, n.value)
    raw = cig.raw
    cigars = [raw[i * stride:(i + 1) * stride].split(b"\0", 1)[0].decode() for i in range(k)]
    return hits[:k].copy(), cigars


def parse_candidate_trace(text: str, target_ids):
    """A candidate trace -> (query ids, per-query (tids, diag scores, cells))
    with targets resolved against target_ids (their order = target index)."""
    raw = text.encode()
    ib = "".join(target_ids).encode()
    io = np.zeros(len(target_ids) + 1, np.uint64)
    io[1:] = np.cumsum([len(t.encode()) for t in target_ids])
    ibuf = np.frombuffer(ib, np.uint8) if ib else np.zeros(1, np.uint8)
    nq, nc = C.c_uint64(0), C.c_uint64(0)
    base = (raw, len(raw), ptr(ibuf), ptr(io), len(target_ids), C.byref(nq), C.byref(nc))
    check(_L().st_parse_candidate_trace(*base, None, 0, None, None, None, None, None))
    qoff = np.zeros(nq.value + 1, np.uint64)
    check(_L().st_parse_candidate_trace(*base, None, 0, ptr(qoff), None, None, None, None))
    qb = C.create_string_buffer(max(int(qoff[-1]), 1))
    coff = np.zeros(nq.value + 1, np.uint64)
    tids = np.zeros(max(nc.value, 1), np.uint32)
    sc = np.zeros(max(nc.value, 1), np.int32)
    ce = np.zeros(max(nc.value, 1), np.uint64)
    check(_L().st_parse_candidate_trace(*base, qb, len(qb), ptr(qoff), ptr(coff), ptr(tids), ptr(sc), ptr(ce)))
    qraw = qb.raw
    qids = [qraw[int(qoff[k]):int(qoff[k + 1])].decode() for k in range(nq.value)]
    per = [(tids[int(coff[k]):int(coff[k + 1])].copy(), sc[int(coff[k]):int(coff[k + 1])].copy(),
            ce[int(coff[k]):int(coff[k + 1])].copy()) for k in range(nq.value)]
    return qids, per


def format_candidate_trace(query_id: str, target_ids, diag_scores, cells) -> str:
    n = len(target_ids)
    tid = (C.c_char_p * max(n, 1))(*[t.encode() for t in target_ids])
    sc = np.ascontiguousarray(diag_scores, dtype=np.int32)
    ce = np.ascontiguousarray(cells, dtype=np.uint64)
    ln = C.c_uint64(0)
    args = (query_id.encode(), tid, ptr(sc), ptr(ce), n)
    check(_L().st_format_candidate_trace(*args, None, 0, C.byref(ln)))
    buf = C.create_string_buffer(ln.value + 1)
    check(_L().st_format_candidate_trace(*args, buf, len(buf), C.byref(ln)))
    return buf.raw[: ln.value].decode()


def format_m8_extra(hit, cigar: str) -> str:
    """Hit::extra_cols: raw, qlen, tlen, qcov, tcov, CIGAR."""
    h = np.zeros(1, HIT_DT)
    h[0] = hit
    buf = C.create_string_buffer(len(cigar) + 128)
    n = C.c_uint64(0)
    check(_L().st_format_m8_extra(ptr(h), cigar.encode(), buf, len(buf), C.byref(n)))
    return buf.value.decode()


def format_m8_block(query_id: str, target_ids, hits, cigars=None, extra: bool = False,
                    header: bool = False) -> str:
    """The lines the reference's writer emits for one query's hits."""
    h = np.ascontiguousarray(hits, dtype=HIT_DT)
    tid = (C.c_char_p * max(len(h), 1))(*[t.encode() for t in target_ids])
    cg = (C.c_char_p * max(len(h), 1))(*[c.encode() for c in cigars]) if cigars is not None else None
    n = C.c_uint64(0)
    args = (query_id.encode(), tid, ptr(h), cg, len(h), 1 if extra else 0, 1 if header else 0)
    check(_L().st_format_m8_block(*args, None, 0, C.byref(n)))
    buf = C.create_string_buffer(n.value + 1)
    check(_L().st_format_m8_block(*args, buf, len(buf), C.byref(n)))
    return buf.raw[: n.value].decode()


def format_m8(query_id: str, target_id: str, hit) -> str:
    h = np.zeros(1, HIT_DT)
    h[0] = hit
    buf = C.create_string_buffer(len(query_id) + len(target_id) + 256)
    n = C.c_uint64(0)
    check(_L().st_format_m8(query_id.encode(), target_id.encode(), ptr(h), buf, len(buf), C.byref(n)))
    return buf.value.decode()


# ---- per-profile statistics (src/evalue.rs:714-749) --------------------------

_SIGS += [
    ("st_calibration_targets", C.c_int, [_P, C.c_uint64, C.c_uint64, C.c_uint64, _P, C.c_uint64, _P]),
    ("st_calibrate_profile", C.c_int, [_P, _P, C.c_int32, C.c_int32, _P, C.c_uint64, C.c_uint64, C.c_uint64, _P,
                                       _P]),
    ("st_simulate_scores", C.c_int, [_P, C.c_int32, C.c_int32, C.c_int, C.c_uint64, C.c_uint64, C.c_uint64,
                                     C.c_uint64, C.c_uint64, _P]),
    ("st_calibrate_sequence", C.c_int, [_P, C.c_int32, C.c_int32, C.c_int, _P, C.c_uint64, C.c_uint64,
                                        C.c_uint64, _P, _P, _P]),
]
EXPORTED = [s[0] for s in _SIGS]
CALIBRATION_SEED = 0x5EED  # SearchParams::calibration_seed (src/search.rs:121)


def calibration_targets(lengths, samples: int, seed: int = CALIBRATION_SEED):
    """(arena, offsets) of calibrate_profile's random targets."""
    ln = np.ascontiguousarray(lengths, dtype=np.uint64)
    off = np.zeros(samples + 1, np.uint64)
    check(_L().st_calibration_targets(ptr(ln), len(ln), samples, seed, None, 0, ptr(off)))
    arena = np.zeros(max(int(off[-1]), 1), np.uint8)
    check(_L().st_calibration_targets(ptr(ln), len(ln), samples, seed, ptr(arena), arena.size, ptr(off)))
    return arena[: int(off[-1])], off


def calibrate_profile(qp, gaps, lengths, samples: int = 2000, seed: int = CALIBRATION_SEED):
    """(lambda, K, (gumbel lambda, mu)) of a profile query (st QueryProfile),
    scored on the GPU; None where the reference returns None."""
    ln = np.ascontiguousarray(lengths, dtype=np.uint64)
    g = np.zeros(1, GAPPED_PARAMS_DT)
    fit = np.zeros(1, GUMBEL_DT)
    rc = _L().st_calibrate_profile(qp.ctx.handle, qp.handle, gaps.open, gaps.extend, ptr(ln), len(ln), samples, seed,
                                   ptr(g), ptr(fit))
    if rc == 1:  # ST_ERR_INVALID: no fit
        return None
    check(rc)
    return float(g["lambda"][0]), float(g["k"][0]), (float(fit["lambda"][0]), float(fit["mu"][0]))


# ---- sequence statistics by simulation (src/evalue.rs:537-700) ---------------

# calibrate_sequence's default lengths (DEFAULT_CALIBRATION_LENGTHS,
# src/evalue.rs:670) and the settings SEQUENCE_TABLE was made with
# (its provenance strings, src/evalue.rs:287, 300)
DEFAULT_CALIBRATION_LENGTHS = (60, 100, 150, 220, 320, 480, 700, 1000)
TABLE_SAMPLES, TABLE_SEED = 20000, 24301


def simulate_scores(ctx, gaps, comp_bias: bool, m: int, n: int, samples: int, seed: int, cell_id: int) -> np.ndarray:
    """simulate_scores (src/evalue.rs:537-565): exact SWG scores (float64) of
    `samples` random background pairs, scored on the GPU."""
    out = np.zeros(max(samples, 1), np.float64)
    check(_L().st_simulate_scores(ctx.handle, gaps.open, gaps.extend, 1 if comp_bias else 0, m, n, samples, seed,
                                  cell_id, ptr(out)))
    return out[:samples]


def calibrate_sequence(ctx, gaps, comp_bias: bool, lengths=DEFAULT_CALIBRATION_LENGTHS, samples: int = 4000,
                       seed: int = 0x5EED):
    """calibrate_sequence (src/evalue.rs:673-700): (params dict, cells, ln K
    rms), or None where the reference returns None."""
    ln = np.ascontiguousarray(lengths, dtype=np.uint64)
    g = np.zeros(1, GAPPED_PARAMS_DT)
    out = np.zeros(max(len(ln), 1), CALIBRATION_CELL_DT)
    rms = C.c_double(0.0)
    rc = _L().st_calibrate_sequence(ctx.handle, gaps.open, gaps.extend, 1 if comp_bias else 0, ptr(ln), len(ln),
                                    samples, seed, ptr(g), ptr(out), C.byref(rms))
    if rc == 1:  # ST_ERR_INVALID: no fit
        return None
    check(rc)
    return _params_tuple(g), out[: len(ln)], float(rms.value)
