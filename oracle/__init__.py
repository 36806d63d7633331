"""CPU oracle for the post-prefilter scoring path — TEST INFRASTRUCTURE ONLY.

ctypes bindings of ``oracle/liboracle.so`` (a plain-C restatement of the
reference's scalar kernels, see ``oracle.h`` for the file:line each function
follows).  Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg may import this package, and only as the checker.  The
product library never links or calls it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
NPROF = 21


def build() -> None:
    """Compile the oracle (and the AVX2 baseline) with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE, "all"], check=True)


def lib() -> C.CDLL:
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        P = C.c_void_p
        L.or_gotoh_score_end.restype = C.c_int32
        L.or_gotoh_score_end.argtypes = [P, C.c_size_t, P, C.c_size_t, C.c_int32, C.c_int32, P, P]
        L.or_gotoh_textbook.restype = C.c_int32
        L.or_gotoh_textbook.argtypes = [P, C.c_size_t, P, C.c_size_t, C.c_int32, C.c_int32, P, P]
        L.or_diag_range.argtypes = [C.c_size_t, C.c_size_t, C.c_int64, P, P]
        L.or_kadane_score.restype = C.c_int32
        L.or_kadane_score.argtypes = [P, P, C.c_int64, C.c_size_t, C.c_size_t]
        L.or_kadane_segment.restype = C.c_int32
        L.or_kadane_segment.argtypes = [P, P, C.c_int64, C.c_size_t, C.c_size_t, P, P]
        L.or_score_and_gate.restype = C.c_size_t
        L.or_score_and_gate.argtypes = [P, C.c_size_t, P, P, P, C.c_size_t, P, C.c_size_t,
                                        C.c_int32, C.c_int32, C.c_int32, C.c_int, P, P, P]
        L.or_size_adjusted_gate.restype = C.c_int32
        L.or_size_adjusted_gate.argtypes = [C.c_int32, C.c_uint64]
        L.or_ung_pair.argtypes = [P, C.c_size_t, P, C.c_size_t, C.c_int32, C.c_int32, C.c_int32,
                                  P, P, P, P]
        L.or_splitmix_next.restype = C.c_uint64
        L.or_splitmix_next.argtypes = [P]
        L.or_background_sequence.argtypes = [P, C.c_size_t, P]
        L.or_blosum62.argtypes = [P, P, P, P]
        L.or_seq_rows.argtypes = [P, C.c_size_t, P]
        L.or_aa2num.argtypes = [P]
        _LIB = L
    return _LIB


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _rows(rows) -> np.ndarray:
    r = np.ascontiguousarray(np.asarray(rows, dtype=np.int32).reshape(-1, NPROF))
    return r


def _seq(t) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(t, dtype=np.uint8).reshape(-1))


# ---- alphabet / matrix -----------------------------------------------------

def aa2num() -> np.ndarray:
    out = np.zeros(256, dtype=np.uint8)
    lib().or_aa2num(_p(out))
    return out


def blosum62():
    """(score int8[21,21], p_back f64[21], R f32[20,20], lambda)."""
    s = np.zeros((21, 21), dtype=np.int8)
    pb = np.zeros(21, dtype=np.float64)
    r = np.zeros((20, 20), dtype=np.float32)
    lam = C.c_double()
    lib().or_blosum62(_p(s), _p(pb), _p(r), C.byref(lam))
    return s, pb, r, lam.value


def seq_rows(q) -> np.ndarray:
    q = _seq(q)
    rows = np.zeros((len(q), NPROF), dtype=np.int32)
    lib().or_seq_rows(_p(q), len(q), _p(rows))
    return rows


def encode(text: str) -> np.ndarray:
    t = aa2num()
    return t[np.frombuffer(text.encode(), dtype=np.uint8)]


# ---- SWG ------------------------------------------------------------------

def gotoh_score_end(rows, t, open_: int, ext: int):
    """(score, q_end, t_end) with -1 ends when the score is 0."""
    r, s = _rows(rows), _seq(t)
    qe, te = C.c_int64(), C.c_int64()
    sc = lib().or_gotoh_score_end(_p(r), len(r), _p(s), len(s), open_, ext,
                                  C.byref(qe), C.byref(te))
    return sc, qe.value, te.value


def gotoh_score(rows, t, open_: int, ext: int) -> int:
    return gotoh_score_end(rows, t, open_, ext)[0]


def span(rows, t, open_: int, ext: int):
    """(score, q_start, q_end, t_start, t_end) from align_canonical's two
    score passes (src/align.rs:847-866): the column-major-first end cell,
    then the same rule on the reversed prefixes rows[..=q_end] and
    t[..=t_end]; start = end - the reversed problem's end.  All cells -1
    when the score is 0.  With a negative gap cost the reversed problem can
    score differently (the reference then searches the start from the end,
    src/align.rs:859-863): the start cells are then -1.  Pinned in
    tests/test_align.py by brute force over every start on small pairs and
    by the reference's alignment tests."""
    r, s = _rows(rows), _seq(t)
    sc, ie, je = gotoh_score_end(r, s, open_, ext)
    if sc <= 0:
        return (0, -1, -1, -1, -1)
    rs, a, b = gotoh_score_end(r[: ie + 1][::-1], s[: je + 1][::-1], open_, ext)
    if rs != sc:
        assert open_ < 0 or ext < 0, (rs, sc)
        return (sc, -1, ie, -1, je)
    return (sc, ie - a, ie, je - b, je)


def gotoh_textbook(rows, t, open_: int, ext: int):
    r, s = _rows(rows), _seq(t)
    qe, te = C.c_int64(), C.c_int64()
    sc = lib().or_gotoh_textbook(_p(r), len(r), _p(s), len(s), open_, ext,
                                 C.byref(qe), C.byref(te))
    return sc, qe.value, te.value


# ---- ungapped / gate ---------------------------------------------------------

def diag_range(m: int, n: int, d: int):
    lo, hi = C.c_size_t(), C.c_size_t()
    lib().or_diag_range(m, n, d, C.byref(lo), C.byref(hi))
    return lo.value, hi.value


def kadane_score(rows, t, d: int, lo: int, hi: int) -> int:
    r, s = _rows(rows), _seq(t)
    return lib().or_kadane_score(_p(r), _p(s), d, lo, hi)


def kadane_segment(rows, t, d: int, lo: int, hi: int):
    r, s = _rows(rows), _seq(t)
    qs, qe = C.c_size_t(), C.c_size_t()
    sc = lib().or_kadane_segment(_p(r), _p(s), d, lo, hi, C.byref(qs), C.byref(qe))
    return sc, qs.value, qe.value


JOB_DT = np.dtype([("tid", "<u4"), ("diag", "<i4"), ("lo", "<u4"), ("hi", "<u4")])
REC_DT = np.dtype([("tid", "<u4"), ("hits", "<u4"), ("j0", "<u4"), ("j1", "<u4")])
CAND_DT = np.dtype([("target_id", "<u4"), ("hits", "<u4"), ("best_diag", "<i8"),
                    ("diag_score", "<i4"), ("t_lo", "<u4"), ("t_hi", "<u4")], align=True)


def score_and_gate(rows, arena, offsets, jobs, recs, open_, ext, min_diag_score, regions=False):
    """(candidates structured array, chains_used, per-job scores)."""
    r = _rows(rows)
    arena = _seq(arena)
    offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
    jobs = np.ascontiguousarray(jobs, dtype=JOB_DT)
    recs = np.ascontiguousarray(recs, dtype=REC_DT)
    out = np.zeros(max(len(recs), 1), dtype=CAND_DT)
    scores = np.zeros(max(len(jobs), 1), dtype=np.int32)
    chains = C.c_uint64(0)
    nc = lib().or_score_and_gate(_p(r), len(r), _p(arena), _p(offsets), _p(jobs), len(jobs),
                                 _p(recs), len(recs), open_, ext, min_diag_score,
                                 1 if regions else 0, _p(out), C.byref(chains), _p(scores))
    return out[:nc].copy(), chains.value, scores[:len(jobs)].copy()


def seed_hits_to_jobs(m, seeds, postings, tlen, max_diagonals, min_hits, diag_hits, diag_window, span):
    """(jobs, recs) of the seed-hit stage of prefilter_query with segment
    evidence: hits packed as (target, diagonal + m, qpos) and sorted
    (src/prefilter.rs:1033-1056), targets with >= min_hits hits, their
    diagonals kept by insertion in count order (top_diagonals,
    src/prefilter.rs:1194-1225), the best with >= diag_hits hits, then at most
    16 jobs per target over diag_range narrowed by diag_window
    (src/prefilter.rs:1432-1487).  seeds: (qpos, cur, end) rows; postings:
    (target << 32) | position.  Pinned in tests/test_oracle.py by a
    brute-force count of hits per (target, diagonal) ranked by sorting."""
    postings = np.asarray(postings, dtype=np.uint64)
    tid, dm, qp = [], [], []
    for q, cur, end in seeds:
        p = postings[int(cur):int(end)]
        tid.append((p >> np.uint64(32)).astype(np.int64))
        dm.append((p & np.uint64(0xFFFFFFFF)).astype(np.int64) + m - int(q))
        qp.append(np.full(len(p), int(q), np.int64))
    if not tid:
        return np.zeros(0, JOB_DT), np.zeros(0, REC_DT)
    tid, dm, qp = np.concatenate(tid), np.concatenate(dm), np.concatenate(qp)
    order = np.lexsort((qp, dm, tid))
    tid, dm, qp = tid[order], dm[order], qp[order]
    jobs, recs = [], []
    i, n = 0, len(tid)
    while i < n:
        j = i
        while j < n and tid[j] == tid[i]:
            j += 1
        t = int(tid[i])
        if j - i >= min_hits:
            top = []  # (count, diag, qmin, qmax)
            a = i
            while a < j:
                b = a
                while b < j and dm[b] == dm[a]:
                    b += 1
                run = (b - a, int(dm[a]) - m, int(qp[a]), int(qp[b - 1]))
                pos = next((k for k, x in enumerate(top) if run[0] > x[0]), len(top))
                if pos < max_diagonals:
                    top.insert(pos, run)
                    if len(top) > max_diagonals:
                        top.pop()
                a = b
            if top and top[0][0] >= diag_hits:
                j0 = len(jobs)
                for cnt, d, qmin, qmax in top[:16]:
                    lo, hi = diag_range(m, int(tlen[t]), d)
                    if diag_window > 0:
                        lo = max(lo, max(qmin - diag_window, 0))
                        hi = min(hi, qmax + span + diag_window)
                    jobs.append((t, d, lo, max(hi, lo)))
                recs.append((t, j - i, j0, len(jobs)))
        i = j
    return np.array(jobs, dtype=JOB_DT), np.array(recs, dtype=REC_DT)


def size_adjusted_gate(gate: int, db_residues: int) -> int:
    return lib().or_size_adjusted_gate(gate, db_residues)


def ung_pair(rows, t, open_, ext, min_diag_score):
    """(ung, best_diag, gate_score, pass) of the all-diagonal evidence."""
    r, s = _rows(rows), _seq(t)
    a, b, c, d = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
    lib().or_ung_pair(_p(r), len(r), _p(s), len(s), open_, ext, min_diag_score,
                      C.byref(a), C.byref(b), C.byref(c), C.byref(d))
    return a.value, b.value, c.value, bool(d.value)


# ---- RNG ----------------------------------------------------------------------

class SplitMix64:
    def __init__(self, seed: int):
        self._s = C.c_uint64(seed)

    def next_u64(self) -> int:
        return lib().or_splitmix_next(C.byref(self._s))

    def background(self, n: int) -> np.ndarray:
        out = np.zeros(n, dtype=np.uint8)
        lib().or_background_sequence(C.byref(self._s), n, _p(out))
        return out


# ---- AVX2 baseline (bench cpu_baseline leg) -------------------------------------

_AVX = None


def avx2_lib() -> C.CDLL:
    global _AVX
    if _AVX is None:
        path = os.path.join(_HERE, "libavx2base.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        P = C.c_void_p
        L.avx2_score_pairs.restype = C.c_double
        L.avx2_score_pairs.argtypes = [C.c_uint32, P, P, C.c_uint64, P, P, P, P, P, C.c_int32, C.c_int32,
                                       C.c_int, P, P]
        L.avx2_score_and_gate.restype = C.c_double
        L.avx2_score_and_gate.argtypes = [C.c_uint32, P, P, P, P, P, P, P, P, P, P, P, C.c_int, C.c_int, P, P, P,
                                          P]
        _AVX = L
    return _AVX


def avx2_score_pairs(rows_list, pair_q, pair_t, arena, offsets, open_, ext, threads=1):
    """(scores, kernels, seconds): the AVX2 u8 -> i16 -> scalar cascade."""
    rows_c = [np.ascontiguousarray(r, dtype=np.int32).reshape(-1, NPROF) for r in rows_list]
    ptrs = (C.c_void_p * len(rows_c))(*[r.ctypes.data for r in rows_c])
    qlen = np.array([len(r) for r in rows_c], dtype=np.uint32)
    pq = np.ascontiguousarray(pair_q, dtype=np.uint32)
    pt = np.ascontiguousarray(pair_t, dtype=np.uint32)
    arena = _seq(arena)
    offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
    lens = (offsets[1:] - offsets[:-1]).astype(np.uint32)
    sc = np.zeros(len(pq), dtype=np.int32)
    kn = np.zeros(len(pq), dtype=np.uint8)
    secs = avx2_lib().avx2_score_pairs(len(rows_c), ptrs, _p(qlen), len(pq), _p(pq), _p(pt), _p(arena),
                                       _p(offsets), _p(lens), open_, ext, threads, _p(sc), _p(kn))
    return sc, kn, secs


def avx2_score_and_gate(rows_list, job_lists, rec_lists, arena, offsets, opens, exts, gates, regions=False,
                        threads=1, want_scores=False):
    """([candidates per query], [chains per query], job scores or None,
    seconds): the reference's AVX2 ungapped stage (32-lane biased-u8 Kadane
    batches over EV_CHUNK chunks, then chain and gate) per query."""
    nq = len(rows_list)
    rows_c = [np.ascontiguousarray(r, dtype=np.int32).reshape(-1, NPROF) for r in rows_list]
    ptrs = (C.c_void_p * max(nq, 1))(*[r.ctypes.data for r in rows_c])
    qlen = np.array([len(r) for r in rows_c], dtype=np.uint32)
    jo = np.zeros(nq + 1, np.uint64)
    jo[1:] = np.cumsum([len(j) for j in job_lists])
    ro = np.zeros(nq + 1, np.uint64)
    ro[1:] = np.cumsum([len(r) for r in rec_lists])
    jobs = np.ascontiguousarray(np.concatenate(job_lists) if nq else np.zeros(0, JOB_DT), dtype=JOB_DT)
    recs = np.ascontiguousarray(np.concatenate(rec_lists) if nq else np.zeros(0, REC_DT), dtype=REC_DT)
    arena = _seq(arena)
    offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
    op = np.ascontiguousarray(opens, dtype=np.int32)
    ex = np.ascontiguousarray(exts, dtype=np.int32)
    gt = np.ascontiguousarray(gates, dtype=np.int32)
    ncand = np.zeros(max(nq, 1), np.uint64)
    chains = np.zeros(max(nq, 1), np.uint64)
    cands = np.zeros(max(len(recs), 1), dtype=CAND_DT)
    js = np.zeros(max(len(jobs), 1), np.int32) if want_scores else None
    secs = avx2_lib().avx2_score_and_gate(nq, ptrs, _p(qlen), _p(jo), _p(jobs), _p(ro), _p(recs), _p(arena),
                                          _p(offsets), _p(op), _p(ex), _p(gt), 1 if regions else 0, threads,
                                          _p(ncand), _p(cands), _p(chains), _p(js) if js is not None else None)
    if secs < 0:
        raise ValueError("avx2_score_and_gate: records must index their query's jobs in order")
    out = [cands[int(ro[q]):int(ro[q]) + int(ncand[q])].copy() for q in range(nq)]
    return out, [int(c) for c in chains[:nq]], (js[: len(jobs)] if js is not None else None), secs
