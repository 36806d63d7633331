"""Seeded synthetic workloads of the scoring path (host only).

Residues: SplitMix64 + BLOSUM62 background (the reference's
BackgroundSampler, src/evalue.rs:52-114).  Homologs: substitutions from
P(a|b) (src/matrix.rs:145-151) plus indel runs.  Candidate picks use numpy's
PCG64 with the same seed.  Every generator is deterministic for a seed.

Workloads (sizes as the benchmark configurations state them):

* ``plain_allpairs``  — 32 queries U(50,500) x 20,000 targets (log-normal
  median 250, sigma 0.5, clipped 30-2000), 5% planted homologs at ~40%
  identity with 5%/residue indels of length 1-10; all pairs.
* ``batched_uneven``  — 1,000 queries (log-normal median 300, sigma 0.5,
  20-3,000) vs 370,000 targets (log-normal median 150, sigma 0.637, 20-5,000,
  about 68 M residues); each query has 3,000 candidates of which 10% are
  planted homologs at 25-60% identity.
* ``long_pairs``       — 64 queries U(2,000-35,000) vs 5,000 targets
  U(1,000-30,000), 10% planted homologs at 50% identity with 1-50 indels.
* ``pssm_search``      — 16 profiles built by the PSSM builder from a3m
  MSAs of 50 members around a random root (length U(100,600)) at 30-70%
  identity with gap runs and insertions, vs 200,000 targets (log-normal
  median 250, sigma 0.5, 30-2,000) with 2% planted family members; gaps
  44/4, gate 30; all pairs.
* ``serving``          — a 1.4 M-target database (log-normal median 170,
  sigma 0.464, 20-10,000, about 265 M residues) and a request pool: 90% plain
  queries (log-normal median 300, 20-5,000), 10% profile queries from such
  MSAs, each with 3,000 candidates (5% planted homologs); gaps 11/1 with the
  size-adjusted gate for plain queries, 44/4 with gate 30 for profiles.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import List

import numpy as np

from ._lib import lib, ptr

_SIG_DONE = False


def _L():
    global _SIG_DONE
    L = lib()
    if not _SIG_DONE:
        P = C.c_void_p
        L.sy_background.argtypes = [P, C.c_uint64, P]
        L.sy_lognormal.argtypes = [P, C.c_uint64, C.c_double, C.c_double, C.c_uint32, C.c_uint32, P]
        L.sy_uniform.argtypes = [P, C.c_uint64, C.c_uint32, C.c_uint32, P]
        L.sy_homolog.restype = C.c_uint64
        L.sy_homolog.argtypes = [P, P, C.c_uint64, C.c_double, C.c_double, C.c_uint32, C.c_uint64, P]
        L.sy_pick.argtypes = [P, C.c_uint64, C.c_uint64, P]
        L.sy_msa_member.restype = C.c_uint64
        L.sy_msa_member.argtypes = [P, P, C.c_uint64, C.c_double, C.c_double, C.c_double, P, C.c_uint64]
        _SIG_DONE = True
    return L


class Stream:
    """A SplitMix64 stream driving the C generators."""

    def __init__(self, seed: int):
        self.s = C.c_uint64(seed & 0xFFFFFFFFFFFFFFFF)

    def background(self, n: int) -> np.ndarray:
        out = np.zeros(n, dtype=np.uint8)
        if n:
            _L().sy_background(C.byref(self.s), n, ptr(out))
        return out

    def lognormal(self, count: int, median: float, sigma: float, lo: int, hi: int) -> np.ndarray:
        out = np.zeros(count, dtype=np.uint32)
        if count:
            _L().sy_lognormal(C.byref(self.s), count, median, sigma, lo, hi, ptr(out))
        return out

    def uniform(self, count: int, lo: int, hi: int) -> np.ndarray:
        out = np.zeros(count, dtype=np.uint32)
        if count:
            _L().sy_uniform(C.byref(self.s), count, lo, hi, ptr(out))
        return out

    def homolog(self, src: np.ndarray, identity: float, indel_rate: float, indel_max: int,
                out_len: int) -> np.ndarray:
        s = np.ascontiguousarray(src, dtype=np.uint8)
        out = np.zeros(out_len, dtype=np.uint8)
        _L().sy_homolog(C.byref(self.s), ptr(s), len(s), identity, indel_rate, indel_max, out_len, ptr(out))
        return out

    def msa_member(self, root: np.ndarray, identity: float, gap_rate: float, insert_rate: float) -> str:
        r = np.ascontiguousarray(root, dtype=np.uint8)
        cap = 8 * len(r) + 64
        buf = C.create_string_buffer(cap)
        n = _L().sy_msa_member(C.byref(self.s), ptr(r), len(r), identity, gap_rate, insert_rate, buf, cap)
        return buf.raw[:n].decode()


@dataclass
class Workload:
    name: str
    queries: List[np.ndarray]
    arena: np.ndarray
    offsets: np.ndarray
    cand_lists: List[np.ndarray]
    gap_open: int = 11
    gap_extend: int = 1
    gate: int = 0
    meta: dict = field(default_factory=dict)

    @property
    def target_lengths(self) -> np.ndarray:
        return (self.offsets[1:] - self.offsets[:-1]).astype(np.int64)

    def target(self, t: int) -> np.ndarray:
        return self.arena[int(self.offsets[t]):int(self.offsets[t + 1])]

    def cells(self) -> int:
        tl = self.target_lengths
        return int(sum(len(q) * int(tl[c].sum()) for q, c in zip(self.queries, self.cand_lists)))

    def pairs(self) -> int:
        return int(sum(len(c) for c in self.cand_lists))


def _build_db(st: Stream, lengths: np.ndarray, homolog_src: dict, ident_fn, indel_rate, indel_max):
    """Targets of the given lengths; targets in homolog_src are homologs of
    the given source sequence (identity from ident_fn(t))."""
    offs = np.zeros(len(lengths) + 1, dtype=np.uint64)
    offs[1:] = np.cumsum(lengths.astype(np.uint64))
    arena = st.background(int(offs[-1]))
    for t, src in homolog_src.items():
        L = int(lengths[t])
        arena[int(offs[t]):int(offs[t]) + L] = st.homolog(src, ident_fn(t), indel_rate, indel_max, L)
    return arena, offs


def plain_allpairs(seed: int = 1, n_queries: int = 32, n_targets: int = 20000) -> Workload:
    st = Stream(seed)
    qlen = st.uniform(n_queries, 50, 500)
    queries = [st.background(int(L)) for L in qlen]
    tlen = st.lognormal(n_targets, 250.0, 0.5, 30, 2000)
    rng = np.random.Generator(np.random.PCG64(seed))
    n_h = n_targets * 5 // 100
    hom = rng.choice(n_targets, size=n_h, replace=False)
    src = {int(t): queries[int(rng.integers(n_queries))] for t in hom}
    arena, offs = _build_db(st, tlen, src, lambda t: 0.40, 0.05, 10)
    cands = [np.arange(n_targets, dtype=np.uint32) for _ in range(n_queries)]
    db_res = int(offs[-1])
    from .core import size_adjusted_gate
    return Workload("plain_allpairs", queries, arena, offs, cands, 11, 1, size_adjusted_gate(42, db_res),
                    {"seed": seed, "homologs": len(src)})


def batched_uneven(seed: int = 1, n_queries: int = 1000, n_targets: int = 370000, n_cand: int = 3000,
                   homolog_frac: float = 0.10) -> Workload:
    st = Stream(seed)
    qlen = st.lognormal(n_queries, 300.0, 0.5, 20, 3000)
    queries = [st.background(int(L)) for L in qlen]
    tlen = st.lognormal(n_targets, 150.0, 0.637, 20, 5000)
    rng = np.random.Generator(np.random.PCG64(seed))
    n_hom = int(round(n_cand * homolog_frac))
    n_hom_total = min(n_hom * n_queries, n_targets)
    per_q = n_hom_total // max(n_queries, 1)
    # targets [0, n_queries * per_q) are homologs, per_q for each query
    src = {}
    idents = {}
    for q in range(n_queries):
        for k in range(per_q):
            t = q * per_q + k
            src[t] = queries[q]
            idents[t] = 0.25 + 0.35 * rng.random()
    arena, offs = _build_db(st, tlen, src, lambda t: idents[t], 0.05, 10)
    cands = []
    for q in range(n_queries):
        own = np.arange(q * per_q, (q + 1) * per_q, dtype=np.uint32)
        rest = rng.choice(n_targets, size=n_cand - per_q, replace=False).astype(np.uint32)
        c = np.concatenate([own, rest])
        rng.shuffle(c)
        cands.append(c)
    from .core import size_adjusted_gate
    return Workload("batched_uneven", queries, arena, offs, cands, 11, 1,
                    size_adjusted_gate(42, int(offs[-1])), {"seed": seed, "homologs_per_query": per_q})


def long_pairs(seed: int = 1, n_queries: int = 64, n_targets: int = 5000, qmax: int = 35000,
               tmax: int = 30000) -> Workload:
    st = Stream(seed)
    qlen = st.uniform(n_queries, min(2000, qmax), qmax)
    queries = [st.background(int(L)) for L in qlen]
    tlen = st.uniform(n_targets, min(1000, tmax), tmax)
    rng = np.random.Generator(np.random.PCG64(seed))
    hom = rng.choice(n_targets, size=n_targets // 10, replace=False)
    src = {int(t): queries[int(rng.integers(n_queries))] for t in hom}
    arena, offs = _build_db(st, tlen, src, lambda t: 0.50, 0.02, 50)
    cands = [np.arange(n_targets, dtype=np.uint32) for _ in range(n_queries)]
    from .core import size_adjusted_gate
    return Workload("long_pairs", queries, arena, offs, cands, 11, 1, size_adjusted_gate(42, int(offs[-1])),
                    {"seed": seed})


def family_msa(st: Stream, rng: np.random.Generator, root_len: int, n_members: int = 50):
    """An a3m MSA around a random root: (text, root codes).  Members keep
    U(30%, 70%) of the root's residues, with gap runs and insertions."""
    root = st.background(int(root_len))
    aa = "ACDEFGHIKLMNPQRSTVWYX"
    lines = [">root", "".join(aa[c] for c in root)]
    for k in range(1, n_members):
        ident = 0.30 + 0.40 * rng.random()
        lines += [f">member_{k}", st.msa_member(root, ident, 0.01, 0.01)]
    return "\n".join(lines) + "\n", root


def pssm_search(seed: int = 1, n_profiles: int = 16, n_targets: int = 200000, n_members: int = 50,
                planted_frac: float = 0.02) -> Workload:
    from .host import Profile
    st = Stream(seed)
    rng = np.random.Generator(np.random.PCG64(seed))
    msas, roots, rows = [], [], []
    for _ in range(n_profiles):
        text, root = family_msa(st, rng, int(st.uniform(1, 100, 600)[0]), n_members)
        msas.append(text)
        roots.append(root)
        rows.append(Profile.from_msa_str(text).rows())
    tlen = st.lognormal(n_targets, 250.0, 0.5, 30, 2000)
    hom = rng.choice(n_targets, size=int(n_targets * planted_frac), replace=False)
    src = {int(t): roots[int(rng.integers(n_profiles))] for t in hom}
    idents = {int(t): 0.30 + 0.40 * rng.random() for t in hom}
    arena, offs = _build_db(st, tlen, src, lambda t: idents[t], 0.02, 10)
    cands = [np.arange(n_targets, dtype=np.uint32) for _ in range(n_profiles)]
    w = Workload("pssm_search", roots, arena, offs, cands, 44, 4, 30, {"seed": seed, "msas": msas})
    w.meta["rows"] = rows
    return w


def _distinct(rng: np.random.Generator, n: int, k: int, exclude=None) -> np.ndarray:
    """k distinct ids in [0, n) not in `exclude`, in random order."""
    got = np.zeros(0, np.int64)
    ex = np.asarray(exclude if exclude is not None else [], dtype=np.int64)
    while len(got) < k:
        draw = rng.integers(0, n, size=int((k - len(got)) * 1.2) + 16)
        got = np.union1d(got, np.setdiff1d(draw, ex))
    pick = rng.permutation(got)[:k]
    return pick.astype(np.uint32)


@dataclass
class ServingWorkload:
    arena: np.ndarray
    offsets: np.ndarray
    rows: List[np.ndarray]      # score rows per request
    cands: List[np.ndarray]
    gaps: List[tuple]           # (open, extend) per request
    gates: List[int]
    profile: List[bool]

    @property
    def target_lengths(self) -> np.ndarray:
        return (self.offsets[1:] - self.offsets[:-1]).astype(np.int64)

    def target(self, t: int) -> np.ndarray:
        return self.arena[int(self.offsets[t]):int(self.offsets[t + 1])]

    def cells(self, k: int) -> int:
        return len(self.rows[k]) * int(self.target_lengths[self.cands[k]].sum())


def serving(seed: int = 1, n_targets: int = 1400000, n_requests: int = 2000, n_cand: int = 3000,
            profile_frac: float = 0.10, n_profiles: int = 16, homolog_frac: float = 0.05) -> ServingWorkload:
    from .core import seq_rows, size_adjusted_gate
    from .host import Profile
    st = Stream(seed)
    rng = np.random.Generator(np.random.PCG64(seed))
    is_prof = rng.random(n_requests) < profile_frac
    qlen = st.lognormal(n_requests, 300.0, 0.5, 20, 5000)
    prof_rows, prof_roots = [], []
    for _ in range(n_profiles):
        text, root = family_msa(st, rng, int(st.uniform(1, 100, 600)[0]))
        prof_rows.append(Profile.from_msa_str(text).rows())
        prof_roots.append(root)
    queries, rows, which = [], [], []
    for k in range(n_requests):
        if is_prof[k]:
            p = int(rng.integers(n_profiles))
            queries.append(prof_roots[p])
            rows.append(prof_rows[p])
        else:
            q = st.background(int(qlen[k]))
            queries.append(q)
            rows.append(seq_rows(q))
    tlen = st.lognormal(n_targets, 170.0, 0.464, 20, 10000)
    per_q = int(round(n_cand * homolog_frac))
    per_q = min(per_q, n_targets // max(n_requests, 1))
    src = {}
    for k in range(n_requests):
        for j in range(per_q):
            src[k * per_q + j] = queries[k]
    idents = {t: 0.25 + 0.35 * rng.random() for t in src}
    arena, offs = _build_db(st, tlen, src, lambda t: idents[t], 0.05, 10)
    gate_seq = size_adjusted_gate(42, int(offs[-1]))
    cands, gaps, gates = [], [], []
    for k in range(n_requests):
        own = np.arange(k * per_q, (k + 1) * per_q, dtype=np.uint32)
        rest = _distinct(rng, n_targets, n_cand - per_q, exclude=own)
        c = np.concatenate([own, rest])
        rng.shuffle(c)
        cands.append(c)
        gaps.append((44, 4) if is_prof[k] else (11, 1))
        gates.append(30 if is_prof[k] else gate_seq)
    return ServingWorkload(arena, offs, rows, cands, gaps, gates, list(map(bool, is_prof)))
