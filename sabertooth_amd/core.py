"""Python mirror of the reference's scoring-path API over the gfx950 library.

Names follow the reference (raw-lab/sabertooth v0.4.0):

* ``GapCosts`` — src/align.rs:196-223
* ``QueryProfile.from_rows / sw_score / sw_score_with`` — src/simd.rs:128-292
* ``TargetDb`` — the resident residue arena + offsets (src/index.rs:10-13)
* ``score_batch`` / ``kadane jobs`` — src/ungapped.rs:285-386
* ``score_and_gate`` / ``Candidate`` — src/prefilter.rs:370-381, 1498-1660
* ``size_adjusted_gate`` — src/prefilter.rs:246-257

Every call runs on the GPU through the C ABI in ``include/sabertooth_amd.h``.
"""
from __future__ import annotations

import ctypes as C
import os
import time
from dataclasses import dataclass
from typing import Iterable, Sequence

import numpy as np

from . import _lib
from ._lib import CAND_DT, JOB_DT, PAIR_RESULT_DT, REC_DT, check, lib, ptr

NPROF = 21


@dataclass(frozen=True)
class GapCosts:
    """A gap of length L costs open + (L-1)*extend (src/align.rs:196-223)."""

    open: int
    extend: int

    @staticmethod
    def sequence_default() -> "GapCosts":
        return GapCosts(11, 1)

    @staticmethod
    def profile_default() -> "GapCosts":
        return GapCosts(44, 4)

    def cost(self, length: int) -> int:
        return 0 if length == 0 else self.open + (length - 1) * self.extend


def blosum62():
    """(score int8[21,21], p_back f64[21], R f32[20,20], lambda)."""
    s = np.zeros((21, 21), dtype=np.int8)
    pb = np.zeros(21, dtype=np.float64)
    r = np.zeros((20, 20), dtype=np.float32)
    lam = C.c_double()
    check(lib().st_blosum62(ptr(s), ptr(pb), ptr(r), C.byref(lam)))
    return s, pb, r, lam.value


def aa2num() -> np.ndarray:
    t = np.zeros(256, dtype=np.uint8)
    check(lib().st_aa2num(ptr(t)))
    return t


def encode(text: str) -> np.ndarray:
    return aa2num()[np.frombuffer(text.encode(), dtype=np.uint8)]


def seq_rows(query) -> np.ndarray:
    q = np.ascontiguousarray(query, dtype=np.uint8)
    rows = np.zeros((len(q), NPROF), dtype=np.int32)
    check(lib().st_seq_rows(ptr(q), len(q), ptr(rows)))
    return rows


def size_adjusted_gate(gate: int, db_residues: int) -> int:
    return int(lib().st_size_adjusted_gate(gate, db_residues))


class Context:
    """One GPU context (device ordinal); owns a stream."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        check(lib().st_ctx_create(device, C.byref(h)))
        self._h = h

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h:
            lib().st_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class TargetDb:
    """Resident target database: residues[offsets[t] .. offsets[t+1])."""

    def __init__(self, ctx: Context, residues: np.ndarray, offsets: np.ndarray):
        self.ctx = ctx
        r = np.ascontiguousarray(residues, dtype=np.uint8)
        o = np.ascontiguousarray(offsets, dtype=np.uint64)
        h = C.c_void_p()
        check(lib().st_db_create(ctx.handle, ptr(r), len(r), ptr(o), len(o) - 1, C.byref(h)))
        self._h = h
        self.lengths = (o[1:] - o[:-1]).astype(np.int64)

    @staticmethod
    def from_stidx(ctx: Context, path: str) -> "TargetDb":
        """The residue and offset sections of a Sabertooth index file
        (<db>.stidx), checked as SeedIndex::open checks them."""
        from . import host
        info = host.stidx_info(path)
        db = TargetDb.__new__(TargetDb)
        db.ctx = ctx
        h = C.c_void_p()
        check(lib().st_db_create_from_stidx(ctx.handle, os.fsencode(path), C.byref(h)))
        db._h = h
        o = host.stidx_targets(path)[1]
        db.lengths = (o[1:] - o[:-1]).astype(np.int64)
        assert len(db.lengths) == info["num_seqs"]
        return db

    @staticmethod
    def from_sequences(ctx: Context, seqs: Sequence[np.ndarray]) -> "TargetDb":
        offs = np.zeros(len(seqs) + 1, dtype=np.uint64)
        if len(seqs):
            offs[1:] = np.cumsum([len(s) for s in seqs])
        arena = np.concatenate([np.asarray(s, dtype=np.uint8) for s in seqs]) if len(seqs) else np.zeros(0, np.uint8)
        return TargetDb(ctx, arena, offs)

    @property
    def handle(self):
        return self._h

    def __len__(self):
        return int(lib().st_db_targets(self._h))

    def residues(self) -> int:
        return int(lib().st_db_residues(self._h))

    def close(self):
        if self._h:
            lib().st_db_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class QueryProfile:
    """A query (sequence rows or PSSM rows) prepared for repeated scoring."""

    def __init__(self, ctx: Context, rows: np.ndarray):
        self.ctx = ctx
        r = np.ascontiguousarray(np.asarray(rows, dtype=np.int32).reshape(-1, NPROF))
        h = C.c_void_p()
        check(lib().st_qprof_create(ctx.handle, ptr(r), len(r), C.byref(h)))
        self._h = h
        self._rows = r

    @staticmethod
    def from_rows(ctx: Context, rows) -> "QueryProfile":
        return QueryProfile(ctx, rows)

    @staticmethod
    def from_sequence(ctx: Context, query) -> "QueryProfile":
        return QueryProfile(ctx, seq_rows(query))

    @property
    def handle(self):
        return self._h

    def query_len(self) -> int:
        return int(lib().st_qprof_len(self._h))

    def rows(self) -> np.ndarray:
        return self._rows

    def sw_score_with(self, db: TargetDb, tids, gaps: GapCosts):
        """(scores int32[], kernels uint8[]) for target ids `tids`."""
        t = np.ascontiguousarray(tids, dtype=np.uint32)
        sc = np.zeros(len(t), dtype=np.int32)
        kn = np.zeros(len(t), dtype=np.uint8)
        check(lib().st_sw_score_batch(self.ctx.handle, self._h, db.handle, ptr(t), len(t),
                                      gaps.open, gaps.extend, ptr(sc), ptr(kn)))
        return sc, kn

    def sw_score_end_with(self, db: TargetDb, tids, gaps: GapCosts) -> np.ndarray:
        """Score and column-major-first end cell per target id
        (sw_score_end_with, src/simd.rs:294-326): SCORE_END_DT records."""
        t = np.ascontiguousarray(tids, dtype=np.uint32)
        out = np.zeros(max(len(t), 1), dtype=_lib.SCORE_END_DT)
        check(lib().st_sw_score_end_batch(self.ctx.handle, self._h, db.handle, ptr(t), len(t),
                                          gaps.open, gaps.extend, ptr(out)))
        return out[: len(t)]

    def sw_score_batch(self, db: TargetDb, tids, gaps: GapCosts) -> np.ndarray:
        return self.sw_score_with(db, tids, gaps)[0]

    def sw_score(self, target, gaps: GapCosts) -> int:
        """Score one target slice (uploads a one-sequence database)."""
        db = TargetDb.from_sequences(self.ctx, [np.asarray(target, dtype=np.uint8)])
        return int(self.sw_score_batch(db, [0], gaps)[0])

    def close(self):
        if self._h:
            lib().st_qprof_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def jobs_array(jobs: Iterable) -> np.ndarray:
    """(tid, diag, lo, hi) tuples -> the C ABI's st_job array."""
    return np.array([tuple(j) for j in jobs], dtype=JOB_DT) if not isinstance(jobs, np.ndarray) \
        else np.ascontiguousarray(jobs, dtype=JOB_DT)


def score_batch(qp: QueryProfile, db: TargetDb, jobs) -> np.ndarray:
    """kadane_score of every job (ungapped::score_batch, src/ungapped.rs:285)."""
    j = jobs_array(jobs)
    out = np.zeros(len(j), dtype=np.int32)
    check(lib().st_ungapped_score_jobs(qp.ctx.handle, qp.handle, db.handle, ptr(j), len(j), ptr(out)))
    return out


class SpansCall:
    """The arguments of st_align_spans_batch packed once (query handles, hit
    offsets and target ids, gap costs) with an output buffer, for callers
    that trace back the same hit lists repeatedly: run() is then the C-ABI
    call alone, as a caller in the reference's language makes it, and
    returns that buffer (each run overwrites it)."""

    def __init__(self, db: TargetDb, queries: Sequence[QueryProfile], hit_lists, gaps: Sequence[GapCosts]):
        nq = len(queries)
        self.db, self.queries, self.nq = db, list(queries), nq
        self.off = np.zeros(nq + 1, dtype=np.uint64)
        if nq:
            self.off[1:] = np.cumsum(np.fromiter((len(h) for h in hit_lists), dtype=np.uint64, count=nq))
        self.tids = np.ascontiguousarray(np.concatenate(hit_lists) if nq else np.zeros(0), dtype=np.uint32)
        self.op = np.fromiter((g.open for g in gaps), dtype=np.int32, count=nq)
        self.ex = np.fromiter((g.extend for g in gaps), dtype=np.int32, count=nq)
        self.qh = np.fromiter((q.handle.value for q in queries), dtype=np.uint64, count=nq)
        # every entry is written by each call
        self.out = np.empty(max(len(self.tids), 1), dtype=_lib.SPAN_DT)

    def run(self, ctx: Context) -> np.ndarray:
        check(lib().st_align_spans_batch(ctx.handle, self.db.handle, ptr(self.qh), self.nq, ptr(self.off),
                                         ptr(self.tids), ptr(self.op), ptr(self.ex), ptr(self.out)))
        return self.out[: len(self.tids)]


def align_spans(ctx: Context, db: TargetDb, queries: Sequence[QueryProfile], hit_lists,
                gaps: Sequence[GapCosts]) -> np.ndarray:
    """Score, end and start of each hit (SPAN_DT, hits in list order): the
    forward and reversed-prefix score passes of align_canonical
    (src/align.rs:847-866) on the device."""
    return SpansCall(db, queries, hit_lists, gaps).run(ctx)


def align_spans_launches(ctx: Context):
    """(launches, n_forward): the DP launches of the last align_spans call on
    ctx (LAUNCH_INFO_DT rows: the forward pass's first, then the reverse
    pass's) with their executed/useful cells and device time."""
    n, nf = C.c_uint64(0), C.c_uint64(0)
    check(lib().st_align_spans_launches(ctx.handle, None, 0, C.byref(n), C.byref(nf)))
    out = np.zeros(max(n.value, 1), dtype=_lib.LAUNCH_INFO_DT)
    check(lib().st_align_spans_launches(ctx.handle, ptr(out), n.value, C.byref(n), C.byref(nf)))
    return out[: n.value], int(nf.value)


SPAN_PHASES = ("forward_plan", "forward_run", "forward_readback", "reverse_prep", "reverse_plan", "reverse_run",
               "reverse_readback", "call")


def align_spans_phases(ctx: Context) -> dict:
    """Host wall-clock phases (ms) of the last align_spans call on ctx."""
    ms = (C.c_double * len(SPAN_PHASES))()
    check(lib().st_align_spans_phases(ctx.handle, ms))
    return {k: float(ms[i]) for i, k in enumerate(SPAN_PHASES)}


class Postings:
    """An index's posting lists on the device: (target << 32) | position."""

    def __init__(self, ctx: Context, postings):
        p = np.ascontiguousarray(postings, dtype=np.uint64)
        h = C.c_void_p()
        check(lib().st_postings_create(ctx.handle, ptr(p), len(p), C.byref(h)))
        self.ctx, self._h = ctx, h

    @staticmethod
    def from_stidx(ctx: Context, path: str) -> "Postings":
        """The posting section of an index file, decoded on the host."""
        p = Postings.__new__(Postings)
        h = C.c_void_p()
        check(lib().st_postings_from_stidx(ctx.handle, os.fsencode(path), C.byref(h)))
        p.ctx, p._h = ctx, h
        return p

    @property
    def handle(self):
        return self._h

    def __len__(self):
        return int(lib().st_postings_count(self._h))

    def close(self):
        if self._h:
            lib().st_postings_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def seed_hits_to_jobs(ctx: Context, db: TargetDb, postings: Postings, m: int, seeds, max_diagonals: int = 4,
                      min_hits: int = 1, diag_hits: int = 1, diag_window: int = 48, span: int = 6):
    """(jobs, recs) of the seed-hit stage of prefilter_query on the device;
    seeds: (qpos, cur, end) rows."""
    sd = np.zeros(max(len(seeds), 1), dtype=_lib.SEED_DT)
    for k, (q, a, b) in enumerate(seeds):
        sd[k]["qpos"], sd[k]["cur"], sd[k]["end"] = q, a, b
    hp = np.zeros(1, dtype=_lib.HIT_PARAMS_DT)
    hp["max_diagonals"], hp["min_hits"], hp["diag_hits"] = max_diagonals, min_hits, diag_hits
    hp["diag_window"], hp["span"] = diag_window, span
    nj, nr = C.c_uint64(0), C.c_uint64(0)
    args = (ctx.handle, db.handle, postings.handle, m, ptr(sd), len(seeds), ptr(hp))
    # every job and every record comes from at least one distinct hit: the
    # hit count bounds both, so one call fills them
    nh = int(sum(int(b) - int(a) for _, a, b in seeds)) if len(seeds) else 0
    jobs = np.zeros(max(nh, 1), dtype=JOB_DT)
    recs = np.zeros(max(nh, 1), dtype=REC_DT)
    check(lib().st_seed_hits_to_jobs(*args, ptr(jobs), len(jobs), C.byref(nj), ptr(recs), len(recs), C.byref(nr)))
    return jobs[: nj.value], recs[: nr.value]


def seed_hits_to_jobs_batch(ctx: Context, db: TargetDb, postings: Postings, ms, seed_lists, max_diagonals: int = 4,
                            min_hits: int = 1, diag_hits: int = 1, diag_window: int = 48, span: int = 6):
    """[(jobs, recs) per query] from one device call over every query."""
    nq = len(ms)
    so, sd = _seed_array(seed_lists)
    nh = int((sd["end"].astype(np.int64) - sd["cur"].astype(np.int64)).sum())
    hp = np.zeros(1, dtype=_lib.HIT_PARAMS_DT)
    hp["max_diagonals"], hp["min_hits"], hp["diag_hits"] = max_diagonals, min_hits, diag_hits
    hp["diag_window"], hp["span"] = diag_window, span
    m = np.ascontiguousarray(ms, dtype=np.uint64)
    # bounds, filled only up to the counts the call returns
    jobs = np.empty(max(nh, 1), dtype=JOB_DT)
    recs = np.empty(max(nh, 1), dtype=REC_DT)
    jo = np.zeros(nq + 1, np.uint64)
    ro = np.zeros(nq + 1, np.uint64)
    nj, nr = C.c_uint64(0), C.c_uint64(0)
    check(lib().st_seed_hits_to_jobs_batch(ctx.handle, db.handle, postings.handle, nq, ptr(m), ptr(so), ptr(sd),
                                           ptr(hp), ptr(jobs), len(jobs), C.byref(nj), ptr(recs), len(recs),
                                           C.byref(nr), ptr(jo), ptr(ro)))
    return [(jobs[int(jo[q]):int(jo[q + 1])], recs[int(ro[q]):int(ro[q + 1])]) for q in range(nq)]


def _seed_array(seed_lists):
    so = np.zeros(len(seed_lists) + 1, np.uint64)
    so[1:] = np.cumsum([len(x) for x in seed_lists]) if len(seed_lists) else []
    sd = np.zeros(max(int(so[-1]), 1), dtype=_lib.SEED_DT)
    if int(so[-1]):
        allv = np.array([t for x in seed_lists for t in x], dtype=np.uint64).reshape(-1, 3)
        sd["qpos"][: len(allv)], sd["cur"][: len(allv)], sd["end"][: len(allv)] = allv[:, 0], allv[:, 1], allv[:, 2]
    return so, sd


def seed_hits_to_candidates_batch(db: TargetDb, postings: Postings, queries: Sequence[QueryProfile], seed_lists,
                                  gaps: Sequence[GapCosts], min_diag_scores, regions: bool = False,
                                  max_diagonals: int = 4, min_hits: int = 1, diag_hits: int = 1,
                                  diag_window: int = 48, span: int = 6, seed_arrays=None):
    """([candidates per query], chains_used): the seed-hit stage and
    score_and_gate in one device call; seed_arrays = (seed_off, SEED_DT
    rows) replaces seed_lists."""
    nq = len(queries)
    so, sd = seed_arrays if seed_arrays is not None else _seed_array(seed_lists)
    so = np.ascontiguousarray(so, dtype=np.uint64)
    sd = np.ascontiguousarray(sd, dtype=_lib.SEED_DT)
    # the C entry point reads nq gap pairs and gates and nq + 1 seed offsets
    if len(gaps) != nq or len(min_diag_scores) != nq or len(so) != nq + 1:
        raise ValueError(f"seed_hits_to_candidates_batch: {nq} queries but {len(gaps)} gap costs, "
                         f"{len(min_diag_scores)} gates and {len(so) - 1} seed lists")
    if len(sd) < int(so[-1]):
        raise ValueError("seed_hits_to_candidates_batch: fewer seed rows than the offsets name")
    hp = np.zeros(1, dtype=_lib.HIT_PARAMS_DT)
    hp["max_diagonals"], hp["min_hits"], hp["diag_hits"] = max_diagonals, min_hits, diag_hits
    hp["diag_window"], hp["span"] = diag_window, span
    qh = (C.c_void_p * max(nq, 1))(*[q.handle for q in queries])
    op = np.array([g.open for g in gaps], np.int32)
    ex = np.array([g.extend for g in gaps], np.int32)
    gt = np.ascontiguousarray(min_diag_scores, dtype=np.int32)
    co = np.zeros(nq + 1, np.uint64)
    n, chains = C.c_uint64(0), C.c_uint64(0)
    if nq == 0:
        return [], 0
    # a candidate is a record, and every record has a distinct seed hit: the
    # hit count bounds the output (np.empty leaves untouched pages unbacked)
    nh = int((sd["end"][: int(so[-1])].astype(np.int64) - sd["cur"][: int(so[-1])].astype(np.int64)).sum())
    out = np.empty(max(nh, 1), dtype=CAND_DT)
    check(lib().st_seed_hits_to_candidates_batch(queries[0].ctx.handle, db.handle, postings.handle, nq, qh, ptr(so),
                                                 ptr(sd), ptr(hp), ptr(op), ptr(ex), ptr(gt), 1 if regions else 0,
                                                 ptr(out), len(out), C.byref(n), ptr(co), C.byref(chains)))
    return [out[int(co[q]):int(co[q + 1])] for q in range(nq)], chains.value


def seed_stage_timing(postings: Postings):
    """(device window, job scan, gate) in ms of the last seeds-to-candidates
    call on these postings (HIP events)."""
    ms = np.zeros(3, np.float64)
    check(lib().st_seed_stage_timing(postings.handle, ptr(ms)))
    return float(ms[0]), float(ms[1]), float(ms[2])


def score_and_gate(qp: QueryProfile, db: TargetDb, jobs, recs, gaps: GapCosts,
                   min_diag_score: int, regions: bool = False):
    """(candidates, chains_used, job_scores) as score_and_gate produces them."""
    j = jobs_array(jobs)
    r = np.ascontiguousarray(recs, dtype=REC_DT) if isinstance(recs, np.ndarray) else \
        np.array([tuple(x) for x in recs], dtype=REC_DT)
    out = np.zeros(max(len(r), 1), dtype=CAND_DT)
    nout = C.c_uint64(0)
    chains = C.c_uint64(0)
    js = np.zeros(max(len(j), 1), dtype=np.int32)
    check(lib().st_score_and_gate(qp.ctx.handle, qp.handle, db.handle, ptr(j), len(j), ptr(r), len(r),
                                  gaps.open, gaps.extend, min_diag_score, 1 if regions else 0, ptr(out),
                                  C.byref(nout), C.byref(chains), ptr(js)))
    return out[: nout.value].copy(), chains.value, js[: len(j)].copy()


class Batch:
    """Many queries x candidate lists, planned once, run on the device."""

    def __init__(self, ctx: Context, db: TargetDb, queries: Sequence[QueryProfile], cand_lists,
                 gaps: Sequence[GapCosts], gates: Sequence[int]):
        nq = len(queries)
        self.ctx, self.db, self.queries = ctx, db, list(queries)
        off = np.zeros(nq + 1, dtype=np.uint64)
        off[1:] = np.cumsum([len(c) for c in cand_lists]) if nq else []
        tids = (np.concatenate([np.asarray(c, dtype=np.uint32) for c in cand_lists])
                if nq else np.zeros(0, np.uint32))
        tids = np.ascontiguousarray(tids, dtype=np.uint32)
        op = np.array([g.open for g in gaps], dtype=np.int32)
        ex = np.array([g.extend for g in gaps], dtype=np.int32)
        gt = np.array(gates, dtype=np.int32)
        qh = (C.c_void_p * max(nq, 1))(*[q.handle.value for q in queries])
        h = C.c_void_p()
        check(lib().st_batch_create(ctx.handle, db.handle, qh, nq, ptr(off), ptr(tids), ptr(op), ptr(ex),
                                    ptr(gt), C.byref(h)))
        self._h = h
        self.offsets = off
        self.npairs = int(off[-1]) if nq else 0

    def run(self, graph: bool = False) -> None:
        """Score the batch; `graph` replays the DP phase from a HIP graph
        captured on the first such call."""
        check((lib().st_batch_run_graph if graph else lib().st_batch_run)(self._h))

    def results(self) -> np.ndarray:
        out = np.zeros(max(self.npairs, 1), dtype=PAIR_RESULT_DT)
        check(lib().st_batch_results(self._h, ptr(out)))
        return out[: self.npairs]

    def cells(self) -> int:
        return int(lib().st_batch_cells(self._h))

    def stats(self):
        """(executed cells incl. padding, units, launches) of the plan."""
        e, u, l = C.c_uint64(0), C.c_uint64(0), C.c_uint64(0)
        check(lib().st_batch_stats(self._h, C.byref(e), C.byref(u), C.byref(l)))
        return int(e.value), int(u.value), int(l.value)

    def launches(self) -> np.ndarray:
        """Per-launch plan figures and device time of the last run."""
        n = C.c_uint64(0)
        check(lib().st_batch_launches(self._h, None, 0, C.byref(n)))
        out = np.zeros(max(n.value, 1), dtype=_lib.LAUNCH_INFO_DT)
        check(lib().st_batch_launches(self._h, ptr(out), n.value, C.byref(n)))
        return out[: n.value]

    def timing(self):
        """(dp_ms, rescore_ms, gate_ms, rescored_pairs) of the last run."""
        ms = np.zeros(3, dtype=np.float64)
        n = C.c_uint64(0)
        check(lib().st_batch_timing(self._h, ptr(ms), C.byref(n)))
        return float(ms[0]), float(ms[1]), float(ms[2]), int(n.value)

    def close(self):
        if self._h:
            lib().st_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Serving loop over a resident database (continuous batching).

    persistent=True (default): requests are planned into work units and
    appended to the queue of one running serving kernel, which merges them
    with the work already on the device; requests it cannot take (int32
    precision, negative gap costs, very long diagonals) go to the worker
    slots.  persistent=False: worker slots take every queued request whose
    DP state fits the memory budget beside what is already running and score
    them as one micro-batch on their own stream.  Either way each request's
    results equal those of ``Batch`` on the same query and candidates.
    """

    def __init__(self, ctx: Context, db: TargetDb, workers: int = 4, max_batch_queries: int = 64,
                 state_budget_bytes: int = 4 << 30, persistent: bool = True, arena_mib: int = 0):
        cfg = np.zeros(1, dtype=_lib.SERVER_CONFIG_DT)
        cfg["workers"], cfg["max_batch_queries"], cfg["state_budget_bytes"] = (workers, max_batch_queries,
                                                                              state_budget_bytes)
        cfg["persistent"] = 1 if persistent else 0
        cfg["arena_mib"] = arena_mib
        h = C.c_void_p()
        check(lib().st_server_create(ctx.handle, db.handle, ptr(cfg), C.byref(h)))
        self._h = h
        self.ctx, self.db = ctx, db
        self._live = {}  # ticket -> (request record, candidates, output) kept alive

    @staticmethod
    def _request(qp: QueryProfile, cands: np.ndarray, out: np.ndarray, gaps: GapCosts, gate: int):
        r = np.zeros(1, dtype=_lib.REQUEST_DT)
        r["query"] = qp.handle.value
        r["gap_open"], r["gap_extend"], r["min_diag_score"] = gaps.open, gaps.extend, gate
        r["cand_tids"] = cands.ctypes.data if len(cands) else 0
        r["n_cands"] = len(cands)
        r["out"] = out.ctypes.data
        return r

    def submit(self, qp: QueryProfile, cands, gaps: GapCosts, gate: int) -> int:
        c = np.ascontiguousarray(cands, dtype=np.uint32)
        out = np.zeros(max(len(c), 1), dtype=PAIR_RESULT_DT)
        r = self._request(qp, c, out, gaps, gate)
        t = C.c_uint64(0)
        check(lib().st_server_submit(self._h, ptr(r), C.byref(t)))
        self._live[t.value] = (r, c, out, qp)
        return int(t.value)

    def poll(self, ticket: int) -> bool:
        d = C.c_int(0)
        check(lib().st_server_poll(self._h, ticket, C.byref(d)))
        return bool(d.value)

    def wait(self, ticket: int, timeout_ms: float = -1.0):
        """(results in candidate order, latency in ms from submit to completion)."""
        lat = C.c_double(0)
        rc = lib().st_server_wait(self._h, ticket, timeout_ms, C.byref(lat))
        if rc != _lib.ST_ERR_TIMEOUT:  # a timed-out ticket stays live
            rec = self._live.pop(ticket, None)
        check(rc)
        _, c, out, _ = rec
        return out[: len(c)].copy(), float(lat.value)

    def run_trace(self, requests, arrival_ms):
        """Open-loop replay: requests[k] = (qp, cands, gaps, gate) submitted at
        arrival_ms[k]; returns (list of result arrays, latencies in ms).
        ``last_trace_ms`` is the server's wall time for the trace: from the
        first submission to the last completion (the request records and
        output arrays are built before it starts)."""
        n = len(requests)
        recs = np.zeros(max(n, 1), dtype=_lib.REQUEST_DT)
        keep = []
        cols = {f: np.zeros(n, dtype=recs.dtype[f]) for f in ("query", "gap_open", "gap_extend", "min_diag_score",
                                                                 "cand_tids", "n_cands", "out")}
        for k, (qp, cands, gaps, gate) in enumerate(requests):
            c = np.ascontiguousarray(cands, dtype=np.uint32)
            out = np.zeros(max(len(c), 1), dtype=PAIR_RESULT_DT)
            cols["query"][k] = qp.handle.value
            cols["gap_open"][k], cols["gap_extend"][k], cols["min_diag_score"][k] = gaps.open, gaps.extend, gate
            cols["cand_tids"][k] = c.ctypes.data if len(c) else 0
            cols["n_cands"][k] = len(c)
            cols["out"][k] = out.ctypes.data
            keep.append((c, out))
        for f, v in cols.items():
            recs[f][:n] = v
        arr = np.ascontiguousarray(arrival_ms, dtype=np.float64)
        lat = np.zeros(max(n, 1), dtype=np.float64)
        t0 = time.perf_counter()
        rc = lib().st_server_run_trace(self._h, ptr(recs), ptr(arr), n, ptr(lat))
        self.last_trace_ms = 1e3 * (time.perf_counter() - t0)
        check(rc)
        return [out[: len(c)].copy() for c, out in keep], lat[:n].copy()

    def stats(self) -> dict:
        s = np.zeros(1, dtype=_lib.SERVER_STATS_DT)
        check(lib().st_server_stats_get(self._h, ptr(s)))
        return {k: s[k][0].item() for k in s.dtype.names}

    def close(self):
        if self._h:
            lib().st_server_destroy(self._h)
            self._h = None
            self._live.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
