"""ctypes binding of the gfx950 library ``libsabertooth_amd.so``.

The product path has no CPU fallback: if the library is missing or no GPU is
usable, the calls raise.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsabertooth_amd.so")

ST_OK = 0
ST_ERR_INVALID = 1
ST_ERR_DEVICE = 2
ST_ERR_NOMEM = 3
ST_ERR_CAPACITY = 4
ST_ERR_CLOSED = 5
ST_ERR_TIMEOUT = 6
KERNEL_I16 = 0
KERNEL_I32 = 1

JOB_DT = np.dtype([("tid", "<u4"), ("diag", "<i4"), ("lo", "<u4"), ("hi", "<u4")])
REC_DT = np.dtype([("tid", "<u4"), ("hits", "<u4"), ("j0", "<u4"), ("j1", "<u4")])
CAND_DT = np.dtype([("target_id", "<u4"), ("hits", "<u4"), ("best_diag", "<i8"),
                    ("diag_score", "<i4"), ("t_lo", "<u4"), ("t_hi", "<u4")], align=True)
PAIR_RESULT_DT = np.dtype([("swg", "<i4"), ("ung", "<i4"), ("best_diag", "<i4"),
                           ("gate_score", "<i4"), ("pass", "u1"), ("kernel", "u1"),
                           ("reserved", "u1", (2,))])
LAUNCH_INFO_DT = np.dtype([("kind", "<i4"), ("rows_per_lane", "<i4"), ("units", "<u8"), ("grid", "<u8"),
                           ("useful_cells", "<u8"), ("exec_cells", "<u8"), ("ms", "<f8")])
SCORE_END_DT = np.dtype([("score", "<i4"), ("reserved", "<i4"), ("query_end", "<i8"), ("target_end", "<i8")])
SEED_DT = np.dtype([("qpos", "<u4"), ("reserved", "<u4"), ("cur", "<u8"), ("end", "<u8")])
HIT_PARAMS_DT = np.dtype([("max_diagonals", "<u4"), ("min_hits", "<u4"), ("diag_hits", "<u4"),
                          ("diag_window", "<u4"), ("span", "<u4"), ("reserved", "<u4")])
SPAN_DT = np.dtype([("score", "<i4"), ("reserved", "<i4"), ("query_start", "<i8"), ("query_end", "<i8"),
                    ("target_start", "<i8"), ("target_end", "<i8")])
SERVER_CONFIG_DT = np.dtype([("workers", "<u4"), ("max_batch_queries", "<u4"), ("state_budget_bytes", "<u8"),
                             ("persistent", "<u4"), ("arena_mib", "<u4")])
REQUEST_DT = np.dtype([("query", "<u8"), ("gap_open", "<i4"), ("gap_extend", "<i4"), ("min_diag_score", "<i4"),
                       ("reserved", "<i4"), ("cand_tids", "<u8"), ("n_cands", "<u8"), ("out", "<u8")])
SERVER_STATS_DT = np.dtype([("batches", "<u8"), ("queries", "<u8"), ("pairs", "<u8"), ("cells", "<u8"),
                            ("rescored", "<u8"), ("max_batch_queries", "<u8"), ("peak_inflight_bytes", "<u8"),
                            ("busy_ms", "<f8"), ("plan_ms", "<f8"), ("device_ms", "<f8"),
                            ("kernel_launches", "<u8"), ("published", "<u8"), ("wait_ms", "<f8"),
                            ("run_ms", "<f8")])


class StError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"sabertooth_amd error {code}: {msg}")
        self.code = code


_L = None

# (name, restype, argtypes)
_P = C.c_void_p
_SIGS = [
    ("st_last_error", C.c_char_p, []),
    ("st_version", C.c_char_p, []),
    ("st_ctx_create", C.c_int, [C.c_int, C.POINTER(_P)]),
    ("st_ctx_destroy", C.c_int, [_P]),
    ("st_blosum62", C.c_int, [_P, _P, _P, _P]),
    ("st_aa2num", C.c_int, [_P]),
    ("st_seq_rows", C.c_int, [_P, C.c_uint64, _P]),
    ("st_db_create", C.c_int, [_P, _P, C.c_uint64, _P, C.c_uint64, C.POINTER(_P)]),
    ("st_db_create_from_stidx", C.c_int, [_P, C.c_char_p, C.POINTER(_P)]),
    ("st_db_destroy", C.c_int, [_P]),
    ("st_db_targets", C.c_uint64, [_P]),
    ("st_db_residues", C.c_uint64, [_P]),
    ("st_qprof_create", C.c_int, [_P, _P, C.c_uint64, C.POINTER(_P)]),
    ("st_qprof_destroy", C.c_int, [_P]),
    ("st_qprof_len", C.c_uint64, [_P]),
    ("st_sw_score_batch", C.c_int, [_P, _P, _P, _P, C.c_uint64, C.c_int32, C.c_int32, _P, _P]),
    ("st_sw_score_end_batch", C.c_int, [_P, _P, _P, _P, C.c_uint64, C.c_int32, C.c_int32, _P]),
    ("st_align_spans_batch", C.c_int, [_P, _P, _P, C.c_uint64, _P, _P, _P, _P, _P]),
    ("st_postings_create", C.c_int, [_P, _P, C.c_uint64, C.POINTER(_P)]),
    ("st_postings_destroy", C.c_int, [_P]),
    ("st_postings_from_stidx", C.c_int, [_P, C.c_char_p, C.POINTER(_P)]),
    ("st_postings_count", C.c_uint64, [_P]),
    ("st_seed_hits_to_jobs", C.c_int, [_P, _P, _P, C.c_uint64, _P, C.c_uint64, _P, _P, C.c_uint64,
                                       C.POINTER(C.c_uint64), _P, C.c_uint64, C.POINTER(C.c_uint64)]),
    ("st_seed_hits_to_jobs_batch", C.c_int, [_P, _P, _P, C.c_uint64, _P, _P, _P, _P, _P, C.c_uint64,
                                             C.POINTER(C.c_uint64), _P, C.c_uint64, C.POINTER(C.c_uint64), _P,
                                             _P]),
    ("st_seed_hits_to_candidates_batch", C.c_int, [_P, _P, _P, C.c_uint64, _P, _P, _P, _P, _P, _P, _P, C.c_int,
                                                   _P, C.c_uint64, C.POINTER(C.c_uint64), _P,
                                                   C.POINTER(C.c_uint64)]),
    ("st_ungapped_score_jobs", C.c_int, [_P, _P, _P, _P, C.c_uint64, _P]),
    ("st_score_and_gate", C.c_int, [_P, _P, _P, _P, C.c_uint64, _P, C.c_uint64, C.c_int32, C.c_int32,
                                    C.c_int32, C.c_int, _P, _P, _P, _P]),
    ("st_score_and_gate_noalloc", C.c_int, [_P, _P, _P, _P, C.c_uint64, _P, C.c_uint64, C.c_int32, C.c_int32,
                                            C.c_int32, C.c_int, _P, _P, _P, _P]),
    ("st_seed_stage_timing", C.c_int, [_P, _P]),
    ("st_size_adjusted_gate", C.c_int32, [C.c_int32, C.c_uint64]),
    ("st_batch_create", C.c_int, [_P, _P, _P, C.c_uint64, _P, _P, _P, _P, _P, C.POINTER(_P)]),
    ("st_batch_run", C.c_int, [_P]),
    ("st_batch_run_graph", C.c_int, [_P]),
    ("st_shape_table_status", C.c_int, []),
    ("st_batch_results", C.c_int, [_P, _P]),
    ("st_batch_timing", C.c_int, [_P, _P, _P]),
    ("st_batch_stats", C.c_int, [_P, _P, _P, _P]),
    ("st_batch_launches", C.c_int, [_P, _P, C.c_uint64, C.POINTER(C.c_uint64)]),
    ("st_align_spans_launches", C.c_int, [_P, _P, C.c_uint64, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    ("st_align_spans_phases", C.c_int, [_P, _P]),
    ("st_batch_cells", C.c_uint64, [_P]),
    ("st_batch_pairs", C.c_uint64, [_P]),
    ("st_batch_destroy", C.c_int, [_P]),
    ("st_server_create", C.c_int, [_P, _P, _P, C.POINTER(_P)]),
    ("st_server_submit", C.c_int, [_P, _P, C.POINTER(C.c_uint64)]),
    ("st_server_poll", C.c_int, [_P, C.c_uint64, C.POINTER(C.c_int)]),
    ("st_server_wait", C.c_int, [_P, C.c_uint64, C.c_double, C.POINTER(C.c_double)]),
    ("st_server_run_trace", C.c_int, [_P, _P, _P, C.c_uint64, _P]),
    ("st_server_stats_get", C.c_int, [_P, _P]),
    ("st_server_destroy", C.c_int, [_P]),
]

EXPORTED = [s[0] for s in _SIGS]


def lib() -> C.CDLL:
    """Load the library (raises if it has not been built)."""
    global _L
    if _L is None:
        if not os.path.exists(LIB_PATH):
            raise StError(ST_ERR_DEVICE, f"{LIB_PATH} is missing: build it with "
                                         "`python -m sabertooth_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, res, args in _SIGS:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _L = L
    return _L


def check(rc: int) -> None:
    if rc != ST_OK:
        raise StError(rc, lib().st_last_error().decode())


def ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)
