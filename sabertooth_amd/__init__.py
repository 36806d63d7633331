"""sabertooth_amd — Sabertooth's post-prefilter scoring path on MI355X (gfx950).

The ungapped diagonal evidence, the gate and the exact Smith-Waterman-Gotoh
score run as hand-written HIP kernels behind the C ABI declared in
``include/sabertooth_amd.h``; this package is a thin Python mirror of the
reference's Rust API over that library (see ``core``).
"""
from ._lib import (CAND_DT, JOB_DT, KERNEL_I16, KERNEL_I32, PAIR_RESULT_DT, REC_DT, SPAN_DT, StError,
                   lib)
from .core import (Batch, Context, GapCosts, QueryProfile, Server, SpansCall, TargetDb, aa2num, align_spans,
                   align_spans_launches, align_spans_phases, blosum62, encode, score_and_gate, score_batch, seq_rows,
                   size_adjusted_gate)

__all__ = [
    "Batch", "Context", "GapCosts", "QueryProfile", "Server", "SpansCall", "TargetDb", "aa2num", "align_spans",
    "align_spans_launches", "align_spans_phases", "blosum62",
    "encode", "score_and_gate", "score_batch", "seq_rows", "size_adjusted_gate", "StError", "lib",
    "CAND_DT", "JOB_DT", "REC_DT", "PAIR_RESULT_DT", "SPAN_DT", "KERNEL_I16", "KERNEL_I32",
]
