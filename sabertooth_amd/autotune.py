"""Launch-shape autotuner for the DP planner.

The planner gives each work unit the (rows per lane R, lanes per group G)
shape with the least modelled device time: executed cells times the shape's
device cost per executed cell.  This module measures those costs on the
device it runs on: for every shape it forces the planner to that shape
(``ST_FORCE_R`` / ``ST_FORCE_G``), runs uniform workloads in which every unit
has the shape (queries of 96, 300 and 1,200 residues against 3,000 targets of
about 150 residues each), and keeps the median over the three query lengths
of the best of three runs.  It writes the table as the planner reads it from
the file ``ST_SHAPE_TABLE`` names: 21 values, rows R = 8, 12, 16 and columns
G = 1, 2, 4, ..., 64, picoseconds per executed cell.

    python -m sabertooth_amd.autotune --out shape_table.txt [--quick]
    ST_SHAPE_TABLE=shape_table.txt python bench.py
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np

ROWS = (8, 12, 16)
GROUPS = (1, 2, 4, 8, 16, 32, 64)
WORKLOADS = ((96, 256), (300, 128), (1200, 32))  # (query residues, queries)


def measure(quick: bool = False, log=print):
    """{(R, G): ps per executed cell} on the current device."""
    import sabertooth_amd as st
    from sabertooth_amd import synth
    s = synth.Stream(5)
    ctx = st.Context(0)
    nt = 40000
    lens = np.full(nt, 150, np.int64) + (np.arange(nt) % 21) - 10
    arena = s.background(int(lens.sum()))
    offs = np.zeros(nt + 1, np.uint64)
    offs[1:] = np.cumsum(lens)
    db = st.TargetDb(ctx, arena, offs)
    rng = np.random.default_rng(1)
    per = {}
    saved = {k: os.environ.get(k) for k in ("ST_FORCE_R", "ST_FORCE_G", "ST_SHAPE_TABLE")}
    os.environ.pop("ST_SHAPE_TABLE", None)
    try:
        for m, nq in WORKLOADS:
            nq = max(4, nq // 4) if quick else nq
            qs = [s.background(m) for _ in range(nq)]
            qps = [st.QueryProfile.from_rows(ctx, st.seq_rows(q)) for q in qs]
            cands = [rng.choice(nt, size=3000, replace=False).astype(np.uint32) for _ in qs]
            for R in ROWS:
                for G in GROUPS:
                    if 22 * (G * R + 2) * 2 > 65536:  # the pass's profile slice must fit 64 KB of LDS
                        continue
                    os.environ["ST_FORCE_R"], os.environ["ST_FORCE_G"] = str(R), str(G)
                    b = st.Batch(ctx, db, qps, cands, [st.GapCosts(11, 1)] * nq, [41] * nq)
                    exec_cells = b.stats()[0]
                    b.run()
                    best = min(_timed(b) for _ in range(3))
                    b.close()
                    ps = best * 1e9 / exec_cells
                    per.setdefault((R, G), []).append(ps)
                    log(f"m {m} R {R} G {G}: {best:.3f} ms, {ps:.3f} ps per executed cell")
            for q in qps:
                q.close()
    finally:
        for k, v in saved.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    db.close()
    return {k: float(np.median(v)) for k, v in per.items()}


def _timed(b) -> float:
    b.run()
    return b.timing()[0]


def table_text(costs, note: str = "") -> str:
    """The planner's table file for measured costs (shapes that were not
    measured get a cost that the planner never prefers)."""
    worst = max(costs.values()) * 100.0 if costs else 100.0
    lines = ["# launch-shape costs: picoseconds per executed DP cell (sabertooth_amd.autotune)"]
    if note:
        lines.append("# " + note)
    lines.append("# rows: R = 8, 12, 16; columns: G = " + " ".join(str(g) for g in GROUPS))
    for R in ROWS:
        lines.append(" ".join(f"{costs.get((R, G), worst):.4f}" for G in GROUPS))
    return "\n".join(lines) + "\n"


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--out", default="shape_table.txt")
    ap.add_argument("--quick", action="store_true", help="a quarter of the queries per workload")
    a = ap.parse_args(argv)
    t0 = time.time()
    costs = measure(a.quick, log=lambda s: print(s, file=sys.stderr, flush=True))
    from sabertooth_amd import _lib
    note = f"measured in {time.time() - t0:.0f} s{' (quick)' if a.quick else ''}; library {_lib.lib().st_version().decode()}"
    text = table_text(costs, note)
    with open(a.out, "w") as f:
        f.write(text)
    print(text, end="")


if __name__ == "__main__":
    main()
