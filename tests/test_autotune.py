"""The planner's launch-shape table: the autotuner's file format and the
library's checks on the file ST_SHAPE_TABLE names (host only)."""
import pytest

import sabertooth_amd as st
from sabertooth_amd import autotune


def _status(monkeypatch, path):
    monkeypatch.setenv("ST_SHAPE_TABLE", str(path))
    rc = st.lib().st_shape_table_status()
    return rc, st.lib().st_last_error().decode()


def test_written_tables_are_accepted(tmp_path, monkeypatch):
    costs = {(R, G): 0.2 + 0.01 * R + 0.001 * G for R in autotune.ROWS for G in autotune.GROUPS}
    p = tmp_path / "t.txt"
    p.write_text(autotune.table_text(costs, "test"))
    assert _status(monkeypatch, p)[0] == 0
    # a shape left unmeasured gets a cost the planner never prefers
    del costs[(16, 64)]
    txt = autotune.table_text(costs)
    assert float(txt.strip().split("\n")[-1].split()[-1]) > 10 * max(costs.values())


def test_malformed_tables_are_refused(tmp_path, monkeypatch):
    monkeypatch.delenv("ST_SHAPE_TABLE", raising=False)
    assert st.lib().st_shape_table_status() == 0
    cases = {"short.txt": "1 2 3\n", "neg.txt": " ".join(["0.3"] * 20 + ["-1"]),
             "word.txt": " ".join(["0.3"] * 20 + ["x"]), "long.txt": " ".join(["0.3"] * 22)}
    for name, text in cases.items():
        p = tmp_path / name
        p.write_text(text)
        rc, msg = _status(monkeypatch, p)
        assert rc == 1, name  # ST_ERR_INVALID
        assert "ST_SHAPE_TABLE" in msg
    rc, msg = _status(monkeypatch, tmp_path / "missing.txt")
    assert rc == 1 and "cannot read" in msg
    # comments and commas are allowed
    p = tmp_path / "ok.txt"
    p.write_text("# c\n" + ",".join(["0.5"] * 7) + "  # row\n" + "\n".join([" ".join(["0.25"] * 7)] * 2))
    assert _status(monkeypatch, p)[0] == 0
