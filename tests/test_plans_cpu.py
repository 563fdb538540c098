"""CPU: the plan-table parser of the native library (no GPU needed) and the
shipped table, if present."""
from __future__ import annotations

import pytest

import shgemm_amd as sg


def _lib_or_skip():
    if not sg.LIB_PATH.exists():
        pytest.skip("native library not built")
    sg.load_library()


def test_parses_good_table(tmp_path):
    _lib_or_skip()
    f = tmp_path / "p.tsv"
    f.write_text("# comment\n\n"
                 "single fp16s N N N 10 10 10 0 1 16.5\n"
                 "single bf16 T N T 7 7 20 2 256\n"
                 "single fp16s N N N 7 7 20 10 256\n"
                 "single fp16s N N N 14 8 14 11 4\n"
                 "grouped fp16 N N N 11 7 12 1 2 3.25\n")
    assert sg.check_plan_file(f) == 5


@pytest.mark.parametrize("line,why", [
    ("single fp16s N N N 10 10 10 12 1", "variant 12"),
    ("grouped fp16s N N N 10 10 10 4 1", "class 4"),
    ("single fp16x N N N 10 10 10 0 1", "mode"),
    ("single fp16s N X N 10 10 10 0 1", "op"),
    ("single fp16s N N N 10 10 10 0 0", "slices 0"),
    ("single fp16s N N N 10 10 10 0 257", "slices 257"),
    ("single fp16s N N N 10 10 0 1", "too few fields"),
    ("batched fp16s N N N 10 10 10 0 1", "kind"),
    ("single fp16s N N N 10 10 -1 0 1", "negative log2"),
])
def test_rejects_bad_rows_with_line_number(tmp_path, line, why):
    _lib_or_skip()
    f = tmp_path / "bad.tsv"
    f.write_text("# header\nsingle fp16s N N N 10 10 10 0 1\n" + line + "\n")
    with pytest.raises(sg.ShgemmError) as ei:
        sg.check_plan_file(f)
    assert ei.value.status == sg.STATUS_EINVAL
    assert "bad.tsv:3" in str(ei.value), why


def test_missing_file(tmp_path):
    _lib_or_skip()
    with pytest.raises(sg.ShgemmError):
        sg.check_plan_file(tmp_path / "nope.tsv")


def test_shipped_table_parses():
    _lib_or_skip()
    shipped = sg.LIB_PATH.parent / "plans_gfx950.tsv"
    if not shipped.exists():
        pytest.skip("no shipped plan table yet")
    n = sg.check_plan_file(shipped)
    assert n > 0
    kinds = {l.split()[0] for l in shipped.read_text().splitlines() if l.strip() and not l.startswith("#")}
    assert kinds <= {"single", "grouped"}
