"""Report reader and comparison (firepony_amd.report_diff) on the golden reports: the
verdicts the reference's diff tool gives (diffreport/diffreport.py) for identical reports,
reordered rows, small and large float differences, integer and format differences."""
import gzip
import os
import subprocess
import sys

import pytest

from firepony_amd import report_diff as rd

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def golden(name="edge_menu_600"):
    with gzip.open(os.path.join(HERE, "golden", f"{name}_report.txt.gz"), "rt") as f:
        return f.read()


def table_lines(text, name):
    """(first row line index, row count) of table `name` in report text."""
    lines = text.split("\n")
    for i, l in enumerate(lines):
        if l.startswith("#:GATKTable:") and lines[i + 1].startswith(f"#:GATKTable:{name}:"):
            return lines, i + 3, int(l.split(":")[3])
    raise KeyError(name)


def edit_field(text, name, row, col, fn):
    lines, first, n = table_lines(text, name)
    f = lines[first + row].split()
    f[col] = fn(f[col])
    lines[first + row] = "  ".join(f)
    return "\n".join(lines)


def test_parse_golden():
    tables = rd.parse(golden())
    assert [t.name for t in tables] == ["Arguments", "Quantized", "RecalTable0", "RecalTable1", "RecalTable2"]
    t1 = tables[3]
    assert t1.column_names == ["ReadGroup", "QualityScore", "EventType", "EmpiricalQuality", "Observations",
                               "Errors"]
    assert t1.formats == [(None, "s"), (None, "s"), (None, "s"), ("4", "f"), (None, "d"), ("2", "f")]
    assert all(isinstance(r[4], int) and isinstance(r[3], float) for r in t1.rows)
    keys = [(r[0], r[1], r[2], r[4]) for r in t1.rows]
    assert keys == sorted(keys)


def test_identical_and_reordered():
    a = golden()
    assert rd.diff(rd.parse(a), rd.parse(a)) == rd.DiffResult(0, ["no errors"])
    lines, first, n = table_lines(a, "RecalTable2")
    rows = lines[first:first + n]
    lines[first:first + n] = rows[::-1]
    assert rd.diff(rd.parse(a), rd.parse("\n".join(lines))).status == 0


def test_float_differences():
    a = golden()
    # 0.1 % of an EmpiricalQuality value is waived, 1 % is an error (diffreport.py:58-61)
    small = edit_field(a, "RecalTable1", 3, 3, lambda v: f"{float(v) * 1.001:.4f}")
    assert rd.diff(rd.parse(a), rd.parse(small)).status == 0
    big = edit_field(a, "RecalTable1", 3, 3, lambda v: f"{float(v) * 1.01 + 0.01:.4f}")
    r = rd.diff(rd.parse(a), rd.parse(big))
    assert r.status == 2 and r.messages[0].startswith("RecalTable1: error = ")
    # magnitudes are compared, so a sign flip alone is not a difference
    neg = edit_field(a, "RecalTable1", 3, 5, lambda v: "-" + v)
    assert rd.diff(rd.parse(a), rd.parse(neg)).status == 0


def test_structural_differences():
    a = golden()
    obs = edit_field(a, "RecalTable2", 0, 6, lambda v: str(int(v) + 1))
    r = rd.diff(rd.parse(a), rd.parse(obs))
    # rows are compared in sorted order, so the row index is the sorted one
    assert r.status == 1 and "table RecalTable2 row" in r.messages[-1] and r.messages[-1].endswith("column 6")
    fmt = a.replace("%.4f:%d:%.2f:;", "%.3f:%d:%.2f:;", 1)
    assert fmt != a
    assert rd.diff(rd.parse(a), rd.parse(fmt)).messages[-1] == "field format specifier mismatch"
    lines, first, n = table_lines(a, "RecalTable1")
    head = lines[first - 3].split(":")
    head[3] = str(n - 1)
    lines[first - 3] = ":".join(head)
    del lines[first + n - 1]
    assert rd.diff(rd.parse(a), rd.parse("\n".join(lines))).status == 1
    # Arguments and Quantized are not compared
    args = a.replace("recalibration_report  null", "recalibration_report  other", 1)
    assert rd.diff(rd.parse(a), rd.parse(args)).status == 0


@pytest.mark.parametrize("bad, why", [
    (lambda t: t.replace("#:GATKReport.v1.1:5", "#:GATKReport.v1.1", 1), "report header"),
    (lambda t: t.replace(":%s:%s:;", ":%s:%x:;", 1), "format specifier"),
    (lambda t: t.replace(":;\n#:GATKTable:Arguments", ":\n#:GATKTable:Arguments", 1), "terminator"),
])
def test_parse_errors(bad, why):
    with pytest.raises(rd.ParseError, match=why):
        rd.parse(bad(golden()))


def test_command_line(tmp_path):
    a = tmp_path / "a.txt"
    b = tmp_path / "b.txt"
    a.write_text(golden())
    b.write_text(edit_field(golden(), "RecalTable0", 0, 2, lambda v: f"{float(v) + 5:.4f}"))
    env = dict(os.environ, PYTHONPATH=ROOT)
    same = subprocess.run([sys.executable, "-m", "firepony_amd.report_diff", str(a), str(a)], env=env,
                          capture_output=True, text=True)
    assert same.returncode == 0 and same.stdout.strip() == "no errors"
    diff = subprocess.run([sys.executable, "-m", "firepony_amd.report_diff", str(a), str(b)], env=env,
                          capture_output=True, text=True)
    assert diff.returncode == 2 and "RecalTable0: error" in diff.stdout
