"""GATKReport v1.1 reader and report comparison, with the semantics of the reference's
report diff tool (diffreport/diffreport.py, gatkreport.py, gatktable.py, textfile.py).

    python -m firepony_amd.report_diff <report-A> <report-B>

Exit status: 0 when every table agrees ("no errors"), 1 on a structural mismatch (a format
specifier or a string/integer field differs), 2 when a floating-point field differs by 0.5 %
or more of the larger magnitude.  Tables are matched by name; Arguments and Quantized are not
compared (diffreport.py:143-149); each table's rows are sorted by its string and integer
columns before comparison (gatktable.py:111-121).  Where the reference's tool fails on an
exception (report B with fewer rows than A), this one reports a mismatch.
"""
import re
import sys
from dataclasses import dataclass, field

_FMT = re.compile(r"^%(\.([0-9]+))?([sfd])$")


class ParseError(ValueError):
    def __init__(self, where, lineno, msg):
        super().__init__(f"Parse error on {where}:{lineno}: {msg}")


class _Lines:
    """Lines with trailing whitespace stripped, counted (textfile.py:2-22)."""

    def __init__(self, text, where):
        self.lines = text.split("\n")
        self.i = 0
        self.where = where

    def readline(self):
        s = self.lines[self.i].rstrip() if self.i < len(self.lines) else ""
        self.i += 1
        return s

    def error(self, msg):
        return ParseError(self.where, self.i, msg)


@dataclass
class Table:
    name: str = ""
    description: str = ""
    formats: list = field(default_factory=list)       # (digits or None, 's' | 'd' | 'f')
    column_names: list = field(default_factory=list)
    rows: list = field(default_factory=list)

    @property
    def num_rows(self):
        return len(self.rows)


def _data_fields(fp, ncols):
    # leading whitespace dropped, whitespace runs are one separator (gatktable.py:6-17)
    f = re.split(r"\s+", re.sub(r"^\s+", "", fp.readline()))
    if len(f) != ncols:
        raise fp.error("invalid number of columns")
    return f


def _parse_table(fp):
    t = Table()
    head = fp.readline().split(":")
    if len(head) < 6:
        raise fp.error("not enough fields in table header")
    if head[0] != "#" or head[1] != "GATKTable":
        raise fp.error("invalid table header")
    if head[-1] != ";":
        raise fp.error("table header terminator not found")
    try:
        ncols, nrows = int(head[2]), int(head[3])
    except ValueError:
        raise fp.error("invalid column or row count in table header")
    for spec in head[4:-1]:
        m = _FMT.match(spec)
        if m is None:
            raise fp.error(f"error parsing field format specifier '{spec}'")
        t.formats.append((m.group(2), m.group(3)))
    desc = fp.readline().split(":")
    if len(desc) != 4 or desc[0] != "#" or desc[1] != "GATKTable":
        raise fp.error("invalid table description")
    t.name, t.description = desc[2], desc[3]
    t.column_names = _data_fields(fp, ncols)
    conv = {"s": str, "d": int, "f": float}
    for _ in range(nrows):
        f = _data_fields(fp, ncols)
        try:
            t.rows.append([conv[c](f[i]) for i, (_, c) in enumerate(t.formats)])
        except ValueError as e:
            raise fp.error(str(e))
    if fp.readline() != "":
        raise fp.error("table delimiter not found")
    keys = [i for i, (_, c) in enumerate(t.formats) if c in "sd"]
    if keys:
        t.rows.sort(key=lambda r: tuple(r[i] for i in keys))
    return t


def parse(text, where="<report>"):
    """Tables of a GATKReport v1.1 text (gatkreport.py:5-26)."""
    fp = _Lines(text, where)
    head = fp.readline().split(":")
    if len(head) != 3 or head[0] != "#" or not head[1].startswith("GATKReport."):
        raise fp.error("invalid report header")
    try:
        n = int(head[2])
    except ValueError:
        raise fp.error("invalid table count in report header")
    return [_parse_table(fp) for _ in range(n)]


def load(path):
    with open(path) as f:
        return parse(f.read(), path)


def _relative_delta(a, b):
    # magnitudes only, as diffreport.py:99-105 computes them
    hi, lo = max(abs(a), abs(b)), min(abs(a), abs(b))
    delta = hi - lo
    return delta, (delta / hi * 100.0 if delta != 0.0 and hi != 0.0 else 0.0)


@dataclass
class DiffResult:
    status: int
    messages: list


def diff(tables_a, tables_b, name_b="B", waive_percent=0.5):
    """Compare two parsed reports (diffreport.py:136-183)."""
    out = []
    status = 0
    by_name = {}
    for t in tables_b:
        by_name.setdefault(t.name, t)
    for ta in tables_a:
        if ta.name in ("Arguments", "Quantized"):
            continue
        tb = by_name.get(ta.name)
        if tb is None:
            continue
        if ta.formats != tb.formats:
            return DiffResult(1, out + ["field format specifier mismatch"])
        if tb.num_rows < ta.num_rows:
            return DiffResult(1, out + [f"Mismatch on file {name_b} table {ta.name}: "
                                        f"{tb.num_rows} rows, expected {ta.num_rows}"])
        exact = [j for j, (_, c) in enumerate(ta.formats) if c != "f"]
        measured = [j for j, (_, c) in enumerate(ta.formats) if c == "f"]
        for i in range(ta.num_rows):
            ra, rb = ta.rows[i], tb.rows[i]
            for j in exact:
                if ra[j] != rb[j]:
                    return DiffResult(1, out + [f"Mismatch on file {name_b} table {ta.name} row {i} column {j}"])
            for j in measured:
                delta, rel = _relative_delta(ra[j], rb[j])
                if not rel < waive_percent:
                    out.append(f"{ta.name}: error = {rel:.8f}% (row {i} col {j} left={ra[j]:.8f} "
                               f"right={rb[j]:.8f} abs delta={delta:.8f})\n"
                               f" left:  {' '.join(str(x) for x in ra)}\n"
                               f" right:  {' '.join(str(x) for x in rb)}")
                    status = 2
    if status == 0:
        out.append("no errors")
    return DiffResult(status, out)


def diff_files(path_a, path_b, waive_percent=0.5):
    return diff(load(path_a), load(path_b), path_b, waive_percent)


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    if len(argv) != 2:
        print("Usage: python -m firepony_amd.report_diff <report-1> <report-2>")
        return 1
    r = diff_files(argv[0], argv[1])
    for m in r.messages:
        print(m)
    return r.status


if __name__ == "__main__":
    sys.exit(main())
