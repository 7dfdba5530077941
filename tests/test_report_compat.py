"""Drop-in compatibility of the report with the reference's float tables and its
rounding switch (CPU, on the oracle).

- The reference accumulates each covariate row's mismatches as a float in its
  sort/reduce order (covariate_table.cu:90-105).  The oracle restates that
  (ACCUM_REF_FLOAT) beside the exact fixed-point sums the engine keeps.  The
  rows, counts and every string and integer column agree exactly; the mismatch
  sums agree to float accumulation noise (< 1e-5 relative), and so do all
  float columns under diffreport's 0.5 % waiver (diffreport/diffreport.py:70-95)
  except where a value lies within that noise of a boundary: an empirical
  quality that steps by one where a row's exact sum is near a half-integer
  (GATK counts errors as uint64(mismatches + 0.5), empirical_quality.cu:182),
  and a printed %.2f / %.4f value whose unrounded values straddle a rounding
  boundary.  Any two float summation orders, the reference's own CPU and GPU
  paths included, differ at those rows; the test shows each such row is one.
- --disable-rounding prints every float column as %.64f in both the header and
  the rows (table_formatter.cu:106-120,189-206): each printed value must read
  back as the table's own double, bit for bit.
"""
import re

import numpy as np
import pytest

from firepony_amd import abi, report_diff, synth
from oracle.pyoracle import ACCUM_EXACT, ACCUM_REF_FLOAT, Oracle


def workload(cfg, n):
    gs, rs, _ = synth.named_spec(cfg)
    g = synth.Genome(gs, rs.num_read_groups)
    step = 1000 if cfg == 0 else 20000
    return g, [synth.Batch(g, rs, f, min(step, n - f)) for f in range(0, n, step)]


def report_of(g, batches, path, **kw):
    o = Oracle(g, **kw)
    for b in batches:
        o.process(b)
    o.postprocess()
    return o, o.report(str(path))


def max_float_deviation(ta, tb):
    worst = 0.0
    for a in ta:
        if a.name in ("Arguments", "Quantized"):
            continue
        b = next(t for t in tb if t.name == a.name)
        for j, (_, c) in enumerate(a.formats):
            if c != "f":
                continue
            for ra, rb in zip(a.rows, b.rows):
                worst = max(worst, report_diff._relative_delta(ra[j], rb[j])[1])
    return worst


def check_float_vs_exact(g, batches, tmp_path, exact_run, label):
    """exact_run(path, disable_rounding) -> (table-getter, report text) of the run with exact sums (the
    oracle's exact mode here, the HIP engine in the GPU test), compared with the oracle's restatement
    of the reference's float accumulation."""
    get_exact, exact = exact_run(tmp_path / "exact.txt", False)
    of, flt = report_of(g, batches, tmp_path / "float.txt", accumulate=ACCUM_REF_FLOAT)
    # tables: same rows and counts; float sums within a few ulps of float accumulation
    flips = 0
    worst_sum = 0.0
    for t in range(4):
        a, f = get_exact(t), of.table(t)
        np.testing.assert_array_equal(a["keys"], f["keys"])
        np.testing.assert_array_equal(a["observations"], f["observations"])
        rel = np.abs(a["mismatches"] - f["mismatches"]) / np.maximum(np.abs(a["mismatches"]), 1e-300)
        worst_sum = max(worst_sum, float(rel.max()))
        assert rel.max() < 1e-5, f"table {t}: float mismatch sums off by {rel.max()}"
        # The empirical quality counts errors as uint64(mismatches + 0.5) (empirical_quality.cu:182):
        # it can move by one only where the exact sum sits within float noise of a half-integer.
        d = np.nonzero(a["empirical_quality"] != f["empirical_quality"])[0]
        for i in d:
            m = float(a["mismatches"][i])
            assert abs(a["empirical_quality"][i] - f["empirical_quality"][i]) == 1.0
            assert abs((m + 0.5) - round(m + 0.5)) < 1e-4 * max(1.0, m), (t, i, m)
        flips += len(d)
    # The reports under diffreport's rule.  Beyond the empirical-quality steps, a printed float
    # column can differ where the unrounded values, within the same noise, straddle a rounding
    # boundary of the printed precision (%.2f errors of a small sum): the %.64f reports of both runs
    # show which.
    _, exact64 = exact_run(tmp_path / "exact64.txt", True)
    _, flt64 = report_of(g, batches, tmp_path / "float64.txt", accumulate=ACCUM_REF_FLOAT,
                         disable_output_rounding=True)
    u_exact = {t.name: t for t in report_diff.parse(exact64)}
    u_float = {t.name: t for t in report_diff.parse(flt64)}
    ta = report_diff.parse(flt, "float")
    tb = report_diff.parse(exact, "exact")
    r = report_diff.diff(ta, tb, "exact")
    worst = max_float_deviation(ta, tb)
    beyond = [m for m in r.messages if m != "no errors"]
    steps = straddles = 0
    for m in beyond:
        assert r.status == 2, m                       # never a string, integer or format mismatch
        tname = m.split(":")[0]
        row = int(re.search(r"row (\d+)", m).group(1))
        col = int(re.search(r"col (\d+)", m).group(1))
        name = u_exact[tname].column_names[col]
        va = float(u_exact[tname].rows[row][col])
        vb = float(u_float[tname].rows[row][col])
        if name == "EmpiricalQuality":
            assert abs(va - vb) == 1.0, m
            steps += 1
        else:
            assert abs(va - vb) <= 1e-5 * max(1.0, abs(va)), m
            straddles += 1
    assert steps <= flips
    print(f"{label}: mismatch sums within {worst_sum:.3g} relative; {flips} empirical-quality steps at "
          f"half-integer sums; {straddles} printed values straddling a rounding boundary; largest printed "
          f"float deviation {worst:.4f} %")


def oracle_exact_run(g, batches):
    def run(path, disable_rounding):
        o, text = report_of(g, batches, path, accumulate=ACCUM_EXACT, disable_output_rounding=disable_rounding)
        return o.table, text
    return run


@pytest.mark.parametrize("cfg,n", [(0, 20000), (1, 60000)])
def test_reference_float_accumulation_within_waiver(tmp_path, cfg, n):
    g, batches = workload(cfg, n)
    check_float_vs_exact(g, batches, tmp_path, oracle_exact_run(g, batches), f"oracle, config {cfg}, {n} reads")


def test_disable_rounding_prints_exact_doubles(tmp_path):
    g, batches = workload(0, 3000)
    o, text = report_of(g, batches, tmp_path / "r64.txt", disable_output_rounding=True)
    _, rounded = report_of(g, batches, tmp_path / "r4.txt")
    t64 = {t.name: t for t in report_diff.parse(text)}
    t4 = {t.name: t for t in report_diff.parse(rounded)}
    for name in ("RecalTable0", "RecalTable1", "RecalTable2"):
        a, b = t64[name], t4[name]
        assert a.column_names == b.column_names and a.num_rows == b.num_rows
        for (d64, c64), (d4, c4) in zip(a.formats, b.formats):
            assert c64 == c4
            assert d64 == ("64" if c4 == "f" else d4), f"{name}: float columns print as %.64f"
    # every printed %.64f value reads back as the table's double
    raw = text.split("\n")
    fmt64 = [ln for ln in raw if ln.startswith("#:GATKTable:") and "%.64f" in ln]
    assert fmt64, "no %.64f table header"
    t1 = o.table(abi.FP_TABLE_QUALITY)
    rows = t64["RecalTable1"].rows
    emp = {(int(r[1]), r[2]): (float(r[3]), int(r[4]), float(r[5])) for r in rows}
    ev = "MID"
    nonzero = t1["observations"] > 0
    for key, obs, mism, q in zip(t1["keys"][nonzero], t1["observations"][nonzero],
                                 t1["mismatches"][nonzero], t1["empirical_quality"][nonzero]):
        got = emp[((int(key) >> 2) & 255, ev[int(key) & 3])]
        assert got == (q, int(obs), mism)
    for ln in raw:
        for tok in ln.split():
            if re.fullmatch(r"-?\d+\.\d+", tok):
                assert len(tok.split(".")[1]) == 64
