"""Comparison of a pipeline run (oracle or engine) with a golden fixture of
tests/golden (written by tests/golden/make_golden.py)."""
import gzip
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
sys.path.insert(0, GOLDEN)

import make_golden  # noqa: E402
from firepony_amd import synth  # noqa: E402


def load(name):
    fx = dict(np.load(os.path.join(GOLDEN, f"{name}.npz")))
    with gzip.open(os.path.join(GOLDEN, f"{name}_report.txt.gz"), "rt") as f:
        fx["report"] = f.read()
    return fx


def check(name, make_runner, tmp_path):
    """Run the fixture's batch through make_runner(genome, options) and compare every array the
    fixture holds: per-read values over the active reads (HMM window and band over the BAQ reads),
    per-base and per-quality values over the active reads' bases, the tables and the report."""
    case = next(c for c in make_golden.cases() if c[0] == name)
    _, g, rs, first, count, opts = case
    b = synth.Batch(g, rs, first, count)
    fx = load(name)
    assert make_golden.input_digest(b) == str(fx["input_sha256"]), "the seeded generator changed its output"
    r = make_runner(g, opts)
    r.process(b)
    st = r.stats()
    np.testing.assert_array_equal(np.array([st.total_reads, st.filtered_reads, st.baq_reads], np.uint64),
                                  fx["stats"], err_msg="read counts")
    got = make_golden.collect(r, b, str(tmp_path / "report.txt"))
    np.testing.assert_array_equal(got["active_reads"], fx["active_reads"], err_msg="active reads")
    act = fx["active_reads"].astype(np.int64)
    baq_reads = act[fx["num_errors"][act] != 0]
    for _, _, key in make_golden.PER_READ[1:]:
        idx = baq_reads if key in ("hmm_window", "bandwidth") else act
        np.testing.assert_array_equal(got[key][idx], fx[key][idx], err_msg=key)
    arr = b.arrays()
    rs_, rl = arr["read_start"].astype(np.int64), arr["read_len"].astype(np.int64)
    qs = arr["qual_start"].astype(np.int64)
    bidx = np.concatenate([np.arange(rs_[x], rs_[x] + rl[x]) for x in act])
    qidx = np.concatenate([np.arange(qs[x], qs[x] + rl[x]) for x in act])
    for _, _, key in make_golden.PER_BASE:
        np.testing.assert_array_equal(got[key][bidx], fx[key][bidx], err_msg=key)
    np.testing.assert_array_equal(got["baq"][qidx], fx["baq"][qidx], err_msg="BAQ tags")
    np.testing.assert_array_equal(got["frac_errors"].reshape(-1, 3)[qidx], fx["frac_errors"].reshape(-1, 3)[qidx],
                                  err_msg="fractional errors")
    for _, tname in make_golden.TABLES:
        for k in ("keys", "observations", "mismatches_fixed_lo", "mismatches_fixed_hi"):
            np.testing.assert_array_equal(got[f"{tname}_{k}"], fx[f"{tname}_{k}"], err_msg=f"{tname} {k}")
    assert got["report"] == fx["report"], "report text"
