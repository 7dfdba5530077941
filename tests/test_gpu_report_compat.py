"""The HIP engine's report beside the reference's float tables and its rounding switch.

- Exact sums (the engine) against the oracle's restatement of the reference's float
  accumulation (covariate_table.cu:90-105), on config-0 and config-1 slices, under the rule of
  tests/test_report_compat.py: identical rows, counts and string/integer columns, float columns
  within diffreport's waiver except at rows sitting on a rounding boundary, each shown to be one.
- --disable-rounding (%.64f, table_formatter.cu:106-120,189-206): the engine's report is byte for
  byte the oracle's, on both engines.
"""
import sys

import pytest

from firepony_amd import abi
from firepony_amd.engine import Engine
from oracle.pyoracle import Oracle

from test_report_compat import check_float_vs_exact, workload

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def engine_run(g, batches, **opts):
    def run(path, disable_rounding):
        e = Engine(g, disable_output_rounding=int(disable_rounding), **opts)
        for b in batches:
            e.process(b)
        e.postprocess()
        text = e.report(str(path))
        tables = {t: e.table(t) for t in range(4)}
        e.close()
        return tables.__getitem__, text
    return run


@pytest.mark.parametrize("cfg,n", [(0, 20000), (1, 60000)])
def test_engine_exact_beside_reference_float(tmp_path, cfg, n):
    say(f"report compatibility: engine (exact sums) vs the reference's float accumulation, config {cfg}, {n} reads")
    g, batches = workload(cfg, n)
    check_float_vs_exact(g, batches, tmp_path, engine_run(g, batches, engine_kind=abi.FP_ENGINE_BATCHED,
                                                          max_read_length=101, baq_waves_per_simd=1),
                         f"engine, config {cfg}, {n} reads")


@pytest.mark.parametrize("kind", [abi.FP_ENGINE_PLAIN, abi.FP_ENGINE_BATCHED])
def test_disable_rounding_report_equals_oracle(tmp_path, kind):
    say(f"report: %.64f output, {'batched' if kind else 'plain'} engine vs oracle")
    g, batches = workload(1, 40000)
    o = Oracle(g, disable_output_rounding=True)
    e = Engine(g, disable_output_rounding=1, engine_kind=kind, **({"max_read_length": 101} if kind else {}))
    for b in batches:
        o.process(b)
        e.process(b)
    o.postprocess()
    e.postprocess()
    ro = o.report(str(tmp_path / "o.txt"))
    re_ = e.report(str(tmp_path / "e.txt"))
    assert "%.64f" in ro
    assert re_ == ro, "%.64f report differs from the oracle's"
    e.close()
