"""The HIP path reproduces the committed golden fixtures (tests/golden) bit
for bit: intermediates, tables and report, with no oracle run at test time."""
import sys

import pytest

from firepony_amd import abi
from firepony_amd.engine import Engine
from golden_check import check

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["config0_first1000", "edge_menu_600"])
@pytest.mark.parametrize("agg", [abi.FP_AGG_DIRECT, abi.FP_AGG_SORT])
def test_engine_matches_golden(tmp_path, name, agg):
    print(f"golden: {name}, aggregation {agg}", flush=True)
    sys.stdout.flush()
    check(name, lambda g, opts: Engine(g, aggregation=agg, **opts), tmp_path)
