"""The oracle reproduces the committed golden fixtures (tests/golden): a
regression pin of the oracle and of the seeded generator."""
import pytest

from golden_check import check
from oracle.pyoracle import Oracle


@pytest.mark.parametrize("name", ["config0_first1000", "edge_menu_600"])
def test_oracle_matches_golden(tmp_path, name):
    check(name, lambda g, opts: Oracle(g, frac_error_f64=bool(opts.get("frac_error_f64", 0))), tmp_path)
