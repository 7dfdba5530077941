"""The oracle's gather (the reference's gather_intermediates with exact sums):
a workload split over two oracles and gathered equals one oracle over all of it."""
import numpy as np

from firepony_amd import abi, synth
from oracle.pyoracle import Oracle


def test_gather_equals_single_pass(tmp_path):
    g, rs, n = synth.small_workload(seed=17, num_reads=2400, edge=0.2, lengths=(76, 101), read_groups=2,
                                    contigs=(80000,), sites=200, clip=0.3, indel=0.2)
    one = Oracle(g)
    a, b = Oracle(g), Oracle(g)
    for i, first in enumerate(range(0, n, 600)):
        batch = synth.Batch(g, rs, first, min(600, n - first))
        one.process(batch)
        (a if i % 2 == 0 else b).process(batch)
    a.gather(b)
    for t in (abi.FP_TABLE_QUALITY, abi.FP_TABLE_CYCLE, abi.FP_TABLE_CONTEXT):
        x, y = one.table(t), a.table(t)
        for k in ("keys", "observations", "mismatches_fixed_lo", "mismatches_fixed_hi"):
            np.testing.assert_array_equal(y[k], x[k])
    s1, s2 = one.stats(), a.stats()
    assert (s1.total_reads, s1.filtered_reads, s1.baq_reads) == (s2.total_reads, s2.filtered_reads, s2.baq_reads)
    one.postprocess()
    a.postprocess()
    assert one.report(str(tmp_path / "one.txt")) == a.report(str(tmp_path / "two.txt"))
