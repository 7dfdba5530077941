"""gather_intermediates on the GPU (fp_engine_gather): engines that each took part of the
batches, merged into one, give the tables and report of a single engine that took them all,
whatever the split, the order of the merges and the engine forms; the reference's multi-device
flow (device/pipeline_interface.cu:52-58,137-160) with exact sums.  Both are checked against the
oracle run over all the batches, and against oracles that each took the same part and were
gathered the same way (fpo_gather)."""
import sys

import numpy as np

from firepony_amd import abi, synth
from firepony_amd.engine import Engine
from oracle.pyoracle import Oracle

import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def same_tables(a, b, what):
    for t in (abi.FP_TABLE_QUALITY, abi.FP_TABLE_CYCLE, abi.FP_TABLE_CONTEXT):
        x, y = a.table(t), b.table(t)
        for k in ("keys", "observations", "mismatches_fixed_lo", "mismatches_fixed_hi"):
            np.testing.assert_array_equal(y[k], x[k], err_msg=f"{what}: table {t} {k}")


def test_gather_equals_one_engine(tmp_path):
    say("gather: three engines (plain, plain, batched) with interleaved batches merged into one")
    g, rs, n = synth.small_workload(seed=17, num_reads=9000, edge=0.1, lengths=(76, 101, 151), read_groups=3,
                                    contigs=(200000, 100000), sites=500, clip=0.3, indel=0.2, paired=1)
    batches = [synth.Batch(g, rs, f, min(1000, n - f)) for f in range(0, n, 1000)]
    one = Engine(g)
    for b in batches:
        one.process(b)
    parts = [Engine(g), Engine(g),
             Engine(g, engine_kind=abi.FP_ENGINE_BATCHED, max_batch_reads=2048, max_read_length=151)]
    for i, b in enumerate(batches):
        parts[i % 3].process(b)
    # merge 2 into 1, then 1 into 0: any order gives the same exact sums
    parts[1].gather(parts[2])
    parts[0].gather(parts[1])
    same_tables(one, parts[0], "gathered")
    assert parts[0].stats().total_reads == one.stats().total_reads == n
    o = Oracle(g)
    for b in batches:
        o.process(b)
    same_tables(o, parts[0], "gathered engines vs the oracle over all batches")
    ops = [Oracle(g) for _ in range(3)]
    for i, b in enumerate(batches):
        ops[i % 3].process(b)
    ops[1].gather(ops[2])
    ops[0].gather(ops[1])
    same_tables(ops[0], parts[0], "gathered engines vs gathered oracles")
    one.postprocess()
    parts[0].postprocess()
    o.postprocess()
    ro = o.report(str(tmp_path / "oracle.txt"))
    assert one.report(str(tmp_path / "one.txt")) == ro
    assert parts[0].report(str(tmp_path / "gathered.txt")) == ro
    for e in parts + [one]:
        e.close()
