"""fp_engine_process_batch_device_noalloc (the entry point that allocates no device memory) against
fp_engine_process_batch_device on the same device batches: two calls in a row, the second smaller,
with bands of every class (15/23/33 cells and wider); per-batch intermediates, tables and report
bit for bit, and the oracle's tables.  Batches beyond the entry point's limits are refused."""
import sys

import numpy as np
import pytest

from firepony_amd import abi, synth
from firepony_amd.engine import Engine, FireponyError
from oracle.pyoracle import Oracle

from devbatch import DeviceCopy

pytestmark = pytest.mark.gpu

TABLES = (abi.FP_TABLE_QUALITY, abi.FP_TABLE_CYCLE, abi.FP_TABLE_CONTEXT)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def same_intermediates(a, z, batch, what):
    act = a.debug(abi.FP_DBG_ACTIVE_READS, np.uint32)
    np.testing.assert_array_equal(z.debug(abi.FP_DBG_ACTIVE_READS, np.uint32), act, err_msg=f"{what}: active reads")
    act = act.astype(np.int64)
    for item, dt in ((abi.FP_DBG_READ_WINDOW, np.uint32), (abi.FP_DBG_NUM_ERRORS, np.uint16),
                     (abi.FP_DBG_HMM_WINDOW, np.uint32), (abi.FP_DBG_BANDWIDTH, np.uint16),
                     (abi.FP_DBG_HQ_WINDOW, np.uint32)):
        np.testing.assert_array_equal(z.debug(item, dt)[act], a.debug(item, dt)[act], err_msg=f"{what}: item {item}")
    arr = batch.arrays()
    qs, rl = arr["qual_start"].astype(np.int64), arr["read_len"].astype(np.int64)
    qidx = np.concatenate([np.arange(qs[r], qs[r] + rl[r]) for r in act])
    np.testing.assert_array_equal(z.debug(abi.FP_DBG_BAQ, np.uint8)[qidx], a.debug(abi.FP_DBG_BAQ, np.uint8)[qidx],
                                  err_msg=f"{what}: BQ tags")


def test_noalloc_equals_allocating_entry_point(tmp_path):
    say("allocation-free entry point vs fp_engine_process_batch_device: 20000 then 6000 reads")
    g, rs, n = synth.small_workload(seed=61, num_reads=26000, lengths=(50, 101, 151, 250), read_groups=3,
                                    contigs=(400000, 200000), sites=500, clip=0.3, indel=0.25)
    rs.indel_max = 24            # deletions up to 24 bp: bands wider than 33 cells (the generic class)
    batches = [synth.Batch(g, rs, 0, 20000), synth.Batch(g, rs, 20000, 6000)]
    a = Engine(g)                # plain engine, allocating entry point
    z = Engine(g)                # plain engine, allocation-free entry point
    o = Oracle(g)
    widest = 0
    for k, b in enumerate(batches):
        d = DeviceCopy(b)
        say(f"  call {k + 1}: {b.view.num_reads} reads")
        a.process_device(d.view)
        a.synchronize()
        z.process_device_noalloc(d.view)
        o.process(b)
        same_intermediates(a, z, b, f"call {k + 1}")
        widest = max(widest, int(z.debug(abi.FP_DBG_BANDWIDTH, np.uint16).max()))
        for t in TABLES:
            x, y = a.table(t), z.table(t)
            for key in ("keys", "observations", "mismatches_fixed_lo", "mismatches_fixed_hi"):
                np.testing.assert_array_equal(y[key], x[key], err_msg=f"after call {k + 1}: table {t} {key}")
        d.close()
    assert widest > 16, "no band wider than 33 cells ran"
    sa, sz = a.stats(), z.stats()
    assert (sz.total_reads, sz.filtered_reads, sz.baq_reads, sz.baq_cells) == \
           (sa.total_reads, sa.filtered_reads, sa.baq_reads, sa.baq_cells)
    for t in TABLES:
        x, y = o.table(t), z.table(t)
        for key in ("keys", "observations", "mismatches_fixed_lo", "mismatches_fixed_hi"):
            np.testing.assert_array_equal(y[key], x[key], err_msg=f"oracle: table {t} {key}")
    a.postprocess()
    z.postprocess()
    o.postprocess()
    rz = z.report(str(tmp_path / "z.txt"))
    assert rz == a.report(str(tmp_path / "a.txt"))
    assert rz == o.report(str(tmp_path / "o.txt"))
    a.close()
    z.close()


def test_noalloc_refuses_what_it_cannot_hold():
    say("allocation-free entry point: refused batches and engines")
    g, rs, n = synth.small_workload(seed=62, num_reads=70000, lengths=(100,), read_groups=1, contigs=(200000,))
    big = synth.Batch(g, rs, 0, 70000)
    d = DeviceCopy(big)
    z = Engine(g)
    with pytest.raises(FireponyError):
        z.process_device_noalloc(d.view)              # more than 65,536 reads: refused on the host
    batched = Engine(g, engine_kind=abi.FP_ENGINE_BATCHED, max_read_length=100)
    with pytest.raises(FireponyError):
        batched.process_device_noalloc(d.view)        # plain engines only
    sort = Engine(g, aggregation=abi.FP_AGG_SORT)
    with pytest.raises(FireponyError):
        sort.process_device_noalloc(d.view)           # direct aggregation only
    for e in (z, batched, sort):
        assert e.stats().total_reads == 0
        e.close()
    d.close()
