"""Error recovery: a batch that fails a check costs only its own staged reads.

A read longer than max_read_length and a read group beyond the tables are
errors.  On the batched engine they surface when the staging buffer holding
the batch runs: that buffer's reads are dropped before they touch the tables,
the error is reported once, and the engine goes on with the next batches or
after fp_engine_reset.  A serving job's batch is checked at its own submit, so
it never reaches a buffer shared with other jobs.  Every case ends with the
engine's tables compared with the oracle over exactly the batches that ran.
"""
import sys

import numpy as np
import pytest

from firepony_amd import abi, synth
from firepony_amd.engine import Engine, FireponyError
from oracle.pyoracle import Oracle

from devbatch import DeviceCopy

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def workload():
    return synth.small_workload(seed=41, num_reads=6000, lengths=(100,), read_groups=2,
                                contigs=(200000,), sites=400, clip=0.1, indel=0.05)


def variant(rs, **fields):
    v = abi.FpsynReadSpec.from_buffer_copy(rs)
    for k, x in fields.items():
        if k == "length":
            v.num_lengths, v.lengths[0] = 1, x
        else:
            setattr(v, k, x)
    return v


def same_as_oracle(g, e, batches, what, counts=True):
    o = Oracle(g)
    for b in batches:
        o.process(b)
    for t in (abi.FP_TABLE_QUALITY, abi.FP_TABLE_CYCLE, abi.FP_TABLE_CONTEXT):
        a, x = o.table(t), e.table(t)
        for k in ("keys", "observations", "mismatches_fixed_lo", "mismatches_fixed_hi"):
            np.testing.assert_array_equal(x[k], a[k], err_msg=f"{what}: table {t} {k}")
    if counts:
        st, so = e.stats(), o.stats()
        assert (st.total_reads, st.filtered_reads, st.baq_reads) == \
            (so.total_reads, so.filtered_reads, so.baq_reads), what
    return o


@pytest.mark.parametrize("kind", ["too_long", "bad_read_group"])
def test_batched_drops_the_failed_buffer_and_goes_on(kind):
    say(f"recovery: batched engine, device batches, {kind}")
    g, rs, n = workload()
    good = [synth.Batch(g, rs, f, 1000) for f in range(0, n, 1000)]
    bad = synth.Batch(g, variant(rs, length=160) if kind == "too_long" else variant(rs, num_read_groups=3, seed=99),
                      0, 500)
    dgood = [DeviceCopy(b) for b in good]
    dbad = DeviceCopy(bad)
    # staging buffers of 16,384 reads: an idle engine runs a buffer once it holds 2,048 reads, so
    # good 0, good 1 and the bad batch share one buffer, dropped as a whole when it runs
    e = Engine(g, engine_kind=abi.FP_ENGINE_BATCHED, max_batch_reads=16384, max_read_length=120,
               baq_waves_per_simd=1)
    e.process_device(dgood[0].view)
    e.process_device(dgood[1].view)
    e.process_device(dbad.view)
    with pytest.raises(FireponyError):
        e.synchronize()
    # the error is not sticky: later batches run and only they reach the tables
    for d in dgood[2:]:
        e.process_device(d.view)
    e.synchronize()
    same_as_oracle(g, e, good[2:], f"{kind}: after the dropped buffer")
    # reset, then the whole good input
    e.reset()
    for d in dgood:
        e.process_device(d.view)
    same_as_oracle(g, e, good, f"{kind}: after reset")
    e.postprocess()
    e.close()
    for d in dgood + [dbad]:
        d.close()


def test_batched_host_batches_bad_read_group_then_reset():
    say("recovery: batched engine, host batches, read group beyond the header, reset, oracle")
    g, rs, n = workload()
    good = [synth.Batch(g, rs, f, 1000) for f in range(0, n, 1000)]
    bad = synth.Batch(g, variant(rs, num_read_groups=3, seed=98), 0, 700)
    e = Engine(g, engine_kind=abi.FP_ENGINE_BATCHED, max_batch_reads=4096, max_read_length=120,
               baq_waves_per_simd=1)
    e.process(good[0])
    e.process(bad)
    with pytest.raises(FireponyError):
        e.synchronize()
    e.reset()
    for b in good:
        e.process(b)
    same_as_oracle(g, e, good, "host batches after reset")
    # a host batch with a read longer than max_read_length is refused at the call itself
    with pytest.raises(FireponyError):
        e.process(synth.Batch(g, variant(rs, length=160), 0, 100))
    same_as_oracle(g, e, good, "after a refused host batch")
    e.close()


def test_plain_engine_bad_read_group_then_more_batches():
    say("recovery: plain engine, read group beyond the header, then more batches, oracle")
    g, rs, n = workload()
    good = [synth.Batch(g, rs, f, 1000) for f in range(0, n, 1000)]
    bad = synth.Batch(g, variant(rs, num_read_groups=3, seed=97), 0, 800)
    e = Engine(g)
    e.process(good[0])
    with pytest.raises(FireponyError):
        e.process(bad)
    for b in good[1:]:
        e.process(b)
    same_as_oracle(g, e, good, "plain engine after a refused batch")
    e.close()


def test_serving_bad_device_batch_fails_its_own_submit():
    say("recovery: serving, one job's device batch with a bad read group; the other job matches the oracle")
    g, rs, n = workload()
    e = Engine(g, engine_kind=abi.FP_ENGINE_BATCHED, max_batch_reads=4096, max_read_length=120,
               serve_rg_slots=8, baq_waves_per_simd=1)
    a = e.open_job(["a0", "a1"])
    b = e.open_job(["b0"])
    good = [synth.Batch(g, rs, f, 1000) for f in range(0, 4000, 1000)]
    dgood = [DeviceCopy(x) for x in good]
    bad = DeviceCopy(synth.Batch(g, variant(rs, num_read_groups=3, seed=96), 0, 600))
    long_ = DeviceCopy(synth.Batch(g, variant(rs, length=160, num_read_groups=1), 0, 300))
    for k, d in enumerate(dgood):
        a.submit_device(d.view)
        if k == 1:
            with pytest.raises(FireponyError):
                b.submit_device(bad.view)       # read group 2 of a 1-group job
            with pytest.raises(FireponyError):
                b.submit_device(long_.view)     # longer than max_read_length
    a.close()
    b.close()
    same_as_oracle(g, a, good, "job a beside a refused job b", counts=False)
    assert a.stats().total_reads == 4000
    assert b.stats().total_reads == 0
    a.release()
    b.release()
    e.close()
    for d in dgood + [bad, long_]:
        d.close()
