"""Batched engine (staged batches, hipGraph replays, autotuned BAQ) against
the plain engine and the oracle.

The batched engine appends input batches to device staging buffers and runs a
full buffer as two captured graphs.  Its tables must equal the plain engine's
bit for bit (counts and exact fixed-point sums) at every checkpoint: several
staging buffers (graph replays) run between checkpoints, with the test's own
eager work and synchronisations in between.
"""
import sys

import numpy as np
import pytest

from firepony_amd import abi, synth
from firepony_amd.engine import Engine
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def same_tables(a, b, what):
    for t in (abi.FP_TABLE_QUALITY, abi.FP_TABLE_CYCLE, abi.FP_TABLE_CONTEXT):
        x, y = a.table(t), b.table(t)
        for k in ("keys", "observations", "mismatches_fixed_lo", "mismatches_fixed_hi"):
            np.testing.assert_array_equal(y[k], x[k], err_msg=f"{what}: table {t} {k}")


def mixed_workload(n=20000, seed=21):
    return synth.small_workload(seed=seed, num_reads=n, lengths=(50, 76, 101, 151, 250), read_groups=8,
                                contigs=(400000, 300000), sites=800, clip=1.0, indel=0.05)


@pytest.mark.parametrize("wpe,ckpt", [(1, 1), (0, 0), (1, 3), (1, 4)])
def test_batched_equals_plain_with_replays(tmp_path, wpe, ckpt):
    say(f"batched vs plain: mixed lengths, staging buffers of 4096 reads, BAQ waves/SIMD {wpe or 'autotuned'}, "
        f"forward checkpoint rows {ckpt or 'autotuned'}")
    g, rs, n = mixed_workload()
    plain = Engine(g, frac_error_f64=1)
    bat = Engine(g, frac_error_f64=1, engine_kind=abi.FP_ENGINE_BATCHED, max_batch_reads=4096,
                 max_read_length=250, baq_waves_per_simd=wpe, baq_checkpoint_rows=ckpt)
    step = 1500
    for first in range(0, n, step):
        b = synth.Batch(g, rs, first, min(step, n - first))
        plain.process(b)
        bat.process(b)
        if first == 9000:
            say("  checkpoint after 10500 reads")
            same_tables(plain, bat, "checkpoint")
    sp, sb = plain.stats(), bat.stats()
    assert (sb.total_reads, sb.filtered_reads, sb.baq_reads, sb.baq_cells) == \
           (sp.total_reads, sp.filtered_reads, sp.baq_reads, sp.baq_cells)
    same_tables(plain, bat, "final")
    plain.postprocess()
    bat.postprocess()
    assert plain.report(str(tmp_path / "p.txt")) == bat.report(str(tmp_path / "b.txt"))
    plain.close()
    bat.close()


def test_batched_oversized_batch_and_oracle(tmp_path):
    say("batched: staged batches around one batch larger than the staging buffer; oracle check")
    g, rs, n = mixed_workload(12000, seed=33)
    o = Oracle(g)
    bat = Engine(g, engine_kind=abi.FP_ENGINE_BATCHED, max_batch_reads=3000, max_read_length=250,
                 baq_waves_per_simd=1)
    for first, cnt in ((0, 2000), (2000, 5000), (7000, 2500), (9500, 2500)):
        b = synth.Batch(g, rs, first, cnt)
        o.process(b)
        bat.process(b)
    o.postprocess()
    bat.postprocess()
    for t in (abi.FP_TABLE_QUALITY, abi.FP_TABLE_CYCLE, abi.FP_TABLE_CONTEXT):
        a, b = o.table(t), bat.table(t)
        for k in ("keys", "observations", "mismatches_fixed_lo", "mismatches_fixed_hi"):
            np.testing.assert_array_equal(b[k], a[k], err_msg=f"table {t} {k}")
    assert o.report(str(tmp_path / "o.txt")) == bat.report(str(tmp_path / "b.txt"))
    bat.close()


def test_batched_device_input_matches_host_input():
    say("batched: device-resident input batches vs host batches")
    from devbatch import DeviceCopy
    g, rs, n = mixed_workload(8000, seed=44)
    host = Engine(g, engine_kind=abi.FP_ENGINE_BATCHED, max_batch_reads=2048, max_read_length=250,
                  baq_waves_per_simd=2)
    dev = Engine(g, engine_kind=abi.FP_ENGINE_BATCHED, max_batch_reads=2048, max_read_length=250,
                 baq_waves_per_simd=2)
    keep = []
    for first in range(0, n, 1000):
        b = synth.Batch(g, rs, first, 1000)
        host.process(b)
        d = DeviceCopy(b)
        keep.append(d)
        dev.process_device(d.view)
    same_tables(host, dev, "device vs host input")
    host.close()
    dev.close()
    for d in keep:
        d.close()


@pytest.mark.parametrize("kind", [abi.FP_ENGINE_BATCHED, abi.FP_ENGINE_PLAIN])
def test_pinned_host_input_against_oracle(tmp_path, kind):
    say(f"{'batched' if kind else 'plain'}: pinned host batches (read by the copy kernel over the host link) vs "
        f"pageable host batches and the oracle")
    from devbatch import PinnedCopy
    g, rs, n = mixed_workload(9000, seed=45)
    o = Oracle(g)
    opts = dict(engine_kind=kind, max_batch_reads=2048, max_read_length=250, baq_waves_per_simd=1) if kind else {}
    pinned = Engine(g, **opts)
    pageable = Engine(g, **opts)
    keep = []
    for first in range(0, n, 1000):
        b = synth.Batch(g, rs, first, 1000)
        o.process(b)
        pageable.process(b)
        p = PinnedCopy(b)
        keep.append(p)
        pinned.process(p)
    same_tables(pageable, pinned, "pinned vs pageable input")
    o.postprocess()
    pinned.postprocess()
    for t in (abi.FP_TABLE_QUALITY, abi.FP_TABLE_CYCLE, abi.FP_TABLE_CONTEXT):
        a, b = o.table(t), pinned.table(t)
        for k in ("keys", "observations", "mismatches_fixed_lo", "mismatches_fixed_hi"):
            np.testing.assert_array_equal(b[k], a[k], err_msg=f"table {t} {k}")
    assert o.report(str(tmp_path / "o.txt")) == pinned.report(str(tmp_path / "p.txt"))
    pinned.close()
    pageable.close()
    for p in keep:
        p.close()
