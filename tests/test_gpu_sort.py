"""Sort aggregation (observation pairs -> LDS-tiled radix sort -> segmented
reduce_by_key) against the direct LDS-privatised tables and the oracle.

Both methods sum exact fixed-point values, so their tables must agree bit for
bit for every radix digit width, on the plain engine and on the batched engine
(whose autotuner times each digit width on a staged batch before it captures
the graph it replays).
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


def mixed_workload(n, seed):
    # lengths up to 300: windows longer than 256 bases take the one-lane emission path
    return synth.small_workload(seed=seed, num_reads=n, lengths=(50, 76, 101, 151, 250, 300), read_groups=6,
                                contigs=(400000, 300000), sites=800, clip=1.0, indel=0.1)


@pytest.mark.parametrize("bits", [4, 5, 6, 7, 8])
def test_sort_equals_direct(bits):
    say(f"sort vs direct aggregation, plain engine, radix digits of {bits} bits")
    g, rs, n = mixed_workload(6000, seed=31)
    direct = Engine(g, frac_error_f64=1)
    srt = Engine(g, frac_error_f64=1, aggregation=abi.FP_AGG_SORT, sort_digit_bits=bits)
    for first in range(0, n, 2000):
        b = synth.Batch(g, rs, first, min(2000, n - first))
        direct.process(b)
        srt.process(b)
    same_tables(direct, srt, f"digits {bits}")
    st = srt.stats()
    assert st.sort_digit_bits == bits
    assert st.covariates_sort_ms > 0


def test_sort_float_values_and_oracle(tmp_path):
    say("sort aggregation with per-base errors narrowed to float, against the oracle")
    g, rs, n = synth.small_workload(seed=7, num_reads=3000, edge=0.35, lengths=(36, 100, 151), read_groups=3,
                                    contigs=(60000, 40000), sites=300, clip=0.3, indel=0.2, paired=1,
                                    n_runs=0.01, multi_sites=0.3)
    o = Oracle(g, frac_error_f64=False)
    e = Engine(g, aggregation=abi.FP_AGG_SORT, sort_digit_bits=6)
    for first in range(0, n, 1000):
        b = synth.Batch(g, rs, first, min(1000, n - first))
        o.process(b)
        e.process(b)
    same_tables(o, e, "oracle")
    o.postprocess()
    e.postprocess()
    assert o.report(str(tmp_path / "o.txt")) == e.report(str(tmp_path / "e.txt"))


def test_sort_batched_autotuned_equals_direct(tmp_path):
    say("sort aggregation on the batched engine: digit width autotuned over staged batches, then graph replays")
    g, rs, n = mixed_workload(24000, seed=33)
    direct = Engine(g, frac_error_f64=1)
    bat = Engine(g, frac_error_f64=1, engine_kind=abi.FP_ENGINE_BATCHED, max_batch_reads=2048,
                 max_read_length=300, baq_waves_per_simd=1, aggregation=abi.FP_AGG_SORT, sort_digit_bits=0)
    for first in range(0, n, 1500):
        b = synth.Batch(g, rs, first, min(1500, n - first))
        direct.process(b)
        bat.process(b)
    st = bat.stats()
    say(f"  digit width chosen: {st.sort_digit_bits}")
    assert 4 <= st.sort_digit_bits <= 8
    same_tables(direct, bat, "batched sort")
    direct.postprocess()
    bat.postprocess()
    assert direct.report(str(tmp_path / "d.txt")) == bat.report(str(tmp_path / "b.txt"))
