"""Engine (HIP, through the C ABI) vs the CPU oracle on seeded workloads.

Stage-by-stage: read filters, windows, error bits, base masks, BAQ tags,
fractional errors, high-quality windows, the raw covariate tables (counts and
exact fixed-point mismatch sums), the postprocessed tables and the report.
Integer and index results must match bit for bit, and so must every floating
point result: fractional errors, mismatch sums, expected errors, estimated and
empirical qualities (their libm terms come from the host's libm, as the
oracle's do).
"""
import os
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


def windows(u32):
    return (u32 & 0xffff).astype(np.int64), (u32 >> 16).astype(np.int64)


def compare_batch(o, e, batch):
    arr = batch.arrays()
    act_o = o.debug(abi.FP_DBG_ACTIVE_READS, np.uint32)
    act_e = e.debug(abi.FP_DBG_ACTIVE_READS, np.uint32)
    np.testing.assert_array_equal(act_e, act_o, err_msg="active read list")
    act = act_o.astype(np.int64)
    for item, dt, name in [(abi.FP_DBG_READ_WINDOW, np.uint32, "read window"),
                           (abi.FP_DBG_NOINS_WINDOW, np.uint32, "no-insertion window"),
                           (abi.FP_DBG_REF_WINDOW, np.uint32, "reference window"),
                           (abi.FP_DBG_NUM_ERRORS, np.uint16, "num_errors"),
                           (abi.FP_DBG_HQ_WINDOW, np.uint32, "high quality window")]:
        a = o.debug(item, dt)[act]
        b = e.debug(item, dt)[act]
        bad = np.nonzero(a != b)[0]
        assert bad.size == 0, f"{name}: {bad.size} reads differ, first read {act[bad[0]]}: oracle {a[bad[0]]:#x} engine {b[bad[0]]:#x}"
    nerr = o.debug(abi.FP_DBG_NUM_ERRORS, np.uint16)
    baq_reads = act[nerr[act] != 0]
    for item, dt, name in [(abi.FP_DBG_HMM_WINDOW, np.uint32, "hmm window"),
                           (abi.FP_DBG_BANDWIDTH, np.uint16, "bandwidth")]:
        a = o.debug(item, dt)[baq_reads]
        b = e.debug(item, dt)[baq_reads]
        bad = np.nonzero(a != b)[0]
        assert bad.size == 0, f"{name}: {bad.size} reads differ, first read {baq_reads[bad[0]]}"
    # per-base and per-quality arrays over active reads
    rs, rl = arr["read_start"].astype(np.int64), arr["read_len"].astype(np.int64)
    qs = arr["qual_start"].astype(np.int64)
    bidx = np.concatenate([np.arange(rs[r], rs[r] + rl[r]) for r in act]) if act.size else np.zeros(0, np.int64)
    qidx = np.concatenate([np.arange(qs[r], qs[r] + rl[r]) for r in act]) if act.size else np.zeros(0, np.int64)
    for item, name in [(abi.FP_DBG_ERROR_FLAGS, "error bits"), (abi.FP_DBG_ACTIVE_MASK, "active mask")]:
        a = o.debug(item, np.uint8)[bidx]
        b = e.debug(item, np.uint8)[bidx]
        bad = np.nonzero(a != b)[0]
        assert bad.size == 0, f"{name}: {bad.size} bases differ, first base {bidx[bad[0]]}: {a[bad[0]]} vs {b[bad[0]]}"
    a = o.debug(abi.FP_DBG_BAQ, np.uint8)[qidx]
    b = e.debug(abi.FP_DBG_BAQ, np.uint8)[qidx]
    bad = np.nonzero(a != b)[0]
    assert bad.size == 0, f"BAQ: {bad.size} positions differ, first quality index {qidx[bad[0]]}: {a[bad[0]]} vs {b[bad[0]]}"
    fo = o.debug(abi.FP_DBG_FRAC_ERRORS, np.float64).reshape(-1, 3)[qidx]
    fe = e.debug(abi.FP_DBG_FRAC_ERRORS, np.float64).reshape(-1, 3)[qidx]
    np.testing.assert_array_equal(fe, fo, err_msg="fractional errors")


def compare_obs_tables(o, e):
    for t in (abi.FP_TABLE_QUALITY, abi.FP_TABLE_CYCLE, abi.FP_TABLE_CONTEXT):
        a, b = o.table(t), e.table(t)
        np.testing.assert_array_equal(b["keys"], a["keys"], err_msg=f"table {t} keys")
        np.testing.assert_array_equal(b["observations"], a["observations"], err_msg=f"table {t} counts")
        np.testing.assert_array_equal(b["mismatches_fixed_lo"], a["mismatches_fixed_lo"], err_msg=f"table {t} sums lo")
        np.testing.assert_array_equal(b["mismatches_fixed_hi"], a["mismatches_fixed_hi"], err_msg=f"table {t} sums hi")
        np.testing.assert_array_equal(b["mismatches"], a["mismatches"], err_msg=f"table {t} sums")


def compare_final(o, e, tmp_path):
    for t in range(4):
        a, b = o.table(t), e.table(t)
        np.testing.assert_array_equal(b["keys"], a["keys"], err_msg=f"final table {t} keys")
        np.testing.assert_array_equal(b["observations"], a["observations"])
        np.testing.assert_array_equal(b["mismatches"], a["mismatches"], err_msg=f"final table {t} mismatches")
        np.testing.assert_array_equal(b["expected_errors"], a["expected_errors"], err_msg=f"final table {t} expected")
        np.testing.assert_array_equal(b["estimated_quality"], a["estimated_quality"],
                                      err_msg=f"final table {t} estimated quality")
        bad = np.nonzero(b["empirical_quality"] != a["empirical_quality"])[0]
        assert bad.size == 0, (f"table {t}: {bad.size} empirical qualities differ; key {a['keys'][bad[0]]:#x} "
                               f"oracle {a['empirical_quality'][bad[0]]} engine {b['empirical_quality'][bad[0]]}")
    ro = o.report(str(tmp_path / "oracle.txt"))
    re_ = e.report(str(tmp_path / "engine.txt"))
    assert ro == re_, "report text differs"


def run_parity(tmp_path, genome, rs, n, batch_size, per_batch=True, **engine_opts):
    f64 = engine_opts.get("frac_error_f64", 0)
    o = Oracle(genome, frac_error_f64=bool(f64))
    e = Engine(genome, **engine_opts)
    for first in range(0, n, batch_size):
        b = synth.Batch(genome, rs, first, min(batch_size, n - first))
        say(f"  batch {first}..{first + b.view.num_reads}")
        o.process(b)
        e.process(b)
        if per_batch:
            compare_batch(o, e, b)
    compare_obs_tables(o, e)
    so, se = o.stats(), e.stats()
    assert (se.total_reads, se.filtered_reads, se.baq_reads) == (so.total_reads, so.filtered_reads, so.baq_reads)
    o.postprocess()
    e.postprocess()
    compare_final(o, e, tmp_path)
    e.close()


@pytest.mark.parametrize("wpe,nostage,ckpt", [(1, 0, 1), (1, 1, 1), (2, 0, 1), (3, 0, 1), (4, 0, 1), (1, 0, 3), (1, 0, 4), (1, 0, 5)])
def test_config0_slice(tmp_path, wpe, nostage, ckpt):
    say(f"parity: config 0 first 4000 reads, BAQ register budget for {wpe} waves/SIMD, LDS staging {1 - nostage}, "
        f"forward checkpoint every {ckpt} rows")
    gs, rs, n = synth.named_spec(0)
    g = synth.Genome(gs, rs.num_read_groups)
    run_parity(tmp_path, g, rs, 4000, 1000, baq_waves_per_simd=wpe, disable_baq_staging=nostage,
               baq_checkpoint_rows=ckpt)


@pytest.mark.parametrize("ckpt", [1, 3, 4])
def test_edge_cases(tmp_path, ckpt):
    say(f"parity: edge-case menu, paired, N runs, multi-base sites, forward checkpoint every {ckpt} rows")
    g, rs, n = synth.small_workload(seed=7, num_reads=3000, edge=0.35, lengths=(36, 100, 151), read_groups=3,
                                    contigs=(60000, 40000), sites=300, clip=0.3, indel=0.2, paired=1,
                                    n_runs=0.01, multi_sites=0.3)
    run_parity(tmp_path, g, rs, n, 1000, baq_checkpoint_rows=ckpt)


@pytest.mark.parametrize("nostage,ckpt", [(0, 1), (1, 1), (0, 3), (0, 4)])
def test_mixed_lengths_f64(tmp_path, nostage, ckpt):
    say(f"parity: mixed lengths, large indels, fp64 fractional errors, LDS staging {1 - nostage}, "
        f"forward checkpoint every {ckpt} rows")
    g, rs, n = synth.small_workload(seed=11, num_reads=3000, lengths=(50, 76, 101, 151, 250), read_groups=8,
                                    contigs=(300000,), sites=600, clip=1.0, indel=0.3)
    run_parity(tmp_path, g, rs, n, 1500, frac_error_f64=1, disable_baq_staging=nostage, baq_checkpoint_rows=ckpt)


@pytest.mark.parametrize("ckpt", [1, 3])
def test_waves_taking_several_chunks(tmp_path, ckpt):
    """One 150k-read batch of the mixed-length workload (50-250 bp): the 15-cell class then has more
    64-read chunks than the band kernel has resident waves (1024 at one wave per SIMD), so each wave
    takes several chunks in turn and reuses its LDS staging area and scratch slot.  Stage by stage
    against the oracle, per forward checkpoint interval."""
    say(f"parity: 150k mixed-length reads in one batch, waves taking several chunks, checkpoint every {ckpt} rows")
    gs, rs, _ = synth.named_spec(2)
    g = synth.Genome(gs, rs.num_read_groups)
    run_parity(tmp_path, g, rs, 150000, 150000, frac_error_f64=1, baq_waves_per_simd=1, baq_checkpoint_rows=ckpt)


def test_dense_known_sites(tmp_path):
    """A known site every 20 reference positions (a third of them multi-base): most reads overlap three
    or more sites, so the prep kernel masks the ranges beyond its two in-register ones by atomic ORs
    into the stored flags.  Stage by stage against the oracle (base flags included)."""
    say("parity: dense known sites (one per 20 bp), reads with three or more masked ranges")
    g, rs, n = synth.small_workload(seed=17, num_reads=3000, edge=0.1, lengths=(101, 151), read_groups=2,
                                    contigs=(40000, 20000), sites=3000, clip=0.3, indel=0.2, multi_sites=0.3)
    run_parity(tmp_path, g, rs, n, 1000)
    run_parity(tmp_path, g, rs, n, 1000, per_batch=False, engine_kind=abi.FP_ENGINE_BATCHED, max_batch_reads=4096)


def test_long_reads(tmp_path):
    say("parity: reads longer than 256 bases next to short ones (per-lane flag path of the prep kernel)")
    g, rs, n = synth.small_workload(seed=13, num_reads=1500, edge=0.1, lengths=(90, 300, 420), read_groups=2,
                                    contigs=(200000,), sites=400, clip=0.3, indel=0.2)
    run_parity(tmp_path, g, rs, n, 750)


def test_genome_3gbp(tmp_path):
    say("parity: 3 Gbp reference (24 contigs) and 1M known sites resident, paired 150 bp reads, 16 read groups")
    gs, rs, n = synth.named_spec(4)
    g = synth.Genome(gs, rs.num_read_groups)
    assert g.reference.total_words == 375_000_000
    run_parity(tmp_path, g, rs, 20000, 10000)


def test_batch_size_invariance(tmp_path):
    say("invariance: one batch vs many batches give identical tables")
    g, rs, n = synth.small_workload(seed=5, num_reads=3000, edge=0.1)
    e1 = Engine(g)
    e1.process(synth.Batch(g, rs, 0, n))
    e2 = Engine(g)
    for first in range(0, n, 700):
        e2.process(synth.Batch(g, rs, first, min(700, n - first)))
    for t in (abi.FP_TABLE_QUALITY, abi.FP_TABLE_CYCLE, abi.FP_TABLE_CONTEXT):
        a, b = e1.table(t), e2.table(t)
        for k in ("keys", "observations", "mismatches_fixed_lo", "mismatches_fixed_hi"):
            np.testing.assert_array_equal(a[k], b[k])
