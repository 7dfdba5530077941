"""The batched and serving workloads by name (the generator's named specs 2 and 3) on the GPU,
against the oracle.

- Mixed-length batched workload (50-250 bp, clips of 0-30 bp, 5 % indel reads of 1-10 bp,
  8 read groups, fp64 fractional errors, the 30 Mbp genome): a 40k-read slice through the
  batched engine, each 20k-read batch run as its own staged buffer and compared with the oracle
  stage by stage (filters, windows, error bits, masks, HMM windows, BQ tags, fractional errors),
  then tables and report; and the workload's full 20M reads, batched == plain, table for table.
- Serving workload: jobs drawn as the serving benchmark draws them (Poisson arrivals, log-uniform
  sizes scaled to what the oracle runs, 1-4 read groups, 100 or 150 bp reads), submitted in
  batches to one serving engine with continuous batching; each job's tables and report equal an
  oracle run of that job alone, and the job's report alone on a serving engine equals its report
  in the mix.
"""
import math
import random
import sys

import numpy as np
import pytest

from firepony_amd import abi, synth
from firepony_amd.engine import Engine
from oracle.pyoracle import Oracle

from test_gpu_parity import compare_batch, compare_final, compare_obs_tables

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


TABLES = (abi.FP_TABLE_QUALITY, abi.FP_TABLE_CYCLE, abi.FP_TABLE_CONTEXT)


def test_config2_slice_stage_by_stage(tmp_path):
    say("named workload 2 (mixed lengths): 40k reads, batched engine, stage by stage against the oracle")
    gs, rs, _ = synth.named_spec(2)
    g = synth.Genome(gs, rs.num_read_groups)
    o = Oracle(g, frac_error_f64=True)
    e = Engine(g, engine_kind=abi.FP_ENGINE_BATCHED, max_batch_reads=32768, max_read_length=250,
               frac_error_f64=1, baq_waves_per_simd=1)
    for first in (0, 20000):
        b = synth.Batch(g, rs, first, 20000, threads=16)
        say(f"  batch {first}..{first + 20000}")
        o.process(b)
        e.process(b)
        e.synchronize()          # the staged buffer runs now: its intermediates are the last batch's
        compare_batch(o, e, b)
    compare_obs_tables(o, e)
    so, se = o.stats(), e.stats()
    assert (se.total_reads, se.filtered_reads, se.baq_reads) == (so.total_reads, so.filtered_reads, so.baq_reads)
    o.postprocess()
    e.postprocess()
    compare_final(o, e, tmp_path)
    e.close()


def test_config2_full_batched_equals_plain():
    gs, rs, total = synth.named_spec(2)
    assert total == 20_000_000
    say(f"named workload 2: all {total} reads, batched == plain, table for table")
    g = synth.Genome(gs, rs.num_read_groups)
    plain = Engine(g, frac_error_f64=1, debug_arrays=0)
    bat = Engine(g, engine_kind=abi.FP_ENGINE_BATCHED, max_read_length=250, frac_error_f64=1, debug_arrays=0)
    step = 200000
    for first in range(0, total, step):
        b = synth.Batch(g, rs, first, step, threads=16)
        plain.process(b)
        bat.process(b)
        if first % 4000000 == 0:
            say(f"  {first + step} reads")
    sp, sb = plain.stats(), bat.stats()
    assert (sb.total_reads, sb.filtered_reads, sb.baq_reads, sb.baq_cells) == \
           (sp.total_reads, sp.filtered_reads, sp.baq_reads, sp.baq_cells)
    assert sb.total_reads == total
    for t in TABLES:
        x, y = plain.table(t), bat.table(t)
        for k in ("keys", "observations", "mismatches_fixed_lo", "mismatches_fixed_hi"):
            np.testing.assert_array_equal(y[k], x[k], err_msg=f"table {t} {k}")
    plain.close()
    bat.close()


def draw_jobs(count, seed, lo=20000, hi=60000):
    """Jobs as the serving benchmark draws them, with sizes log-uniform in [lo, hi]."""
    gs, rs0, _ = synth.named_spec(3)
    rnd = random.Random(seed)
    jobs, t = [], 0.0
    for k in range(count):
        t += rnd.expovariate(20.0)
        n = int(math.exp(rnd.uniform(math.log(lo), math.log(hi))))
        nrg = rnd.randint(1, 4)
        L = rnd.choice((100, 150))
        rs = abi.FpsynReadSpec.from_buffer_copy(rs0)
        rs.seed = seed * 1000 + k
        rs.num_lengths, rs.lengths[0] = 1, L
        rs.num_read_groups = nrg
        jobs.append({"arrival": t, "reads": n, "nrg": nrg, "rs": rs,
                     "names": [f"job{k}_rg{i}" for i in range(nrg)]})
    return gs, jobs


def serving_engine(g):
    return Engine(g, engine_kind=abi.FP_ENGINE_BATCHED, max_batch_reads=65536, max_read_length=150,
                  serve_rg_slots=16, serve_hbm_budget_mib=1200, baq_waves_per_simd=1)


def test_config3_jobs_against_oracle(tmp_path):
    say("named workload 3 (serving): 4 jobs, continuous batching, each job vs an oracle run alone")
    gs, jobs = draw_jobs(4, 0xF1E90003)
    g = synth.Genome(gs, 1)
    B = 8192
    for j in jobs:
        j["batches"] = [synth.Batch(g, j["rs"], f, min(B, j["reads"] - f), threads=16)
                        for f in range(0, j["reads"], B)]
    # each job alone through the oracle, with its own header
    ref = []
    for k, j in enumerate(jobs):
        gj = synth.Genome(gs, j["nrg"])
        o = Oracle(gj)
        o.set_read_group_names(j["names"])
        for b in j["batches"]:
            o.process(b)
        o.postprocess()
        ref.append((o, o.report(str(tmp_path / f"oracle{k}.txt"))))
    # all jobs in one serving engine: admitted as they arrive (the budget holds 11 read groups),
    # one batch of every open job per round, jobs closing while others still submit
    e = serving_engine(g)
    waiting = list(range(len(jobs)))
    active, done = [], {}
    while waiting or active:
        while waiting:
            h = e.open_job(jobs[waiting[0]]["names"])
            if h is None:
                break
            jobs[waiting[0]]["job"], jobs[waiting[0]]["next"] = h, 0
            active.append(waiting.pop(0))
        assert active, "no job admitted although none is open"
        for k in list(active):
            j = jobs[k]
            j["job"].submit(j["batches"][j["next"]])
            j["next"] += 1
            if j["next"] == len(j["batches"]):
                j["job"].close()
                done[k] = j["job"].report(str(tmp_path / f"mix{k}.txt"))
                o = ref[k][0]
                for t in TABLES:
                    a, x = o.table(t), j["job"].table(t)
                    for key in ("keys", "observations", "mismatches_fixed_lo", "mismatches_fixed_hi"):
                        np.testing.assert_array_equal(x[key], a[key], err_msg=f"job {k}: table {t} {key}")
                assert j["job"].stats().total_reads == j["reads"]
                j["job"].release()
                active.remove(k)
                say(f"  job {k} ({j['reads']} reads, {j['nrg']} read groups) closed")
    e.close()
    for k in range(len(jobs)):
        assert done[k] == ref[k][1], f"job {k}: report in the mix differs from the oracle's"
    # one job alone on a serving engine gives the same report as in the mix
    e = serving_engine(g)
    h = e.open_job(jobs[1]["names"])
    for b in jobs[1]["batches"]:
        h.submit(b)
    h.close()
    assert h.report(str(tmp_path / "alone1.txt")) == done[1], "job 1 alone differs from job 1 in the mix"
    h.release()
    e.close()
