"""Serving engine: jobs with their own read groups share one batched engine.

Each job's covariate tables live in its own read-group slots, so its report
must be byte-identical whether it runs alone, interleaved with other jobs
(continuous batching: reads of several jobs in the same staging buffers and
graph replays, jobs opening and closing at different times), or on an
ordinary engine set up with the job's own header, and that engine's report is
the oracle's for the same reads.
"""
import ctypes as C
import sys

import numpy as np
import pytest

from firepony_amd import abi, synth
from firepony_amd.engine import Engine, FireponyError
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu

GENOME_SEED = 77
CONTIGS = (300000, 200000)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def genome(nrg):
    g, _, _ = synth.small_workload(seed=GENOME_SEED, contigs=CONTIGS, sites=600, read_groups=nrg)
    return g


def job_spec(seed, nrg, n):
    _, rs, _ = synth.small_workload(seed=GENOME_SEED, contigs=CONTIGS, sites=600, read_groups=nrg,
                                    lengths=(100, 150), clip=0.05, indel=0.03)
    rs.seed = seed
    return rs, n


JOBS = [(501, 1, 6000), (502, 2, 9000), (503, 4, 7000)]


def batches(g, spec, step=1000):
    rs, n = spec
    return [synth.Batch(g, rs, f, min(step, n - f)) for f in range(0, n, step)]


def serving_engine(g, slots=8):
    return Engine(g, engine_kind=abi.FP_ENGINE_BATCHED, max_batch_reads=4096, max_read_length=150,
                  serve_rg_slots=slots, baq_waves_per_simd=1)


def test_job_result_alone_equals_in_mix(tmp_path):
    say("serving: 3 jobs alone vs interleaved vs ordinary engines")
    g = genome(1)
    specs = [job_spec(seed, nrg, n) for seed, nrg, n in JOBS]
    names = [[f"job{k}_rg{i}" for i in range(nrg)] for k, (_, nrg, _) in enumerate(JOBS)]

    # each job on an ordinary (non-serving) engine with its own header
    ref = []
    for k, (seed, nrg, n) in enumerate(JOBS):
        gk = genome(nrg)
        gk_names = (C.c_char_p * nrg)(*[x.encode() for x in names[k]])
        gk.header.read_group_names = gk_names
        e = Engine(gk)
        o = Oracle(gk)
        for b in batches(gk, specs[k]):
            e.process(b)
            o.process(b)
        e.postprocess()
        o.postprocess()
        ref.append(e.report(str(tmp_path / f"ref{k}.txt")))
        assert ref[-1] == o.report(str(tmp_path / f"oracle{k}.txt")), f"job {k}: ordinary engine vs oracle"
        e.close()

    # alone on a serving engine, one after another
    e = serving_engine(g)
    alone = []
    for k in range(len(JOBS)):
        j = e.open_job(names[k])
        for b in batches(g, specs[k]):
            j.submit(b)
        j.close()
        alone.append(j.report(str(tmp_path / f"alone{k}.txt")))
        j.release()
    e.close()

    # in a mix: all open, batches interleaved round-robin, job 0 closes while the others still submit
    e = serving_engine(g)
    jobs = [e.open_job(names[k]) for k in range(len(JOBS))]
    queues = [batches(g, specs[k]) for k in range(len(JOBS))]
    mixed = [None] * len(JOBS)
    while any(queues):
        for k in range(len(JOBS)):
            if queues[k]:
                jobs[k].submit(queues[k].pop(0))
                if not queues[k]:
                    jobs[k].close()
                    mixed[k] = jobs[k].report(str(tmp_path / f"mix{k}.txt"))
                    jobs[k].release()
    e.close()

    for k in range(len(JOBS)):
        assert alone[k] == ref[k], f"job {k}: serving engine alone differs from an ordinary engine"
        assert mixed[k] == alone[k], f"job {k}: result in the mix differs from alone"


def test_admission_by_slot_budget():
    say("serving: admission by the read-group slot budget")
    g = genome(1)
    e = serving_engine(g, slots=4)
    a = e.open_job(["a0", "a1", "a2"])
    assert a is not None
    assert e.open_job(["b0", "b1"]) is None           # 1 slot left: refused now
    b1 = e.open_job(["c0"])
    assert b1 is not None
    spec = job_spec(601, 3, 3000)
    for bt in batches(g, spec):
        a.submit(bt)
    a.close()
    a.release()
    b = e.open_job(["b0", "b1"])                        # room again after the release
    assert b is not None
    bad = job_spec(602, 3, 1000)                        # read group 2 does not exist in a 2-group job
    with pytest.raises(FireponyError):
        for bt in batches(g, bad):
            b.submit(bt)
    b.release()
    b1.release()
    e.close()


def test_admission_by_hbm_budget(tmp_path):
    """Jobs are admitted by their HBM footprint (resident tables of their read groups plus a
    postprocess reserve, about 104 MiB per read group) against serve_hbm_budget_mib; a job admitted
    after another's release runs to the oracle's report."""
    say("serving: admission by the HBM byte budget")
    g = genome(1)
    e = Engine(g, engine_kind=abi.FP_ENGINE_BATCHED, max_batch_reads=4096, max_read_length=150,
               serve_rg_slots=8, serve_hbm_budget_mib=250, baq_waves_per_simd=1)
    a = e.open_job(["a0"])
    b = e.open_job(["b0"])
    assert a is not None and b is not None             # 2 x 104 MiB of 250
    assert e.open_job(["c0"]) is None                   # a third would need 312 MiB: refused now
    with pytest.raises(FireponyError):
        e.open_job(["x0", "x1", "x2"])                  # larger than the whole budget: an error
    a.release()
    c = e.open_job(["c0"])                              # room again after the release
    assert c is not None
    spec = job_spec(611, 1, 3000)
    for bt in batches(g, spec):
        c.submit(bt)
    c.close()
    names = (C.c_char_p * 1)(b"c0")   # the job's read-group name, for the oracle's report
    g.header.read_group_names = names
    o2 = Oracle(g)
    for bt in batches(g, spec):
        o2.process(bt)
    o2.postprocess()
    assert c.report(str(tmp_path / "c.txt")) == o2.report(str(tmp_path / "o.txt")), "budget-admitted job vs oracle"
    c.release()
    b.release()
    e.close()
