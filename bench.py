"""Benchmark: reads/s of the BQSR+BAQ device pipeline on one MI355X.

Workload (the reference's GPU-at-scale case): 30 Mbp genome in 3 contigs,
10M single-end 101 bp reads over 4 read groups (substitution rates
0.5/1.0/1.5/2.0 %), 3 % indel reads, 5 % soft clips, 30k known sites, BAQ
band 7, seeded synthetic data, delivered as batches of 20,000 reads (the
reference's GPU batch size, firepony.cu:148) from pinned host memory.

A step is 500,000 reads: 25 input batches handed to fp_engine_process_batch
(host arrays; the engine copies them to HBM), so the default 20 steps are the
workload's full 10M reads.  The timed region covers the steps' batches, H2D
copies included, plus postprocess, which is how the reference defines reads/s
(firepony.cu:348-376, 222; the worker's H2D copy, pipeline_interface.cu:214).
More than 20 steps wrap around the input; warm-up steps run before it on the
same engine, which is then reset.

Passes:
  value     the batched engine on host batches: staged into 2M-read device
            buffers, each buffer run as captured hipGraph replays, the BAQ
            variant chosen by the engine's autotuner in an untimed pass first;
  device    the same on batches already resident in HBM (no PCIe);
  plain     one launch per stage per 20,000-read batch, host-fed and resident;
  profile   the batched engine with an event between kernels (no graphs), on
            the resident batches at 1M-read launches: per-kernel times and the
            roofline of the dominant kernel (the 15-cell BAQ band kernel),
            whose traffic comes from the committed rocprofv3 FETCH_SIZE /
            WRITE_SIZE summary taken at the same launch size.
The mixed-length workload (20M reads of 50-250 bp, 8 read groups, 5 % indels
of 1-10 bp, clips on every read) is timed plain and batched on resident
batches, the 3 Gbp genome batched, and the serving workload (a Poisson stream
of 100 jobs of 50k-5M reads, each with its own read groups, submitted from
pinned host memory in 64k-read batches to one serving engine in real time)
reports throughput and p50/p99 job latency under `served`.  The CPU oracle (a
restatement of the reference's host path) is timed on a bounded sample of the
same reads (400k), one oracle per host thread (up to 16) and an exact gather,
as the reference's host pipelines run.  The sort-aggregation pass (covariate
pairs, radix sort, reduce_by_key) reports its reads/s and the sort/reduce
stages' HBM rate.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import ctypes as C
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from firepony_amd import abi, synth  # noqa: E402
from firepony_amd.engine import Engine  # noqa: E402

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
PCIE_PEAK_GBS = 64.0             # PCIe 5.0 x16, one direction: the host link of the H2D copies
BAQ_FLOPS_PER_CELL = 43          # forward 17 + rescale 3, backward 16 + rescale 3, MAP 4
STEP_READS = 500000              # one step: 25 input batches of 20,000 reads


def log(*a):
    print(*a, file=sys.stderr, flush=True)


class DeviceBatch:
    """A generated batch whose arrays live in HBM (torch tensors as plumbing)."""

    FIELDS = [("chromosome", "u16", "chromosome"), ("alignment_start", "u32", "alignment_start"),
              ("alignment_stop", "u32", "alignment_stop"), ("mate_chromosome", "u32", "mate_chromosome"),
              ("mate_alignment_start", "u32", "mate_alignment_start"),
              ("inferred_insert_size", "i32", "inferred_insert_size"), ("cigars", "u32", "cigars"),
              ("cigar_start", "u32", "cigar_start"), ("cigar_len", "u32", "cigar_len"),
              ("reads", "u32", "reads"), ("read_start", "u32", "read_start"), ("read_len", "u32", "read_len"),
              ("qualities", "u8", "qualities"), ("qual_start", "u32", "qual_start"),
              ("qual_len", "u32", "qual_len"), ("flags", "u16", "flags"), ("mapq", "u8", "mapq"),
              ("read_group", "u32", "read_group")]

    def __init__(self, host_batch, torch):
        arrs = host_batch.arrays()
        self.keep = []
        v = abi.FpBatchView()
        hv = host_batch.view
        v.num_reads = hv.num_reads
        v.num_cigar_ops = hv.num_cigar_ops
        v.num_read_bases = hv.num_read_bases
        v.num_qualities = hv.num_qualities
        ptype = {"u8": abi.u8p, "u16": abi.u16p, "u32": abi.u32p, "i32": abi.i32p}
        sview = {"u8": np.uint8, "u16": np.int16, "u32": np.int32, "i32": np.int32}
        for name, kind, field in self.FIELDS:
            a = np.ascontiguousarray(arrs[name]).view(sview[kind])
            if a.size == 0:
                a = np.zeros(1, sview[kind])
            t = torch.from_numpy(a).to("cuda:0")
            self.keep.append(t)
            setattr(v, field, C.cast(C.c_void_p(t.data_ptr()), ptype[kind]))
        self.view = v
        self.num_reads = hv.num_reads
        self.num_bases = int(arrs["read_len"].sum())


class PinnedBatch:
    """A generated batch whose arrays live in pinned host memory (torch tensors as plumbing): the
    host batch handed to fp_engine_process_batch, which copies it to HBM."""

    def __init__(self, host_batch, torch):
        arrs = host_batch.arrays()
        self.keep = []
        v = abi.FpBatchView()
        hv = host_batch.view
        v.num_reads = hv.num_reads
        v.num_cigar_ops = hv.num_cigar_ops
        v.num_read_bases = hv.num_read_bases
        v.num_qualities = hv.num_qualities
        ptype = {"u8": abi.u8p, "u16": abi.u16p, "u32": abi.u32p, "i32": abi.i32p}
        sview = {"u8": np.uint8, "u16": np.int16, "u32": np.int32, "i32": np.int32}
        self.nbytes = 0
        for name, kind, field in DeviceBatch.FIELDS:
            a = np.ascontiguousarray(arrs[name]).view(sview[kind])
            if a.size == 0:
                a = np.zeros(1, sview[kind])
            t = torch.from_numpy(a).pin_memory()
            self.keep.append(t)
            self.nbytes += a.nbytes
            setattr(v, field, C.cast(C.c_void_p(t.data_ptr()), ptype[kind]))
        self.view = v
        self.num_reads = hv.num_reads


def load_traffic():
    """Per-dispatch HBM bytes by kernel from the committed counter summary
    (lines 'traffic <kernel> <fetch KiB> <write KiB> <fetch scale> <write scale>')."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_config1_batched.txt")
    out = {}
    if not os.path.exists(path):
        return out
    with open(path) as f:
        for line in f:
            p = line.split()
            if len(p) == 6 and p[0] == "traffic":
                out[p[1]] = (float(p[2]) * float(p[4]) + float(p[3]) * float(p[5])) * 1024.0
    return out


def roofline(stats):
    """Roofline of the kernel with the most device time, the 15-cell BAQ band kernel: its FP64 work
    (43 flops per band cell over the cells the device counted) per launch over its average launch
    time (HIP events around each launch, profile pass at 1M-read launches), and its HBM traffic per
    launch from the committed rocprofv3 counter summary taken at that launch size."""
    launches = max(int(stats.baq15_launches), 1)
    avg_s = stats.baq15_kernel_ms / 1e3 / launches
    flops = BAQ_FLOPS_PER_CELL * stats.baq15_cells / launches
    ach = flops / avg_s / 1e12
    traffic = load_traffic().get("baq15")
    out = {"bound": "valu_fp64", "achieved": ach, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
           "frac": ach / FP64_PEAK_TFLOPS, "traffic": traffic, "avg_launch_ms": avg_s * 1e3,
           "launches": launches, "cells_per_launch": stats.baq15_cells / launches,
           "kernel": f"the 15-cell BAQ class ({'fpd::k_baq_ckpt<15, %d, ...>' % stats.baq_checkpoint_rows if stats.baq_checkpoint_rows > 1 else 'fpd::k_baq_band<15, ...>'}), 1M-read launches",
           "note": "vector FP64 (the recurrence has no dense-contraction form, so no MFMA); achieved = 43 "
                   "flops x band cells per launch / average launch time"}
    if traffic:
        rate = traffic / avg_s / 1e9
        out.update(traffic_rate_gbs=rate, traffic_frac_of_peak=rate / HBM_PEAK_GBS)
        if rate > HBM_PEAK_GBS:
            out["traffic_warning"] = "traffic over launch time exceeds the HBM peak: profile and run disagree"
    return out


def end_to_end(value, reads, bytes_per_read, cells_per_read, h2d_gbs):
    """The path's own bound per read: compulsory bytes once through HBM (the batch arrays plus the 4-bit
    reference window of the HMM), the BAQ's FP64 work at the vector FP64 peak and, for host-fed
    batches, the batch bytes over the host link."""
    s1 = bytes_per_read + 54.0                  # batch arrays + ceil((span + 6) / 2) reference nibbles
    s2 = BAQ_FLOPS_PER_CELL * cells_per_read    # flops per read
    device_s = s1 / (HBM_PEAK_GBS * 1e9) + s2 / (FP64_PEAK_TFLOPS * 1e12)
    link_s = bytes_per_read / (PCIE_PEAK_GBS * 1e9)
    bound = 1.0 / max(device_s, link_s)
    return {"reads_per_s_bound": bound, "frac": value / bound, "device_bound": 1.0 / device_s,
            "pcie_bound": 1.0 / link_s, "bytes_per_read": bytes_per_read, "flops_per_read": s2,
            "h2d_copy_gbs_measured": h2d_gbs,
            "note": "bound = 1 / max(S1/HBM + S2/FP64, batch bytes / PCIe 5.0 x16); the direct "
                    "aggregation has no sort traffic"}


def cpu_baseline(genome, rs, sample_reads, threads):
    """The CPU oracle (a restatement of the reference host path) on the first sample_reads reads of the
    workload in batches of 1,000 (firepony.cu:185-190): `threads` host threads, each with its own
    oracle taking whole batches in turn, then the exact table gather and postprocess on one of them,
    as the reference's host pipelines do (pipeline_interface.cu:52-58,195-224)."""
    from concurrent.futures import ThreadPoolExecutor
    from oracle.pyoracle import Oracle
    batches = [synth.Batch(genome, rs, f, min(1000, sample_reads - f), threads=threads)
               for f in range(0, sample_reads, 1000)]
    oracles = [Oracle(genome) for _ in range(threads)]

    def work(t):
        for b in batches[t::threads]:
            oracles[t].process(b)   # ctypes releases the interpreter lock during the call

    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=threads) as ex:
        list(ex.map(work, range(threads)))
    for o in oracles[1:]:
        oracles[0].gather(o)
    oracles[0].postprocess()
    dt = time.perf_counter() - t0
    return {"value": sample_reads / dt, "unit": "reads/s", "cores": threads, "kind": "port",
            "sample": f"first {sample_reads} reads of the workload, batches of 1000, {threads} threads, "
                      f"{dt:.2f} s", "host": host_cpu()}


def host_cpu():
    """CPU model and logical CPU count of this host (lscpu's 'Model name' and 'CPU(s)')."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return f"{model}, {os.cpu_count()} logical CPUs"


def upload(genome, rs, total, batch, torch, pinned=False, device=True):
    """Generate the workload's first `total` reads as batches of `batch`: device-resident copies and/or
    pinned host copies."""
    dev, pin = [], []
    nthreads = min(16, os.cpu_count() or 4)
    for first in range(0, total, batch):
        hb = synth.Batch(genome, rs, first, min(batch, total - first), threads=nthreads)
        if device:
            dev.append(DeviceBatch(hb, torch))
        if pinned:
            pin.append(PinnedBatch(hb, torch))
        del hb
    torch.cuda.synchronize()
    return (dev, pin) if pinned else dev


def h2d_rate(torch, mib=1024):
    """Measured copy rate from pinned host memory to HBM (GB/s), one large copy."""
    src = torch.empty(mib << 20, dtype=torch.uint8).pin_memory()
    dst = torch.empty(mib << 20, dtype=torch.uint8, device="cuda:0")
    dst.copy_(src, non_blocking=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    dst.copy_(src, non_blocking=True)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    del src, dst
    return (mib << 20) / dt / 1e9


def stepped_pass(genome, batches, steps, warmup, torch, host, per_step, **opts):
    """`warmup` untimed steps, a reset, then exactly `steps` timed steps of `per_step` input batches each
    (wrapping around the input), plus postprocess.  host: the batches are pinned host arrays handed to
    fp_engine_process_batch; else device arrays handed to fp_engine_process_batch_device.
    Returns (seconds, engine stats, reads timed)."""
    e = Engine(genome, device=0, debug_arrays=0, **opts)
    nb = len(batches)

    def run(k):
        n = 0
        for i in range(k * per_step, (k + 1) * per_step):
            b = batches[i % nb]
            if host:
                e.process(b)
            else:
                e.process_device(b.view)
            n += b.num_reads
        return n

    for k in range(warmup):
        run(k)
    if warmup:
        e.synchronize()
        e.reset()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    reads = 0
    for k in range(steps):
        reads += run(k)
    e.postprocess()
    e.synchronize()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    st = e.stats()
    e.close()
    return dt, st, reads


def timed_pass(genome, dev, warmup, torch, **opts):
    """warmup untimed passes over the input (reset after), then every batch plus postprocess timed;
    returns (seconds, engine stats)."""
    e = Engine(genome, device=0, debug_arrays=0, **opts)
    for _ in range(warmup):
        for d in dev:
            e.process_device(d.view)
        e.synchronize()
        e.reset()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for d in dev:
        e.process_device(d.view)
    e.postprocess()
    e.synchronize()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    st = e.stats()
    e.close()
    return dt, st


def served(args, wpe, torch):
    """The serving workload: a Poisson stream (mean 20 jobs/s) of jobs of 50k-5M reads (log-uniform; each
    rounded up to whole 64k-read batches), each with 1-4 read groups and 100 or 150 bp reads, submitted
    from pinned host memory in batches of 64k reads to one serving engine in real time.  Jobs are
    admitted by the engine's HBM budget (resident tables of their read groups); a job's latency runs from
    its arrival to its postprocessed tables and includes the H2D copies of its reads.  The batches come
    from a pool of 16 distinct 64k-read batches per (read length, read-group count); a job takes its
    pool's batches in turn from its own starting point."""
    import random
    gs, rs0, _ = synth.named_spec(3)
    genome = synth.Genome(gs, 1)
    rnd = random.Random(0xF1E90003)
    B = 65536
    jobs, t = [], 0.0
    for k in range(args.served_jobs):
        t += rnd.expovariate(20.0)
        n = int(math.exp(rnd.uniform(math.log(50000), math.log(5000000))))
        jobs.append({"arrival": t, "nbatches": (n + B - 1) // B, "nrg": rnd.randint(1, 4),
                     "L": rnd.choice((100, 150)), "start": rnd.randrange(1 << 16)})
    pools = {}
    t0 = time.perf_counter()
    for key in sorted({(j["L"], j["nrg"]) for j in jobs}):
        rs = abi.FpsynReadSpec.from_buffer_copy(rs0)
        rs.seed = 0xF1E90003 * 1000 + key[0] * 10 + key[1]
        rs.num_lengths, rs.lengths[0] = 1, key[0]
        rs.num_read_groups = key[1]
        _, pools[key] = upload(genome, rs, 16 * B, B, torch, pinned=True, device=False)
    total = sum(j["nbatches"] * B for j in jobs)
    log(f"served: {len(jobs)} jobs, {total} reads; pools of 16 x 64k reads generated in "
        f"{time.perf_counter() - t0:.1f} s")

    budget_mib = 4096
    e = Engine(genome, device=0, debug_arrays=0, engine_kind=abi.FP_ENGINE_BATCHED, max_read_length=150,
               serve_rg_slots=64, serve_hbm_budget_mib=budget_mib, baq_waves_per_simd=wpe)
    for k in range(2):   # warm-up: capture each staging buffer's graphs
        w = e.open_job(["warm"] * jobs[k]["nrg"])
        w.submit(pools[(jobs[k]["L"], jobs[k]["nrg"])][0])
        w.close()
        w.release()
    torch.cuda.synchronize()

    pending = list(range(len(jobs)))
    waiting, active = [], []
    done, deferred = {}, set()
    t0 = time.perf_counter()
    while pending or waiting or active:
        now = time.perf_counter() - t0
        while pending and jobs[pending[0]]["arrival"] <= now:
            waiting.append(pending.pop(0))
        while waiting:
            j = jobs[waiting[0]]
            h = e.open_job([f"j{waiting[0]}_rg{i}" for i in range(j["nrg"])])
            if h is None:
                deferred.add(waiting[0])
                break
            j["job"], j["next"] = h, 0
            active.append(waiting.pop(0))
        if not active:
            if pending:
                time.sleep(max(0.0, jobs[pending[0]]["arrival"] - (time.perf_counter() - t0)))
            continue
        # continuous batching: one batch of every active job per round
        for k in list(active):
            j = jobs[k]
            pool = pools[(j["L"], j["nrg"])]
            j["job"].submit(pool[(j["start"] + j["next"]) % len(pool)])
            j["next"] += 1
            if j["next"] == j["nbatches"]:
                j["job"].close()
                done[k] = time.perf_counter() - t0
                j["job"].release()
                active.remove(k)
    e.close()
    lat = sorted(done[k] - jobs[k]["arrival"] for k in done)
    makespan = max(done.values()) - jobs[0]["arrival"]
    pct = lambda q: lat[min(len(lat) - 1, int(math.ceil(q * len(lat))) - 1)]
    res = {"jobs": len(jobs), "reads": total, "value": total / makespan, "unit": "reads/s",
           "p50_latency_s": pct(0.50), "p99_latency_s": pct(0.99), "max_latency_s": lat[-1],
           "makespan_s": makespan, "arrival_span_s": jobs[-1]["arrival"] - jobs[0]["arrival"],
           "offered_reads_per_s": total / max(jobs[-1]["arrival"] - jobs[0]["arrival"], 1e-9),
           "arrival_rate_jobs_per_s": 20, "batch_reads": B, "read_group_slots": 64,
           "hbm_budget_mib": budget_mib, "jobs_deferred_by_budget": len(deferred),
           "input": "pinned host batches, H2D inside each job's latency"}
    del pools
    torch.cuda.empty_cache()
    return res


def sort_pass(args, genome, dev, wpe, ckpt, batched_opts, profile_opts, reads, torch):
    """The batched engine with the sort aggregation (covariate pairs -> LDS-tiled radix sort ->
    segmented reduce_by_key): reads/s with the digit width autotuned during the warm-up pass, then a
    per-stage pass at that width for the sort/reduce time and its HBM rate."""
    sopts = dict(batched_opts, aggregation=abi.FP_AGG_SORT)
    dt, st = timed_pass(genome, dev, 1, torch, baq_waves_per_simd=wpe, baq_checkpoint_rows=ckpt, **sopts)
    bits = int(st.sort_digit_bits)
    # the per-stage pass at 1M-read staged batches, the batch size of the committed traffic profile
    _, pst = timed_pass(genome, dev, 0, torch, baq_waves_per_simd=wpe, baq_checkpoint_rows=ckpt, profile_stages=1,
                        **dict(profile_opts, aggregation=abi.FP_AGG_SORT, sort_digit_bits=bits))
    batches = max(int(pst.cov_launches), 1)
    sort_s = pst.covariates_sort_ms / 1e3
    # algorithmic bytes: first radix pass reads every emitted slot's key twice and value once and
    # writes the real pairs; each further pass reads key (twice) and value and writes both; the
    # reduce reads the sorted pairs once.  Passes are averaged over the three tables.
    slots, pairs = float(pst.sort_slots), float(pst.sort_pairs)
    passes = float(pst.sort_passes) / (3.0 * batches)
    alg = slots * 16 + pairs * 12 + pairs * 28 * max(passes - 1, 0) + pairs * 12
    traffic = load_traffic().get("sort")
    res = {"value": reads / dt, "unit": "reads/s", "digit_bits": bits, "sort_reduce_ms": pst.covariates_sort_ms,
           "emit_ms": pst.cov_kernel_ms, "pairs": int(pairs), "slots": int(slots), "avg_passes": passes,
           "hbm": {"achieved": alg / sort_s / 1e9 if sort_s > 0 else None, "peak": HBM_PEAK_GBS,
                   "unit": "GB/s", "frac": (alg / sort_s / 1e9 / HBM_PEAK_GBS) if sort_s > 0 else None,
                   "bytes": "algorithmic (pair reads and writes of the radix passes and the reduce)"}}
    if traffic and sort_s > 0:
        meas = traffic * batches / sort_s / 1e9
        res["hbm"]["measured"] = meas
        res["hbm"]["measured_frac"] = meas / HBM_PEAK_GBS
    log(f"sort aggregation: {reads / dt:.4g} reads/s, digits {bits}, sort+reduce {pst.covariates_sort_ms:.1f} ms, "
        f"{res['hbm']['achieved']} GB/s algorithmic")
    return res


def kernel_ms(st):
    return {"prep": st.prep_kernel_ms, "baq_hmm": st.baq_hmm_kernel_ms, "cov": st.cov_kernel_ms,
            "postprocess": st.postprocessing_ms}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20,
                    help="timed steps of 500,000 reads (25 input batches); 20 = the workload's 10M reads")
    ap.add_argument("--warmup", type=int, default=5, help="untimed steps before the timed ones (same engine)")
    ap.add_argument("--batch-reads", type=int, default=20000, help="reads per input batch")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--profile-only", action="store_true",
                    help="only the per-kernel pass of the batched engine (eager stages, no graphs): the run "
                         "the committed rocprofv3 summaries profile")
    ap.add_argument("--aggregation", choices=("direct", "sort"), default="direct",
                    help="covariate aggregation of the profile-only pass")
    ap.add_argument("--no-sort", action="store_true", help="skip the sort-aggregation pass")
    ap.add_argument("--batched-only", action="store_true",
                    help="only the batched engine's timed passes (host-fed and HBM-resident), for kernel-overlap traces")
    ap.add_argument("--workload", choices=("config1", "mixed"), default="config1",
                    help="workload of the profile-only pass (mixed: the first --mixed-reads reads of the "
                         "50-250 bp set)")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--mixed-reads", type=int, default=20000000,
                    help="reads of the mixed-length pass (the workload's full 20M by default)")
    ap.add_argument("--no-genome3g", action="store_true")
    ap.add_argument("--genome3g-reads", type=int, default=4000000)
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--served-jobs", type=int, default=100)
    ap.add_argument("--cpu-sample", type=int, default=400000, help="reads of the CPU baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--stage-reads", type=int, default=1 << 21,
                    help="reads per staging buffer of the HBM-resident, mixed-length and sort batched passes "
                         "(0 = the engine's default, 2^20); the per-kernel profile and roofline passes keep "
                         "the default")
    ap.add_argument("--host-stage-reads", type=int, default=1 << 20,
                    help="reads per staging buffer of the host-fed batched pass (the value): the staging size "
                         "at which it measured fastest (0 = --stage-reads)")
    ap.add_argument("--staging-buffers", type=int, default=0,
                    help="staging buffers of the batched passes, 2 or 3 (0 = the engine's default, 3)")
    ap.add_argument("--wpe", type=int, default=0, help="BAQ waves per SIMD (0 = autotuned)")
    ap.add_argument("--ckpt", type=int, default=0,
                    help="rows per BAQ forward checkpoint: 1 (every row spilled) or 3-5 (0 = autotuned; the "
                         "profile-only pass uses 3, the autotuner's choice on config 1)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (reference point for HBM byte counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = args.copy_calibration << 20
        src = torch.empty(n, dtype=torch.uint8, device="cuda:0").fill_(1)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": n, "launches": 3}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not run here (the path shards by batch with one "
              f"table merge at the end; see the sharding paragraph of the design notes)", file=sys.stderr)
        sys.exit(2)
    if args.steps < 1 or args.warmup < 0:
        print("--steps must be >= 1 and --warmup >= 0", file=sys.stderr)
        sys.exit(2)

    import torch
    mixed_profile = args.profile_only and args.workload == "mixed"
    gs, rs, total = synth.named_spec(2 if mixed_profile else 1)
    genome = synth.Genome(gs, rs.num_read_groups)
    if mixed_profile:
        total = min(total, args.mixed_reads)
    B = args.batch_reads
    per_step = max(1, STEP_READS // B)
    # per-kernel passes (profile, roofline) stage 2^20 reads per launch, as the committed profiles do
    profile_opts = dict(engine_kind=abi.FP_ENGINE_BATCHED, max_read_length=101)
    batched_opts = dict(profile_opts)
    if args.stage_reads:
        batched_opts.update(max_batch_reads=args.stage_reads)
    if args.staging_buffers:
        batched_opts.update(staging_buffers=args.staging_buffers)
    if args.profile_only:
        log(f"generating {(total + B - 1) // B} input batches of {B} reads")
        dev = upload(genome, rs, total, B, torch)
        reads = sum(d.num_reads for d in dev)
        popts = dict(profile_opts)
        if mixed_profile:
            popts.update(max_read_length=250, frac_error_f64=1)
        if args.aggregation == "sort":
            popts.update(aggregation=abi.FP_AGG_SORT, sort_digit_bits=8)
        pdt, pst = timed_pass(genome, dev, 0, torch, baq_waves_per_simd=args.wpe or 1,
                              baq_checkpoint_rows=args.ckpt or 3, profile_stages=1, **popts)
        log(f"profile: {reads} reads in {pdt:.3f} s; prep {pst.prep_kernel_ms:.1f} ms, baq_hmm "
            f"{pst.baq_hmm_kernel_ms:.1f} ms ({pst.baq_hmm_launches} launches), 15-cell band "
            f"{pst.baq15_kernel_ms:.1f} ms ({pst.baq15_launches} launches), cov {pst.cov_kernel_ms:.1f} ms")
        print(json.dumps({"profile_only": True, "reads": reads, "seconds": pdt, "kernel_ms": kernel_ms(pst),
                          "baq15_avg_launch_ms": pst.baq15_kernel_ms / max(pst.baq15_launches, 1),
                          "baq15_cells": pst.baq15_cells, "sort_ms": pst.covariates_sort_ms,
                          "sort_pairs": pst.sort_pairs}), flush=True)
        return

    log(f"generating {(total + B - 1) // B} input batches of {B} reads (HBM and pinned host copies)")
    t0 = time.perf_counter()
    dev, pin = upload(genome, rs, total, B, torch, pinned=True)
    log(f"generated + uploaded in {time.perf_counter() - t0:.1f} s")
    K = len(dev)
    reads_all = sum(d.num_reads for d in dev)
    bytes_per_read = sum(p.nbytes for p in pin) / reads_all
    h2d = h2d_rate(torch)
    log(f"pinned H2D copy rate {h2d:.1f} GB/s; {bytes_per_read:.1f} batch bytes per read")

    # BAQ variant: the batched engine's autotuner, in an untimed pass over the first 8M reads
    wpe, ckpt = args.wpe, args.ckpt
    if not wpe:
        _, tst = timed_pass(genome, dev[:400], 0, torch, **batched_opts)
        wpe, ckpt = int(tst.baq_waves_per_simd), int(tst.baq_checkpoint_rows)
    log(f"BAQ variant: {wpe} waves per SIMD, checkpoint rows {ckpt or 1}")

    # value: host-fed batched engine, exactly --steps steps
    host_opts = dict(batched_opts)
    if args.host_stage_reads:
        host_opts.update(max_batch_reads=args.host_stage_reads)
    dt, st, reads = stepped_pass(genome, pin, args.steps, args.warmup, torch, True, per_step,
                                 baq_waves_per_simd=wpe, baq_checkpoint_rows=ckpt, **host_opts)
    value = reads / dt
    log(f"batched, host batches: {reads} reads in {dt:.3f} s -> {value:.4g} reads/s")
    ddt, _, dreads = stepped_pass(genome, dev, args.steps, args.warmup, torch, False, per_step,
                                  baq_waves_per_simd=wpe, baq_checkpoint_rows=ckpt, **batched_opts)
    log(f"batched, HBM-resident batches: {dreads / ddt:.4g} reads/s")
    if args.batched_only:
        print(json.dumps({"batched_only": True, "reads": reads, "seconds": dt, "value": value,
                          "device_resident": dreads / ddt}), flush=True)
        return
    out_plain = None
    if not args.no_plain:
        pdt, pst_, preads = stepped_pass(genome, pin, args.steps, 1, torch, True, per_step,
                                         baq_waves_per_simd=wpe)
        pdt2, _, _ = stepped_pass(genome, dev, args.steps, 1, torch, False, per_step, baq_waves_per_simd=wpe)
        out_plain = {"value": preads / pdt, "seconds": pdt, "device_resident": preads / pdt2,
                     "kernel_ms": kernel_ms(pst_)}
        log(f"plain: host batches {preads / pdt:.4g} reads/s, HBM-resident {preads / pdt2:.4g} reads/s")

    # per-kernel profile at 1M-read launches over all 10M reads: the roofline of the dominant kernel
    _, pst = timed_pass(genome, dev, 0, torch, baq_waves_per_simd=wpe, baq_checkpoint_rows=ckpt,
                        profile_stages=1, **profile_opts)
    roof = roofline(pst)
    cells_per_read = pst.baq_cells / reads_all
    log(f"profile: prep {pst.prep_kernel_ms:.1f} ms, baq_hmm {pst.baq_hmm_kernel_ms:.1f} ms, 15-cell band "
        f"{pst.baq15_kernel_ms:.1f} ms over {pst.baq15_launches} launches, cov {pst.cov_kernel_ms:.1f} ms; "
        f"baq reads {pst.baq_reads}, cells {pst.baq_cells}")
    out = {
        "metric": "reads/s BQSR+BAQ at 1 GPU, batched vs plain; p50/p99 job latency",
        "value": value,
        "unit": "reads/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": f"synthetic (seeded generator, 30 Mbp genome, 101 bp reads): {reads} reads timed, "
                f"{per_step} input batches of {B} reads per step from pinned host memory",
        "config": {"workload": "30 Mbp / 3 contigs, 10M SE 101 bp reads, 4 read groups, 30k known sites, "
                               "BAQ band 7, input batches of 20,000 reads", "reads": reads, "batch_reads": B,
                   "step_reads": per_step * B,
                   "engine": f"batched ({(args.host_stage_reads or args.stage_reads or (1 << 20)) / 2**20:g}M-read "
                             "staging buffers, hipGraph replays), host batches through fp_engine_process_batch; "
                             "per-kernel profile and roofline at 1M-read launches", "baq_waves_per_simd": wpe,
                   "baq_checkpoint_rows": ckpt or 1},
        "device_resident": {"value": dreads / ddt, "seconds": ddt,
                            "note": f"the same steps with the batches already in HBM (no PCIe), "
                                    f"{(args.stage_reads or (1 << 20)) / 2**20:g}M-read staging buffers"},
        "kernel_ms": kernel_ms(pst),
        "roofline": roof,
        "end_to_end": end_to_end(value, reads, bytes_per_read, cells_per_read, h2d),
    }
    if out_plain:
        out["plain"] = out_plain
        out["batched_over_plain"] = value / out_plain["value"]
    if not args.no_sort:
        out["sort_aggregation"] = sort_pass(args, genome, dev, wpe, ckpt, batched_opts, profile_opts, reads_all, torch)
    del dev, pin
    torch.cuda.empty_cache()

    if not args.no_mixed:
        gs2, rs2, total2 = synth.named_spec(2)
        genome2 = synth.Genome(gs2, rs2.num_read_groups)
        n2 = min(total2, args.mixed_reads)
        log(f"mixed lengths: generating {n2} reads in batches of {B}")
        dev2 = upload(genome2, rs2, n2, B, torch)
        mopts = dict(engine_kind=abi.FP_ENGINE_BATCHED, max_read_length=250, frac_error_f64=1)
        if args.stage_reads:
            mopts.update(max_batch_reads=args.stage_reads)
        if args.staging_buffers:
            mopts.update(staging_buffers=args.staging_buffers)
        pdt2, _ = timed_pass(genome2, dev2, 1, torch, baq_waves_per_simd=1, frac_error_f64=1)
        bdt2, bst2 = timed_pass(genome2, dev2, 1, torch, **mopts)
        out["mixed_length"] = {"reads": n2, "workload": "50-250 bp reads, 8 read groups, 5 % indel reads of "
                                                        "1-10 bp, clips of 0-30 bp on every read, fp64 "
                                                        "fractional errors; the 20M-read set, HBM-resident batches",
                               "plain": n2 / pdt2, "batched": n2 / bdt2, "batched_over_plain": pdt2 / bdt2,
                               "baq_waves_per_simd": int(bst2.baq_waves_per_simd),
                               "baq_checkpoint_rows": int(bst2.baq_checkpoint_rows)}
        log(f"mixed lengths: plain {n2 / pdt2:.4g} reads/s, batched {n2 / bdt2:.4g} reads/s")
        del dev2
        torch.cuda.empty_cache()

    if not args.no_genome3g:
        gs3, rs3, total3 = synth.named_spec(4)
        t0 = time.perf_counter()
        genome3 = synth.Genome(gs3, rs3.num_read_groups)
        n3 = min(total3, args.genome3g_reads)
        log(f"3 Gbp genome in {time.perf_counter() - t0:.1f} s; generating {n3} paired 150 bp reads")
        dev3 = upload(genome3, rs3, n3, B, torch)
        bdt3, bst3 = timed_pass(genome3, dev3, 1, torch, engine_kind=abi.FP_ENGINE_BATCHED,
                                max_read_length=150)
        out["genome_3gbp"] = {"reads": n3, "value": n3 / bdt3, "unit": "reads/s",
                              "workload": "3 Gbp / 24 contigs resident in HBM (1.5 GB packed) with 1M known "
                                          "sites, paired 150 bp reads, 16 read groups, 4 % indels, clips 0-40; "
                                          "first reads of the 600M-read set, HBM-resident batches",
                              "baq_waves_per_simd": int(bst3.baq_waves_per_simd),
                              "baq_checkpoint_rows": int(bst3.baq_checkpoint_rows)}
        log(f"3 Gbp genome: batched {n3 / bdt3:.4g} reads/s")
        del dev3, genome3
        torch.cuda.empty_cache()

    if not args.no_served:
        out["served"] = served(args, wpe, torch)
        log(f"served: {out['served']['value']:.4g} reads/s, p50 {out['served']['p50_latency_s']:.3f} s, "
            f"p99 {out['served']['p99_latency_s']:.3f} s")

    if not args.no_cpu_baseline:
        log("timing the CPU oracle baseline")
        threads = max(1, min(16, os.cpu_count() or 1))
        out["cpu_baseline"] = cpu_baseline(genome, rs, min(args.cpu_sample, total), threads)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
