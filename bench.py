"""Benchmark: reads/s of the BQSR+BAQ device pipeline on one MI355X.

Workload: 30 Mbp genome in 3 contigs, 10M single-end 101 bp reads over 4 read
groups (substitution rates 0.5/1.0/1.5/2.0 %), 3 % indel reads, 5 % soft
clips, 30k known sites, processed in batches of 20,000 reads (the reference's
GPU batch size, firepony.cu:148).  All batches are uploaded to HBM before the
timed region; the timed region covers every batch plus postprocess, which is
how the reference defines reads/s (firepony.cu:348-376, 222).

    python bench.py [--gpus 1] [--steps K] [--warmup W]

A step is one batch.  The JSON line reports the engine's whole-job reads/s,
the dominant kernel's roofline fraction (HIP events on the engine's stream),
and the CPU oracle timed on a bounded sample of the same workload.
"""
import argparse
import ctypes as C
import json
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
BAQ_FLOPS_PER_CELL = 43          # forward 17 + rescale 3, backward 16 + rescale 3, MAP 4


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


def roofline(stats, total_reads, bases):
    """Roofline of the kernel with the most device time."""
    k = {
        "baq_hmm": (stats.baq_hmm_kernel_ms, stats.baq_hmm_launches),
        "cov": (stats.cov_kernel_ms, stats.cov_launches),
        "prep": (stats.prep_kernel_ms, stats.prep_launches),
    }
    name = max(k, key=lambda x: k[x][0])
    ms, launches = k[name]
    avg_s = ms / 1e3 / max(launches, 1)
    if name == "baq_hmm":
        flops = BAQ_FLOPS_PER_CELL * stats.baq_cells / max(launches, 1)
        ach = flops / avg_s / 1e12
        return name, {"bound": "mfma", "achieved": ach, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                      "frac": ach / FP64_PEAK_TFLOPS, "traffic": None}
    # HBM-bound kernels: compulsory bytes of the inputs they read and outputs they write
    per_read = {"prep": 49 + 13 * 4 + 101 + 4 * 1.2 + 54, "cov": 101 + 13 * 4 + 101 + 101 + 8}[name]
    byts = per_read * total_reads / max(launches, 1)
    ach = byts / avg_s / 1e9
    return name, {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                  "frac": ach / HBM_PEAK_GBS, "traffic": None}


def cpu_baseline(genome, rs, sample_reads):
    """The CPU oracle (a restatement of the reference host path) on the first
    sample_reads reads, single thread, batches of 1,000 (firepony.cu:185-190)."""
    from oracle.pyoracle import Oracle
    o = Oracle(genome)
    batches = [synth.Batch(genome, rs, f, min(1000, sample_reads - f), threads=4)
               for f in range(0, sample_reads, 1000)]
    t0 = time.perf_counter()
    for b in batches:
        o.process(b)
    o.postprocess()
    dt = time.perf_counter() - t0
    return {"value": sample_reads / dt, "unit": "reads/s", "cores": 1, "kind": "port",
            "sample": f"first {sample_reads} reads of the workload, batches of 1000, single thread, "
                      f"{dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch-reads", type=int, default=20000)
    ap.add_argument("--cpu-sample", type=int, default=100000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--wpe", type=int, default=0, help="BAQ waves per SIMD (0 = engine default)")
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
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported (the path would shard by batch, "
              f"see the sharding paragraph of the design notes)", file=sys.stderr)
        sys.exit(2)

    import torch
    gs, rs, total = synth.named_spec(1)
    genome = synth.Genome(gs, rs.num_read_groups)
    B = args.batch_reads
    K = min(args.steps, (total + B - 1) // B)
    if B != 20000 and args.steps == 500:
        K = (total + B - 1) // B
    log(f"generating {K} batches of {B} reads")
    t0 = time.perf_counter()
    dev = []
    nthreads = min(16, os.cpu_count() or 4)
    for i in range(K):
        hb = synth.Batch(genome, rs, i * B, min(B, total - i * B), threads=nthreads)
        dev.append(DeviceBatch(hb, torch))
        del hb
    torch.cuda.synchronize()
    log(f"generated + uploaded in {time.perf_counter() - t0:.1f} s")

    e = Engine(genome, device=0, baq_waves_per_simd=args.wpe, debug_arrays=0)
    for i in range(args.warmup):
        e.process_device(dev[i % K].view)
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
    reads = sum(d.num_reads for d in dev)
    bases = sum(d.num_bases for d in dev)
    st = e.stats()
    value = reads / dt
    kname, roof = roofline(st, reads, bases)
    log(f"plain engine: {reads} reads in {dt:.3f} s -> {value:.4g} reads/s; kernels: prep {st.prep_kernel_ms:.1f} ms, "
        f"baq_hmm {st.baq_hmm_kernel_ms:.1f} ms ({st.baq_hmm_launches} launches), cov {st.cov_kernel_ms:.1f} ms, "
        f"postprocess {st.postprocessing_ms:.1f} ms; baq reads {st.baq_reads}, cells {st.baq_cells}")

    out = {
        "metric": "reads/s BQSR+BAQ at 1 GPU, batched vs plain; p50/p99 job latency",
        "value": value,
        "unit": "reads/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3 / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded generator, configs[1] shape)",
        "config": {"workload": "30 Mbp / 3 contigs, 10M SE 101 bp reads, 4 read groups, 30k known sites, "
                               f"batches of {B} reads, BAQ band 7", "reads": reads, "batch_reads": B,
                   "engine": "plain"},
        "roofline": dict(roof, kernel=kname),
    }
    if not args.no_cpu_baseline:
        log("timing the CPU oracle baseline")
        out["cpu_baseline"] = cpu_baseline(genome, rs, min(args.cpu_sample, total))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
