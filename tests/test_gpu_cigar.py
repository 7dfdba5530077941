"""Reads with long CIGARs on the GPU vs the oracle.

The prep and covariate kernels keep the first 64 CIGAR ops of a read in registers, one op per
lane, and read any further ops from memory; the covariate kernel also software-pipelines reads of
up to 128 window bases and takes longer ones one at a time.  These batches mix reads of 63, 64,
65 and up to 160 ops with plain reads, at lengths on either side of 128 and 256, forward and
reverse, with mismatches, Q2 runs and known sites, so every one of those paths meets the oracle:
stage by stage on the plain engine, tables and report on the batched one.
"""
import sys

import numpy as np
import pytest

from firepony_amd import abi
from firepony_amd.engine import Engine
from oracle.pyoracle import Oracle

from handbatch import HandBatch, HandGenome, random_genome_string
from test_gpu_parity import compare_batch, compare_final, compare_obs_tables

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


OTHER = {"A": "C", "C": "G", "G": "T", "T": "A"}


def indel_read(contig, rng, length, nops):
    """A read of `length` bases whose CIGAR has exactly `nops` ops: soft clips at both ends around
    short M runs separated by single-base insertions and deletions."""
    for _ in range(100):
        ops = [("S", 2)]
        used = 2
        while len(ops) < nops - 2 and used < length - 6:
            m = int(rng.integers(1, 3))
            ops.append(("M", m))
            used += m
            if len(ops) == nops - 2:
                break
            if rng.random() < 0.5:
                ops.append(("I", 1))
                used += 1
            else:
                ops.append(("D", 1))
        tail = length - used - 2
        if tail < 1 or len(ops) != nops - 2:
            continue
        ops += [("M", tail), ("S", 2)]
        span = sum(n for op, n in ops if op in "MD")
        start = int(rng.integers(10, len(contig) - span - 10))
        seq, pos = [], start
        for op, n in ops:
            if op == "M":
                for _ in range(n):
                    b = contig[pos]
                    seq.append(OTHER[b] if rng.random() < 0.03 else b)
                    pos += 1
            elif op == "D":
                pos += n
            else:
                seq += list(rng.choice(list("ACGT"), size=n))
        qual = [int(q) for q in rng.integers(5, 41, size=length)]
        if rng.random() < 0.3:
            t = int(rng.integers(1, 12))
            qual[-t:] = [2] * t
        return {"seq": "".join(seq), "qual": qual, "cigar": "".join(f"{n}{op}" for op, n in ops), "start": start,
                "flags": int(rng.choice([0, 16, 1 | 64, 1 | 128 | 16])), "rg": int(rng.integers(0, 2))}
    raise AssertionError("no read layout found")


def plain_read(contig, rng, length):
    start = int(rng.integers(10, len(contig) - length - 10))
    seq = [OTHER[b] if rng.random() < 0.02 else b for b in contig[start:start + length]]
    return {"seq": "".join(seq), "qual": [int(q) for q in rng.integers(5, 41, size=length)],
            "cigar": f"{length}M", "start": start, "flags": int(rng.choice([0, 16])), "rg": int(rng.integers(0, 2))}


def batches(seed):
    rng = np.random.default_rng(seed)
    contig = random_genome_string(rng, 30000)
    sites = sorted({(0, int(p), int(p)) for p in rng.integers(50, 29900, size=300)})
    g = HandGenome([contig], sites=sites, read_groups=("rg0", "rg1"))
    out = []
    for _ in range(3):
        reads = []
        for length, nops in [(101, 63), (101, 64), (101, 65), (101, 70), (101, 90), (120, 110), (128, 120),
                             (129, 121), (200, 130), (250, 160), (256, 150), (300, 160)]:
            reads.append(indel_read(contig, rng, length, nops))
            reads += [plain_read(contig, rng, int(rng.choice([60, 101, 150]))) for _ in range(3)]
        order = rng.permutation(len(reads))
        out.append(HandBatch([reads[i] for i in order]))
    return g, out


def test_long_cigars_plain_engine_stage_by_stage(tmp_path):
    say("long CIGARs (63-160 ops), plain engine vs oracle, stage by stage")
    g, bs = batches(31)
    for b in bs:
        assert int(b.cigar_len.max()) > 64
    o, e = Oracle(g), Engine(g)
    for i, b in enumerate(bs):
        say(f"  batch {i}: {b.view.num_reads} reads, up to {int(b.cigar_len.max())} ops")
        o.process(b)
        e.process(b)
        compare_batch(o, e, b)
    compare_obs_tables(o, e)
    o.postprocess()
    e.postprocess()
    compare_final(o, e, tmp_path)
    e.close()


def test_long_cigars_batched_engine(tmp_path):
    say("long CIGARs (63-160 ops), batched engine vs oracle: tables and report")
    g, bs = batches(32)
    o = Oracle(g)
    e = Engine(g, engine_kind=abi.FP_ENGINE_BATCHED, max_batch_reads=256, max_read_length=300)
    for b in bs:
        o.process(b)
        e.process(b)
    e.synchronize()
    compare_obs_tables(o, e)
    o.postprocess()
    e.postprocess()
    compare_final(o, e, tmp_path)
    e.close()


def layout_read(contig, rng):
    """A short read with a random layout of soft clips, insertions and deletions around one long
    M run: clips of 0-3 bases at either end, then 1-4 ops of 1-3 bases drawn from M, I and D on
    each side of the run, so insertions and deletions sit next to the clips and to each other."""
    ops = []
    lead = int(rng.integers(0, 4))
    if lead:
        ops.append(("S", lead))
    for _ in range(int(rng.integers(1, 5))):
        ops.append((str(rng.choice(list("MID"))), int(rng.integers(1, 4))))
    ops.append(("M", int(rng.integers(8, 30))))
    for _ in range(int(rng.integers(1, 5))):
        ops.append((str(rng.choice(list("MID"))), int(rng.integers(1, 4))))
    trail = int(rng.integers(0, 4))
    if trail:
        ops.append(("S", trail))
    span = sum(n for op, n in ops if op in "MD")
    start = int(rng.integers(20, len(contig) - span - 20))
    seq, pos = [], start
    for op, n in ops:
        if op == "M":
            for _ in range(n):
                b = contig[pos]
                seq.append(OTHER[b] if rng.random() < 0.05 else b)
                pos += 1
        elif op == "D":
            pos += n
        else:
            seq += list(rng.choice(list("ACGT"), size=n))
    length = len(seq)
    qual = [int(q) for q in rng.integers(5, 41, size=length)]
    return {"seq": "".join(seq), "qual": qual, "cigar": "".join(f"{n}{op}" for op, n in ops), "start": start,
            "flags": int(rng.choice([0, 16])), "rg": int(rng.integers(0, 2))}


def test_short_cigar_layouts_stage_by_stage(tmp_path):
    """Windows, known-site masks and error bits of reads whose clips, insertions and deletions sit
    side by side (the walks that place an event inside its CIGAR op by arithmetic), with a known
    site at every third reference position, plain engine vs oracle stage by stage."""
    say("short CIGAR layouts around clips and indels, dense known sites, plain engine vs oracle")
    rng = np.random.default_rng(41)
    contig = random_genome_string(rng, 4000)
    sites = [(0, p, p) for p in range(10, 3990, 3)]
    g = HandGenome([contig], sites=sites, read_groups=("rg0", "rg1"))
    o, e = Oracle(g), Engine(g)
    for i in range(3):
        b = HandBatch([layout_read(contig, rng) for _ in range(400)])
        say(f"  batch {i}: {b.view.num_reads} reads")
        o.process(b)
        e.process(b)
        compare_batch(o, e, b)
    compare_obs_tables(o, e)
    o.postprocess()
    e.postprocess()
    compare_final(o, e, tmp_path)
    e.close()
