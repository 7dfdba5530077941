"""Edge cases and full-size properties of the HIP path.

- empty batches and batches whose reads are all filtered leave the tables as they were;
- the batched engine takes reads of exactly max_read_length and refuses longer ones with an error;
- at the full size of the 10M-read workload the batched engine (staging buffers, two streams,
  graph replays) and the plain engine give identical tables: the sums are exact, so any lost,
  doubled or raced update would show;
- 2M reads of the same workload against the oracle fed the same batches (16 host threads);
- batches whose BAQ reads all have bands wider than 15 cells (the 15-cell kernel has nothing to do
  while the wider classes run on their own stream), and batches with only wider-than-33 bands.
"""
import sys
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from firepony_amd import abi, synth
from firepony_amd.engine import Engine, FireponyError
from oracle.pyoracle import Oracle

from handbatch import HandBatch, HandGenome, random_genome_string

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


TABLES = (abi.FP_TABLE_QUALITY, abi.FP_TABLE_CYCLE, abi.FP_TABLE_CONTEXT)


def same_tables(a, b, what):
    for t in TABLES:
        x, y = a.table(t), b.table(t)
        for k in ("keys", "observations", "mismatches_fixed_lo", "mismatches_fixed_hi"):
            np.testing.assert_array_equal(y[k], x[k], err_msg=f"{what}: table {t} {k}")


def hand_reads(contig, n, rng, mapq=60, length=60):
    reads = []
    for i in range(n):
        start = int(rng.integers(0, len(contig) - length - 1))
        seq = list(contig[start:start + length])
        if i % 3 == 0:
            seq[length // 2] = "A" if seq[length // 2] != "A" else "C"
        reads.append({"seq": "".join(seq), "qual": 30, "cigar": f"{length}M", "start": start, "mapq": mapq})
    return reads


@pytest.mark.parametrize("kind", [abi.FP_ENGINE_PLAIN, abi.FP_ENGINE_BATCHED])
def test_empty_and_filtered_batches(tmp_path, kind):
    say(f"empty and all-filtered batches, engine kind {kind}")
    rng = np.random.default_rng(3)
    contig = random_genome_string(rng, 5000)
    g = HandGenome([contig], sites=[(0, 100, 100)])
    good = HandBatch(hand_reads(contig, 40, rng))
    filtered = HandBatch(hand_reads(contig, 30, rng, mapq=0))      # MAPQ 0: every read filtered
    empty = HandBatch([])
    opts = dict(engine_kind=kind)
    if kind == abi.FP_ENGINE_BATCHED:
        opts.update(max_batch_reads=256, max_read_length=64)
    e = Engine(g, **opts)
    ref = Engine(g, **opts)
    for b in (empty, filtered, good, empty, filtered):
        e.process(b)
    ref.process(good)
    same_tables(ref, e, "with empty and filtered batches")
    st = e.stats()
    assert st.total_reads == 40 + 2 * 30
    assert st.filtered_reads == 2 * 30
    # only filtered reads: empty tables, and postprocess and the report still work
    only = Engine(g, **opts)
    only.process(filtered)
    only.process(empty)
    for t in TABLES:
        assert only.table(t)["keys"].size == 0
    only.postprocess()
    o = Oracle(g)
    o.process(filtered)
    o.postprocess()
    assert only.report(str(tmp_path / "e.txt")) == o.report(str(tmp_path / "o.txt"))
    for x in (e, ref, only):
        x.close()


def test_read_length_limit_of_the_batched_engine():
    say("batched engine: reads of exactly max_read_length pass, longer reads are refused")
    rng = np.random.default_rng(4)
    contig = random_genome_string(rng, 4000)
    g = HandGenome([contig])
    exact = HandBatch(hand_reads(contig, 20, rng, length=80))
    longer = HandBatch(hand_reads(contig, 5, rng, length=81))
    e = Engine(g, engine_kind=abi.FP_ENGINE_BATCHED, max_batch_reads=128, max_read_length=80)
    e.process(exact)
    e.synchronize()
    plain = Engine(g)
    plain.process(exact)
    same_tables(plain, e, "reads of max_read_length")
    with pytest.raises(FireponyError, match="max_read_length"):
        e.process(longer)
        e.synchronize()
    e.close()
    plain.close()


def test_full_size_batched_equals_plain():
    say("10M reads of the 101 bp workload: batched engine vs plain engine, exact tables")
    gs, rs, n = synth.named_spec(1)
    g = synth.Genome(gs, rs.num_read_groups)
    plain = Engine(g)
    bat = Engine(g, engine_kind=abi.FP_ENGINE_BATCHED, max_read_length=101)
    step = 100000
    with ThreadPoolExecutor(max_workers=1) as ex:
        fut = ex.submit(synth.Batch, g, rs, 0, step, 16)
        for first in range(0, n, step):
            b = fut.result()
            if first + step < n:
                fut = ex.submit(synth.Batch, g, rs, first + step, min(step, n - first - step), 16)
            plain.process(b)
            bat.process(b)
            if first % 2000000 == 0:
                say(f"  {first + b.view.num_reads} reads")
    bat.synchronize()
    sp, sb = plain.stats(), bat.stats()
    assert sp.total_reads == sb.total_reads == n
    assert (sp.filtered_reads, sp.baq_reads, sp.baq_cells) == (sb.filtered_reads, sb.baq_reads, sb.baq_cells)
    same_tables(plain, bat, "10M reads")
    plain.close()
    bat.close()


def test_two_million_reads_against_the_oracle(tmp_path):
    say("2M reads of the 101 bp workload: batched engine vs the oracle (16 host threads, exact gather)")
    gs, rs, n = synth.named_spec(1)
    g = synth.Genome(gs, rs.num_read_groups)
    n = 2000000
    step = 20000
    batches = [synth.Batch(g, rs, f, step, 16) for f in range(0, n, step)]
    bat = Engine(g, engine_kind=abi.FP_ENGINE_BATCHED, max_read_length=101)
    for b in batches:
        bat.process(b)
    bat.synchronize()
    say("  engine done; oracle on 16 threads")
    threads = 16
    oracles = [Oracle(g) for _ in range(threads)]

    def work(t):
        for b in batches[t::threads]:
            oracles[t].process(b)

    with ThreadPoolExecutor(max_workers=threads) as ex:
        list(ex.map(work, range(threads)))
    for o in oracles[1:]:
        oracles[0].gather(o)
    same_tables(oracles[0], bat, "2M reads vs oracle")
    oracles[0].postprocess()
    bat.postprocess()
    assert oracles[0].report(str(tmp_path / "o.txt")) == bat.report(str(tmp_path / "e.txt"))
    bat.close()


def long_indel_reads(contig, rng, n):
    """101 bp reads with one 9-15 bp deletion or insertion and a few mismatches: their BAQ bands
    are wider than the 15-cell class."""
    other = {"A": "C", "C": "G", "G": "T", "T": "A"}
    reads = []
    for i in range(n):
        L, dl, a = 101, int(rng.integers(9, 16)), int(rng.integers(30, 60))
        start = int(rng.integers(10, len(contig) - L - dl - 10))
        if i % 2 == 0:
            ref = contig[start:start + L + dl]
            seq, cig = ref[:a] + ref[a + dl:a + dl + (L - a)], f"{a}M{dl}D{L - a}M"
        else:
            ref = contig[start:start + L - dl]
            seq = ref[:a] + "".join(rng.choice(list("ACGT"), size=dl)) + ref[a:a + (L - a - dl)]
            cig = f"{a}M{dl}I{L - a - dl}M"
        seq = "".join(other[c] if rng.random() < 0.02 else c for c in seq)
        reads.append({"seq": seq, "qual": [int(q) for q in rng.integers(10, 41, size=L)], "cigar": cig,
                      "start": start, "flags": int(rng.choice([0, 16]))})
    return reads


@pytest.mark.parametrize("kind", [abi.FP_ENGINE_PLAIN, abi.FP_ENGINE_BATCHED])
def test_batches_without_the_15_cell_class(tmp_path, kind):
    say(f"BAQ bands of 17-33 cells only, then wider-than-33 only (no 15-cell reads), engine kind {kind}")
    rng = np.random.default_rng(9)
    contig = random_genome_string(rng, 20000)
    g = HandGenome([contig])
    cand = long_indel_reads(contig, rng, 400)
    probe = Oracle(g)
    probe.process(HandBatch(cand))
    w = 2 * probe.debug(abi.FP_DBG_BANDWIDTH, np.uint16).astype(np.int64) + 1
    nerr = probe.debug(abi.FP_DBG_NUM_ERRORS, np.uint16)
    mid = [cand[i] for i in range(len(cand)) if nerr[i] and 17 <= w[i] <= 33]
    wide = [cand[i] for i in range(len(cand)) if nerr[i] and w[i] > 33]
    assert len(mid) > 20 and len(wide) > 20
    batches = [HandBatch(mid), HandBatch(wide), HandBatch(mid[:7])]
    o = Oracle(g)
    opts = dict(engine_kind=kind)
    if kind == abi.FP_ENGINE_BATCHED:
        opts.update(max_batch_reads=128, max_read_length=101)
    e = Engine(g, **opts)
    for b in batches:
        o.process(b)
        e.process(b)
    e.synchronize()
    same_tables(o, e, "wide bands only")
    o.postprocess()
    e.postprocess()
    assert e.report(str(tmp_path / "e.txt")) == o.report(str(tmp_path / "o.txt"))
    e.close()
