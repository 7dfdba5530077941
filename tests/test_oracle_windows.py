"""Pins for the oracle's read windows, error bits, HMM windows, bandwidths and
BQ tags, against an independent model (tests/refmodel.py) written from the
reference's source text.

- Every CIGAR of one to three ops of M/I/D/S/H with lengths 1..6, every
  four-op CIGAR with lengths 1..3 and a sample of four-op CIGARs with lengths
  up to 6, each on a forward and a reverse read whose aligned bases all
  mismatch: the active list (filters), the clipped, no-insertion and reference
  windows (cigar.cu:347-534), the SNP / insertion / deletion bits and
  num_errors (cigar.cu:616-737), the HMM window and bandwidth (baq.cu:89-187).
- Random reads of 20-45 bases with indels, clips, adapter-clipped mates, reads
  at the contig start (the clamped HMM window) and at the contig end (reads
  dropped from BAQ): every BQ tag (flat, capped, recoded; baq.cu:842-1081),
  with the HMM evaluated by an unbanded forward-backward of the model.
  The cap's fallback to the original qualities (baq.cu:1048-1055) needs a
  soft-clipped base inside the clipped window of a read that has an error in
  that window; soft-clip removal moves each window end out of an S run unless
  the whole window lies inside one, where no error can be counted, so a read
  that reaches BAQ never takes it (a search over paired, adapter-clipped reads
  with trailing and leading clips found none).
- Base masks: non-ACGT bases and qualities below 6 (read_filters.cu:376-438).
- Expected errors: obs * 10^(-q/10) per row (expected_error.h:37-76).
"""
import itertools

import numpy as np
import pytest

from firepony_amd import abi
from oracle import pyoracle
from oracle.pyoracle import Oracle

import refmodel as rm
from handbatch import HandBatch, HandGenome, random_genome_string

CODE = {"A": 1, "C": 2, "G": 4, "T": 8, "N": 15}
LET = {1: "A", 2: "C", 4: "G", 8: "T", 15: "N"}
OPCH = "MIDNSHP=X"


def cigar_str(ops):
    return "".join(f"{n}{OPCH[op]}" for op, n in ops)


def make_read(contig, ops, start, flags=0, rng=None, mismatch_all=True, qual=30, **extra):
    """Bases: every aligned base mismatches the reference (or matches, mismatch_all=False); inserted
    and clipped bases are arbitrary."""
    L = rm.read_length(ops)
    ev, rc, fc = rm.expand(ops)
    seq = ["A"] * L
    for e, c, f in zip(ev, rc, fc):
        if c == rm.NONE16:
            continue
        if e == rm.M_ and start + f < len(contig):
            b = contig[start + f]
            seq[c] = ("C" if b == "A" else "A") if mismatch_all else b
        elif rng is not None:
            seq[c] = "ACGT"[int(rng.integers(4))]
    d = dict(seq="".join(seq), cigar=cigar_str(ops), start=start, flags=flags, qual=qual)
    d.update(extra)
    return d


def model_read(r, contig):
    ops = [(rm.OP[c], int(n)) for n, c in __import__("re").findall(r"(\d+)([MIDNSHP=X])", r["cigar"])]
    span = sum(n for op, n in ops if op in (0, 2, 3, 7, 8))
    q = r["qual"]
    q = [q] * len(r["seq"]) if isinstance(q, int) else list(q)
    ref_at = lambda p: CODE[contig[p]] if 0 <= p < len(contig) else 15
    return dict(ops=ops, start=r["start"], flags=r.get("flags", 0), mapq=r.get("mapq", 60),
                isize=r.get("isize", 0), mate_start=r.get("mate_start", 0),
                stop=r["start"] + max(span, 1) - 1, seq=[CODE[c] for c in r["seq"]], qual=q,
                ref_at=ref_at, contig_len=len(contig))


def oracle_run(genome, reads):
    o = Oracle(genome)
    b = HandBatch(reads)
    o.process(b)
    return o, b


def small_cigars():
    ops = [0, 1, 2, 4, 5]   # M I D S H
    out = []
    for k in (1, 2, 3):
        for kinds in itertools.product(ops, repeat=k):
            for lens in itertools.product(range(1, 7), repeat=k):
                out.append(list(zip(kinds, lens)))
    for kinds in itertools.product(ops, repeat=4):
        for lens in itertools.product(range(1, 4), repeat=4):
            out.append(list(zip(kinds, lens)))
    rng = np.random.default_rng(17)
    for _ in range(20000):
        out.append([(int(rng.choice(ops)), int(rng.integers(1, 7))) for _ in range(4)])
    return [c for c in out if rm.read_length(c) > 0]


@pytest.fixture(scope="module")
def contig():
    return random_genome_string(np.random.default_rng(23), 60000)


def test_windows_and_error_bits_brute_force(contig):
    g = HandGenome([contig])
    cigars = small_cigars()
    reads = []
    for k, ops in enumerate(cigars):
        for flags in (0, 16):
            reads.append(make_read(contig, ops, 500 + (k * 13) % 50000, flags))
    o, b = oracle_run(g, reads)
    act = o.debug(abi.FP_DBG_ACTIVE_READS, np.uint32).astype(np.int64)
    models = [model_read(r, contig) for r in reads]
    exp_active = [i for i, m in enumerate(models) if rm.is_active(m, len(contig))]
    # BAQ drops reads whose HMM window passes the contig end; none sit near it here
    assert list(act) == exp_active, "active read list (read filters)"
    rw = o.debug(abi.FP_DBG_READ_WINDOW, np.uint32)
    nw = o.debug(abi.FP_DBG_NOINS_WINDOW, np.uint32)
    fw = o.debug(abi.FP_DBG_REF_WINDOW, np.uint32)
    ne = o.debug(abi.FP_DBG_NUM_ERRORS, np.uint16)
    hw = o.debug(abi.FP_DBG_HMM_WINDOW, np.uint32)
    bw = o.debug(abi.FP_DBG_BANDWIDTH, np.uint16)
    fl = o.debug(abi.FP_DBG_ERROR_FLAGS, np.uint8)
    checked = baq_checked = 0
    for i in exp_active:
        m = models[i]
        try:
            w = rm.windows(m, None)
        except rm.Undefined:
            continue
        what = f"read {i} {reads[i]['cigar']} flags {reads[i]['flags']}"
        assert (int(rw[i]) & 0xFFFF, int(rw[i]) >> 16) == (w["x"], w["y"]), f"read window, {what}"
        assert (int(nw[i]) & 0xFFFF, int(nw[i]) >> 16) == (w["nix"], w["niy"]), f"no-insertion window, {what}"
        assert (int(fw[i]) & 0xFFFF, int(fw[i]) >> 16) == (w["fx"], w["fy"]), f"reference window, {what}"
        assert int(ne[i]) == w["nerr"], f"num_errors, {what}"
        L = rm.read_length(m["ops"])
        bits = fl[b.read_start[i]:b.read_start[i] + L]
        assert {j for j in range(L) if bits[j] & 1} == w["snp"], f"SNP bits, {what}"
        assert {j for j in range(L) if bits[j] & 2} == w["ins"], f"insertion bits, {what}"
        assert {j for j in range(L) if bits[j] & 4} == w["dele"], f"deletion bits, {what}"
        checked += 1
        if w["nerr"]:
            h = int(hw[i])
            s16 = lambda v: ((v + 0x8000) & 0xFFFF) - 0x8000
            assert (s16(h & 0xFFFF), s16(h >> 16)) == (w["hx"], w["hy"]), f"HMM window, {what}"
            assert int(bw[i]) == w["bw"], f"bandwidth, {what}"
            baq_checked += 1
    assert checked > 10000 and baq_checked > 5000, (checked, baq_checked)


def random_reads(contig, rng, n):
    """Reads of 20-45 bases: indels inside, clips at either end, some paired with adapter clipping,
    a few at the contig start and end."""
    reads = []
    clen = len(contig)
    for k in range(n):
        body = [(0, int(rng.integers(6, 14)))]
        for _ in range(int(rng.integers(0, 3))):
            body += [(int(rng.choice([1, 2])), int(rng.integers(1, 4))), (0, int(rng.integers(5, 14)))]
        ops = list(body)
        if rng.random() < 0.3:
            ops = [(4, int(rng.integers(1, 6)))] + ops
        if rng.random() < 0.3:
            ops = ops + [(4, int(rng.integers(1, 6)))]
        if rng.random() < 0.1:
            ops = [(5, int(rng.integers(1, 4)))] + ops
        span = sum(l for op, l in ops if op in (0, 2))
        where = rng.random()
        if where < 0.08:
            start = int(rng.integers(0, 3))                        # HMM window clamped at the contig start
        elif where < 0.16:
            start = clen - span - int(rng.integers(0, 4))           # window past the end: dropped from BAQ
        else:
            start = int(rng.integers(50, clen - 200))
        flags = 16 if rng.random() < 0.5 else 0
        extra = {}
        if rng.random() < 0.25:
            # paired, mates on opposite strands, the mate's start inside this read: adapter clipping
            flags |= 1 | 2 | 64 | (0 if flags & 16 else 32)
            t = int(rng.integers(span // 3, span + 5))
            if flags & 16:
                extra = dict(mate_start=start + int(rng.integers(1, max(2, span - 2))), isize=-t)
            else:
                extra = dict(mate_start=start + t - span + int(rng.integers(0, 10)), isize=t)
        qual = [int(x) for x in rng.integers(2, 42, size=rm.read_length(ops))]
        r = make_read(contig, ops, start, flags, rng=rng, mismatch_all=False, qual=qual, **extra)
        # a few substitutions on aligned bases
        seq = list(r["seq"])
        for _ in range(int(rng.integers(0, 4))):
            j = int(rng.integers(len(seq)))
            seq[j] = "ACGT"[(("ACGT".index(seq[j]) if seq[j] in "ACGT" else 0) + 1) % 4]
        r["seq"] = "".join(seq)
        reads.append(r)
    return reads


def test_baq_tags_against_independent_model():
    rng = np.random.default_rng(31)
    contig = random_genome_string(rng, 3000)
    g = HandGenome([contig])
    reads = random_reads(contig, rng, 700)
    o, b = oracle_run(g, reads)
    act = set(int(x) for x in o.debug(abi.FP_DBG_ACTIVE_READS, np.uint32))
    tags = o.debug(abi.FP_DBG_BAQ, np.uint8)
    qs = b.qual_start
    compared = {"flat": 0, "baq": 0, "dropped": 0, "clamped": 0, "fallback": 0}
    for i, r in enumerate(reads):
        m = model_read(r, contig)
        if not rm.is_active(m, len(contig)):
            assert i not in act
            continue
        try:
            w = rm.windows(m, None)
        except rm.Undefined:
            continue
        if w["dropped"]:
            assert i not in act, f"read {i} should leave the active list (HMM window past the contig end)"
            compared["dropped"] += 1
            continue
        assert i in act
        exp = rm.baq_tags(m, w, rm.hmm_independent)
        if exp is None:
            continue   # a MAP quality within rounding of an integer boundary
        L = rm.read_length(m["ops"])
        got = [int(t) for t in tags[qs[i]:qs[i] + L]]
        assert got == exp, f"BQ tags of read {i} {r['cigar']} flags {r['flags']} start {r['start']}"
        if w["nerr"] == 0:
            compared["flat"] += 1
        else:
            compared["baq"] += 1
            if w["hx"] < 0 and r["start"] < abs(w["hx"]):
                compared["clamped"] += 1
            if exp[w["x"]:w["y"] + 1] == [64] * (w["y"] - w["x"] + 1) and w["nerr"]:
                compared["fallback"] += 1
    assert compared["baq"] > 300 and compared["flat"] > 20, compared
    assert compared["dropped"] > 5 and compared["clamped"] > 5, compared


def test_base_masks_non_acgt_and_low_quality():
    rng = np.random.default_rng(5)
    contig = random_genome_string(rng, 2000)
    g = HandGenome([contig])
    seq = list(contig[300:360])
    quals = [30] * 60
    for j in (3, 20, 41):
        seq[j] = "N"
    for j, q in ((7, 5), (8, 6), (30, 2), (31, 0), (50, 41)):
        quals[j] = q
    o, b = oracle_run(g, [dict(seq="".join(seq), cigar="60M", start=300, qual=quals)])
    mask = o.debug(abi.FP_DBG_ACTIVE_MASK, np.uint8)[:60]
    # masked: non-ACGT bases and qualities below MIN_USABLE_Q_SCORE = 6 (read_filters.cu:403)
    exp = [0 if (seq[j] not in "ACGT" or quals[j] < 6) else 1 for j in range(60)]
    assert list(mask) == exp


@pytest.mark.parametrize("table", [abi.FP_TABLE_QUALITY, abi.FP_TABLE_CYCLE, abi.FP_TABLE_CONTEXT])
def test_expected_errors_closed_form(table):
    from firepony_amd import synth
    gs, rs, n = synth.named_spec(0)
    gen = synth.Genome(gs, rs.num_read_groups)
    o = Oracle(gen)
    o.process(synth.Batch(gen, rs, 0, 1500))
    o.postprocess()
    t = o.table(table)
    shift = {abi.FP_TABLE_QUALITY: 2, abi.FP_TABLE_CYCLE: 12, abi.FP_TABLE_CONTEXT: 13}[table]
    q = (t["keys"].astype(np.int64) >> shift) & 255
    # expected = observations * 10^(-q/10), the double product of the count and the quality's
    # error probability (expected_error.h:47-50)
    exp = t["observations"].astype(np.float64) * np.array([10.0 ** (qq / -10.0) for qq in q])
    np.testing.assert_array_equal(t["expected_errors"], exp)
    # the read-group table sums them over qualities in key order (read_group_table.cu:62-93)
    t0 = o.table(abi.FP_TABLE_READ_GROUP)
    if table == abi.FP_TABLE_QUALITY:
        for key, e0 in zip(t0["keys"], t0["expected_errors"]):
            sel = (t["keys"] & ~np.uint32(255 << 2)) == key
            acc = None
            for v in exp[sel]:
                acc = v if acc is None else acc + v
            assert e0 == acc


@pytest.mark.parametrize("clen", [31, 32, 33, 34, 35, 36])
def test_hmm_window_at_both_contig_ends(clen):
    """Reads that start within 3 bases of a short contig's start, so the HMM window is clamped there,
    and whose window may also pass the contig's end: the end check adds the clamp's -1
    (baq.cu:127-158), and a read kept there runs its HMM on the shortened window."""
    rng = np.random.default_rng(clen)
    contig = random_genome_string(rng, clen)
    g = HandGenome([contig])
    reads = []
    for start in (0, 1, 2):
        for L in (26, 28, 30):
            if start + L > clen:
                continue
            r = make_read(contig, [(0, L)], start, rng=rng, mismatch_all=False, qual=30)
            seq = list(r["seq"])
            seq[L // 2] = "A" if seq[L // 2] != "A" else "G"          # one error: the read needs BAQ
            r["seq"] = "".join(seq)
            reads.append(r)
    o, b = oracle_run(g, reads)
    act = set(int(x) for x in o.debug(abi.FP_DBG_ACTIVE_READS, np.uint32))
    tags = o.debug(abi.FP_DBG_BAQ, np.uint8)
    kept = dropped = 0
    for i, r in enumerate(reads):
        m = model_read(r, contig)
        w = rm.windows(m, None)
        if w["dropped"]:
            assert i not in act, f"read {i} at {r['start']} ({r['cigar']}) must leave the active list"
            dropped += 1
            continue
        assert i in act, f"read {i} at {r['start']} ({r['cigar']}) must stay active"
        exp = rm.baq_tags(m, w, rm.hmm_independent)
        if exp is not None:
            L = rm.read_length(m["ops"])
            assert [int(t) for t in tags[b.qual_start[i]:b.qual_start[i] + L]] == exp, f"BQ tags of read {i}"
        kept += 1
    assert kept + dropped == len(reads)
