"""Pins for the CPU oracle, independent of its own code.

The reference holds no golden vectors or tests (its CI only compiles), so each
oracle stage is pinned here by what the definitions fix: closed forms of the
GATK covariates and filters on hand-built reads, an unbanded forward-backward
HMM written from the model's definition, an independent Bayesian posterior,
hand-worked fractional-error blocks and the GATKReport grammar.
"""
import math

import numpy as np
import pytest

from firepony_amd import abi
from oracle import pyoracle
from oracle.pyoracle import Oracle

from handbatch import HandBatch, HandGenome, random_genome_string
from refmodel import hmm_independent

COMP = {"A": "T", "C": "G", "G": "C", "T": "A"}


def revcomp(s):
    return "".join(COMP[c] for c in reversed(s))


def run(genome, reads, **kw):
    o = Oracle(genome, **kw)
    b = HandBatch(reads)
    o.process(b)
    return o, b


def decode_rows(t, table_id):
    out = []
    for k, n in zip(t["keys"], t["observations"]):
        k = int(k)
        ev = k & 3
        if table_id == abi.FP_TABLE_QUALITY:
            out.append((ev, (k >> 2) & 255, (k >> 10) & 255, int(n)))
        elif table_id == abi.FP_TABLE_CYCLE:
            raw = (k >> 2) & 1023
            cyc = -(raw >> 1) if raw & 1 else raw >> 1
            out.append((ev, cyc, (k >> 12) & 255, int(n)))
        else:
            raw = (k >> 2) & 2047
            size, bits = raw & 15, raw >> 4
            s = "".join("ACGT"[(bits >> (2 * j)) & 3] for j in range(size))
            out.append((ev, s, (k >> 13) & 255, int(n)))
    return out


@pytest.fixture(scope="module")
def genome():
    rng = np.random.default_rng(3)
    return HandGenome([random_genome_string(rng, 4000)])


def contig(genome):
    # decode contig 0 back to a string
    w = genome.words
    n = int(genome.lengths[0])
    inv = {1: "A", 2: "C", 4: "G", 8: "T", 15: "N"}
    return "".join(inv[int((w[i >> 3] >> ((i & 7) * 4)) & 15)] for i in range(n))


# ---------------------------------------------------------------------------
# covariates

@pytest.mark.parametrize("flags,expect", [
    (0, lambda L, i: i + 1),                 # forward: sequencing cycle 1..L
    (16, lambda L, i: L - i),                # reverse: cycles run backwards along the alignment
    (1 | 128, lambda L, i: -(i + 1)),        # second of pair: negative cycles
    (1 | 128 | 16, lambda L, i: -(L - i)),
])
def test_cycle_covariate_closed_form(genome, flags, expect):
    ref = contig(genome)
    L = 30
    o, _ = run(genome, [dict(seq=ref[100:100 + L], cigar=f"{L}M", start=100, flags=flags, qual=30)])
    rows = decode_rows(o.table(abi.FP_TABLE_CYCLE), abi.FP_TABLE_CYCLE)
    got_m = sorted(c for ev, c, q, n in rows if ev == 0)
    assert got_m == sorted(expect(L, i) for i in range(L))
    # indel keys exist only 4 bases away from either end (CUSHION_FOR_INDELS)
    got_i = sorted(c for ev, c, q, n in rows if ev == 1)
    assert got_i == sorted(expect(L, i) for i in range(4, L - 4))
    assert all(n == 1 for *_, n in rows)


def test_quality_table_counts(genome):
    ref = contig(genome)
    quals = [10 + (i % 7) for i in range(40)]
    o, _ = run(genome, [dict(seq=ref[500:540], cigar="40M", start=500, qual=quals, rg=0)])
    rows = decode_rows(o.table(abi.FP_TABLE_QUALITY), abi.FP_TABLE_QUALITY)
    m = {q: n for ev, q, rg, n in rows if ev == 0}
    assert m == {q: quals.count(q) for q in set(quals)}
    assert [(q, n) for ev, q, rg, n in rows if ev == 1] == [(45, 40)]
    assert [(q, n) for ev, q, rg, n in rows if ev == 2] == [(45, 40)]


@pytest.mark.parametrize("flags", [0, 16])
def test_context_covariate_closed_form(genome, flags):
    ref = contig(genome)
    seq = ref[900:950]
    o, _ = run(genome, [dict(seq=seq, cigar="50M", start=900, flags=flags, qual=30)])
    rows = decode_rows(o.table(abi.FP_TABLE_CONTEXT), abi.FP_TABLE_CONTEXT)
    # contexts are the k-mers ending at each base in sequencing order
    s = seq if flags == 0 else revcomp(seq)
    exp_m, exp_i = {}, {}
    for i in range(len(s)):
        if i >= 1:
            exp_m[s[i - 1:i + 1]] = exp_m.get(s[i - 1:i + 1], 0) + 1
        if i >= 2:
            exp_i[s[i - 2:i + 1]] = exp_i.get(s[i - 2:i + 1], 0) + 1
    assert {c: n for ev, c, q, n in rows if ev == 0} == exp_m
    assert {c: n for ev, c, q, n in rows if ev == 1} == exp_i


# ---------------------------------------------------------------------------
# CIGAR expansion, error bits, windows

def bits(o, b, item):
    return o.debug(item, np.uint8)


def test_snp_bits_closed_form(genome):
    ref = contig(genome)
    seq = list(ref[1000:1060])
    mism = [3, 17, 40, 59]
    for p in mism:
        seq[p] = "A" if seq[p] != "A" else "C"
    o, b = run(genome, [dict(seq="".join(seq), cigar="60M", start=1000)])
    flags = o.debug(abi.FP_DBG_ERROR_FLAGS, np.uint8)[:60]
    assert list(np.nonzero(flags & 1)[0]) == mism
    assert int(o.debug(abi.FP_DBG_NUM_ERRORS, np.uint16)[0]) == len(mism)


@pytest.mark.parametrize("rev", [False, True])
def test_indel_marks_closed_form(genome, rev):
    ref = contig(genome)
    # 20M 3I 20M: insertion starts at read base 20
    seq_i = ref[1500:1520] + "GGG" + ref[1520:1540]
    # 20M 2D 20M: deletion follows read base 19
    seq_d = ref[1600:1620] + ref[1622:1642]
    f = 16 if rev else 0
    o, b = run(genome, [dict(seq=seq_i, cigar="20M3I20M", start=1500, flags=f),
                        dict(seq=seq_d, cigar="20M2D20M", start=1600, flags=f)])
    fl = o.debug(abi.FP_DBG_ERROR_FLAGS, np.uint8)
    rs = b.read_start
    ins = list(np.nonzero(fl[rs[0]:rs[0] + 43] & 2)[0])
    dels = list(np.nonzero(fl[rs[1]:rs[1] + 40] & 4)[0])
    # marked on the base preceding the event in sequencing order
    assert ins == ([20 + 3] if rev else [19])
    assert dels == ([20] if rev else [19])


def test_soft_clip_window(genome):
    ref = contig(genome)
    seq = "TTTTT" + ref[2000:2040] + "GGG"
    o, _ = run(genome, [dict(seq=seq, cigar="5S40M3S", start=2000)])
    w = int(o.debug(abi.FP_DBG_READ_WINDOW, np.uint32)[0])
    assert (w & 0xffff, w >> 16) == (5, 44)


@pytest.mark.parametrize("rev", [False, True])
def test_adapter_clipping_closed_form(genome, rev):
    ref = contig(genome)
    L, start, t = 50, 2500, 30
    if not rev:
        # mate ends inside this read: adaptor boundary = start + |isize| + 1
        r = dict(seq=ref[start:start + L], cigar=f"{L}M", start=start, flags=1 | 2 | 64 | 32,
                 mate_start=start + t - L + 10, isize=t)
        exp = (0, t)
    else:
        mate = start + 12
        r = dict(seq=ref[start:start + L], cigar=f"{L}M", start=start, flags=1 | 2 | 64 | 16,
                 mate_start=mate, isize=-t)
        exp = (mate - start, L - 1)
    o, _ = run(genome, [r])
    w = int(o.debug(abi.FP_DBG_READ_WINDOW, np.uint32)[0])
    assert (w & 0xffff, w >> 16) == exp


def test_known_site_mask_closed_form():
    rng = np.random.default_rng(9)
    ref = random_genome_string(rng, 3000)
    sites = sorted({int(x) for x in rng.integers(100, 2900, 120)})
    feats = [(0, s, s + (int(rng.integers(0, 3)) if i % 5 == 0 else 0)) for i, s in enumerate(sites)]
    g = HandGenome([ref], sites=feats)
    reads = []
    for k in range(40):
        st = int(rng.integers(50, 2850))
        reads.append(dict(seq=ref[st:st + 80], cigar="80M", start=st))
    o, b = run(g, reads)
    mask = o.debug(abi.FP_DBG_ACTIVE_MASK, np.uint8)
    for k, r in enumerate(reads):
        for i in range(80):
            pos = r["start"] + i
            covered = any(s <= pos <= e for _, s, e in feats)
            assert mask[b.read_start[k] + i] == (0 if covered else 1), (k, i)


@pytest.mark.parametrize("flags,mapq,cigar,kept", [
    (0, 60, "40M", True), (0, 0, "40M", False), (0, 255, "40M", False),
    (1024, 60, "40M", False), (512, 60, "40M", False), (4, 60, "40M", False), (256, 60, "40M", False),
    (2048, 60, "40M", True),                      # supplementary reads are kept
    (0, 60, "20M3N20M", False),                   # N operator
    (0, 60, "20M5S15M", False),                   # soft clip between aligned ops
    (0, 60, "5H5S30M", True), (0, 60, "5S5H30M", False),
    (0, 60, "1M39S", False),                      # aligned span of one base
])
def test_read_filter_truth_table(genome, flags, mapq, cigar, kept):
    from handbatch import parse_cigar
    ref = contig(genome)
    qlen = sum(n for op, n in parse_cigar(cigar) if op in (0, 1, 4, 7, 8))
    o, _ = run(genome, [dict(seq=ref[3000:3000 + qlen], cigar=cigar, start=3000, flags=flags, mapq=mapq)])
    act = o.debug(abi.FP_DBG_ACTIVE_READS, np.uint32)
    assert (act.size == 1) == kept


# ---------------------------------------------------------------------------
# fractional errors: GATK calculateFractionalErrorArray, worked by hand

@pytest.mark.parametrize("baq,err,exp", [
    ([64, 64, 64], [0, 1, 0], [0, 1, 0]),
    # block of two non-64 tags at 2..3, terminated by 64 at 4: covers 1..4
    ([64, 64, 70, 70, 64, 64], [0, 1, 0, 0, 1, 0], [0, 0.5, 0.5, 0.5, 0.5, 0]),
    # block running to the end of the read covers start-1 .. end
    ([64, 64, 64, 70, 70], [1, 0, 0, 0, 1], [1, 0, 1 / 3, 1 / 3, 1 / 3]),
    # block at the very start covers 0 .. terminator
    ([70, 64, 64], [1, 0, 1], [0.5, 0.5, 1]),
    # two blocks sharing a boundary base: the later block wins on the shared base
    ([64, 70, 64, 70, 64], [1, 0, 0, 0, 1], [1 / 3, 1 / 3, 1 / 3, 1 / 3, 1 / 3]),
])
def test_fractional_errors_hand_cases(baq, err, exp):
    L = len(baq)
    b = np.array(baq, np.uint8)
    e = np.array(err, np.uint8)
    out = np.zeros(L, np.float64)
    pyoracle.lib().fpo_fractional_errors(b.ctypes.data_as(abi.u8p), e.ctypes.data_as(abi.u8p), L,
                                         out.ctypes.data_as(abi.f64p))
    np.testing.assert_allclose(out, exp, rtol=0, atol=1e-15)


# ---------------------------------------------------------------------------
# BAQ: an unbanded glocal HMM written from the model definition

@pytest.mark.parametrize("seed,lq,bw,full", [(1, 25, 7, False), (2, 30, 40, True), (3, 40, 10, False),
                                             (4, 18, 30, True)])
def test_baq_hmm_against_independent(seed, lq, bw, full):
    rng = np.random.default_rng(seed)
    lr = lq + 6
    ref = rng.choice([1, 2, 4, 8], size=lr).astype(np.uint8)
    query = ref[3:3 + lq].copy()
    for p in rng.choice(lq, size=3, replace=False):
        query[p] = [1, 2, 4, 8][(int(np.log2(query[p])) + 1) % 4]
    if seed == 3:
        query = np.concatenate([query[:15], query[17:], ref[3 + lq:3 + lq + 2]]).astype(np.uint8)  # 2 bp shift
    qual = rng.integers(8, 41, size=lq).astype(np.uint8)
    out_q = np.zeros(lq, np.uint8)
    out_s = np.zeros(lq, np.uint32)
    pyoracle.lib().fpo_hmm_glocal(ref.ctypes.data_as(abi.u8p), lr, query.ctypes.data_as(abi.u8p),
                                  qual.ctypes.data_as(abi.u8p), lq, bw, out_q.ctypes.data_as(abi.u8p),
                                  out_s.ctypes.data_as(abi.u32p))
    states, post = hmm_independent([int(x) for x in ref], [int(x) for x in query], [int(x) for x in qual], bw)
    assert list(out_s) == states
    for i, p in enumerate(post):
        real = -4.343 * math.log(1 - p) + 0.499 if p < 1 else float("inf")
        if abs(real - round(real)) < 1e-6:
            continue   # within rounding of an integer boundary: either side is correct
        k = int(real) if real < 2 ** 31 else -2 ** 31
        exp = 99 if k > 100 else (4 if k < 4 else k)
        assert out_q[i] == exp, (i, p, real)


# ---------------------------------------------------------------------------
# empirical quality: the posterior over Q bins, written independently

def empirical_independent(obs, mism, expected, rounding):
    est = -10.0 * math.log10(expected / obs)
    n = obs + 2
    k = int(mism + 0.5) + 1
    cap = 2 ** 31 - 2   # GATK caps the binomial at the largest int observation count
    if n > cap:
        k = round(k * (cap / n))
        n = cap
    vals = []
    for q in range(61):
        d = q - est
        if rounding:
            d = math.floor(d + 0.5) if d >= 0 else -math.floor(-d + 0.5)
        d = min(abs(int(d)), 40)
        g = 0.9 * math.exp(-d * d / 0.5)
        prior = math.log10(g) if g > 0 else -math.inf   # underflows to 0 beyond ~19 Q
        if q == 0:
            lik = -math.inf
        else:
            p = 10 ** (-q / 10.0)
            lik = (math.lgamma(n + 1) - math.lgamma(k + 1) - math.lgamma(n - k + 1)) / math.log(10) \
                + k * math.log10(p) + (n - k) * math.log10(1 - p)
        vals.append(prior + lik)
    order = sorted(range(61), key=lambda q: -vals[q])
    return order[0], vals[order[0]] - vals[order[1]], est


@pytest.mark.parametrize("obs", [3, 40, 1000, 250000, 30000000, 5000000000])
@pytest.mark.parametrize("rate", [0.0, 0.0007, 0.01, 0.2])
@pytest.mark.parametrize("qrep", [12, 30, 45])
def test_empirical_quality_against_independent(obs, rate, qrep):
    expected = obs * 10 ** (-qrep / 10.0)
    mism = obs * rate
    for rounding in (0, 1):
        got = pyoracle.lib().fpo_empirical_quality(obs, mism, expected, rounding, None)
        q, gap, _ = empirical_independent(obs, mism, expected, rounding)
        if gap < 1e-6 * max(1.0, abs(q)):
            continue
        assert got == q, (obs, rate, qrep, rounding, got, q)


def test_round_n():
    lib = pyoracle.lib()
    for v in [0.125, 0.135, 2.5, 19756.665, 1.00005, -0.125, 42.0]:
        for n in (2, 4):
            s = v * 10 ** n
            exp = (math.floor(s + 0.5) if s >= 0 else -math.floor(-s + 0.5)) / 10 ** n
            assert lib.fpo_round_n(v, n) == exp


def test_map_quality_inf_case():
    # posterior exactly 1: log(0) = -inf, the x86 conversion gives INT_MIN -> 4
    assert pyoracle.lib().fpo_baq_map_quality(1.0) == 4
    assert pyoracle.lib().fpo_baq_map_quality(0.5) == 4
    assert pyoracle.lib().fpo_baq_map_quality(1 - 1e-9) == 90


# ---------------------------------------------------------------------------
# report grammar (diffreport/gatktable.py semantics)

def parse_report(text):
    lines = text.split("\n")
    assert lines[0] == "#:GATKReport.v1.1:5"
    tables, i = {}, 1
    while i < len(lines):
        if not lines[i].startswith("#:GATKTable:"):
            i += 1
            continue
        head = lines[i].split(":")
        ncols, nrows = int(head[2]), int(head[3])
        fmts = head[4:4 + ncols]
        assert head[4 + ncols] == ";"
        name = lines[i + 1].split(":")[2]
        cols = lines[i + 2].split()
        assert len(cols) == ncols
        rows = [lines[i + 3 + r].split() for r in range(nrows)]
        assert all(len(r) == ncols for r in rows), name
        assert lines[i + 3 + nrows] == ""
        tables[name] = (fmts, cols, rows)
        i += 3 + nrows
    return tables


def test_report_grammar(tmp_path):
    from firepony_amd import synth
    gs, rs, n = synth.named_spec(0)
    g = synth.Genome(gs, rs.num_read_groups)
    o = Oracle(g)
    o.process(synth.Batch(g, rs, 0, 2000))
    o.postprocess()
    tables = parse_report(o.report(str(tmp_path / "r.txt")))
    assert set(tables) == {"Arguments", "Quantized", "RecalTable0", "RecalTable1", "RecalTable2"}
    assert len(tables["Arguments"][2]) == 18
    assert len(tables["Quantized"][2]) == 94
    fmts, cols, rows = tables["RecalTable1"]
    assert cols == ["ReadGroup", "QualityScore", "EventType", "EmpiricalQuality", "Observations", "Errors"]
    assert fmts == ["%s", "%s", "%s", "%.4f", "%d", "%.2f"]
    t1 = o.table(abi.FP_TABLE_QUALITY)
    assert len(rows) == int((t1["observations"] > 0).sum())
    # RecalTable0 observations equal the per-event sums of RecalTable1
    rg = {(r[0], r[1]): int(r[4]) for r in tables["RecalTable0"][2]}
    for ev in "MID":
        assert rg[("rg0", ev)] == sum(int(r[4]) for r in rows if r[2] == ev)
