"""Input loaders (include/firepony_amd_io.h) against files written from synthetic workloads
(round trips: the loaded batch must equal the generator's batch field for field) and against
hand-written records whose meaning the reference's loader code (loader/alignments.cu,
loader/reference.cu, loader/variants.cu) and the SAM/VCF specifications fix."""
import gzip

import numpy as np
import pytest

from firepony_amd import report_diff as rd
from firepony_amd import synth
from firepony_amd.loaders import Alignments, FileInputs, KnownSites, LoaderError, Reference

import fileio

PER_READ = ["alignment_start", "alignment_stop", "mate_alignment_start", "inferred_insert_size",
            "cigar_len", "read_len", "qual_len", "flags", "mapq"]


def workload(seed=5, n=1500, **kw):
    args = dict(seed=seed, num_reads=n, lengths=(36, 76, 101, 151), read_groups=3, contigs=(50000, 30000, 20000),
                sites=300, clip=0.3, indel=0.2, paired=1, multi_sites=0.3)
    args.update(kw)
    return synth.small_workload(**args)


def write_inputs(tmp_path, g, rs, n, fmt="sam", fai=False, vcf_gz=False, **kw):
    names = [f"chr{c}" for c in range(int(g.reference.num_sequences))]
    lengths = [int(g.reference.length[c]) for c in range(len(names))]
    fasta = str(tmp_path / "ref.fa")
    fileio.write_fasta(fasta, g.reference, names, fai=fai)
    vcf = str(tmp_path / ("sites.vcf.gz" if vcf_gz else "sites.vcf"))
    fileio.write_vcf(vcf, g.reference, g.variants, names, compress=vcf_gz)
    b = synth.Batch(g, rs, 0, n)
    a = b.arrays()
    rg_lines = fileio.default_rg_lines(int(rs.num_read_groups))
    if fmt == "bam":
        path = str(tmp_path / "reads.bam")
        fileio.write_bam(path, a, names, lengths, rg_lines, **kw)
    else:
        path = str(tmp_path / ("reads.sam.gz" if fmt == "sam.gz" else "reads.sam"))
        fileio.write_sam(path, a, names, lengths, rg_lines, compress=(fmt == "sam.gz"), **kw)
    return path, fasta, vcf, b


def load_all(aln, batch_size):
    """Every batch of an Alignments handle, its arrays copied."""
    out = []
    while True:
        b = aln.next_batch(batch_size)
        if b is None:
            return out
        out.append({k: v.copy() for k, v in b.arrays().items()})


def check_round_trip(batches, want, names_by_id, num_contigs):
    n = len(want["chromosome"])
    assert sum(len(b["chromosome"]) for b in batches) == n
    cat = {k: np.concatenate([b[k] for b in batches]) for k in batches[0]}
    for k in PER_READ:
        np.testing.assert_array_equal(cat[k], want[k], err_msg=k)
    ch = want["chromosome"].astype(np.uint32)
    np.testing.assert_array_equal(cat["chromosome"], np.where(ch < num_contigs, ch, 0xFFFF).astype(np.uint16))
    mc = want["mate_chromosome"]
    np.testing.assert_array_equal(cat["mate_chromosome"], np.where(mc < num_contigs, mc, 0xFFFFFFFF))
    # concatenated op, base and quality streams, and each batch's offsets as running sums
    np.testing.assert_array_equal(cat["cigars"], want["cigars"])
    np.testing.assert_array_equal(cat["reads"], want["reads"])
    np.testing.assert_array_equal(cat["qualities"], want["qualities"])
    for b in batches:
        np.testing.assert_array_equal(b["cigar_start"], np.concatenate([[0], np.cumsum(b["cigar_len"])[:-1]]))
        padded = (b["read_len"].astype(np.int64) + 7) // 8 * 8
        np.testing.assert_array_equal(b["read_start"], np.concatenate([[0], np.cumsum(padded)[:-1]]))
        np.testing.assert_array_equal(b["qual_start"], np.concatenate([[0], np.cumsum(b["qual_len"])[:-1]]))
    got_rg = [names_by_id[i] for i in cat["read_group"]]
    assert got_rg == [f"rg{i}" for i in want["read_group"]]
    # ids in order of first appearance
    first = list(dict.fromkeys(got_rg))
    assert first == names_by_id[:len(first)]


@pytest.mark.parametrize("fmt", ["sam", "sam.gz", "bam"])
def test_round_trip(tmp_path, fmt):
    g, rs, n = workload()
    path, fasta, vcf, b = write_inputs(tmp_path, g, rs, n, fmt=fmt)
    ref = Reference(fasta)
    aln = Alignments(path, ref)
    batches = load_all(aln, 577)
    _, names, max_rg = aln.header()
    assert max_rg == 3
    assert aln.counts() == (n, 0)
    check_round_trip(batches, b.arrays(), names, len(ref))


def test_round_trip_crlf_and_extra_tags(tmp_path):
    g, rs, n = workload(seed=6, n=400)
    path, fasta, _, b = write_inputs(tmp_path, g, rs, n, fmt="sam", crlf=True)
    ref = Reference(fasta)
    aln = Alignments(path, ref)
    check_round_trip(load_all(aln, 1000), b.arrays(), aln.header()[1], len(ref))
    bam = str(tmp_path / "t.bam")
    a = b.arrays()
    names = [ref.name(i) for i in range(len(ref))]
    tags = b"XAi" + (7).to_bytes(4, "little") + b"XBB" + b"s" + (3).to_bytes(4, "little") + bytes(6) + b"XZZabc\0"
    # 300-byte BGZF blocks: more blocks than one parallel step inflates, records split across blocks
    fileio.write_bam(bam, a, names, [int(g.reference.length[i]) for i in range(len(names))],
                     fileio.default_rg_lines(3), extra_tags=tags, block=300)
    aln2 = Alignments(bam, ref)
    check_round_trip(load_all(aln2, 128), a, aln2.header()[1], len(ref))


def test_reference_round_trip(tmp_path):
    g, rs, n = workload(n=10)
    for fai in (False, True):
        _, fasta, _, _ = write_inputs(tmp_path, g, rs, n, fai=fai)
        ref = Reference(fasta)
        v, w = ref.view, g.reference
        assert v.num_sequences == w.num_sequences
        for c in range(v.num_sequences):
            assert v.length[c] == w.length[c]
            assert fileio.contig_bases(v, c) == fileio.contig_bases(w, c)
            assert ref.lookup(f"chr{c}") == c
            assert ref.name(c) == f"chr{c}"
        assert ref.lookup("chrX") is None


@pytest.mark.parametrize("gz", [False, True])
def test_known_sites_round_trip(tmp_path, gz):
    g, rs, n = workload(n=10)
    _, fasta, vcf, _ = write_inputs(tmp_path, g, rs, n, vcf_gz=gz)
    ref = Reference(fasta)
    ks = KnownSites(vcf, ref)
    v, w = ks.view, g.variants
    assert v.num_sequences == w.num_sequences and v.num_features == w.num_features
    nf = int(v.num_features)
    got = {k: np.ctypeslib.as_array(getattr(v, k), shape=(nf,)).copy() for k in ("feature_start", "feature_stop")}
    for k in got:
        np.testing.assert_array_equal(got[k], np.ctypeslib.as_array(getattr(w, k), shape=(nf,)))
    offs = [int(v.feature_offset[c]) for c in range(v.num_sequences + 1)]
    assert offs == [int(w.feature_offset[c]) for c in range(w.num_sequences + 1)]
    mel = np.ctypeslib.as_array(v.max_end_point_left, shape=(nf,))
    for c in range(v.num_sequences):
        seg = got["feature_stop"][offs[c]:offs[c + 1]]
        np.testing.assert_array_equal(mel[offs[c]:offs[c + 1]], np.maximum.accumulate(seg) if len(seg) else seg)
    assert ks.counts() == (nf, 0)


# ---------------------------------------------------------------------------
# hand-written inputs

FASTA = ">c1 first contig\nACGTacgtNNRY\nAC\r\n>c2\nGGGG\n"


def write(tmp_path, name, text, mode="w"):
    p = tmp_path / name
    if name.endswith(".gz"):
        with gzip.open(p, "wt", newline="") as f:
            f.write(text)
    else:
        with open(p, mode, newline="") as f:
            f.write(text)
    return str(p)


def bases_of(view, c):
    return fileio.contig_bases(view, c)


def test_fasta_semantics(tmp_path):
    ref = Reference(write(tmp_path, "a.fa", FASTA))
    assert len(ref) == 2
    # lower case a..t upper-cased, every other letter N, CR dropped, name up to the first space
    assert bases_of(ref.view, 0) == "ACGTACGTNNNNAC"
    assert bases_of(ref.view, 1) == "GGGG"
    assert ref.name(0) == "c1" and ref.lookup("c1") == 0 and ref.lookup("c2") == 1
    # without an index, loading stops at the first empty sequence (which stays defined)
    ref = Reference(write(tmp_path, "b.fa", ">e1\nACGT\n>e2\n>e3\nTTTT\n"))
    assert len(ref) == 2 and ref.view.length[1] == 0 and ref.lookup("e3") is None
    # with an index, only the contigs it names are loaded
    p = write(tmp_path, "c.fa", ">e1\nACGT\n>e2\n>e3\nTTTT\n")
    write(tmp_path, "c.fa.fai", "e1\t4\t4\t4\t5\ne3\t4\t18\t4\t5\n")
    ref = Reference(p)
    assert [ref.name(i) for i in range(len(ref))] == ["e1", "e3"]
    assert bases_of(ref.view, 1) == "TTTT"
    with pytest.raises(LoaderError):
        Reference(write(tmp_path, "d.fa", "ACGT\n"))


def test_vcf_semantics(tmp_path):
    ref = Reference(write(tmp_path, "a.fa", FASTA))
    vcf = write(tmp_path, "s.vcf.gz", "##fileformat=VCFv4.2\n#CHROM\tPOS\tID\tREF\tALT\tQUAL\tFILTER\tINFO\n"
                "c1\t2\t.\tCGTA\tC\t.\t.\t.\n"
                "c1\t3\t.\tG\tT\t.\t.\t.\n"
                "cX\t1\t.\tA\tT\t.\t.\t.\n"
                "c1\t9\t.\tN\t<DEL>\t.\t.\tSVTYPE=DEL;END=12\n"
                "c2\t4\t.\tG\tA\t.\t.\tDP=3\n")
    ks = KnownSites(vcf, ref)
    v = ks.view
    assert ks.counts() == (5, 1)
    assert [int(v.feature_offset[i]) for i in range(3)] == [0, 3, 4]
    assert [int(v.feature_start[i]) for i in range(4)] == [1, 2, 8, 3]
    assert [int(v.feature_stop[i]) for i in range(4)] == [4, 2, 11, 3]
    assert [int(v.max_end_point_left[i]) for i in range(4)] == [4, 4, 11, 3]


SQ = "@SQ\tSN:c1\tLN:14\n@SQ\tSN:c2\tLN:4\n@SQ\tSN:c9\tLN:100\n"


def rec(name, flag, rname, pos, cigar, seq, qual, rnext="*", pnext=0, tlen=0, mapq=60, tags="RG:Z:a"):
    return f"{name}\t{flag}\t{rname}\t{pos}\t{mapq}\t{cigar}\t{rnext}\t{pnext}\t{tlen}\t{seq}\t{qual}" + \
        (f"\t{tags}" if tags else "") + "\n"


def load_one(tmp_path, text, name="r.sam"):
    ref = Reference(write(tmp_path, "a.fa", FASTA))
    aln = Alignments(write(tmp_path, name, text), ref)
    b = aln.next_batch(100)
    return aln, b.arrays()


def test_read_group_header_rules(tmp_path):
    # name = PU if read, else ID; a line's last field is never read (loader/alignments.cu:62-118)
    hdr = SQ + ("@RG\tID:a\tPU:unit1\tSM:s\n"      # PU read -> unit1
                "@RG\tID:b\tSM:s\tPU:unit2\n"      # PU is the last field -> name b
                "@RG\tID:c\n"                      # ID is the last field -> ID ""
                "@RG\tID:d\tPU:unit1\tLB:x\n")     # a second ID with name unit1
    body = (rec("r1", 0, "c1", 1, "4M", "ACGT", "IIII", tags="RG:Z:b") +
            rec("r2", 0, "c1", 1, "4M", "ACGT", "IIII", tags="RG:Z:a") +
            rec("r3", 0, "c1", 1, "4M", "ACGT", "IIII", tags="RG:Z:d") +
            rec("r4", 0, "c1", 1, "4M", "ACGT", "IIII", tags="RG:Z:c") +
            rec("r5", 0, "c1", 1, "4M", "ACGT", "IIII", tags="XX:i:1") +
            rec("r6", 0, "c1", 1, "4M", "ACGT", "IIII", tags="RG:A:a") +
            rec("r7", 0, "c1", 1, "4M", "ACGT", "IIII", tags="NM:i:0\tRG:Z:b\tRG:Z:a"))
    aln, a = load_one(tmp_path, hdr + body)
    _, names, max_rg = aln.header()
    assert names == ["b", "unit1"]
    assert max_rg == 3                     # unit1, b, "" (the ID-only line)
    assert list(a["read_group"]) == [0, 1, 1, 0xFFFFFFFF, 0xFFFFFFFF, 0xFFFFFFFF, 0]
    assert aln.counts() == (7, 1)          # RG:Z:c names no parsed @RG ID


def test_record_fields(tmp_path):
    hdr = SQ + "@RG\tID:a\tSM:s\n"
    body = (rec("m", 0x63, "c1", 3, "2S3M1I2M2D1M4N1M", "acgtNRYNK.", "!#+5?IIIII", "=", 20, -30) +
            rec("u", 0x10, "c1", 5, "*", "ACG", "*") +                  # CIGAR * -> unmapped
            rec("b", 0, "c2", 2, "2M1B1M", "ACG", "III") +             # B -> X
            rec("i", 0, "c1", 7, "2S2I", "ACGT", "IIII") +             # no reference-consuming op
            rec("z", 0, "*", 0, "*", "*", "*", "c2", 3) +              # no contig, no bases
            rec("q", "0x10", "c9x", 1, "4M", "ACGT", "IIII", "c2", 1, mapq=255) +  # RNAME not in @SQ
            rec("h", 2048 + 4096, "c2", 1, "4M", "ACGT", "IIII", "c9x", 1))
    aln, a = load_one(tmp_path, hdr + body)
    assert list(a["chromosome"]) == [0, 0, 1, 0, 0xFFFF, 0xFFFF, 1]
    assert list(a["alignment_start"]) == [2, 4, 1, 6, 0xFFFFFFFF, 0, 0]
    # bam_endpos - 1: reference span of a mapped read's CIGAR, else pos
    assert list(a["alignment_stop"]) == [2 + 13 - 1, 4, 1 + 3 - 1, 6 - 1, 0xFFFFFFFF, 3, 3]
    assert list(a["flags"]) == [0x63, 0x10 | 4, 0, 0, 4, 0x10, 2048]
    assert list(a["mapq"]) == [60, 60, 60, 60, 60, 255, 60]
    assert list(a["mate_chromosome"]) == [0, 0xFFFFFFFF, 0xFFFFFFFF, 0xFFFFFFFF, 1, 1, 0xFFFFFFFF]
    assert list(a["mate_alignment_start"])[:2] == [19, 0xFFFFFFFF]
    assert list(a["inferred_insert_size"])[0] == -30
    cig, seq, qual = fileio.read_fields(a, 0)
    assert cig == [(2, 4), (3, 0), (1, 1), (2, 0), (2, 2), (1, 0), (4, 3), (1, 0)]
    assert seq == [1, 2, 4, 8, 15, 5, 10, 15, 12, 15]
    assert list(qual) == [0, 2, 10, 20, 30, 40, 40, 40, 40, 40]
    assert fileio.read_fields(a, 1)[2].tolist() == [255, 255, 255]
    assert fileio.read_fields(a, 2)[0] == [(2, 0), (1, 8), (1, 0)]
    assert a["read_len"][4] == 0 and a["qual_len"][4] == 0 and a["cigar_len"][4] == 0
    # every read starts on a word boundary
    assert list(a["read_start"]) == [0, 16, 24, 32, 40, 40, 48]


@pytest.mark.parametrize("body, why", [
    (rec("x", 0, "c1", 1, "5M", "ACGT", "IIII"), "SEQ and CIGAR"),
    (rec("x", 0, "c1", 1, "4M", "ACGT", "III"), "QUAL"),
    (rec("x", 0, "c1", 1, "4Q", "ACGT", "IIII"), "CIGAR operator"),
    (rec("x", 0, "c2", "x", "4M", "ACGT", "IIII"), "POS"),
    ("x\t0\tc1\t1\n", "fields"),
])
def test_malformed_records(tmp_path, body, why):
    ref = Reference(write(tmp_path, "a.fa", FASTA))
    aln = Alignments(write(tmp_path, "r.sam", SQ + "@RG\tID:a\tSM:s\n" + body), ref)
    with pytest.raises(LoaderError, match=why):
        aln.next_batch(10)


def test_contig_missing_from_reference(tmp_path):
    # a read on a contig the header names but the reference lacks is an error (loader/alignments.cu:262-273)
    ref = Reference(write(tmp_path, "a.fa", FASTA))
    aln = Alignments(write(tmp_path, "r.sam", SQ + rec("x", 0, "c9", 1, "4M", "ACGT", "IIII")), ref)
    with pytest.raises(LoaderError, match="c9 not found"):
        aln.next_batch(10)


def test_batches_and_file_inputs(tmp_path):
    g, rs, n = workload(seed=8, n=900)
    path, fasta, vcf, b = write_inputs(tmp_path, g, rs, n, fmt="bam", fai=True, vcf_gz=True)
    inp = FileInputs(path, fasta, vcf)
    assert inp.header.num_read_groups == 3 and inp.header.num_chromosomes == 3
    got = [{k: v.copy() for k, v in bb.arrays().items()} for bb in inp.batches(250)]
    assert [len(x["chromosome"]) for x in got] == [250, 250, 250, 150]
    check_round_trip(got, b.arrays(), inp.read_group_names(), int(inp.reference.num_sequences))


def test_oracle_on_files_equals_generator(tmp_path):
    # the pipeline fed from files computes the tables it computes from the generator's batches;
    # read-group ids differ (first appearance vs generator order), so tables are compared by row
    from oracle.pyoracle import Oracle
    g, rs, n = workload(seed=9, n=1200)
    path, fasta, vcf, b = write_inputs(tmp_path, g, rs, n, fmt="bam", fai=True)
    inp = FileInputs(path, fasta, vcf)
    of = Oracle(inp)
    for bb in inp.batches(500):
        of.process(bb)
    of.set_read_group_names(inp.read_group_names())
    of.postprocess()
    og = Oracle(g)
    og.process(b)
    og.postprocess()
    ta = rd.parse(of.report(str(tmp_path / "f.txt")))
    tb = rd.parse(og.report(str(tmp_path / "g.txt")))
    assert [t.name for t in ta] == [t.name for t in tb]
    for x, y in zip(ta, tb):
        assert x.rows == y.rows, x.name
    assert of.stats().total_reads == n


def test_loader_library_exports():
    # every function include/firepony_amd_io.h declares is exported, and the binding names them all
    import ctypes as C
    import os
    import re
    from firepony_amd import loaders
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    text = open(os.path.join(root, "include", "firepony_amd_io.h")).read()
    names = set(re.findall(r"^\s*(?:fp_status|void|uint32_t|const char \*)\s*\*?\s*(fp_\w+)\s*\(", text, re.M))
    assert len(names) == 15
    L = C.CDLL(loaders.LIB_PATH)
    for n in names:
        assert hasattr(L, n), n
    assert set(loaders.EXPORTED) == names
