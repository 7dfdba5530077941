"""Hand-built genomes and batches for known-answer tests.

Reads are given as Python values (bases as strings, CIGAR strings, flags) and
packed into the boundary layout of include/firepony_amd.h.
"""
import ctypes as C
import re

import numpy as np

from firepony_amd import abi

CODE = {"A": 1, "C": 2, "G": 4, "T": 8, "N": 15}
OPS = {"M": 0, "I": 1, "D": 2, "N": 3, "S": 4, "H": 5, "P": 6, "=": 7, "X": 8}


def pack_bases(seq):
    n = (len(seq) + 7) // 8 * 8
    words = np.zeros(n // 8, np.uint32)
    for i, ch in enumerate(seq):
        words[i >> 3] |= np.uint32(CODE[ch]) << np.uint32((i & 7) * 4)
    return words


class HandGenome:
    """contigs: list of base strings; sites: list of (contig, start, stop)."""

    def __init__(self, contigs, sites=(), read_groups=("rg0",)):
        self._keep = []
        words, offs, lens = [], [], []
        off = 0
        for c in contigs:
            w = pack_bases(c)
            words.append(w)
            offs.append(off)
            lens.append(len(c))
            off += len(w)
        self.words = np.concatenate(words) if words else np.zeros(1, np.uint32)
        self.word_offset = np.array(offs, np.uint64)
        self.lengths = np.array(lens, np.uint32)
        self.reference = abi.FpReferenceView()
        self.reference.num_sequences = len(contigs)
        self.reference.word_offset = self.word_offset.ctypes.data_as(abi.u64p)
        self.reference.length = self.lengths.ctypes.data_as(abi.u32p)
        self.reference.words = self.words.ctypes.data_as(abi.u32p)
        self.reference.total_words = len(self.words)

        per = [[] for _ in contigs]
        for c, s, e in sites:
            per[c].append((s, e))
        fs, fe, fo = [], [], [0]
        for p in per:
            p.sort()
            fs += [s for s, _ in p]
            fe += [e for _, e in p]
            fo.append(len(fs))
        self.fstart = np.array(fs or [0], np.uint32)
        self.fstop = np.array(fe or [0], np.uint32)
        self.foff = np.array(fo, np.uint64)
        self.variants = abi.FpVariantsView()
        self.variants.num_sequences = len(contigs)
        self.variants.feature_offset = self.foff.ctypes.data_as(abi.u64p)
        self.variants.feature_start = self.fstart.ctypes.data_as(abi.u32p)
        self.variants.feature_stop = self.fstop.ctypes.data_as(abi.u32p)
        self.variants.max_end_point_left = None
        self.variants.num_features = len(fs)

        self.names = (C.c_char_p * len(read_groups))(*[n.encode() for n in read_groups])
        self.header = abi.FpHeaderView()
        self.header.num_chromosomes = len(contigs)
        self.header.chromosome_lengths = self.lengths.ctypes.data_as(abi.u32p)
        self.header.num_read_groups = len(read_groups)
        self.header.read_group_names = C.cast(self.names, C.POINTER(C.c_char_p))


def parse_cigar(s):
    return [(OPS[op], int(n)) for n, op in re.findall(r"(\d+)([MIDNSHP=X])", s)]


def ref_span(cigar):
    return sum(n for op, n in cigar if op in (0, 2, 3, 7, 8))


class HandBatch:
    """reads: list of dicts with keys seq, qual (list or int), cigar, start,
    chrom=0, flags=0, mapq=60, rg=0, mate_start, isize, mate_chrom."""

    def __init__(self, reads):
        n = len(reads)
        self.chromosome = np.zeros(n, np.uint16)
        self.alignment_start = np.zeros(n, np.uint32)
        self.alignment_stop = np.zeros(n, np.uint32)
        self.mate_chromosome = np.full(n, 0xffffffff, np.uint32)
        self.mate_alignment_start = np.full(n, 0xffffffff, np.uint32)
        self.isize = np.zeros(n, np.int32)
        self.flags = np.zeros(n, np.uint16)
        self.mapq = np.zeros(n, np.uint8)
        self.read_group = np.zeros(n, np.uint32)
        cig, cs, cl, rsv, rl, qsv, ql = [], [], [], [], [], [], []
        bases, quals = [], []
        roff = 0
        for i, r in enumerate(reads):
            c = parse_cigar(r["cigar"])
            self.chromosome[i] = r.get("chrom", 0)
            self.alignment_start[i] = r["start"]
            self.alignment_stop[i] = r["start"] + max(ref_span(c), 1) - 1
            self.flags[i] = r.get("flags", 0)
            self.mapq[i] = r.get("mapq", 60)
            self.read_group[i] = r.get("rg", 0)
            if "mate_start" in r:
                self.mate_alignment_start[i] = r["mate_start"]
                self.mate_chromosome[i] = r.get("mate_chrom", r.get("chrom", 0))
            self.isize[i] = r.get("isize", 0)
            cs.append(len(cig))
            cl.append(len(c))
            cig += [(n & 0xffffff) | (op << 24) for op, n in c]
            seq = r["seq"]
            q = r.get("qual", 30)
            q = [q] * len(seq) if isinstance(q, int) else list(q)
            rsv.append(roff)
            rl.append(len(seq))
            padded = (len(seq) + 7) // 8 * 8
            bases.append(seq + "A" * (padded - len(seq)))
            roff += padded
            qsv.append(len(quals))
            ql.append(len(q))
            quals += q
        self.cigars = np.array(cig or [0], np.uint32)
        self.ncig = len(cig)
        self.cigar_start = np.array(cs, np.uint32)
        self.cigar_len = np.array(cl, np.uint32)
        self.reads = pack_bases("".join(bases)) if bases else np.zeros(1, np.uint32)
        self.nbases = roff
        self.read_start = np.array(rsv, np.uint32)
        self.read_len = np.array(rl, np.uint32)
        self.quals = np.array(quals or [0], np.uint8)
        self.nquals = len(quals)
        self.qual_start = np.array(qsv, np.uint32)
        self.qual_len = np.array(ql, np.uint32)
        v = abi.FpBatchView()
        v.num_reads = n
        v.chromosome = self.chromosome.ctypes.data_as(abi.u16p)
        v.alignment_start = self.alignment_start.ctypes.data_as(abi.u32p)
        v.alignment_stop = self.alignment_stop.ctypes.data_as(abi.u32p)
        v.mate_chromosome = self.mate_chromosome.ctypes.data_as(abi.u32p)
        v.mate_alignment_start = self.mate_alignment_start.ctypes.data_as(abi.u32p)
        v.inferred_insert_size = self.isize.ctypes.data_as(abi.i32p)
        v.cigars = self.cigars.ctypes.data_as(abi.u32p)
        v.num_cigar_ops = self.ncig
        v.cigar_start = self.cigar_start.ctypes.data_as(abi.u32p)
        v.cigar_len = self.cigar_len.ctypes.data_as(abi.u32p)
        v.reads = self.reads.ctypes.data_as(abi.u32p)
        v.num_read_bases = self.nbases
        v.read_start = self.read_start.ctypes.data_as(abi.u32p)
        v.read_len = self.read_len.ctypes.data_as(abi.u32p)
        v.qualities = self.quals.ctypes.data_as(abi.u8p)
        v.num_qualities = self.nquals
        v.qual_start = self.qual_start.ctypes.data_as(abi.u32p)
        v.qual_len = self.qual_len.ctypes.data_as(abi.u32p)
        v.flags = self.flags.ctypes.data_as(abi.u16p)
        v.mapq = self.mapq.ctypes.data_as(abi.u8p)
        v.read_group = self.read_group.ctypes.data_as(abi.u32p)
        self.view = v

    def arrays(self):
        return {"read_start": self.read_start, "read_len": self.read_len, "qual_start": self.qual_start}


def random_genome_string(rng, n):
    return "".join(rng.choice(list("ACGT"), size=n))
