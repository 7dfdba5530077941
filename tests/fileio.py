"""Writers of FASTA, .fai, VCF, SAM and BAM files from a synthetic workload, for the loader
tests.  They follow the SAM/BAM and VCF specifications (BAM in BGZF blocks with the
specification's empty end-of-file block), independently of the loaders they test."""
import gzip
import struct
import zlib

import numpy as np

NT16 = "=ACMGRSVTWYHKDBN"
CIGAR_OPS = "MIDNSHP=XB"
BGZF_EOF = bytes.fromhex("1f8b08040000000000ff0600424302001b0003000000000000000000")


def contig_bases(ref_view, c):
    """Bases of contig c of an fp_reference_view as a string."""
    n = int(ref_view.length[c])
    off = int(ref_view.word_offset[c])
    nw = (n + 7) // 8
    words = np.ctypeslib.as_array(ref_view.words, shape=(int(ref_view.total_words),))[off:off + nw]
    codes = ((words[:, None] >> (np.arange(8, dtype=np.uint32) * 4)) & 15).reshape(-1)[:n]
    return "".join(NT16[x] for x in codes)


def write_fasta(path, ref_view, names, line_width=60, fai=False):
    offsets = []
    with open(path, "w") as f:
        pos = 0
        for c, name in enumerate(names):
            head = f">{name} synthetic contig {c}\n"
            f.write(head)
            pos += len(head)
            s = contig_bases(ref_view, c)
            offsets.append((name, len(s), pos))
            for i in range(0, len(s), line_width):
                line = s[i:i + line_width] + "\n"
                f.write(line)
                pos += len(line)
    if fai:
        with open(path + ".fai", "w") as f:
            for name, n, off in offsets:
                f.write(f"{name}\t{n}\t{off}\t{line_width}\t{line_width + 1}\n")


def write_vcf(path, ref_view, var_view, names, compress=False):
    bases = [contig_bases(ref_view, c) for c in range(len(names))]
    lines = ["##fileformat=VCFv4.1"]
    lines += [f"##contig=<ID={n},length={len(b)}>" for n, b in zip(names, bases)]
    lines.append("#CHROM\tPOS\tID\tREF\tALT\tQUAL\tFILTER\tINFO")
    for c in range(int(var_view.num_sequences)):
        a, b = int(var_view.feature_offset[c]), int(var_view.feature_offset[c + 1])
        for i in range(a, b):
            s, e = int(var_view.feature_start[i]), int(var_view.feature_stop[i])
            ref = bases[c][s:e + 1]
            alt = "A" if ref[0] != "A" else "C"
            lines.append(f"{names[c]}\t{s + 1}\trs{i}\t{ref}\t{alt}\t.\tPASS\t.")
    text = "\n".join(lines) + "\n"
    if compress:
        with gzip.open(path, "wt") as f:
            f.write(text)
    else:
        with open(path, "w") as f:
            f.write(text)


def read_fields(arrays, i):
    """Per-read fields of a batch (abi.view_arrays) as Python values."""
    cs, cl = int(arrays["cigar_start"][i]), int(arrays["cigar_len"][i])
    cig = [(int(w) & 0xFFFFFF, int(w) >> 24 & 15) for w in arrays["cigars"][cs:cs + cl]]
    rs, rl = int(arrays["read_start"][i]), int(arrays["read_len"][i])
    words = arrays["reads"]
    seq = [int(words[(rs + k) >> 3]) >> (((rs + k) & 7) * 4) & 15 for k in range(rl)]
    qs, ql = int(arrays["qual_start"][i]), int(arrays["qual_len"][i])
    return cig, seq, arrays["qualities"][qs:qs + ql]


def header_text(names, lengths, rg_lines):
    h = "@HD\tVN:1.4\tSO:unsorted\n"
    h += "".join(f"@SQ\tSN:{n}\tLN:{l}\n" for n, l in zip(names, lengths))
    h += "".join(l + "\n" for l in rg_lines)
    return h


def default_rg_lines(num_rg):
    return [f"@RG\tID:rg{i}\tSM:sample\tPL:illumina" for i in range(num_rg)]


def sam_records(arrays, names):
    n = len(arrays["chromosome"])
    nc = len(names)
    for i in range(n):
        cig, seq, qual = read_fields(arrays, i)
        ch = int(arrays["chromosome"][i])
        rname = names[ch] if ch < nc else "*"
        pos = (int(arrays["alignment_start"][i]) + 1) & 0xFFFFFFFF
        mc = int(arrays["mate_chromosome"][i])
        rnext = "*" if mc >= nc else ("=" if mc == ch else names[mc])
        pnext = (int(arrays["mate_alignment_start"][i]) + 1) & 0xFFFFFFFF
        cigar = "".join(f"{l}{CIGAR_OPS[o]}" for l, o in cig) or "*"
        s = "".join(NT16[x] for x in seq) or "*"
        q = "".join(chr(int(x) + 33) for x in qual) or "*"
        rg = int(arrays["read_group"][i])
        yield (f"r{i}\t{int(arrays['flags'][i])}\t{rname}\t{np.int32(np.uint32(pos))}\t{int(arrays['mapq'][i])}\t"
               f"{cigar}\t{rnext}\t{np.int32(np.uint32(pnext))}\t{int(arrays['inferred_insert_size'][i])}\t{s}\t{q}"
               f"\tXA:i:7\tRG:Z:rg{rg}\tNM:i:0")


def write_sam(path, arrays, names, lengths, rg_lines, compress=False, crlf=False):
    eol = "\r\n" if crlf else "\n"
    text = header_text(names, lengths, rg_lines).replace("\n", eol)
    text += "".join(r + eol for r in sam_records(arrays, names))
    if compress:
        with gzip.open(path, "wt", newline="") as f:
            f.write(text)
    else:
        with open(path, "w", newline="") as f:
            f.write(text)


def bgzf_blocks(data, block=65280):
    out = bytearray()
    for a in range(0, len(data), block):
        chunk = data[a:a + block]
        co = zlib.compressobj(6, zlib.DEFLATED, -15)
        cdata = co.compress(chunk) + co.flush()
        bsize = 18 + len(cdata) + 8 - 1
        out += struct.pack("<BBBBIBBHBBHH", 31, 139, 8, 4, 0, 0, 255, 6, 66, 67, 2, bsize)
        out += cdata
        out += struct.pack("<II", zlib.crc32(chunk) & 0xFFFFFFFF, len(chunk))
    out += BGZF_EOF
    return bytes(out)


def bam_record(arrays, i, nc, extra_tags=b""):
    cig, seq, qual = read_fields(arrays, i)
    ch = int(arrays["chromosome"][i])
    mc = int(arrays["mate_chromosome"][i])
    name = f"r{i}".encode() + b"\0"
    rec = struct.pack("<iiBBHHHiiii",
                      ch if ch < nc else -1, int(np.int32(np.uint32(arrays["alignment_start"][i]))),
                      len(name), int(arrays["mapq"][i]), 4680, len(cig), int(arrays["flags"][i]), len(seq),
                      mc if mc < nc else -1, int(np.int32(np.uint32(arrays["mate_alignment_start"][i]))),
                      int(arrays["inferred_insert_size"][i]))
    rec += name
    rec += b"".join(struct.pack("<I", (l << 4) | o) for l, o in cig)
    packed = bytearray((len(seq) + 1) // 2)
    for k, x in enumerate(seq):
        packed[k >> 1] |= x << (4 if k % 2 == 0 else 0)
    rec += bytes(packed) + bytes(int(x) for x in qual)
    rec += extra_tags + b"RGZrg%d\0" % int(arrays["read_group"][i])
    return struct.pack("<i", len(rec)) + rec


def write_bam(path, arrays, names, lengths, rg_lines, extra_tags=b"", block=65280):
    text = header_text(names, lengths, rg_lines).encode()
    data = bytearray(b"BAM\1" + struct.pack("<i", len(text)) + text + struct.pack("<i", len(names)))
    for n, l in zip(names, lengths):
        nb = n.encode() + b"\0"
        data += struct.pack("<i", len(nb)) + nb + struct.pack("<I", l)
    for i in range(len(arrays["chromosome"])):
        data += bam_record(arrays, i, len(names), extra_tags)
    with open(path, "wb") as f:
        f.write(bgzf_blocks(bytes(data), block))
