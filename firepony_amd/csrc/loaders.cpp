// Host-side input loaders (include/firepony_amd_io.h): FASTA reference, VCF
// known sites, SAM/BAM alignments -> the engine's boundary views.
//
// Files are read through zlib's gz* stream, which passes plain files through
// and inflates gzip and BGZF (concatenated gzip members) alike, so SAM, SAM.gz,
// BAM, VCF and VCF.gz share one reader.  The SAM and BAM record grammar is the
// SAM/BAM specification's; the reference parses it with htslib
// (nsubtil/htslib@fix-signed-one-bit-bitfield, an htslib 1.2 fork, not vendored):
// the htslib behaviours the reference depends on are restated where they are
// used, each with the reference call site.
#include "firepony_amd_io.h"

#include <zlib.h>

#include <algorithm>
#include <atomic>
#include <cctype>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>

namespace {

thread_local std::string g_err;

fp_status fail(fp_status s, const std::string &msg)
{
    g_err = msg;
    return s;
}

// string_database::hash (string_database.cu:83-94): h = h*101 + c over the
// bytes up to the first NUL, each char sign-extended as the reference's
// (uint32)(char) conversion does on x86.
uint32_t name_hash(const char *s)
{
    uint32_t h = 0;
    for (; *s; s++) h = h * 101u + (uint32_t)(int32_t)(signed char)*s;
    return h;
}

// string_database (string_database.cu:33-76): ids in insertion order, looked up
// by the hash of the name (two names with one hash share an id, as there).
struct NameDb {
    std::unordered_map<uint32_t, uint32_t> by_hash;
    std::vector<std::string> names;

    uint32_t lookup(const char *s) const
    {
        auto it = by_hash.find(name_hash(s));
        return it == by_hash.end() ? UINT32_MAX : it->second;
    }
    uint32_t insert(const std::string &s)
    {
        const uint32_t h = name_hash(s.c_str());
        auto it = by_hash.find(h);
        if (it != by_hash.end()) return it->second;
        const uint32_t id = (uint32_t)names.size();
        by_hash.emplace(h, id);
        names.push_back(s);
        return id;
    }
};

// BGZF (the blocked gzip of BAM and bgzip files): independent deflate blocks of at most 64 KiB,
// each a gzip member whose extra field "BC" gives its size (SAM/BAM specification, section 4.1).
// Blocks are inflated in parallel, a chunk of them at a time, and appended in order.
class Bgzf {
public:
    ~Bgzf() { if (f_) std::fclose(f_); }

    // true if path is BGZF (the file is then kept open)
    bool open(const char *path)
    {
        f_ = std::fopen(path, "rb");
        if (!f_) return false;
        uint8_t h[18];
        const size_t got = std::fread(h, 1, 18, f_);
        const bool bgzf = got == 18 && h[0] == 31 && h[1] == 139 && h[2] == 8 && (h[3] & 4) &&
                          h[10] == 6 && h[11] == 0 && h[12] == 'B' && h[13] == 'C' && h[14] == 2 && h[15] == 0;
        if (!bgzf) { std::fclose(f_); f_ = nullptr; return false; }
        std::fseek(f_, 0, SEEK_SET);
        unsigned t = std::thread::hardware_concurrency();
        threads_ = t == 0 ? 4 : (t > 16 ? 16 : t);
        return true;
    }

    static constexpr size_t CHUNK = 256;                    // blocks per parallel step
    static constexpr size_t MAX_CHUNK_BYTES = CHUNK * 65536; // inflated bytes of one step at most

    // inflates up to one chunk of blocks to out + end (room for MAX_CHUNK_BYTES); false at the end
    // of the file (err set on a malformed or truncated block)
    bool next(char *out, size_t &end, bool &err)
    {
        if (done_) return false;
        // read whole blocks, up to CHUNK of them
        struct Block { size_t coff, clen; uint32_t isize; size_t ooff; };
        std::vector<Block> blocks;
        comp_.clear();
        size_t total_out = 0;
        while (blocks.size() < CHUNK) {
            uint8_t h[18];
            const size_t got = std::fread(h, 1, 18, f_);
            if (got == 0) { done_ = true; break; }
            if (got < 18 || h[0] != 31 || h[1] != 139 || h[2] != 8 || !(h[3] & 4)) { err = true; done_ = true; return false; }
            const uint32_t xlen = h[10] | (h[11] << 8);
            // the BC subfield is the first of the extra field in files written to the specification
            if (xlen < 6 || h[12] != 'B' || h[13] != 'C') { err = true; done_ = true; return false; }
            const uint32_t bsize = (uint32_t)(h[16] | (h[17] << 8)) + 1;
            if (bsize < 18 + xlen - 6 + 8 + 2) { err = true; done_ = true; return false; }
            const size_t rest = bsize - 18;
            const size_t at = comp_.size();
            comp_.resize(at + rest);
            if (std::fread(comp_.data() + at, 1, rest, f_) != rest) { err = true; done_ = true; return false; }
            const uint8_t *tail = comp_.data() + at + rest - 4;
            const uint32_t isize = tail[0] | (tail[1] << 8) | (tail[2] << 16) | ((uint32_t)tail[3] << 24);
            if (isize > 65536) { err = true; done_ = true; return false; }
            const size_t skip = xlen - 6;   // any further extra subfields
            blocks.push_back({at + skip, rest - skip - 8, isize, total_out});
            total_out += isize;
        }
        if (blocks.empty()) return false;
        std::atomic<bool> bad{false};
        auto work = [&](size_t t) {
            z_stream z;
            for (size_t i = t; i < blocks.size(); i += threads_) {
                const Block &bk = blocks[i];
                std::memset(&z, 0, sizeof z);
                if (inflateInit2(&z, -15) != Z_OK) { bad = true; return; }
                z.next_in = comp_.data() + bk.coff;
                z.avail_in = (uInt)bk.clen;
                z.next_out = (Bytef *)out + end + bk.ooff;
                z.avail_out = bk.isize;
                const int rc = inflate(&z, Z_FINISH);
                if (rc != Z_STREAM_END || z.total_out != bk.isize) bad = true;
                inflateEnd(&z);
            }
        };
        const size_t nt = std::min<size_t>(threads_, blocks.size());
        std::vector<std::thread> pool;
        for (size_t t = 1; t < nt; t++) pool.emplace_back(work, t);
        work(0);
        for (auto &th : pool) th.join();
        if (bad) { err = true; done_ = true; return false; }
        end += total_out;
        return true;
    }

private:
    FILE *f_ = nullptr;
    unsigned threads_ = 1;
    bool done_ = false;
    std::vector<uint8_t> comp_;
};

// Buffered byte/line reader: BGZF through Bgzf, anything else (plain, gzip) through zlib's gz stream.
class Reader {
public:
    ~Reader() { if (f_) gzclose(f_); }

    bool open(const char *path)
    {
        path_ = path;
        reserve(4u << 20);
        if (bgzf_.open(path)) { use_bgzf_ = true; return true; }
        f_ = gzopen(path, "rb");
        if (!f_) return false;
        gzbuffer(f_, 1u << 20);
        return true;
    }
    const std::string &path() const { return path_; }
    bool error() const { return err_; }

    // makes at least n bytes available if the stream holds them; returns the count available
    size_t want(size_t n)
    {
        if (e_ - b_ >= n) return e_ - b_;
        if (b_) {
            std::memmove(buf_.get(), buf_.get() + b_, e_ - b_);
            e_ -= b_;
            b_ = 0;
        }
        reserve(n);
        while (e_ < n && !eof_) {
            if (use_bgzf_) {
                reserve(e_ + Bgzf::MAX_CHUNK_BYTES);
                if (!bgzf_.next(buf_.get(), e_, err_)) eof_ = true;
                continue;
            }
            const size_t room = std::min<size_t>(cap_ - e_, 1u << 30);
            const int r = gzread(f_, buf_.get() + e_, (unsigned)room);
            if (r < 0) { err_ = true; eof_ = true; break; }
            if (r == 0) { eof_ = true; break; }
            e_ += (size_t)r;
        }
        return e_ - b_;
    }
    const char *ptr() const { return buf_.get() + b_; }
    void skip(size_t n) { b_ += n; }

    // next line without its '\n' and without one trailing '\r'; false at end of stream
    bool line(std::string &out)
    {
        size_t scan = 0;
        for (;;) {
            const char *p = (const char *)std::memchr(buf_.get() + b_ + scan, '\n', e_ - b_ - scan);
            if (p) {
                const size_t n = (size_t)(p - (buf_.get() + b_));
                take(out, n);
                b_ += 1;
                return true;
            }
            scan = e_ - b_;
            if (want(scan + 1) == scan) {
                if (scan == 0) return false;
                take(out, scan);
                return true;
            }
        }
    }
    // the line starting c bytes into the unread data, without consuming it: its length (without
    // '\n' and one trailing '\r') and the offset after it; false at the end of the stream
    bool line_at(size_t c, size_t &len, size_t &next)
    {
        size_t scan = c;
        for (;;) {
            const size_t have = e_ - b_;
            if (scan < have) {
                const char *p = (const char *)std::memchr(buf_.get() + b_ + scan, '\n', have - scan);
                if (p) {
                    const size_t end = (size_t)(p - (buf_.get() + b_));
                    len = end - c;
                    next = end + 1;
                    if (len && buf_[b_ + c + len - 1] == '\r') len--;
                    return true;
                }
                scan = have;
            }
            if (want(have + 1) == have) {
                if (have <= c) return false;
                len = have - c;
                next = have;
                if (len && buf_[b_ + c + len - 1] == '\r') len--;
                return true;
            }
        }
    }
    // peeks the next byte without consuming it; -1 at end of stream
    int peek()
    {
        if (want(1) < 1) return -1;
        return (unsigned char)buf_[b_];
    }

private:
    void take(std::string &out, size_t n)
    {
        out.assign(buf_.get() + b_, n);
        b_ += n;
        if (!out.empty() && out.back() == '\r') out.pop_back();
    }

    // capacity for n unread bytes, keeping the unread ones (allocated uninitialised, and grown
    // rarely: fresh pages cost more to fault in than the bytes cost to inflate)
    void reserve(size_t n)
    {
        if (n <= cap_) return;
        const size_t nc = std::max(n, 2 * cap_);
        std::unique_ptr<char[]> nb(new char[nc]);
        if (e_ > b_) std::memcpy(nb.get(), buf_.get() + b_, e_ - b_);
        e_ -= b_;
        b_ = 0;
        buf_ = std::move(nb);
        cap_ = nc;
    }

    gzFile f_ = nullptr;
    Bgzf bgzf_;
    bool use_bgzf_ = false;
    std::string path_;
    std::unique_ptr<char[]> buf_;
    size_t cap_ = 0;
    size_t b_ = 0, e_ = 0;
    bool eof_ = false, err_ = false;
};

std::vector<std::string> split(const std::string &s, char sep)
{
    std::vector<std::string> out;
    size_t a = 0;
    for (;;) {
        const size_t p = s.find(sep, a);
        out.push_back(s.substr(a, p == std::string::npos ? std::string::npos : p - a));
        if (p == std::string::npos) break;
        a = p + 1;
    }
    return out;
}

// IUPAC16 code of an upper-cased reference base: A C G T, else N (loader/reference.cu:288-325)
inline uint32_t ref_code(char c)
{
    switch (c) {
    case 'A': case 'a': return 1;
    case 'C': case 'c': return 2;
    case 'G': case 'g': return 4;
    case 'T': case 't': return 8;
    default: return 15;
    }
}

// seq_nt16_table of the SAM specification: "=ACMGRSVTWYHKDBN", either case, every other
// character N.
struct Nt16 {
    uint8_t t[256];
    Nt16()
    {
        std::memset(t, 15, sizeof t);
        const char *codes = "=ACMGRSVTWYHKDBN";
        for (int i = 0; i < 16; i++) {
            t[(unsigned char)codes[i]] = (uint8_t)i;
            t[(unsigned char)std::tolower(codes[i])] = (uint8_t)i;
        }
    }
};
const Nt16 NT16;

// CIGAR operation letters in BAM code order; B (9) is turned into X by the reference
// (loader/alignments.cu:200-212).
int cigar_code(char c)
{
    static const char *ops = "MIDNSHP=XB";
    const char *p = std::strchr(ops, c);
    return (c && p) ? (int)(p - ops) : -1;
}

// bam_cigar_type bits of the SAM specification: bit 0 consumes query, bit 1 consumes reference.
inline uint32_t cigar_type(uint32_t op)
{
    static const uint32_t t[16] = {3, 1, 2, 2, 1, 0, 0, 3, 3, 0, 0, 0, 0, 0, 0, 0};
    return t[op & 15];
}

} // namespace

// ===========================================================================
// reference

struct fp_sequence_db {
    NameDb names;
    std::vector<uint64_t> word_offset;
    std::vector<uint32_t> length;
    std::vector<uint32_t> words;
};

namespace {

// faidx index names (loader/reference.cu:74-108): the reference keys them by the hash of the first
// whitespace-separated field.
bool load_fai(const std::string &path, std::unordered_set<uint32_t> &names)
{
    FILE *f = std::fopen(path.c_str(), "r");
    if (!f) return false;
    char line[65536];
    while (std::fgets(line, sizeof line, f)) {
        char *p = line;
        while (*p == ' ' || *p == '\t') p++;
        char *q = p;
        while (*q && *q != ' ' && *q != '\t' && *q != '\n' && *q != '\r') q++;
        *q = 0;
        names.insert(name_hash(p));
    }
    std::fclose(f);
    return true;
}

} // namespace

extern "C" fp_status fp_reference_load(const char *fasta_path, fp_sequence_db **out)
{
    if (!fasta_path || !out) return fail(FP_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    Reader rd;
    if (!rd.open(fasta_path)) return fail(FP_ERR_IO, std::string("cannot open ") + fasta_path);
    std::unordered_set<uint32_t> indexed;
    const bool have_index = load_fai(std::string(fasta_path) + ".fai", indexed);

    fp_sequence_db *db = new fp_sequence_db();
    std::string line;
    bool have_header = rd.line(line);
    while (have_header) {
        if (line.empty() || line[0] != '>') {
            // load_next_sequence accepts only a '>' header here (loader/reference.cu:240-244)
            if (db->length.empty()) {
                delete db;
                return fail(FP_ERR_FORMAT, std::string(fasta_path) + ": not a FASTA file");
            }
            break;
        }
        // name: after '>', up to the first space (loader/reference.cu:246-252)
        std::string name = line.substr(1);
        const size_t sp = name.find(' ');
        if (sp != std::string::npos) name.resize(sp);
        const bool take = !have_index || indexed.count(name_hash(name.c_str()));
        const uint64_t woff = db->words.size();
        uint64_t n = 0;
        have_header = false;
        while (rd.line(line)) {
            if (!line.empty() && line[0] == '>') { have_header = true; break; }
            if (!take) continue;
            for (char c : line) {
                if ((n & 7) == 0) db->words.push_back(0);
                db->words.back() |= ref_code(c) << ((n & 7) * 4);
                n++;
            }
        }
        if (rd.error()) {
            delete db;
            return fail(FP_ERR_IO, std::string(fasta_path) + ": read error");
        }
        if (take) {
            if (n > UINT32_MAX) {
                delete db;
                return fail(FP_ERR_UNSUPPORTED, "contig " + name + " is longer than 2^32-1 bases");
            }
            const uint32_t before = (uint32_t)db->names.names.size();
            const uint32_t id = db->names.insert(name);
            if (id == before) {
                db->word_offset.push_back(woff);
                db->length.push_back((uint32_t)n);
            } else {
                // a second sequence of the same name hash: the first stays (string_database::insert)
                db->words.resize(woff);
            }
            // whole-file load stops at the first empty sequence (loader/reference.cu:123-127,324-326)
            if (!have_index && n == 0) break;
        }
    }
    *out = db;
    return FP_OK;
}

extern "C" fp_status fp_reference_get_view(const fp_sequence_db *db, fp_reference_view *out)
{
    if (!db || !out) return fail(FP_ERR_INVALID_ARG, "null argument");
    out->num_sequences = (uint32_t)db->length.size();
    out->word_offset = db->word_offset.data();
    out->length = db->length.data();
    out->words = db->words.data();
    out->total_words = db->words.size();
    return FP_OK;
}

extern "C" uint32_t fp_reference_lookup(const fp_sequence_db *db, const char *name)
{
    if (!db || !name) return UINT32_MAX;
    return db->names.lookup(name);
}

extern "C" fp_status fp_reference_name(const fp_sequence_db *db, uint32_t id, char *buf, size_t buf_len)
{
    if (!db || !buf || !buf_len) return fail(FP_ERR_INVALID_ARG, "null argument");
    if (id >= db->names.names.size()) return fail(FP_ERR_INVALID_ARG, "no such contig id");
    const std::string &s = db->names.names[id];
    const size_t n = std::min(s.size(), buf_len - 1);
    std::memcpy(buf, s.data(), n);
    buf[n] = 0;
    return FP_OK;
}

extern "C" void fp_reference_free(fp_sequence_db *db) { delete db; }

// ===========================================================================
// known sites

struct fp_variant_db {
    std::vector<uint64_t> offset;
    std::vector<uint32_t> start, stop, max_end;
    uint64_t records = 0, skipped = 0;
};

extern "C" fp_status fp_variants_load(const char *vcf_path, const fp_sequence_db *db, fp_variant_db **out)
{
    if (!vcf_path || !db || !out) return fail(FP_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    Reader rd;
    if (!rd.open(vcf_path)) return fail(FP_ERR_IO, std::string("cannot open ") + vcf_path);
    const uint32_t nseq = (uint32_t)db->length.size();
    std::vector<std::vector<uint32_t>> fs(nseq), fe(nseq);
    fp_variant_db *v = new fp_variant_db();
    std::string line;
    uint64_t lineno = 0;
    auto bad = [&](const char *why) {
        delete v;
        return fail(FP_ERR_FORMAT, std::string(vcf_path) + ":" + std::to_string(lineno) + ": " + why);
    };
    while (rd.line(line)) {
        lineno++;
        if (line.empty() || line[0] == '#') continue;
        const std::vector<std::string> f = split(line, '\t');
        if (f.size() < 8) return bad("fewer than 8 columns");
        char *end = nullptr;
        errno = 0;
        const long long pos1 = std::strtoll(f[1].c_str(), &end, 10);
        if (errno || end == f[1].c_str() || *end || pos1 < 1 || pos1 > (long long)UINT32_MAX)
            return bad("bad POS");
        if (f[3].empty()) return bad("empty REF");
        const uint32_t pos = (uint32_t)(pos1 - 1);
        // bcf1_t.rlen: the REF length, or END - POS + 1 when INFO holds END
        int64_t rlen = (int64_t)f[3].size();
        for (const std::string &kv : split(f[7], ';')) {
            if (kv.compare(0, 4, "END=") == 0) {
                const long long endv = std::strtoll(kv.c_str() + 4, &end, 10);
                if (end != kv.c_str() + 4 && !*end) rlen = endv - (long long)pos;
                break;
            }
        }
        v->records++;
        // records on contigs the reference does not hold are skipped (loader/variants.cu:104-110)
        const uint32_t id = db->names.lookup(f[0].c_str());
        if (id == UINT32_MAX || id >= nseq) {
            v->skipped++;
            continue;
        }
        fs[id].push_back(pos);
        fe[id].push_back((uint32_t)((int64_t)pos + rlen - 1));
    }
    if (rd.error()) return bad("read error");
    v->offset.assign(nseq + 1, 0);
    for (uint32_t s = 0; s < nseq; s++) {
        v->offset[s + 1] = v->offset[s] + fs[s].size();
        v->start.insert(v->start.end(), fs[s].begin(), fs[s].end());
        v->stop.insert(v->stop.end(), fe[s].begin(), fe[s].end());
        // inclusive max-scan of the stops (loader/variants.cu:125-134)
        uint32_t m = 0;
        for (size_t i = 0; i < fe[s].size(); i++) {
            m = i ? std::max(m, fe[s][i]) : fe[s][i];
            v->max_end.push_back(m);
        }
    }
    *out = v;
    return FP_OK;
}

extern "C" fp_status fp_variants_get_view(const fp_variant_db *v, fp_variants_view *out)
{
    if (!v || !out) return fail(FP_ERR_INVALID_ARG, "null argument");
    out->num_sequences = (uint32_t)(v->offset.size() - 1);
    out->feature_offset = v->offset.data();
    out->feature_start = v->start.data();
    out->feature_stop = v->stop.data();
    out->max_end_point_left = v->max_end.data();
    out->num_features = v->start.size();
    return FP_OK;
}

extern "C" fp_status fp_variants_counts(const fp_variant_db *v, uint64_t *records, uint64_t *skipped)
{
    if (!v) return fail(FP_ERR_INVALID_ARG, "null argument");
    if (records) *records = v->records;
    if (skipped) *skipped = v->skipped;
    return FP_OK;
}

extern "C" void fp_variants_free(fp_variant_db *v) { delete v; }

// ===========================================================================
// alignments

namespace {

struct BatchStore {
    std::vector<uint16_t> chromosome;
    std::vector<uint32_t> alignment_start, alignment_stop, mate_chromosome, mate_alignment_start;
    std::vector<int32_t> inferred_insert_size;
    std::vector<uint32_t> cigars, cigar_start, cigar_len;
    std::vector<uint32_t> reads, read_start, read_len;
    uint64_t read_bases = 0;
    std::vector<uint8_t> qualities;
    std::vector<uint32_t> qual_start, qual_len;
    std::vector<uint16_t> flags;
    std::vector<uint8_t> mapq;
    std::vector<uint32_t> read_group;

    void clear()
    {
        chromosome.clear(); alignment_start.clear(); alignment_stop.clear(); mate_chromosome.clear();
        mate_alignment_start.clear(); inferred_insert_size.clear(); cigars.clear(); cigar_start.clear();
        cigar_len.clear(); reads.clear(); read_start.clear(); read_len.clear(); read_bases = 0;
        qualities.clear(); qual_start.clear(); qual_len.clear(); flags.clear(); mapq.clear();
        read_group.clear();
    }
    void view(fp_batch_view *o) const
    {
        std::memset(o, 0, sizeof *o);
        o->num_reads = (uint32_t)chromosome.size();
        o->chromosome = chromosome.data();
        o->alignment_start = alignment_start.data();
        o->alignment_stop = alignment_stop.data();
        o->mate_chromosome = mate_chromosome.data();
        o->mate_alignment_start = mate_alignment_start.data();
        o->inferred_insert_size = inferred_insert_size.data();
        o->cigars = cigars.data();
        o->num_cigar_ops = cigars.size();
        o->cigar_start = cigar_start.data();
        o->cigar_len = cigar_len.data();
        o->reads = reads.data();
        o->num_read_bases = read_bases;
        o->read_start = read_start.data();
        o->read_len = read_len.data();
        o->qualities = qualities.data();
        o->num_qualities = qualities.size();
        o->qual_start = qual_start.data();
        o->qual_len = qual_len.data();
        o->flags = flags.data();
        o->mapq = mapq.data();
        o->read_group = read_group.data();
    }
};

// One alignment record in BAM's in-memory terms (bam1_core_t), whichever file form it came from.
struct Record {
    int32_t tid = -1, pos = -1, mtid = -1, mpos = -1, isize = 0;
    uint32_t flag = 0, mapq = 0;
    std::vector<uint32_t> cigar;   // BAM words: len << 4 | op
    std::vector<uint8_t> seq;      // SAM: one nt16 code per base
    std::vector<uint8_t> qual;     // SAM
    // BAM: the record's own packed bases (two per byte, first base in the high nibble) and
    // qualities, in the reader's buffer until the record is appended
    const uint8_t *bseq = nullptr, *bqual = nullptr;
    uint32_t lseq = 0;
    bool has_rg = false;
    std::string rg;
};

// read_group(header_line) (loader/alignments.cu:62-118): the record is split at tabs and every
// field that is FOLLOWED by a tab is read as TG:value; the last field of the line is not read.
void parse_rg_line(const std::string &line, std::string &id, std::string &pu)
{
    id.clear();
    pu.clear();
    size_t a = 0;
    for (;;) {
        const size_t p = line.find('\t', a);
        if (p == std::string::npos) break;
        const std::string tok = line.substr(a, p - a);
        a = p + 1;
        if (tok.size() < 3) continue;
        if (tok[0] == 'I' && tok[1] == 'D') id = tok.substr(3);
        if (tok[0] == 'P' && tok[1] == 'U') pu = tok.substr(3);
    }
}

} // namespace

struct fp_alignment_file {
    Reader rd;
    const fp_sequence_db *db = nullptr;
    bool bam = false;
    uint64_t records = 0, unknown_rg = 0, lineno = 0;
    std::string header_text;
    std::vector<std::string> target_name;
    std::vector<uint32_t> target_len;
    std::unordered_map<std::string, uint32_t> sam_target;   // SAM: RNAME -> tid
    std::vector<uint32_t> target_contig;                    // tid -> contig id of db, or UINT32_MAX
    std::unordered_map<std::string, uint32_t> rg_id_to_name;   // @RG ID -> index into rg_names
    std::vector<std::string> rg_names;                         // name = PU if read, else ID
    uint32_t max_read_groups = 0;
    NameDb read_groups;                                     // alignment_header_host::read_groups_db
    std::vector<const char *> rg_name_ptrs;
    BatchStore batch[2];   // alternate, so a batch stays valid while the next one loads
    unsigned cur = 0;
    unsigned threads = 1;
    // per parsing thread: its records' arrays and read-group names
    std::vector<BatchStore> part;
    std::vector<std::vector<uint32_t>> part_rg;
    std::vector<size_t> span_off, span_len;      // records of the batch being read: byte spans
    std::vector<uint64_t> span_line;             // SAM: line numbers
    std::string line;                            // header lines
};

namespace {

// a malformed header: the file and (SAM) the line
fp_status fmt_error(fp_alignment_file *f, const std::string &why)
{
    const std::string where = f->bam ? " header" : ":" + std::to_string(f->lineno);
    return fail(FP_ERR_FORMAT, f->rd.path() + where + ": " + why);
}

// alignment_file::init (loader/alignments.cu:151-198): read-group map from every "@RG" of the
// header text (found anywhere, up to the next newline).
void build_rg_map(fp_alignment_file *f)
{
    const std::string &h = f->header_text;
    std::unordered_set<uint32_t> distinct;
    size_t at = 0;
    for (;;) {
        at = h.find("@RG", at);
        if (at == std::string::npos) break;
        size_t end = h.find('\n', at + 1);
        const std::string rec = h.substr(at, end == std::string::npos ? std::string::npos : end - at);
        std::string id, pu;
        parse_rg_line(rec, id, pu);
        f->rg_names.push_back(pu.empty() ? id : pu);
        f->rg_id_to_name[id] = (uint32_t)f->rg_names.size() - 1;
        if (end == std::string::npos) break;
        at = end + 1;
    }
    for (auto &kv : f->rg_id_to_name) distinct.insert(name_hash(f->rg_names[kv.second].c_str()));
    f->max_read_groups = (uint32_t)distinct.size();
}

void map_targets(fp_alignment_file *f)
{
    f->target_contig.resize(f->target_name.size());
    for (size_t i = 0; i < f->target_name.size(); i++)
        f->target_contig[i] = f->db->names.lookup(f->target_name[i].c_str());
}

fp_status open_sam_header(fp_alignment_file *f)
{
    while (f->rd.peek() == '@') {
        f->rd.line(f->line);
        f->lineno++;
        f->header_text += f->line;
        f->header_text += '\n';
        if (f->line.compare(0, 4, "@SQ\t") == 0) {
            std::string sn;
            long long ln = -1;
            for (const std::string &t : split(f->line, '\t')) {
                if (t.compare(0, 3, "SN:") == 0) sn = t.substr(3);
                if (t.compare(0, 3, "LN:") == 0) ln = std::strtoll(t.c_str() + 3, nullptr, 10);
            }
            if (sn.empty() || ln < 0) return fmt_error(f, "@SQ line without SN or LN");
            if (!f->sam_target.count(sn)) {
                f->sam_target[sn] = (uint32_t)f->target_name.size();
                f->target_name.push_back(sn);
                f->target_len.push_back((uint32_t)ln);
            }
        }
    }
    return FP_OK;
}

template <class T> T rd_le(const char *p)
{
    T v;
    std::memcpy(&v, p, sizeof v);
    return v;
}

fp_status open_bam_header(fp_alignment_file *f)
{
    Reader &r = f->rd;
    if (r.want(12) < 12) return fmt_error(f, "truncated BAM header");
    r.skip(4);
    const int32_t l_text = rd_le<int32_t>(r.ptr());
    r.skip(4);
    if (l_text < 0 || r.want((size_t)l_text + 4) < (size_t)l_text + 4) return fmt_error(f, "truncated BAM header");
    f->header_text.assign(r.ptr(), (size_t)l_text);
    r.skip((size_t)l_text);
    const int32_t n_ref = rd_le<int32_t>(r.ptr());
    r.skip(4);
    if (n_ref < 0) return fmt_error(f, "negative reference count");
    for (int32_t i = 0; i < n_ref; i++) {
        if (r.want(4) < 4) return fmt_error(f, "truncated BAM header");
        const int32_t l_name = rd_le<int32_t>(r.ptr());
        r.skip(4);
        if (l_name < 1 || r.want((size_t)l_name + 4) < (size_t)l_name + 4) return fmt_error(f, "truncated BAM header");
        f->target_name.emplace_back(r.ptr(), strnlen(r.ptr(), (size_t)l_name));
        r.skip((size_t)l_name);
        f->target_len.push_back(rd_le<uint32_t>(r.ptr()));
        r.skip(4);
    }
    return FP_OK;
}

inline int32_t parse_i32(const char *s, const char *e, bool &ok, int base = 10)
{
    char *end = nullptr;
    const long long v = std::strtoll(s, &end, base);
    ok = end == e && s != e;
    return (int32_t)v;
}

// a record the grammar rejects: the reason, for the caller to place (file and line or record)
inline fp_status bad(std::string &why, const char *what)
{
    why = what;
    return FP_ERR_FORMAT;
}

// One SAM text line into rec (the SAM specification's mandatory fields and the RG tag).
// Where htslib's sam_parse1 (1.2) gives a field a defined meaning, the same meaning is used:
// RNAME "*" or not among the @SQ names -> tid -1; CIGAR "*" -> no ops and the read flagged
// unmapped; RNEXT "=" -> tid; SEQ "*" -> no bases; QUAL "*" -> 0xFF per base; FLAG in any
// strtol base.  A record htslib rejects is an error here (the reference ends the batch there).
fp_status parse_sam(const fp_alignment_file *f, const std::string &l, Record &rec, std::string &why)
{
    const char *p = l.c_str();
    const char *fld[11], *fend[11];
    const char *q = p;
    for (int i = 0; i < 11; i++) {
        fld[i] = q;
        const char *t = std::strchr(q, '\t');
        if (!t) {
            if (i < 10) return bad(why, "fewer than 11 fields");
            t = q + std::strlen(q);
        }
        fend[i] = t;
        q = *t ? t + 1 : t;
    }
    const char *tags = (*fend[10] == '\t') ? fend[10] + 1 : nullptr;
    auto field = [&](int i) { return std::string(fld[i], fend[i]); };
    bool ok;
    rec.flag = (uint32_t)parse_i32(fld[1], fend[1], ok, 0);
    if (!ok) return bad(why, "bad FLAG");
    const std::string rname = field(2);
    if (rname == "*") {
        rec.tid = -1;
    } else {
        if (f->target_name.empty()) return bad(why, "missing SAM header");
        auto it = f->sam_target.find(rname);
        rec.tid = it == f->sam_target.end() ? -1 : (int32_t)it->second;
    }
    rec.pos = parse_i32(fld[3], fend[3], ok) - 1;
    if (!ok) return bad(why, "bad POS");
    rec.mapq = (uint32_t)parse_i32(fld[4], fend[4], ok) & 0xff;
    if (!ok) return bad(why, "bad MAPQ");
    rec.cigar.clear();
    if (fend[5] - fld[5] == 1 && *fld[5] == '*') {
        rec.flag |= 4;
    } else {
        const char *c = fld[5];
        while (c < fend[5]) {
            char *e = nullptr;
            const unsigned long long len = std::strtoull(c, &e, 10);
            if (e == c || e >= fend[5]) return bad(why, "bad CIGAR");
            const int op = cigar_code(*e);
            if (op < 0) return bad(why, "unrecognized CIGAR operator");
            rec.cigar.push_back((uint32_t)(len << 4) | (uint32_t)op);
            c = e + 1;
        }
    }
    const std::string rnext = field(6);
    if (rnext == "=") {
        rec.mtid = rec.tid;
    } else if (rnext == "*") {
        rec.mtid = -1;
    } else {
        auto it = f->sam_target.find(rnext);
        rec.mtid = it == f->sam_target.end() ? -1 : (int32_t)it->second;
    }
    rec.mpos = parse_i32(fld[7], fend[7], ok) - 1;
    if (!ok) return bad(why, "bad PNEXT");
    rec.isize = parse_i32(fld[8], fend[8], ok);
    if (!ok) return bad(why, "bad TLEN");
    rec.bseq = rec.bqual = nullptr;
    rec.seq.clear();
    const size_t lseq = (fend[9] - fld[9] == 1 && *fld[9] == '*') ? 0 : (size_t)(fend[9] - fld[9]);
    if (lseq) {
        uint64_t qlen = 0;
        for (uint32_t w : rec.cigar)
            if (cigar_type(w & 15) & 1) qlen += w >> 4;
        if (!rec.cigar.empty() && qlen != lseq) return bad(why, "SEQ and CIGAR are inconsistent");
        rec.seq.resize(lseq);
        for (size_t i = 0; i < lseq; i++) rec.seq[i] = NT16.t[(unsigned char)fld[9][i]];
    }
    rec.qual.assign(lseq, 0xff);
    if (!(fend[10] - fld[10] == 1 && *fld[10] == '*')) {
        if ((size_t)(fend[10] - fld[10]) != lseq) return bad(why, "SEQ and QUAL are of different length");
        for (size_t i = 0; i < lseq; i++) rec.qual[i] = (uint8_t)(fld[10][i] - 33);
    }
    // first RG tag (bam_aux_get); a non-Z RG is treated as absent
    rec.has_rg = false;
    for (const char *t = tags; t && *t;) {
        const char *e = std::strchr(t, '\t');
        const size_t n = e ? (size_t)(e - t) : std::strlen(t);
        if (n >= 5 && t[0] == 'R' && t[1] == 'G' && t[2] == ':') {
            if (t[3] == 'Z' && t[4] == ':') {
                rec.has_rg = true;
                rec.rg.assign(t + 5, n - 5);
            }
            break;
        }
        t = e ? e + 1 : nullptr;
    }
    return FP_OK;
}

// One BAM record (SAM specification 4.2) into rec.
fp_status parse_bam(const char *p, size_t n, Record &rec, std::string &why)
{
    if (n < 32) return bad(why, "record shorter than its fixed fields");
    rec.tid = rd_le<int32_t>(p);
    rec.pos = rd_le<int32_t>(p + 4);
    const uint32_t l_read_name = (uint8_t)p[8];
    rec.mapq = (uint8_t)p[9];
    const uint32_t n_cigar = rd_le<uint16_t>(p + 12);
    rec.flag = rd_le<uint16_t>(p + 14);
    const int32_t l_seq = rd_le<int32_t>(p + 16);
    rec.mtid = rd_le<int32_t>(p + 20);
    rec.mpos = rd_le<int32_t>(p + 24);
    rec.isize = rd_le<int32_t>(p + 28);
    if (l_seq < 0) return bad(why, "negative sequence length");
    const uint64_t need = 32ull + l_read_name + 4ull * n_cigar + (uint64_t)(l_seq + 1) / 2 + (uint64_t)l_seq;
    if (need > n) return bad(why, "record shorter than its fields");
    const char *c = p + 32 + l_read_name;
    rec.cigar.resize(n_cigar);
    for (uint32_t i = 0; i < n_cigar; i++) rec.cigar[i] = rd_le<uint32_t>(c + 4 * i);
    const uint8_t *s = (const uint8_t *)(c + 4 * n_cigar);
    const uint8_t *qs = s + (l_seq + 1) / 2;
    rec.bseq = s;
    rec.bqual = qs;
    rec.lseq = (uint32_t)l_seq;
    // aux fields: first RG (bam_aux_get), its string if of type Z (bam_aux2Z)
    rec.has_rg = false;
    const char *a = (const char *)(qs + l_seq), *ae = p + n;
    while (a + 3 <= ae) {
        const char t0 = a[0], t1 = a[1], ty = a[2];
        const char *v = a + 3;
        size_t sz;
        switch (ty) {
        case 'A': case 'c': case 'C': sz = 1; break;
        case 's': case 'S': sz = 2; break;
        case 'i': case 'I': case 'f': sz = 4; break;
        case 'd': sz = 8; break;
        case 'Z': case 'H': {
            const char *z = (const char *)std::memchr(v, 0, (size_t)(ae - v));
            if (!z) return bad(why, "unterminated string tag");
            sz = (size_t)(z - v) + 1;
            break;
        }
        case 'B': {
            if (v + 5 > ae) return bad(why, "truncated array tag");
            const char sub = v[0];
            const uint32_t cnt = rd_le<uint32_t>(v + 1);
            const size_t es = (sub == 'c' || sub == 'C') ? 1 : (sub == 's' || sub == 'S') ? 2 : 4;
            sz = 5 + (size_t)cnt * es;
            break;
        }
        default: return bad(why, "unknown tag type");
        }
        if (v + sz > ae) return bad(why, "truncated tag");
        if (t0 == 'R' && t1 == 'G') {
            if (ty == 'Z') {
                rec.has_rg = true;
                rec.rg.assign(v, sz - 1);
            }
            break;
        }
        a = v + sz;
    }
    return FP_OK;
}

// bam_endpos of htslib 1.2: pos + reference length of the CIGAR for a mapped read with CIGAR
// ops, else pos + 1.  The reference stores bam_endpos - 1 (loader/alignments.cu:302-309).
int32_t bam_endpos(const Record &r)
{
    if (!(r.flag & 4) && !r.cigar.empty()) {
        int64_t l = 0;
        for (uint32_t w : r.cigar)
            if (cigar_type(w & 15) & 2) l += w >> 4;
        return (int32_t)(r.pos + l);
    }
    return r.pos + 1;
}

// rg_name of a record: an index into fp_alignment_file::rg_names, or one of these
constexpr uint32_t RGN_ABSENT = 0xffffffffu;    // no RG tag (or one not of type Z)
constexpr uint32_t RGN_UNKNOWN = 0xfffffffeu;   // an RG tag naming no @RG record

// alignment_file::next_batch body for one record (loader/alignments.cu:250-404), into b; the read
// group is left as rg_name, and becomes an id by first appearance when the batch is assembled
fp_status append_local(const fp_alignment_file *f, const Record &r, BatchStore &b, uint32_t &rg_name,
                       std::string &why)
{
    const uint32_t ntarget = (uint32_t)f->target_name.size();
    // chromosome: contig id of the record's reference name; a named contig missing from the
    // reference is an error, no contig at all loads as 0xFFFF for the filters to drop
    uint16_t chrom = 0xffff;
    if (r.tid >= 0 && (uint32_t)r.tid < ntarget) {
        const uint32_t id = f->target_contig[(uint32_t)r.tid];
        if (id == UINT32_MAX) {
            why = "sequence " + f->target_name[(uint32_t)r.tid] + " not found in reference file";
            return FP_ERR_FORMAT;
        }
        if (id >= 0xffff) { why = "more than 65535 reference contigs"; return FP_ERR_UNSUPPORTED; }
        chrom = (uint16_t)id;
    }
    b.chromosome.push_back(chrom);
    b.alignment_start.push_back((uint32_t)r.pos);
    b.alignment_stop.push_back((uint32_t)(bam_endpos(r) - 1));
    uint32_t mchrom = UINT32_MAX;
    if (r.mtid >= 0 && (uint32_t)r.mtid < ntarget) mchrom = f->target_contig[(uint32_t)r.mtid];
    b.mate_chromosome.push_back(mchrom);
    b.mate_alignment_start.push_back((uint32_t)r.mpos);
    b.inferred_insert_size.push_back(r.isize);
    // cigar_op{len:24, op:4}: B becomes X (loader/alignments.cu:200-212, 318-334)
    b.cigar_start.push_back((uint32_t)b.cigars.size());
    b.cigar_len.push_back((uint32_t)r.cigar.size());
    for (uint32_t w : r.cigar) {
        uint32_t op = w & 15;
        if (op == 9) op = 8;
        b.cigars.push_back(((w >> 4) & 0xffffffu) | (op << 24));
    }
    // bases: 4-bit codes, each read starting on a word, padded to 8 (loader/alignments.cu:337-367)
    const uint32_t n = r.bseq ? r.lseq : (uint32_t)r.seq.size();
    b.read_start.push_back((uint32_t)b.read_bases);
    b.read_len.push_back(n);
    const uint64_t padded = ((uint64_t)n + 7) / 8 * 8;
    const size_t w0 = b.reads.size();
    b.reads.resize(w0 + padded / 8, 0);
    if (r.bseq) {
        // BAM byte k holds bases 2k (high nibble) and 2k+1 (low); the batch's byte k (little-endian
        // words) holds them the other way round, so each byte has its nibbles swapped
        uint8_t *dst = (uint8_t *)(b.reads.data() + w0);
        const uint32_t nb = (n + 1) / 2;
        for (uint32_t k = 0; k < nb; k++) dst[k] = (uint8_t)((r.bseq[k] >> 4) | (r.bseq[k] << 4));
        if (n & 1) dst[nb - 1] &= 15;   // the unused half of BAM's last byte stays padding
    } else {
        for (uint32_t i = 0; i < n; i++) b.reads[w0 + i / 8] |= (uint32_t)r.seq[i] << ((i & 7) * 4);
    }
    b.read_bases += padded;
    b.qual_start.push_back((uint32_t)b.qualities.size());
    b.qual_len.push_back(n);
    if (r.bqual) b.qualities.insert(b.qualities.end(), r.bqual, r.bqual + n);
    else b.qualities.insert(b.qualities.end(), r.qual.begin(), r.qual.end());
    // the twelve SAM flag bits map one to one onto AlignmentFlags (loader/alignments.cu:214-255)
    b.flags.push_back((uint16_t)(r.flag & 0xfff));
    b.mapq.push_back((uint8_t)r.mapq);
    // read group: RG tag -> @RG name (loader/alignments.cu:383-403)
    rg_name = RGN_ABSENT;
    if (r.has_rg) {
        auto it = f->rg_id_to_name.find(r.rg);
        rg_name = it == f->rg_id_to_name.end() ? RGN_UNKNOWN : it->second;
    }
    b.read_group.push_back(UINT32_MAX);
    return FP_OK;
}

} // namespace

extern "C" fp_status fp_alignments_open(const char *path, const fp_sequence_db *db, fp_alignment_file **out)
{
    if (!path || !db || !out) return fail(FP_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    fp_alignment_file *f = new fp_alignment_file();
    f->db = db;
    if (!f->rd.open(path)) {
        delete f;
        return fail(FP_ERR_IO, std::string("cannot open ") + path);
    }
    f->bam = f->rd.want(4) >= 4 && std::memcmp(f->rd.ptr(), "BAM\1", 4) == 0;
    fp_status rc = f->bam ? open_bam_header(f) : open_sam_header(f);
    if (rc != FP_OK) {
        delete f;
        return rc;
    }
    build_rg_map(f);
    map_targets(f);
    // parsing threads: FP_IO_THREADS if set, else the host's CPUs (at most 16)
    const unsigned hw = std::thread::hardware_concurrency();
    f->threads = hw == 0 ? 4 : (hw > 16 ? 16 : hw);
    if (const char *env = std::getenv("FP_IO_THREADS")) {
        const long v = std::strtol(env, nullptr, 10);
        if (v >= 1 && v <= 64) f->threads = (unsigned)v;
    }
    *out = f;
    return FP_OK;
}

// alignment_file::next_batch (loader/alignments.cu:238-407).  The records' byte spans are found in
// order; their fields are then parsed on several threads, each into its own arrays, which are joined
// in record order, where the read groups get their ids by first appearance.
extern "C" fp_status fp_alignments_next_batch(fp_alignment_file *f, uint32_t batch_size, fp_batch_view *out)
{
    if (!f || !out) return fail(FP_ERR_INVALID_ARG, "null argument");
    f->cur ^= 1;
    BatchStore &b = f->batch[f->cur];
    b.clear();
    Reader &rd = f->rd;
    f->span_off.clear();
    f->span_len.clear();
    f->span_line.clear();
    size_t c = 0;
    while (f->span_off.size() < batch_size) {
        if (f->bam) {
            if (rd.want(c + 4) < c + 4) break;
            const int32_t bs = rd_le<int32_t>(rd.ptr() + c);
            if (bs < 0 || rd.want(c + 4 + (size_t)bs) < c + 4 + (size_t)bs) {
                b.view(out);
                return fail(FP_ERR_FORMAT, rd.path() + " record " + std::to_string(f->records + f->span_off.size() + 1) +
                                               ": truncated record");
            }
            f->span_off.push_back(c + 4);
            f->span_len.push_back((size_t)bs);
            c += 4 + (size_t)bs;
        } else {
            size_t len, next;
            if (!rd.line_at(c, len, next)) break;
            f->lineno++;
            if (len) {
                f->span_off.push_back(c);
                f->span_len.push_back(len);
                f->span_line.push_back(f->lineno);
            }
            c = next;
        }
    }
    if (rd.error()) {
        b.view(out);
        return fail(FP_ERR_IO, rd.path() + ": read error");
    }
    const size_t nrec = f->span_off.size();
    const unsigned T = nrec >= 4096 ? f->threads : 1u;
    if (f->part.size() < T) { f->part.resize(T); f->part_rg.resize(T); }
    std::vector<size_t> fail_at(T, SIZE_MAX);
    std::vector<fp_status> fail_rc(T, FP_OK);
    std::vector<std::string> fail_why(T);
    const char *base = rd.ptr();
    auto work = [&](unsigned t) {
        const size_t k0 = nrec * t / T, k1 = nrec * (t + 1) / T;
        BatchStore &lb = f->part[t];
        std::vector<uint32_t> &lrg = f->part_rg[t];
        lb.clear();
        lrg.clear();
        Record rec;
        std::string line, why;
        for (size_t k = k0; k < k1; k++) {
            fp_status rc;
            if (f->bam) {
                rc = parse_bam(base + f->span_off[k], f->span_len[k], rec, why);
            } else {
                line.assign(base + f->span_off[k], f->span_len[k]);
                rc = parse_sam(f, line, rec, why);
            }
            uint32_t rgn = RGN_ABSENT;
            if (rc == FP_OK) rc = append_local(f, rec, lb, rgn, why);
            if (rc != FP_OK) { fail_at[t] = k; fail_rc[t] = rc; fail_why[t] = why; return; }
            lrg.push_back(rgn);
        }
    };
    {
        std::vector<std::thread> pool;
        for (unsigned t = 1; t < T; t++) pool.emplace_back(work, t);
        work(0);
        for (auto &th : pool) th.join();
    }
    rd.skip(c);
    for (unsigned t = 0; t < T; t++) {
        if (fail_at[t] == SIZE_MAX) continue;
        const size_t k = fail_at[t];
        const std::string where = f->bam ? " record " + std::to_string(f->records + k + 1)
                                         : ":" + std::to_string(f->span_line[k]);
        b.clear();
        b.view(out);
        return fail(fail_rc[t], rd.path() + where + ": " + fail_why[t]);
    }
    // join the threads' arrays in record order
    for (unsigned t = 0; t < T; t++) {
        const BatchStore &L = f->part[t];
        const uint32_t co = (uint32_t)b.cigars.size(), qo = (uint32_t)b.qualities.size();
        const uint32_t bo = (uint32_t)b.read_bases;
        auto cat = [](auto &dst, const auto &src) { dst.insert(dst.end(), src.begin(), src.end()); };
        cat(b.chromosome, L.chromosome);
        cat(b.alignment_start, L.alignment_start);
        cat(b.alignment_stop, L.alignment_stop);
        cat(b.mate_chromosome, L.mate_chromosome);
        cat(b.mate_alignment_start, L.mate_alignment_start);
        cat(b.inferred_insert_size, L.inferred_insert_size);
        cat(b.cigar_len, L.cigar_len);
        cat(b.read_len, L.read_len);
        cat(b.qual_len, L.qual_len);
        cat(b.flags, L.flags);
        cat(b.mapq, L.mapq);
        for (uint32_t x : L.cigar_start) b.cigar_start.push_back(x + co);
        for (uint32_t x : L.read_start) b.read_start.push_back(x + bo);
        for (uint32_t x : L.qual_start) b.qual_start.push_back(x + qo);
        cat(b.cigars, L.cigars);
        cat(b.reads, L.reads);
        cat(b.qualities, L.qualities);
        b.read_bases += L.read_bases;
        // read groups: ids by first appearance (loader/alignments.cu:396-401)
        for (uint32_t rgn : f->part_rg[t]) {
            uint32_t id = UINT32_MAX;
            if (rgn == RGN_UNKNOWN) f->unknown_rg++;
            else if (rgn != RGN_ABSENT) id = f->read_groups.insert(f->rg_names[rgn]);
            b.read_group.push_back(id);
        }
    }
    f->records += nrec;
    b.view(out);
    return FP_OK;
}

extern "C" fp_status fp_alignments_header(const fp_alignment_file *cf, fp_header_view *out, uint32_t *max_read_groups)
{
    if (!cf || !out) return fail(FP_ERR_INVALID_ARG, "null argument");
    fp_alignment_file *f = const_cast<fp_alignment_file *>(cf);
    f->rg_name_ptrs.clear();
    for (const std::string &s : f->read_groups.names) f->rg_name_ptrs.push_back(s.c_str());
    out->num_chromosomes = (uint32_t)f->target_len.size();
    out->chromosome_lengths = f->target_len.data();
    out->num_read_groups = (uint32_t)f->rg_name_ptrs.size();
    out->read_group_names = f->rg_name_ptrs.data();
    if (max_read_groups) *max_read_groups = f->max_read_groups;
    return FP_OK;
}

extern "C" fp_status fp_alignments_counts(const fp_alignment_file *f, uint64_t *records, uint64_t *unknown_read_group)
{
    if (!f) return fail(FP_ERR_INVALID_ARG, "null argument");
    if (records) *records = f->records;
    if (unknown_read_group) *unknown_read_group = f->unknown_rg;
    return FP_OK;
}

extern "C" void fp_alignments_close(fp_alignment_file *f) { delete f; }

extern "C" const char *fp_io_last_error(void) { return g_err.c_str(); }
