/*
 * Seeded synthetic BQSR workloads (see synth.h).
 *
 * Read model (per read, from its own xoshiro256** stream):
 *   read group uniform; length uniform over spec.lengths; strand 50/50;
 *   optional soft clip on one random end (non-reference bases); optional
 *   single indel (I or D, equal odds) placed inside the aligned part;
 *   M bases copy the genome and are substituted with the read group's rate;
 *   qualities P(q) ~ exp(-((q-32)/6)^2 / 2) on 2..41; optional Q2 tail on the
 *   last bases in sequencing order; MAPQ 60.
 * A fraction of reads can instead be drawn from an edge-case menu (filtered
 * flags, hard clips, N bases, =/X ops, contig ends, adapter read-through ...)
 * to exercise every branch of the filters and windows.
 */
#include "synth.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

namespace {

struct Rng {
    uint64_t s[4];
    static uint64_t splitmix(uint64_t &x)
    {
        uint64_t z = (x += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    }
    explicit Rng(uint64_t seed)
    {
        uint64_t x = seed;
        for (int i = 0; i < 4; i++) s[i] = splitmix(x);
    }
    static uint64_t rotl(uint64_t v, int k) { return (v << k) | (v >> (64 - k)); }
    uint64_t next()
    {
        const uint64_t result = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return result;
    }
    double u01() { return (next() >> 11) * 0x1.0p-53; }
    uint32_t below(uint64_t n) { return n ? (uint32_t)(((unsigned __int128)next() * n) >> 64) : 0; }
    uint32_t range(uint32_t lo, uint32_t hi) { return lo + below((uint64_t)hi - lo + 1); }   // inclusive
    double normal()
    {
        double u1 = u01(), u2 = u01();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2 * M_PI * u2);
    }
};

static const uint8_t ACGT[4] = {1, 2, 4, 8};
enum { OP_M = 0, OP_I = 1, OP_D = 2, OP_N = 3, OP_S = 4, OP_H = 5, OP_P = 6, OP_EQ = 7, OP_X = 8 };
enum { F_PAIRED = 1, F_PROPER = 2, F_UNMAP = 4, F_MUNMAP = 8, F_REV = 16, F_MREV = 32, F_R1 = 64,
       F_R2 = 128, F_SEC = 256, F_QCFAIL = 512, F_DUP = 1024 };

static inline uint32_t cig(uint32_t op, uint32_t len) { return (len & 0xffffff) | (op << 24); }

} // namespace

struct fpsyn_genome {
    fpsyn_genome_spec spec;
    std::vector<uint32_t> words;
    std::vector<uint64_t> word_offset;
    std::vector<uint32_t> length;
    std::vector<uint64_t> cum_length;      // prefix of contig lengths
    std::vector<uint64_t> feature_offset;
    std::vector<uint32_t> feature_start, feature_stop;
    std::vector<std::string> rg_name_store;
    std::vector<const char *> rg_names;

    uint8_t base(uint32_t c, uint64_t i) const
    {
        return (words[word_offset[c] + (i >> 3)] >> ((i & 7) * 4)) & 0xf;
    }
};

struct fpsyn_batch {
    uint32_t num_reads = 0;
    std::vector<uint16_t> chromosome, flags;
    std::vector<uint32_t> alignment_start, alignment_stop, mate_chromosome, mate_alignment_start;
    std::vector<int32_t> isize;
    std::vector<uint32_t> cigars, cigar_start, cigar_len;
    std::vector<uint32_t> reads, read_start, read_len;
    std::vector<uint8_t> quals, mapq;
    std::vector<uint32_t> qual_start, qual_len, read_group;
};

namespace {

struct QualTable {
    double cdf[40];
    QualTable()
    {
        double s = 0;
        for (int q = 2; q <= 41; q++) { s += std::exp(-0.5 * ((q - 32) / 6.0) * ((q - 32) / 6.0)); cdf[q - 2] = s; }
        for (int i = 0; i < 40; i++) cdf[i] /= s;
    }
    uint8_t draw(Rng &r) const
    {
        const double u = r.u01();
        int i = 0;
        while (i < 39 && cdf[i] < u) i++;
        return (uint8_t)(i + 2);
    }
};

static const QualTable g_qual;

struct ReadOut {
    uint16_t chromosome = 0, flags = 0;
    uint32_t start = 0, stop = 0, mchrom = 0xffffffffu, mstart = 0xffffffffu, rg = 0;
    int32_t isize = 0;
    uint8_t mapq = 60;
    std::vector<uint32_t> ops;
    std::vector<uint8_t> bases, quals;
};

static uint32_t ref_span(const std::vector<uint32_t> &ops)
{
    uint32_t s = 0;
    for (uint32_t o : ops) {
        const uint32_t op = o >> 24, len = o & 0xffffff;
        if (op == OP_M || op == OP_D || op == OP_N || op == OP_EQ || op == OP_X) s += len;
    }
    return s;
}

static uint32_t pick_contig(const fpsyn_genome &g, Rng &rng, uint32_t span, uint32_t *start)
{
    const uint64_t total = g.cum_length.back();
    for (int attempt = 0; attempt < 64; attempt++) {
        const uint64_t p = (uint64_t)(rng.u01() * (double)total);
        uint32_t c = (uint32_t)(std::upper_bound(g.cum_length.begin(), g.cum_length.end(), p) - g.cum_length.begin()) - 1;
        if (c >= g.length.size()) c = (uint32_t)g.length.size() - 1;
        if (g.length[c] < span + 1) continue;
        *start = rng.below(g.length[c] - span + 1);
        return c;
    }
    *start = 0;
    return 0;
}

// build one read: CIGAR over a clean aligned read, then bases and qualities
static void make_read(const fpsyn_genome &g, const fpsyn_read_spec &sp, uint64_t index, ReadOut &o)
{
    Rng rng(sp.seed * 0x2545f4914f6cdd1dull ^ (index + 1) * 0x9e3779b97f4a7c15ull);
    o = ReadOut();
    o.rg = sp.num_read_groups ? rng.below(sp.num_read_groups) : 0;
    const double p_mis = sp.rg_mismatch[o.rg];
    uint32_t L = sp.lengths[sp.num_lengths > 1 ? rng.below(sp.num_lengths) : 0];
    const bool rev = rng.u01() < 0.5;

    int edge = -1;
    if (sp.edge_case_frac > 0 && rng.u01() < sp.edge_case_frac) edge = (int)rng.below(27);

    // ---- CIGAR
    uint32_t cl = 0, cr = 0;
    if (sp.clip_read_frac > 0 && rng.u01() < sp.clip_read_frac) {
        const uint32_t c = rng.range(sp.clip_min, sp.clip_max);
        if (rng.u01() < 0.5) cl = c; else cr = c;
    }
    if (cl + cr + 12 > L) { cl = 0; cr = 0; }
    uint32_t ilen = 0, itype = OP_M;
    if (sp.indel_read_frac > 0 && rng.u01() < sp.indel_read_frac) {
        ilen = rng.range(sp.indel_min, sp.indel_max);
        itype = rng.u01() < 0.5 ? OP_I : OP_D;
    }
    if (edge == 20) { ilen = rng.range(12, 20); itype = OP_D; }
    uint32_t a = L - cl - cr, b = 0;   // M before / after the indel
    if (ilen) {
        const uint32_t consumed = itype == OP_I ? ilen : 0;
        const uint32_t aligned = L - cl - cr;
        if (aligned >= consumed + 10 + 10) {
            // position inside [10, aligned - consumed - 10] of the aligned part
            const uint32_t pos = rng.range(10, aligned - consumed - 10);
            a = pos;
            b = aligned - consumed - pos;
        } else {
            ilen = 0;
        }
    }
    std::vector<uint32_t> &ops = o.ops;
    if (cl) ops.push_back(cig(OP_S, cl));
    ops.push_back(cig(OP_M, a));
    if (ilen) { ops.push_back(cig(itype, ilen)); ops.push_back(cig(OP_M, b)); }
    if (cr) ops.push_back(cig(OP_S, cr));

    // edge cases that rewrite the CIGAR
    switch (edge) {
    case 6:   // N operator (filtered)
        ops.clear(); ops.push_back(cig(OP_M, L / 2)); ops.push_back(cig(OP_N, 3)); ops.push_back(cig(OP_M, L - L / 2));
        break;
    case 7:   // misplaced soft clip (filtered)
        ops.clear(); ops.push_back(cig(OP_M, L / 2)); ops.push_back(cig(OP_S, 5)); ops.push_back(cig(OP_M, L - L / 2 - 5));
        break;
    case 8: ops.insert(ops.begin(), cig(OP_H, 5)); break;                 // leading hard clip
    case 9: ops.push_back(cig(OP_H, 4)); break;                           // trailing hard clip
    case 13:  // '=' and 'X' ops in place of M runs
        for (size_t k = 0, seen = 0; k < ops.size(); k++)
            if ((ops[k] >> 24) == OP_M) { ops[k] = cig(seen++ == 0 ? OP_EQ : OP_X, ops[k] & 0xffffff); }
        break;
    case 14: ops.clear(); ops.push_back(cig(OP_I, 3)); ops.push_back(cig(OP_M, L - 3)); break;
    case 15: ops.clear(); ops.push_back(cig(OP_M, L - 3)); ops.push_back(cig(OP_I, 3)); break;
    case 19: L = 1; ops.clear(); ops.push_back(cig(OP_M, 1)); break;
    case 21: ops.clear(); ops.push_back(cig(OP_S, L - 8)); ops.push_back(cig(OP_M, 2)); ops.push_back(cig(OP_S, 6)); break;
    case 23: ops.clear(); ops.push_back(cig(OP_S, 5)); ops.push_back(cig(OP_D, 3)); ops.push_back(cig(OP_M, L - 5)); break;
    case 24: ops.clear(); ops.push_back(cig(OP_M, L - 8)); ops.push_back(cig(OP_I, 3)); ops.push_back(cig(OP_S, 5)); break;
    default: break;
    }

    // ---- placement
    const uint32_t span = std::max<uint32_t>(ref_span(ops), 1);
    uint32_t start;
    uint32_t c = pick_contig(g, rng, span + 2, &start);
    if (edge == 16) start = rng.below(3);
    if (edge == 17) start = g.length[c] - span;
    if (start + span > g.length[c]) start = g.length[c] - span;
    o.chromosome = (uint16_t)c;
    o.start = start;
    o.stop = start + span - 1;

    // ---- bases (H consumes reference in firepony's coordinate expansion but
    // not here: bases follow the SAM meaning of each op)
    uint64_t rpos = start;
    for (uint32_t w : ops) {
        const uint32_t op = w >> 24, len = w & 0xffffff;
        switch (op) {
        case OP_M: case OP_EQ: case OP_X:
            for (uint32_t k = 0; k < len; k++, rpos++) {
                uint8_t gb = rpos < g.length[c] ? g.base(c, rpos) : 15;
                uint8_t rb = gb == 15 ? ACGT[rng.below(4)] : gb;
                if (rng.u01() < p_mis) {
                    uint8_t nb;
                    do { nb = ACGT[rng.below(4)]; } while (nb == gb);
                    rb = nb;
                }
                o.bases.push_back(rb);
            }
            break;
        case OP_I: case OP_S:
            for (uint32_t k = 0; k < len; k++) o.bases.push_back(ACGT[rng.below(4)]);
            break;
        case OP_D: case OP_N:
            rpos += len;
            break;
        default: break;
        }
    }
    L = (uint32_t)o.bases.size();

    // ---- qualities
    o.quals.resize(L);
    for (uint32_t k = 0; k < L; k++) o.quals[k] = g_qual.draw(rng);
    if (sp.q2_base_frac > 0)
        for (uint32_t k = 0; k < L; k++) if (rng.u01() < sp.q2_base_frac) o.quals[k] = 2;
    if (sp.q2_tail_frac > 0 && rng.u01() < sp.q2_tail_frac && L > 1) {
        const uint32_t T = std::min<uint32_t>(rng.range(1, std::max<uint32_t>(sp.q2_tail_max, 1)), L - 1);
        for (uint32_t k = 0; k < T; k++) o.quals[rev ? k : L - 1 - k] = 2;
    }

    // ---- flags and mate fields
    o.flags = rev ? F_REV : 0;
    if (sp.paired || edge == 18 || edge == 25 || edge == 26) {
        const bool second = (edge == 25 || edge == 26) ? true : (index & 1);
        o.flags |= F_PAIRED | F_PROPER | (second ? F_R2 : F_R1);
        if (edge == 26) o.flags |= F_REV;
        const bool r = o.flags & F_REV;
        if (!r) o.flags |= F_MREV;
        int32_t t = (int32_t)std::lround(sp.insert_mean + sp.insert_sd * rng.normal());
        if (sp.insert_mean <= 0) t = (int32_t)L + 200;
        if (edge == 18) t = (int32_t)span - (int32_t)rng.range(3, 20);
        if (t < 2) t = 2;
        o.mchrom = c;
        if (!r) { o.mstart = (uint32_t)std::max<int64_t>(0, (int64_t)start + t - (int64_t)L); o.isize = t; }
        else { o.mstart = (uint32_t)std::max<int64_t>(0, (int64_t)o.stop - t + 1); o.isize = -t; }
    }

    switch (edge) {
    case 0: o.flags |= F_UNMAP; break;
    case 1: o.mapq = 0; break;
    case 2: o.mapq = 255; break;
    case 3: o.flags |= F_DUP; break;
    case 4: o.flags |= F_QCFAIL; break;
    case 5: o.flags |= F_SEC; break;
    case 10: o.rg = 0xffffffffu; break;
    case 11: for (int k = 0; k < 3; k++) o.bases[rng.below(L)] = 15; break;
    case 12: for (int k = 0; k < 6; k++) o.quals[rng.below(L)] = (uint8_t)rng.below(6); break;
    case 22: for (uint32_t k = 0; k < std::min<uint32_t>(4, L); k++) o.quals[rev ? L - 1 - k : k] = 2; break;
    default: break;
    }
}

struct Chunk {
    std::vector<ReadOut> reads;
};

} // namespace

extern "C" {

fpsyn_genome *fpsyn_genome_create(const fpsyn_genome_spec *spec)
{
    if (!spec || spec->num_contigs == 0 || spec->num_contigs > 32) return nullptr;
    fpsyn_genome *g = new fpsyn_genome();
    g->spec = *spec;
    uint64_t woff = 0;
    g->cum_length.push_back(0);
    for (uint32_t c = 0; c < spec->num_contigs; c++) {
        g->word_offset.push_back(woff);
        g->length.push_back((uint32_t)spec->contig_length[c]);
        woff += (spec->contig_length[c] + 7) / 8;
        g->cum_length.push_back(g->cum_length.back() + spec->contig_length[c]);
    }
    g->words.assign(woff, 0);
    std::vector<std::thread> th;
    for (uint32_t c = 0; c < spec->num_contigs; c++) {
        th.emplace_back([g, c, spec]() {
            Rng rng(spec->seed ^ (0xc0117ull + c * 0x9e3779b97f4a7c15ull));
            const uint64_t n = g->length[c];
            uint32_t *w = &g->words[g->word_offset[c]];
            for (uint64_t i = 0; i < n; i += 32) {
                uint64_t bits = rng.next();
                for (uint64_t k = i; k < std::min<uint64_t>(i + 32, n); k++, bits >>= 2)
                    w[k >> 3] |= (uint32_t)ACGT[bits & 3] << ((k & 7) * 4);
            }
            if (spec->n_run_fraction > 0) {
                uint64_t target = (uint64_t)(spec->n_run_fraction * n), done = 0;
                while (done < target) {
                    const uint64_t len = rng.range(50, 500), at = rng.below(n);
                    for (uint64_t k = at; k < std::min<uint64_t>(n, at + len); k++) {
                        w[k >> 3] &= ~(0xfu << ((k & 7) * 4));
                        w[k >> 3] |= 15u << ((k & 7) * 4);
                    }
                    done += len;
                }
            }
        });
    }
    for (auto &t : th) t.join();

    // known sites: distinct uniform positions, sorted per contig
    Rng rng(spec->seed ^ 0x517e5ull);
    std::vector<std::vector<std::pair<uint32_t, uint32_t>>> per(spec->num_contigs);
    const uint64_t total = g->cum_length.back();
    std::vector<uint64_t> pos;
    pos.reserve(spec->num_sites);
    for (uint32_t i = 0; i < spec->num_sites; i++) pos.push_back((uint64_t)(rng.u01() * (double)total));
    std::sort(pos.begin(), pos.end());
    pos.erase(std::unique(pos.begin(), pos.end()), pos.end());
    for (uint64_t p : pos) {
        uint32_t c = (uint32_t)(std::upper_bound(g->cum_length.begin(), g->cum_length.end(), p) - g->cum_length.begin()) - 1;
        const uint32_t off = (uint32_t)(p - g->cum_length[c]);
        uint32_t stop = off;
        if (spec->multi_base_site_frac > 0 && rng.u01() < spec->multi_base_site_frac)
            stop = std::min<uint32_t>(off + rng.range(1, 5), g->length[c] - 1);
        per[c].push_back({off, stop});
    }
    g->feature_offset.push_back(0);
    for (uint32_t c = 0; c < spec->num_contigs; c++) {
        for (auto &f : per[c]) { g->feature_start.push_back(f.first); g->feature_stop.push_back(f.second); }
        g->feature_offset.push_back(g->feature_start.size());
    }
    return g;
}

void fpsyn_genome_destroy(fpsyn_genome *g) { delete g; }

void fpsyn_genome_reference(const fpsyn_genome *g, fp_reference_view *o)
{
    o->num_sequences = (uint32_t)g->length.size();
    o->word_offset = g->word_offset.data();
    o->length = g->length.data();
    o->words = g->words.data();
    o->total_words = g->words.size();
}

void fpsyn_genome_variants(const fpsyn_genome *g, fp_variants_view *o)
{
    o->num_sequences = (uint32_t)g->length.size();
    o->feature_offset = g->feature_offset.data();
    o->feature_start = g->feature_start.data();
    o->feature_stop = g->feature_stop.data();
    o->max_end_point_left = nullptr;
    o->num_features = g->feature_start.size();
}

void fpsyn_genome_header(fpsyn_genome *g, uint32_t num_read_groups, fp_header_view *o)
{
    g->rg_name_store.clear();
    g->rg_names.clear();
    for (uint32_t i = 0; i < num_read_groups; i++) g->rg_name_store.push_back("rg" + std::to_string(i));
    for (auto &s : g->rg_name_store) g->rg_names.push_back(s.c_str());
    o->num_chromosomes = (uint32_t)g->length.size();
    o->chromosome_lengths = g->length.data();
    o->num_read_groups = num_read_groups;
    o->read_group_names = g->rg_names.data();
}

fpsyn_batch *fpsyn_batch_create(const fpsyn_genome *g, const fpsyn_read_spec *sp, uint64_t first,
                                uint32_t n, uint32_t threads)
{
    if (!g || !sp || sp->num_lengths == 0) return nullptr;
    threads = std::max<uint32_t>(1, std::min<uint32_t>(threads, 64));
    std::vector<Chunk> chunks(threads);
    const uint32_t per = (n + threads - 1) / threads;
    std::vector<std::thread> th;
    for (uint32_t t = 0; t < threads; t++) {
        th.emplace_back([&, t]() {
            const uint32_t a = std::min(n, t * per), e = std::min(n, a + per);
            chunks[t].reads.resize(e - a);
            for (uint32_t i = a; i < e; i++) make_read(*g, *sp, first + i, chunks[t].reads[i - a]);
        });
    }
    for (auto &x : th) x.join();

    fpsyn_batch *b = new fpsyn_batch();
    b->num_reads = n;
    uint64_t nops = 0, nbases = 0, nq = 0;
    for (auto &ch : chunks)
        for (auto &r : ch.reads) {
            nops += r.ops.size();
            nbases += (r.bases.size() + 7) / 8 * 8;
            nq += r.quals.size();
        }
    b->chromosome.resize(n); b->flags.resize(n); b->alignment_start.resize(n); b->alignment_stop.resize(n);
    b->mate_chromosome.resize(n); b->mate_alignment_start.resize(n); b->isize.resize(n);
    b->cigars.resize(nops); b->cigar_start.resize(n); b->cigar_len.resize(n);
    b->reads.assign(nbases / 8, 0); b->read_start.resize(n); b->read_len.resize(n);
    b->quals.resize(nq); b->qual_start.resize(n); b->qual_len.resize(n);
    b->mapq.resize(n); b->read_group.resize(n);
    uint64_t co = 0, ro = 0, qo = 0;
    uint32_t i = 0;
    for (auto &ch : chunks)
        for (auto &r : ch.reads) {
            b->chromosome[i] = r.chromosome; b->flags[i] = r.flags;
            b->alignment_start[i] = r.start; b->alignment_stop[i] = r.stop;
            b->mate_chromosome[i] = r.mchrom; b->mate_alignment_start[i] = r.mstart; b->isize[i] = r.isize;
            b->cigar_start[i] = (uint32_t)co; b->cigar_len[i] = (uint32_t)r.ops.size();
            std::copy(r.ops.begin(), r.ops.end(), b->cigars.begin() + co);
            co += r.ops.size();
            b->read_start[i] = (uint32_t)ro; b->read_len[i] = (uint32_t)r.bases.size();
            for (size_t k = 0; k < r.bases.size(); k++)
                b->reads[(ro + k) >> 3] |= (uint32_t)r.bases[k] << (((ro + k) & 7) * 4);
            ro += (r.bases.size() + 7) / 8 * 8;
            b->qual_start[i] = (uint32_t)qo; b->qual_len[i] = (uint32_t)r.quals.size();
            std::copy(r.quals.begin(), r.quals.end(), b->quals.begin() + qo);
            qo += r.quals.size();
            b->mapq[i] = r.mapq; b->read_group[i] = r.rg;
            i++;
        }
    return b;
}

void fpsyn_batch_view(const fpsyn_batch *b, fp_batch_view *o)
{
    memset(o, 0, sizeof(*o));
    o->num_reads = b->num_reads;
    o->chromosome = b->chromosome.data();
    o->alignment_start = b->alignment_start.data();
    o->alignment_stop = b->alignment_stop.data();
    o->mate_chromosome = b->mate_chromosome.data();
    o->mate_alignment_start = b->mate_alignment_start.data();
    o->inferred_insert_size = b->isize.data();
    o->cigars = b->cigars.data();
    o->num_cigar_ops = b->cigars.size();
    o->cigar_start = b->cigar_start.data();
    o->cigar_len = b->cigar_len.data();
    o->reads = b->reads.data();
    o->num_read_bases = b->reads.size() * 8;
    o->read_start = b->read_start.data();
    o->read_len = b->read_len.data();
    o->qualities = b->quals.data();
    o->num_qualities = b->quals.size();
    o->qual_start = b->qual_start.data();
    o->qual_len = b->qual_len.data();
    o->flags = b->flags.data();
    o->mapq = b->mapq.data();
    o->read_group = b->read_group.data();
}

void fpsyn_batch_destroy(fpsyn_batch *b) { delete b; }

/* Named workloads.  Seeds are 0xF1E90000 + config. */
int fpsyn_named_spec(int config, fpsyn_genome_spec *gs, fpsyn_read_spec *rs, uint64_t *num_reads)
{
    memset(gs, 0, sizeof(*gs));
    memset(rs, 0, sizeof(*rs));
    const uint64_t seed = 0xF1E90000ull + (uint64_t)config;
    gs->seed = seed;
    rs->seed = seed;
    rs->q2_tail_frac = 0.05;
    rs->q2_tail_max = 20;
    switch (config) {
    case 0:   // 1 Mbp, 20k x 100 bp, 1 RG, 2% indels 1-5 bp, no clips, 1k sites
        gs->num_contigs = 1; gs->contig_length[0] = 1000000; gs->num_sites = 1000;
        rs->num_lengths = 1; rs->lengths[0] = 100;
        rs->num_read_groups = 1; rs->rg_mismatch[0] = 0.01;
        rs->indel_read_frac = 0.02; rs->indel_min = 1; rs->indel_max = 5;
        *num_reads = 20000;
        return 0;
    case 1:   // 30 Mbp / 3 contigs, 10M x 101 bp, 4 RGs, 3% indels, 5% clips 0-10, 30k sites
        gs->num_contigs = 3;
        for (int c = 0; c < 3; c++) gs->contig_length[c] = 10000000;
        gs->num_sites = 30000;
        rs->num_lengths = 1; rs->lengths[0] = 101;
        rs->num_read_groups = 4;
        for (int r = 0; r < 4; r++) rs->rg_mismatch[r] = 0.005 * (r + 1);
        rs->indel_read_frac = 0.03; rs->indel_min = 1; rs->indel_max = 5;
        rs->clip_read_frac = 0.05; rs->clip_min = 0; rs->clip_max = 10;
        *num_reads = 10000000;
        return 0;
    case 2:   // mixed lengths, 20M reads, 8 RGs, 5% indels 1-10, clips 0-30 on all reads
        gs->num_contigs = 3;
        for (int c = 0; c < 3; c++) gs->contig_length[c] = 10000000;
        gs->num_sites = 30000;
        rs->num_lengths = 5;
        rs->lengths[0] = 50; rs->lengths[1] = 76; rs->lengths[2] = 101; rs->lengths[3] = 151; rs->lengths[4] = 250;
        rs->num_read_groups = 8;
        for (int r = 0; r < 8; r++) rs->rg_mismatch[r] = 0.005 + 0.015 * r / 7.0;
        rs->indel_read_frac = 0.05; rs->indel_min = 1; rs->indel_max = 10;
        rs->clip_read_frac = 1.0; rs->clip_min = 0; rs->clip_max = 30;
        *num_reads = 20000000;
        return 0;
    case 3:   // serving job template (job size and RG count are drawn per job)
        gs->num_contigs = 3;
        for (int c = 0; c < 3; c++) gs->contig_length[c] = 10000000;
        gs->num_sites = 30000;
        rs->num_lengths = 2; rs->lengths[0] = 100; rs->lengths[1] = 150;
        rs->num_read_groups = 4;
        for (int r = 0; r < 4; r++) rs->rg_mismatch[r] = 0.005 * (r + 1);
        rs->indel_read_frac = 0.03; rs->indel_min = 1; rs->indel_max = 5;
        rs->clip_read_frac = 0.05; rs->clip_min = 0; rs->clip_max = 10;
        *num_reads = 64 * 1024;
        return 0;
    case 4:   // 3 Gbp / 24 contigs, 600M PE 150 bp, 16 RGs
        gs->num_contigs = 24;
        for (int c = 0; c < 24; c++) gs->contig_length[c] = 125000000;
        gs->num_sites = 1000000;
        rs->num_lengths = 1; rs->lengths[0] = 150;
        rs->num_read_groups = 16;
        for (int r = 0; r < 16; r++) rs->rg_mismatch[r] = 0.005 + 0.015 * r / 15.0;
        rs->indel_read_frac = 0.04; rs->indel_min = 1; rs->indel_max = 10;
        rs->clip_read_frac = 0.5; rs->clip_min = 0; rs->clip_max = 40;
        rs->q2_base_frac = 0.02;
        rs->paired = 1; rs->insert_mean = 350; rs->insert_sd = 50;
        *num_reads = 600000000ull;
        return 0;
    default:
        return -1;
    }
}

} // extern "C"
