/*
 * K_prep: per-read preparation, one wave per 64 reads.
 *
 * Covers, for each read, what the reference spreads over ~40 launches
 * (device/read_filters.cu, snp_filter.cu, cigar.cu, baq.cu:89-204):
 *   read filters -> alignment window -> malformed-read filter ->
 *   CIGAR event walk (windows, adapter and soft-clip clipping) ->
 *   error bits + num_errors -> per-base masks -> known-site mask ->
 *   HMM reference window + band width.
 *
 * CIGAR events are never materialised: an EventCursor walks (op, offset)
 * pairs and yields each event's class, read coordinate and reference
 * coordinate (the values cigar_coordinates_expand would have stored,
 * cigar.cu:538-609).  The CIGAR walks run one lane per read; the per-base
 * flags of the wave's reads are built in LDS (lanes over one read's bases at a
 * time, coalesced loads) and written to HBM once, as whole words.
 */
#include "fp_device.h"

#include <algorithm>

namespace fpd {

// the event sequence of one read
struct EventCursor {
    const uint32_t *ops;
    uint32_t nops;
    int32_t oi;        // current op index
    uint32_t k;        // offset inside the op
    uint16_t ro0, fo0; // read / reference coordinate at the op's first event
    uint32_t ev0;      // event index (relative to the read) of the op's first event

    __device__ uint32_t code() const { return op_code(ops[oi]); }
    __device__ uint32_t len() const { return op_len(ops[oi]); }
    __device__ uint32_t type() const { return op_event(code()); }
    __device__ uint16_t rc() const { return op_reads(code()) ? (uint16_t)(ro0 + k) : NONE16; }
    __device__ uint16_t fc() const { return op_refs(code()) ? (uint16_t)(fo0 + k) : NONE16; }
    __device__ uint32_t ev() const { return ev0 + k; }

    // position on the first event; false if the read has no events
    __device__ bool first(const uint32_t *o, uint32_t n)
    {
        ops = o; nops = n; oi = 0; k = 0; ro0 = 0; fo0 = 0; ev0 = 0;
        while (oi < (int32_t)nops && op_len(ops[oi]) == 0) oi++;
        return oi < (int32_t)nops;
    }
    // position on the last event
    __device__ bool last(const uint32_t *o, uint32_t n)
    {
        if (!first(o, n)) return false;
        int32_t lastop = -1;
        for (int32_t j = 0; j < (int32_t)n; j++) if (op_len(o[j])) lastop = j;
        while (oi < lastop) advance_op();
        k = len() - 1;
        return true;
    }
    __device__ void advance_op()
    {
        const uint32_t c = code(), l = len();
        if (op_reads(c)) ro0 = (uint16_t)(ro0 + l);
        if (op_refs(c)) fo0 = (uint16_t)(fo0 + l);
        ev0 += l;
        oi++;
        while (oi < (int32_t)nops && op_len(ops[oi]) == 0) oi++;
        k = 0;
    }
    // step to the next event; false (cursor unchanged) at the last event
    __device__ bool next()
    {
        if (k + 1 < len()) { k++; return true; }
        int32_t j = oi + 1;
        while (j < (int32_t)nops && op_len(ops[j]) == 0) j++;
        if (j >= (int32_t)nops) return false;
        advance_op();
        return true;
    }
    // step forward to the first event at or after the cursor whose reference coordinate is at least
    // fs (events without one skipped), a whole op at a time; false (cursor on the last op) if none
    __device__ bool seek_ref(int fs)
    {
        for (;;) {
            const uint32_t c = code(), l = len();
            if (op_refs(c) && (int)fo0 + (int)l - 1 >= fs) {
                const int kk = fs - (int)fo0;
                if (kk > (int)k) k = (uint32_t)kk;
                return true;
            }
            int32_t j = oi + 1;
            while (j < (int32_t)nops && op_len(ops[j]) == 0) j++;
            if (j >= (int32_t)nops) return false;
            advance_op();
        }
    }
    // step to the previous event; false (cursor unchanged) at the first event
    __device__ bool prev()
    {
        if (k > 0) { k--; return true; }
        int32_t j = oi - 1;
        while (j >= 0 && op_len(ops[j]) == 0) j--;
        if (j < 0) return false;
        const uint32_t c = op_code(ops[j]), l = op_len(ops[j]);
        if (op_reads(c)) ro0 = (uint16_t)(ro0 - l);
        if (op_refs(c)) fo0 = (uint16_t)(fo0 - l);
        ev0 -= l;
        oi = j;
        k = l - 1;
        return true;
    }
};

// the read-coordinate event: the event carrying read coordinate x, if any
__device__ bool seek_read_coord(EventCursor &c, const uint32_t *ops, uint32_t nops, uint16_t x)
{
    if (!c.first(ops, nops)) return false;
    for (;;) {
        const uint32_t cd = c.code();
        if (op_reads(cd)) {
            const uint16_t lo = c.ro0;
            const uint32_t l = c.len();
            if ((uint16_t)(x - lo) < l) { c.k = (uint16_t)(x - lo); return true; }
        }
        int32_t j = c.oi + 1;
        while (j < (int32_t)nops && op_len(ops[j]) == 0) j++;
        if (j >= (int32_t)nops) return false;
        c.advance_op();
    }
}

struct PrepArgs {
    BatchDev b;
    GenomeDev g;
    ReadStateDev s;
    uint32_t *batch;      // the batch's counter words (BW_ACTIVE .. BW_BAD_RG)
    uint32_t num_rg;      // read groups of the tables (active reads beyond them are counted as errors)
};

// filter_mapq + filter_if_any_set + filter_if_cigar_malformed (read_filters.cu:43-258),
// plus the query-length / op-code preconditions shared with the oracle
__device__ bool passes_invalid_filters(const BatchDev &b, uint32_t r)
{
    const uint8_t mq = b.mapq[r];
    if (mq == 0 || mq == 255) return false;
    if (b.flags[r] & (F_DUPLICATE | F_QC_FAIL | F_UNMAP | F_SECONDARY)) return false;
    // CIGAR clip layout must match H* S* (X+ S* H*)?, no N
    const uint32_t *ops = b.cigars + b.cigar_start[r];
    const uint32_t n = b.cigar_len[r];
    int st = 0;   // 0 lead-hard, 1 lead-soft, 2 body, 3 trail-soft, 4 trail-hard
    uint64_t qlen = 0;
    for (uint32_t i = 0; i < n; i++) {
        const uint32_t op = op_code(ops[i]);
        if (op == OP_N) return false;
        if (op > OP_X) return false;
        if (op == OP_H) {
            if (st == 1) return false;
            if (st == 2 || st == 3) st = 4;
        } else if (op == OP_S) {
            if (st == 0) st = 1;
            else if (st == 2) st = 3;
            else if (st == 4) return false;
        } else {
            if (st <= 1) st = 2;
            else if (st >= 3) return false;
        }
        if (op_reads(op)) qlen += op_len(ops[i]);
    }
    return qlen == b.read_len[r];
}

// per-read windows (cigar.cu:98-535); meaningful for reads that pass the filters
struct Windows {
    uint16_t rwx, rwy;   // read window, clipped
    uint16_t nix, niy;   // read window without leading / trailing insertions
    uint16_t fwx, fwy;   // reference window
};

// filter_invalid_reads + filter_malformed_reads (read_filters.cu:43-154); false: the read is filtered
__device__ __forceinline__ bool prep_filter(const PrepArgs &a, uint32_t r)
{
    const BatchDev &b = a.b;
    const uint32_t L = b.read_len[r];
    const uint32_t *ops = b.cigars + b.cigar_start[r];
    const uint32_t nops = b.cigar_len[r];
    const uint16_t flags = b.flags[r];
    const uint32_t astart = b.alignment_start[r];
    bool ok = passes_invalid_filters(b, r);

    // ---- alignment window (snp_filter.cu:44-147): offset of the last aligned base
    uint32_t wx = 0xffffffffu, wy = 0xffffffffu;
    if (ok) {
        uint16_t off = 0;
        bool have = false;
        uint16_t last = 0;
        for (uint32_t i = 0; i < nops; i++) {
            const uint32_t op = op_code(ops[i]), l = op_len(ops[i]);
            if (op == OP_M || op == OP_EQ || op == OP_X) {
                if (l) { last = (uint16_t)(off + l - 1); have = true; }
                off = (uint16_t)(off + l);
            } else if (op == OP_I || op == OP_N) {
                if (l) { last = off; have = true; }
            } else if (op == OP_S) {
                // soft-clipped bases carry no offset; the scan skips them
            } else {
                off = (uint16_t)(off + l);
            }
        }
        if (have) { wx = astart; wy = astart + last; }
    }
    // ---- filter_malformed_reads (read_filters.cu:78-154)
    if (ok) {
        if (!(flags & F_UNMAP)) {
            if (astart == 0xffffffffu) ok = false;
            else if (wy <= wx) ok = false;
            else if (b.chromosome[r] >= a.g.num_contigs) ok = false;
            else if (wx >= a.g.length[b.chromosome[r]]) ok = false;
            else if (nops == 0) ok = false;
        }
        if (b.read_group[r] == 0xffffffffu) ok = false;
        if (b.qual_len[r] != L) ok = false;
        if (L == 0) ok = false;
    }
    return ok;
}

__device__ __forceinline__ Windows prep_windows(const BatchDev &b, uint32_t r)
{
    const uint32_t L = b.read_len[r];
    const uint32_t *ops = b.cigars + b.cigar_start[r];
    const uint32_t nops = b.cigar_len[r];
    const uint16_t flags = b.flags[r];
    const uint32_t astart = b.alignment_start[r];

    // ---- read window + adapter clipping (cigar.cu:98-326)
    uint16_t rwx = 0, rwy = (uint16_t)(L - 1);
    EventCursor cur;
    {
        // getAdaptorBoundary
        uint32_t boundary = 0xffffffffu;
        const int32_t isz = b.isize[r];
        bool well = true;
        if (isz == 0) well = false;
        else if (!(flags & F_PAIRED)) well = false;
        else if ((flags & F_UNMAP) || (flags & F_MATE_UNMAP)) well = false;
        else if (((flags & F_REVERSE) != 0) == ((flags & F_MATE_REVERSE) != 0)) well = false;
        else if (flags & F_REVERSE) well = b.alignment_stop[r] > b.mate_alignment_start[r];
        else well = astart <= b.mate_alignment_start[r] + isz;
        if (well) {
            if (flags & F_REVERSE) boundary = b.mate_alignment_start[r] - 1;
            else boundary = astart + (uint32_t)(isz < 0 ? -isz : isz) + 1;
        }
        if (boundary != 0xffffffffu) {
            // current reference window of [rwx, rwy]
            uint16_t cx = 0, cy = 0;
            bool invalid = false;
            EventCursor c;
            if (seek_read_coord(c, ops, nops, rwx)) {
                bool found = true;
                while (c.fc() == NONE16) { if (!c.next()) { found = false; break; } }
                if (!found) invalid = true;
                else cx = c.fc();
            }
            if (!invalid && seek_read_coord(c, ops, nops, rwy)) {
                while (c.fc() == NONE16 && c.ev() > 0) c.prev();
                if (c.ev() == 0) invalid = true;
                else cy = c.fc();
            }
            if (invalid) { cx = NONE16; cy = NONE16; }
            if (!(boundary < astart + cx || boundary > astart + cy)) {
                // getReadCoordinateForReferenceCoordinate
                const bool left = (flags & F_REVERSE) != 0;
                int coord = -1;
                if (c.first(ops, nops)) {
                    for (;;) {
                        if (astart + (uint32_t)c.fc() == boundary) {
                            uint16_t rc = c.rc();
                            if (rc == NONE16) {
                                if (!left) {
                                    for (;;) {
                                        rc = c.rc();
                                        if (rc != NONE16) break;
                                        if (!c.next()) break;
                                    }
                                } else {
                                    for (;;) {
                                        if (!c.prev()) { rc = NONE16; break; }
                                        rc = c.rc();
                                        if (rc != NONE16) break;
                                    }
                                }
                            }
                            coord = rc == NONE16 ? -1 : rc;
                            break;
                        }
                        if (!c.next()) break;
                    }
                }
                if (left) {
                    const uint16_t stop = (uint16_t)coord;
                    const uint16_t v = (uint16_t)(stop + 1);
                    rwx = rwx > v ? rwx : v;
                } else {
                    const uint16_t start = (uint16_t)coord;
                    const uint16_t v = (uint16_t)(start - 1);
                    rwy = rwy < v ? rwy : v;
                }
            }
        }
    }
    // ---- remove_soft_clips (cigar.cu:329-389)
    {
        uint32_t ci = 0, ro = 0;
        while (ci < nops && ro < L) {
            const uint32_t op = op_code(ops[ci]), l = op_len(ops[ci]);
            if (op == OP_H || op == OP_S) {
                if (ro + l > rwx) { const uint16_t v = (uint16_t)(ro + l); rwx = v < rwy ? v : rwy; }
                ro += l; ci++;
            } else break;
        }
        if (nops > 0) {
            ci = nops - 1;
            ro = L - 1;
            while (ro > rwx) {
                const uint32_t op = op_code(ops[ci]), l = op_len(ops[ci]);
                if (op == OP_H || op == OP_S) {
                    if (ro - l < rwy) { const uint16_t v = (uint16_t)(ro - l); rwy = v > rwx ? v : rwx; }
                    ro -= l;
                    if (ci == 0) break;
                    ci--;
                } else break;
            }
        }
    }
    // ---- no-insertions window (cigar.cu:392-451)
    // the first event with a read coordinate of at least rwx that is not an insertion, and the last
    // with one of at most rwy: inside an op the coordinate moves by one per event, so whole ops are
    // passed over and the event is found by arithmetic in the op that holds it
    uint16_t nix = 0, niy = 0;
    if (cur.first(ops, nops)) {
        for (;;) {
            const uint32_t cd = cur.code(), l = cur.len();
            if (op_reads(cd) && op_event(cd) != EV_I && (int)cur.ro0 + (int)l - 1 >= (int)rwx) {
                const int kk = (int)rwx - (int)cur.ro0;
                nix = (uint16_t)(cur.ro0 + (kk > 0 ? (uint32_t)kk : 0u));
                break;
            }
            int32_t j = cur.oi + 1;
            while (j < (int32_t)nops && op_len(ops[j]) == 0) j++;
            if (j >= (int32_t)nops) break;
            cur.advance_op();
        }
    }
    if (cur.last(ops, nops)) {
        for (;;) {
            const uint32_t cd = cur.code();
            if (op_reads(cd) && op_event(cd) != EV_I && (int)cur.ro0 <= (int)rwy) {
                const int kk = (int)rwy - (int)cur.ro0;
                niy = (uint16_t)(cur.ro0 + (kk < (int)cur.k ? (uint32_t)kk : cur.k));
                break;
            }
            cur.k = 0;   // prev() from the op's first event moves to the previous op's last
            if (!cur.prev()) break;
        }
    }
    // ---- reference window (cigar.cu:454-535)
    uint16_t fwx = 0, fwy = 0;
    {
        bool invalid = false;
        if (seek_read_coord(cur, ops, nops, nix)) {
            bool found = true;
            while (cur.fc() == NONE16) { if (!cur.next()) { found = false; break; } }
            if (!found) { invalid = true; fwx = NONE16; fwy = NONE16; }
            else fwx = cur.fc();
        }
        if (!invalid && seek_read_coord(cur, ops, nops, niy)) {
            while (cur.fc() == NONE16 && cur.ev() > 0) cur.prev();
            if (cur.ev() == 0) { fwx = NONE16; fwy = NONE16; }
            else fwy = cur.fc();
        }
    }
    return Windows{rwx, rwy, nix, niy, fwx, fwy};
}

// error vectors (cigar.cu:612-733) of one read by one lane: orf(read offset, flag) ORs a flag in;
// returns num_errors
template <class OrFlag>
__device__ uint16_t error_walk_lane(const PrepArgs &a, uint32_t r, const Windows &w, OrFlag orf)
{
    const BatchDev &b = a.b;
    const uint32_t L = b.read_len[r];
    const uint32_t rs = b.read_start[r];
    const uint32_t *ops = b.cigars + b.cigar_start[r];
    const uint32_t nops = b.cigar_len[r];
    const uint32_t astart = b.alignment_start[r];
    const uint32_t chrom = b.chromosome[r];
    const bool neg = (b.flags[r] & F_REVERSE) != 0;
    const uint16_t rwx = w.rwx, rwy = w.rwy, fwx = w.fwx, fwy = w.fwy;
    uint16_t nerr = 0;
    {
        uint16_t bp = 0;      // current_bp_idx
        uint16_t ro = 0, fo = 0;
        for (uint32_t i = 0; i < nops; i++) {
            const uint32_t op = op_code(ops[i]), l = op_len(ops[i]);
            if (op == OP_M) {
                for (uint32_t k = 0; k < l; k++) {
                    bp = (uint16_t)(ro + k);
                    const uint8_t rb = read_base(b.reads, (uint64_t)rs + bp);
                    const uint8_t fb = contig_base(a.g, chrom, astart + (uint16_t)(fo + k));
                    if (fb != rb) {
                        orf(bp, BF_SNP);
                        if (bp >= rwx && bp <= rwy) nerr++;
                    }
                }
            } else if (op == OP_I) {
                if (l) {
                    bp = ro;
                    if (bp >= rwx && bp <= rwy) {
                        const int off = neg ? (int)bp + (int)l : (int)bp - 1;
                        if (off >= 0 && off <= rwy) { orf((uint32_t)off, BF_INS); nerr++; }
                    }
                }
            } else if (op == OP_D) {
                if (l) {
                    const uint16_t cri = fo;
                    if (cri >= fwx && cri <= fwy) {
                        if (!neg) { orf(bp, BF_DEL); nerr++; }
                        else {
                            const uint16_t off = (uint16_t)(bp + 1);
                            if (off < L) { orf(off, BF_DEL); nerr++; }
                        }
                    }
                }
            }
            if (op_reads(op)) ro = (uint16_t)(ro + l);
            if (op_refs(op)) fo = (uint16_t)(fo + l);
        }
    }
    return nerr;
}

// known-site masking (snp_filter.cu:162-395): mask(d0, d1) masks read offsets d0..d1 (inside the
// read window)
template <class Mask>
__device__ __forceinline__ void known_sites(const PrepArgs &a, uint32_t r, const Windows &w, Mask mask)
{
    const BatchDev &b = a.b;
    const uint32_t *ops = b.cigars + b.cigar_start[r];
    const uint32_t nops = b.cigar_len[r];
    const uint32_t astart = b.alignment_start[r];
    const uint32_t chrom = b.chromosome[r];
    const uint16_t rwx = w.rwx, rwy = w.rwy, fwx = w.fwx, fwy = w.fwy;
    // ---- known sites (snp_filter.cu:162-395)
    if (chrom < a.g.num_variant_contigs) {
        const uint64_t f0 = a.g.feature_offset[chrom];
        const uint32_t nf = (uint32_t)(a.g.feature_offset[chrom + 1] - f0);
        const uint32_t *fstart = a.g.feature_start + f0;
        const uint32_t *fstop = a.g.feature_stop + f0;
        const uint32_t *mel = a.g.max_end_left + f0;
        const uint32_t qx = astart + fwx, qy = astart + fwy;
        // the contig's bucket index brackets both searches: the bound for a position in bucket j lies
        // between the bucket's entry and the next one (the last entry stands for +infinity)
        const uint64_t b0 = a.g.bucket_offset[chrom];
        const uint32_t last = (uint32_t)(a.g.bucket_offset[chrom + 1] - b0) - 2;
        const uint32_t jx = qx >> SITE_BUCKET_BITS < last ? qx >> SITE_BUCKET_BITS : last;
        const uint32_t jy = qy >> SITE_BUCKET_BITS < last ? qy >> SITE_BUCKET_BITS : last;
        // lower_bound over max_end_point_left (primitives/algorithms.h:39-70)
        uint32_t vx;
        {
            uint32_t lo = a.g.site_lb[b0 + jx], cnt = a.g.site_lb[b0 + jx + 1] - lo;
            while (cnt > 0) {
                const uint32_t h = cnt / 2;
                if (mel[lo + h] < qx) { lo = lo + h + 1; cnt -= h + 1; }
                else cnt = h;
            }
            vx = lo;
        }
        if (vx < nf) {
            // upper_bound over feature_start (primitives/algorithms.h:77-100)
            uint32_t lo = a.g.site_sb[b0 + jy], cnt = a.g.site_sb[b0 + jy + 1] - lo;
            while (cnt > 0) {
                const uint32_t h = cnt / 2;
                if (!(qy < fstart[lo + h])) { lo = lo + h + 1; cnt -= h + 1; }
                else cnt = h;
            }
            uint32_t vy = lo;
            if (vy > vx) {
                if (vy >= nf) vy = nf - 1;
                for (uint32_t f = vx; f <= vy; f++) {
                    const int fs = (int)(fstart[f] - astart);
                    const int fe = (int)(fstop[f] - astart);
                    EventCursor ev;
                    // the first event whose reference coordinate reaches the feature
                    if (!ev.first(ops, nops) || !ev.seek_ref(fs)) continue;
                    const uint16_t refc = ev.fc();
                    int left = (int)refc - fs;
                    EventCursor es = ev;
                    bool es_end = false;   // es one past the last event
                    if (left > 0) {
                        while (left && es.ev() > 0) {
                            const uint32_t t = es.type();
                            if (t == EV_S) break;
                            if (t == EV_M || t == EV_D) left--;
                            es.prev();
                        }
                        for (;;) {
                            if (es.ev() == 0) break;
                            EventCursor p = es;
                            p.prev();
                            if (p.type() != EV_I) break;
                            es = p;
                        }
                        while (es.rc() == NONE16) {
                            if (!es.next()) { es_end = true; break; }
                        }
                    } else {
                        if (es.fc() <= fwy) {
                            while (es.ev() > 0 && es.type() == EV_D) es.prev();
                        }
                    }
                    int right = fe - (int)refc;
                    while (right > 0) {
                        if (!ev.next()) break;
                        const uint32_t t = ev.type();
                        if (t == EV_S) break;
                        if (t == EV_M || t == EV_D) right--;
                    }
                    while (ev.ev() > 0 && ev.rc() == NONE16) ev.prev();
                    if (es_end || ev.ev() < es.ev()) continue;
                    uint32_t r0 = es.rc(), r1 = ev.rc();
                    if ((r0 < rwx && r1 < rwx) || (r0 > rwy && r1 > rwy)) continue;
                    r0 = r0 > rwx ? r0 : rwx; r0 = r0 < rwy ? r0 : rwy;
                    r1 = r1 > rwx ? r1 : rwx; r1 = r1 < rwy ? r1 : rwy;
                    mask(r0, r1);
                }
            }
        }
    }

}

// HMM reference window + band width (baq.cu:89-187); RS_ACTIVE, or RS_BAQ_DROPPED when the window
// runs past the contig
__device__ __forceinline__ uint8_t hmm_window(const PrepArgs &a, uint32_t r, const Windows &w, uint16_t nerr)
{
    const BatchDev &b = a.b;
    const ReadStateDev &s = a.s;
    const uint32_t L = b.read_len[r];
    const uint32_t astart = b.alignment_start[r];
    const uint32_t chrom = b.chromosome[r];
    const uint16_t rwx = w.rwx, rwy = w.rwy, nix = w.nix, niy = w.niy, fwx = w.fwx, fwy = w.fwy;
    // ---- HMM window + band width (baq.cu:89-187)
    {
        const bool needs = nerr != 0;
        const bool lclip = rwx != 0;
        const bool rclip = rwy != L - 1;
        const int16_t lins = (int16_t)(lclip ? 0 : (int)nix - (int)rwx);
        const int16_t rins = (int16_t)(rclip ? 0 : (int)rwy - (int)niy);
        const int16_t hx = (int16_t)((int)fwx - lins - 3);
        const int16_t hy = (int16_t)((int)fwy + rins + 3);
        int lshift = 0;
        const uint32_t ahx = hx < 0 ? (uint32_t)(-(int)hx) : (uint32_t)hx;
        if (hx < 0 && astart < ahx) lshift = -1;
        if (needs && (uint32_t)(astart + (int)hy + lshift) >= a.g.length[chrom]) {
            s.hmmw[r] = pack2((uint16_t)-1, (uint16_t)-1);
            s.bandwidth[r] = 0;
            return RS_BAQ_DROPPED;
        }
        s.hmmw[r] = pack2((uint16_t)hx, (uint16_t)hy);
        const int refLen = (int)hy - (int)hx + 1;
        const int qLen = (int)rwy - (int)rwx + 1;
        const int d = refLen - qLen;
        const uint32_t ad = d < 0 ? (uint32_t)(-d) : (uint32_t)d;
        uint16_t bw = (uint16_t)(refLen > qLen ? refLen : qLen);
        if (7u < ad) bw = (uint16_t)(ad + 3);
        if (bw > 7) bw = 7;
        if (bw < ad) bw = (uint16_t)ad;
        s.bandwidth[r] = bw;
    }
    return RS_ACTIVE;
}

// One read by one lane, every per-base flag read-modify-written in HBM: the path of waves that hold
// a read longer than the LDS rows.  Returns the read's status.
__device__ __forceinline__ uint8_t prep_read_lane(const PrepArgs &a, uint32_t r, uint16_t &nerr)
{
    const BatchDev &b = a.b;
    const ReadStateDev &s = a.s;
    const uint32_t L = b.read_len[r];
    const uint32_t rs = b.read_start[r];
    const uint32_t qs = b.qual_start[r];
    const uint32_t padded = (L + 7) & ~7u;
    uint8_t *fl = s.base_flags + rs;
    s.rwc[r] = 0; s.noins[r] = 0; s.refw[r] = 0; s.num_errors[r] = 0;
    s.hmmw[r] = 0; s.bandwidth[r] = 0; s.hqw[r] = 0;
    for (uint32_t i = 0; i < padded; i++) fl[i] = 0;
    nerr = 0;
    if (!prep_filter(a, r)) { s.status[r] = RS_FILTERED; return RS_FILTERED; }
    const Windows w = prep_windows(b, r);
    s.rwc[r] = pack2(w.rwx, w.rwy);
    s.noins[r] = pack2(w.nix, w.niy);
    s.refw[r] = pack2(w.fwx, w.fwy);
    nerr = error_walk_lane(a, r, w, [&](uint32_t i, uint8_t v) { fl[i] |= v; });
    s.num_errors[r] = nerr;
    // per-base filters (read_filters.cu:376-438)
    for (uint32_t i = 0; i < L; i++) {
        const uint8_t bpv = read_base(b.reads, (uint64_t)rs + i);
        const bool regular = bpv == B_A || bpv == B_C || bpv == B_G || bpv == B_T;
        if (!regular || b.quals[qs + i] < 6) fl[i] |= BF_MASKED;
    }
    known_sites(a, r, w, [&](uint32_t d0, uint32_t d1) { for (uint32_t d = d0; d <= d1; d++) fl[d] |= BF_MASKED; });
    const uint8_t st = hmm_window(a, r, w, nerr);
    s.status[r] = st;
    return st;
}

__device__ inline uint32_t rl(uint32_t v, int j) { return (uint32_t)__builtin_amdgcn_readlane((int)v, j); }

// one read of the per-base pass, its fields wave-uniform
struct PrepRead {
    uint32_t rs, qs, ast, clen, cwo, nops, rwx, rwy, fwx, fwy, L, pj;
    bool neg;
    const uint32_t *ops;
};

// the per-base loads of one read: bases, qualities and the CIGAR ops, one op per lane
template <int NC>
struct BaseLoads {
    uint32_t rb[NC], q[NC], op;
};

template <int NC>
__device__ inline BaseLoads<NC> load_bases(const BatchDev &b, const PrepRead &p)
{
    const uint32_t lane = threadIdx.x & 63;
    BaseLoads<NC> x;
    x.op = lane < p.nops ? p.ops[lane] : 0u;
#pragma unroll
    for (int c = 0; c < NC; c++) {
        const uint32_t i = lane + 64 * c;
        x.rb[c] = 0;
        x.q[c] = 255;
        if (i < p.L) {
            x.rb[c] = (b.reads[(p.rs + i) >> 3] >> (((p.rs + i) & 7) * 4)) & 15;
            x.q[c] = b.quals[p.qs + i];
        }
    }
    return x;
}

// The error walk of one read with the lanes' bases side by side (positions lane + 64c): the
// current_bp_idx (bp) and the read / reference offsets are the same in every lane, and the walk
// only reads registers.  Sets the mask, insertion and deletion flags and each aligned base's
// reference offset (rf, -1 when not aligned); returns the insertions and deletions counted.
template <int NC>
__device__ uint32_t error_walk(const PrepRead &p, const BaseLoads<NC> &x, uint8_t *f, int32_t *rf)
{
    const uint32_t lane = threadIdx.x & 63;
#pragma unroll
    for (int c = 0; c < NC; c++) {
        const uint32_t i = lane + 64 * c;
        f[c] = 0;
        rf[c] = -1;
        if (i < p.L) {
            const uint32_t rb = x.rb[c];
            const bool regular = rb == B_A || rb == B_C || rb == B_G || rb == B_T;
            if (!regular || x.q[c] < 6) f[c] |= BF_MASKED;
        }
    }
    uint32_t cnt = 0;
    uint16_t ro = 0, fo = 0, bp = 0;
    for (uint32_t oi = 0; oi < p.nops; oi++) {
        const uint32_t wop = oi < 64 ? rl(x.op, (int)oi) : p.ops[oi];
        const uint32_t op = op_code(wop), l = op_len(wop);
        if (op == OP_M) {
#pragma unroll
            for (int c = 0; c < NC; c++) {
                const uint32_t i = lane + 64 * c;
                if (i >= ro && i < (uint32_t)ro + l) rf[c] = (uint16_t)(fo + (i - ro));
            }
            if (l) bp = (uint16_t)(ro + l - 1);
        } else if (op == OP_I) {
            if (l) {
                bp = ro;
                if (bp >= p.rwx && bp <= p.rwy) {
                    const int off = p.neg ? (int)bp + (int)l : (int)bp - 1;
                    if (off >= 0 && off <= (int)p.rwy) {
#pragma unroll
                        for (int c = 0; c < NC; c++)
                            if ((int)(lane + 64 * c) == off) f[c] |= BF_INS;
                        cnt++;
                    }
                }
            }
        } else if (op == OP_D) {
            if (l) {
                const uint16_t cri = fo;
                if (cri >= p.fwx && cri <= p.fwy) {
                    if (!p.neg) {
#pragma unroll
                        for (int c = 0; c < NC; c++)
                            if (lane + 64 * c == bp) f[c] |= BF_DEL;
                        cnt++;
                    } else {
                        const uint16_t off = (uint16_t)(bp + 1);
                        if (off < p.L) {
#pragma unroll
                            for (int c = 0; c < NC; c++)
                                if (lane + 64 * c == off) f[c] |= BF_DEL;
                            cnt++;
                        }
                    }
                }
            }
        }
        if (op_reads(op)) ro = (uint16_t)(ro + l);
        if (op_refs(op)) fo = (uint16_t)(fo + l);
    }
    return cnt;
}

// The per-base pass over the wave's reads that passed the filters (read_filters.cu:376-438,
// cigar.cu:612-733), one read at a time across the wave: the read's flags, with the known-site
// ranges of its lane (ks_n of them, up to two packed as d0 | d1 << 16 in ks0 and ks1) masked in,
// written once to HBM (padded to a whole number of 8-byte words), and its error count (SNPs inside
// the read window plus the walk's insertions and deletions) into nerr of the read's lane.
// Software-pipelined: the next read's bases, qualities and ops are in flight while the current
// read's reference words return and are compared, so a read costs about one round trip to memory.
template <int NC, class ReadOf>
__device__ __forceinline__ void prep_pass(const PrepArgs &a, uint64_t okmask, ReadOf read_of, uint32_t ks_n, uint32_t ks0,
                          uint32_t ks1, uint16_t &nerr)
{
    const uint32_t lane = threadIdx.x & 63;
    uint64_t m = okmask;
    if (!m) return;
    int j = __builtin_ctzll(m);
    PrepRead p = read_of(j);
    BaseLoads<NC> cur = load_bases<NC>(a.b, p);
    for (;;) {
        uint8_t f[NC];
        int32_t rf[NC];
        uint32_t cnt = error_walk<NC>(p, cur, f, rf);
        // the reference words of the aligned bases (past the contig: N)
        uint32_t fw[NC];
#pragma unroll
        for (int c = 0; c < NC; c++) {
            fw[c] = 0;
            const uint32_t pos = p.ast + (uint32_t)rf[c];
            if (rf[c] >= 0 && pos < p.clen) fw[c] = a.g.words[p.cwo + (pos >> 3)];
        }
        m &= m - 1;
        const bool more = m != 0;
        const int jn = more ? __builtin_ctzll(m) : j;
        const PrepRead pn = read_of(jn);
        const BaseLoads<NC> nxt = load_bases<NC>(a.b, pn);   // in flight during the comparisons
        const uint32_t kn = rl(ks_n, j), k0 = rl(ks0, j), k1 = rl(ks1, j);
        uint8_t *row = a.s.base_flags + p.rs;
        const bool words = (p.rs & 3) == 0;   // 4-byte stores of 4 lanes' flags, else one byte per lane
#pragma unroll
        for (int c = 0; c < NC; c++) {
            const uint32_t i = lane + 64 * c;
            bool snp_in = false;
            if (rf[c] >= 0) {
                const uint32_t pos = p.ast + (uint32_t)rf[c];
                const uint32_t fb = pos < p.clen ? (fw[c] >> ((pos & 7) * 4)) & 15 : (uint32_t)B_N;
                if (fb != cur.rb[c]) {
                    f[c] |= BF_SNP;
                    snp_in = i >= p.rwx && i <= p.rwy;
                }
            }
            cnt += (uint32_t)__popcll(__ballot(snp_in));
            const bool ks = (kn > 0 && i >= lo16(k0) && i <= hi16(k0)) || (kn > 1 && i >= lo16(k1) && i <= hi16(k1));
            if (ks) f[c] |= BF_MASKED;
            // lanes 4k gather positions 4k .. 4k + 3 (pj is a multiple of 8, so a word below pj is whole)
            const uint32_t g2 = (uint32_t)f[c] | ((uint32_t)__shfl_down((int)f[c], 1) << 8);
            const uint32_t g4 = g2 | ((uint32_t)__shfl_down((int)g2, 2) << 16);
            if (words) {
                if ((lane & 3) == 0 && i < p.pj) *(uint32_t *)(row + i) = g4;
            } else if (i < p.pj) {
                row[i] = f[c];
            }
        }
        if ((int)lane == j) nerr = (uint16_t)cnt;
        if (!more) break;
        p = pn;
        cur = nxt;
        j = jn;
    }
}

// One wave per 64 reads.  The CIGAR walks (filters, windows, known sites, HMM window) run one lane
// per read; the per-base work (base masks, SNP / insertion / deletion flags) runs one read at a time
// across the wave, lanes over bases, so its loads and its flag stores are coalesced.  The known-site
// walk runs first and keeps each read's masked ranges in its lane (two of them; a read with more
// ORs the rest into its stored flags afterwards, the only read-modify-write in HBM), so a read's
// flags are complete when the pass writes them.  No LDS: the wave count per SIMD is set by the
// registers alone.  A wave that holds a read longer than 256 bases takes prep_read_lane for all its
// reads.
__global__ void __launch_bounds__(64) k_prep(PrepArgs a)
{
    const BatchDev &b = a.b;
    const ReadStateDev &s = a.s;
    const uint32_t lane = threadIdx.x;
    const uint32_t n = live_reads(b);
    if (blockIdx.x * 64u >= n) return;
    const uint32_t r = blockIdx.x * 64u + lane;
    const bool valid = r < n;
    const uint32_t L = valid ? b.read_len[r] : 0u;
    const uint32_t rs = valid ? b.read_start[r] : 0u;

    uint8_t st = RS_FILTERED;
    uint16_t nerr = 0;
    if (!wave_all(L <= 256)) {
        if (valid) st = prep_read_lane(a, r, nerr);
    } else {
        const bool ok = valid && prep_filter(a, r);
        Windows w{};
        if (ok) w = prep_windows(b, r);
        const uint32_t qs = valid ? b.qual_start[r] : 0u;
        const uint32_t cs = valid ? b.cigar_start[r] : 0u;
        const uint32_t nops = valid ? b.cigar_len[r] : 0u;
        const uint32_t astart = valid ? b.alignment_start[r] : 0u;
        const uint32_t chrom = valid ? b.chromosome[r] : 0u;
        const uint32_t neg = valid ? ((b.flags[r] & F_REVERSE) != 0) : 0u;
        const uint64_t vmask = __ballot(valid), okmask = __ballot(ok);

        // the contig of each lane's read: its length and first reference word, taken by readlane below
        const bool has_contig = valid && chrom < a.g.num_contigs;
        const uint32_t clen = has_contig ? a.g.length[chrom] : 0u;
        const uint32_t cwo = has_contig ? a.g.word_offset[chrom] : 0u;

        // flags of valid reads that the filters dropped are all zero
        for (uint64_t m = vmask & ~okmask; m; m &= m - 1) {
            const int j = __builtin_ctzll(m);
            const uint32_t pj = (rl(L, j) + 7) & ~7u;
            uint8_t *row = s.base_flags + rl(rs, j);
            for (uint32_t i = lane; i < pj; i += 64) row[i] = 0;
        }
        // known-site ranges of each lane's read, first (read offsets d0 | d1 << 16)
        uint32_t ks_n = 0, ks0 = 0, ks1 = 0;
        if (ok) {
            known_sites(a, r, w, [&](uint32_t d0, uint32_t d1) {
                const uint32_t v = d0 | (d1 << 16);
                ks1 = ks_n == 1 ? v : ks1;   // selects: an indexed store would go through scratch memory
                ks0 = ks_n == 0 ? v : ks0;
                ks_n++;
            });
        }
        auto read_of = [&](int j) {
            PrepRead p;
            p.rs = rl(rs, j); p.qs = rl(qs, j); p.ast = rl(astart, j);
            p.clen = rl(clen, j); p.cwo = rl(cwo, j);
            p.nops = rl(nops, j); p.neg = rl(neg, j) != 0;
            p.ops = b.cigars + rl(cs, j);
            p.rwx = rl(w.rwx, j); p.rwy = rl(w.rwy, j); p.fwx = rl(w.fwx, j); p.fwy = rl(w.fwy, j);
            p.L = rl(L, j); p.pj = (p.L + 7) & ~7u;
            return p;
        };
        if (!wave_any(ok && L > 128)) prep_pass<2>(a, okmask, read_of, ks_n, ks0, ks1, nerr);
        else prep_pass<4>(a, okmask, read_of, ks_n, ks0, ks1, nerr);

        // a read with more than two known-site ranges ORs the rest into its stored flags (atomically, on
        // the aligned words that hold them, after the wave's stores have completed)
        if (wave_any(ks_n > 2)) {
            __builtin_amdgcn_s_waitcnt(0);
            __threadfence();
            if (ks_n > 2) {
                uint32_t k = 0;
                known_sites(a, r, w, [&](uint32_t d0, uint32_t d1) {
                    if (k++ < 2) return;
                    for (uint32_t d = d0; d <= d1; d++) {
                        const uintptr_t addr = (uintptr_t)(s.base_flags + rs + d);
                        atomicOr((uint32_t *)(addr & ~(uintptr_t)3), (uint32_t)BF_MASKED << (8 * (addr & 3)));
                    }
                });
            }
        }
        if (ok) st = hmm_window(a, r, w, nerr);
        if (valid) {
            s.rwc[r] = ok ? pack2(w.rwx, w.rwy) : 0u;
            s.noins[r] = ok ? pack2(w.nix, w.niy) : 0u;
            s.refw[r] = ok ? pack2(w.fwx, w.fwy) : 0u;
            s.num_errors[r] = nerr;
            s.hqw[r] = 0;
            s.status[r] = st;
            if (!ok) { s.hmmw[r] = 0; s.bandwidth[r] = 0; }
        }
    }

    // counters: wave totals, one atomic each
    uint32_t mx = L;
    for (int o = 32; o > 0; o >>= 1) { const uint32_t v = (uint32_t)__shfl_xor((int)mx, o); mx = v > mx ? v : mx; }
    const uint64_t m_act = __ballot(valid && st == RS_ACTIVE);
    const uint64_t m_filt = __ballot(valid && st == RS_FILTERED);
    const uint64_t m_baq = __ballot(valid && st == RS_ACTIVE && nerr != 0);
    const uint64_t m_rg = __ballot(valid && st == RS_ACTIVE && b.read_group[valid ? r : 0] >= a.num_rg);
    if (lane == 0) {
        if (m_act) atomicAdd(&a.batch[BW_ACTIVE], (uint32_t)__popcll(m_act));
        if (m_filt) atomicAdd(&a.batch[BW_FILTERED], (uint32_t)__popcll(m_filt));
        if (m_baq) atomicAdd(&a.batch[BW_BAQ], (uint32_t)__popcll(m_baq));
        if (m_rg) atomicAdd(&a.batch[BW_BAD_RG], (uint32_t)__popcll(m_rg));
        if (mx) atomicMax(&a.batch[BW_LONGEST], mx);
    }
}

} // namespace fpd

// host launcher
extern "C" hipError_t fpd_launch_prep(const fpd::BatchDev *b, const fpd::GenomeDev *g,
                                      const fpd::ReadStateDev *s, uint32_t *batch, uint32_t num_rg,
                                      uint32_t max_read_len, hipStream_t st)
{
    fpd::PrepArgs a{};
    a.b = *b; a.g = *g; a.s = *s; a.batch = batch; a.num_rg = num_rg;
    const uint32_t n = b->num_reads;
    if (n == 0) return hipSuccess;
    (void)max_read_len;
    hipLaunchKernelGGL(fpd::k_prep, dim3((n + 63) / 64), dim3(64), 0, st, a);
    return hipGetLastError();
}

namespace fpd {
// append of a batch into a staging buffer: its start offsets become offsets into the staged arrays,
// and its read groups move to the owning job's table slots (ids at or beyond rg_limit become invalid,
// which the read-group check reports)
__global__ void __launch_bounds__(256) k_rebase(uint32_t n, uint32_t *cigar_start, uint32_t *read_start,
                                               uint32_t *qual_start, uint32_t *read_group, uint32_t cig_off,
                                               uint32_t base_off, uint32_t qual_off, uint32_t rg_off,
                                               uint32_t rg_limit)
{
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    cigar_start[r] += cig_off;
    read_start[r] += base_off;
    qual_start[r] += qual_off;
    if (rg_off || rg_limit != 0xffffffffu) {
        const uint32_t g = read_group[r];
        read_group[r] = g < rg_limit ? g + rg_off : 0xffffffffu;
    }
}
} // namespace fpd

namespace fpd {
// Append of a batch in one launch: the 18 batch arrays get runs of workgroups in proportion to their
// bytes (first[i] .. first[i + 1] - 1 copy array i); the three start-offset arrays and the read
// groups are rebased on the way (as k_rebase does), everything else is copied in 16-, 4- or 1-byte
// units as alignment allows.  Each lane keeps APPEND_UNROLL loads in flight: for a batch read from
// pinned host memory over the host link, a few hundred waves then keep the link busy, where one
// load per lane would need thousands of resident waves (which crowd out the kernels running beside
// the copy).  The kernel keeps to 24 registers: the checkpointed BAQ kernel holds 488 of a SIMD's
// 512 with one wave, and a copy of a host batch that did not fit beside it would wait for it.
static constexpr int APPEND_UNROLL = 2;
struct AppendArgs {
    const uint8_t *src[18];
    uint8_t *dst[18];
    uint64_t bytes[18];
    uint32_t first[19];
    uint32_t cig_off, base_off, qual_off, rg_off, rg_limit;
};

// d[r] = f(s[r]) for r < n, APPEND_UNROLL loads in flight per lane
template <class T, class F>
__device__ inline void append_copy(const T *s, T *d, uint64_t n, uint64_t tid, uint64_t stride, F f)
{
    for (uint64_t r0 = tid; r0 < n; r0 += APPEND_UNROLL * stride) {
        T v[APPEND_UNROLL];
#pragma unroll
        for (int u = 0; u < APPEND_UNROLL; u++) {
            const uint64_t r = r0 + u * stride;
            if (r < n) v[u] = s[r];
        }
#pragma unroll
        for (int u = 0; u < APPEND_UNROLL; u++) {
            const uint64_t r = r0 + u * stride;
            if (r < n) d[r] = f(v[u]);
        }
    }
}

__global__ void __launch_bounds__(256) __attribute__((amdgpu_num_vgpr(24))) k_append(AppendArgs a)
{
    int i = 0;
    while (i < 17 && a.first[i + 1] <= blockIdx.x) i++;
    const uint64_t nb = a.bytes[i];
    const uint64_t tid = (uint64_t)(blockIdx.x - a.first[i]) * blockDim.x + threadIdx.x;
    const uint64_t stride = (uint64_t)(a.first[i + 1] - a.first[i]) * blockDim.x;
    const uint8_t *s = a.src[i];
    uint8_t *d = a.dst[i];
    auto same = [](auto v) { return v; };
    if (i == 7 || i == 10 || i == 13 || i == 17) {
        const bool remap = a.rg_off || a.rg_limit != 0xffffffffu;
        const uint32_t off = i == 7 ? a.cig_off : (i == 10 ? a.base_off : (i == 13 ? a.qual_off : 0u));
        const uint32_t rg_off = a.rg_off, rg_limit = a.rg_limit;
        if (i == 17) {
            append_copy((const uint32_t *)s, (uint32_t *)d, nb / 4, tid, stride, [&](uint32_t v) {
                return remap ? (v < rg_limit ? v + rg_off : 0xffffffffu) : v;
            });
        } else {
            append_copy((const uint32_t *)s, (uint32_t *)d, nb / 4, tid, stride, [&](uint32_t v) { return v + off; });
        }
        return;
    }
    const uintptr_t al = (uintptr_t)s | (uintptr_t)d;
    if ((al & 15) == 0) {
        append_copy((const uint4 *)s, (uint4 *)d, nb / 16, tid, stride, same);
        for (uint64_t r = nb / 16 * 16 + tid; r < nb; r += stride) d[r] = s[r];
        return;
    }
    if ((al & 3) == 0) {
        append_copy((const uint32_t *)s, (uint32_t *)d, nb / 4, tid, stride, same);
        for (uint64_t r = nb / 4 * 4 + tid; r < nb; r += stride) d[r] = s[r];
        return;
    }
    // source and destination aligned differently (qualities appended at any byte offset): whole
    // destination words built from four source bytes, the unaligned head and tail byte by byte
    const uint64_t lead = (4 - ((uintptr_t)d & 3)) & 3;
    const uint64_t head = lead < nb ? lead : nb;
    for (uint64_t r = tid; r < head; r += stride) d[r] = s[r];
    const uint8_t *sw = s + head;
    uint32_t *dw = (uint32_t *)(d + head);
    const uint64_t words = (nb - head) / 4;
    for (uint64_t r0 = tid; r0 < words; r0 += APPEND_UNROLL * stride) {
        uint32_t w[APPEND_UNROLL];
#pragma unroll
        for (int u = 0; u < APPEND_UNROLL; u++) {
            const uint64_t r = r0 + u * stride;
            if (r < words) {
                const uint8_t *p = sw + 4 * r;
                w[u] = (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
            }
        }
#pragma unroll
        for (int u = 0; u < APPEND_UNROLL; u++) {
            const uint64_t r = r0 + u * stride;
            if (r < words) dw[r] = w[u];
        }
    }
    for (uint64_t r = head + 4 * words + tid; r < nb; r += stride) d[r] = s[r];
}
} // namespace fpd

namespace fpd {
// a job's device batch before it joins a shared staging buffer: the longest read and the reads whose
// read group lies at or beyond the job's own (out[0] max read_len, out[1] such reads; out zeroed by
// the caller)
__global__ void __launch_bounds__(256) k_check_batch(BatchDev b, uint32_t rg_limit, uint32_t *out)
{
    uint32_t mx = 0, bad = 0;
    for (uint32_t r = blockIdx.x * 256 + threadIdx.x; r < b.num_reads; r += gridDim.x * 256) {
        const uint32_t L = b.read_len[r];
        mx = L > mx ? L : mx;
        bad += b.read_group[r] >= rg_limit ? 1u : 0u;
    }
    for (int o = 32; o > 0; o >>= 1) {
        const uint32_t v = (uint32_t)__shfl_xor((int)mx, o);
        mx = v > mx ? v : mx;
        bad += (uint32_t)__shfl_xor((int)bad, o);
    }
    if ((threadIdx.x & 63) == 0) {
        if (mx) atomicMax(&out[0], mx);
        if (bad) atomicAdd(&out[1], bad);
    }
}
} // namespace fpd

extern "C" hipError_t fpd_check_batch(const fpd::BatchDev *b, uint32_t rg_limit, uint32_t *out, hipStream_t st)
{
    if (b->num_reads == 0) return hipSuccess;
    const uint32_t blocks = std::min<uint32_t>((b->num_reads + 255) / 256, 1024u);
    hipLaunchKernelGGL(fpd::k_check_batch, dim3(blocks), dim3(256), 0, st, *b, rg_limit, out);
    return hipGetLastError();
}

// block_bytes: bytes of an array per workgroup (at most 64 workgroups per array)
extern "C" hipError_t fpd_append(const void *const *src, void *const *dst, const uint64_t *bytes, uint32_t cig_off,
                                 uint32_t base_off, uint32_t qual_off, uint32_t rg_off, uint32_t rg_limit,
                                 uint64_t block_bytes, hipStream_t st)
{
    fpd::AppendArgs a{};
    uint32_t blocks = 0;
    for (int i = 0; i < 18; i++) {
        a.src[i] = (const uint8_t *)src[i];
        a.dst[i] = (uint8_t *)dst[i];
        a.bytes[i] = src[i] && dst[i] ? bytes[i] : 0;
        a.first[i] = blocks;
        if (a.bytes[i]) blocks += (uint32_t)std::min<uint64_t>(64, (a.bytes[i] + block_bytes - 1) / block_bytes);
    }
    a.first[18] = blocks;
    if (blocks == 0) return hipSuccess;
    a.cig_off = cig_off; a.base_off = base_off; a.qual_off = qual_off; a.rg_off = rg_off; a.rg_limit = rg_limit;
    hipLaunchKernelGGL(fpd::k_append, dim3(blocks), dim3(256), 0, st, a);
    return hipGetLastError();
}

extern "C" hipError_t fpd_rebase(uint32_t n, uint32_t *cigar_start, uint32_t *read_start, uint32_t *qual_start,
                                 uint32_t *read_group, uint32_t cig_off, uint32_t base_off, uint32_t qual_off,
                                 uint32_t rg_off, uint32_t rg_limit, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(fpd::k_rebase, dim3((n + 255) / 256), dim3(256), 0, st, n, cigar_start, read_start, qual_start,
                       read_group, cig_off, base_off, qual_off, rg_off, rg_limit);
    return hipGetLastError();
}
