/*
 * K_prep: per-read preparation, one lane per read.
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
 * cigar.cu:538-609).  All per-base output of a read is written by the lane
 * that owns the read, so there are no read-modify-write hazards between reads.
 */
#include "fp_device.h"

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
    uint32_t *counters;   // [0] active reads, [1] filtered, [2] baq reads, [3] max read length
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

__global__ void __launch_bounds__(256) k_prep(PrepArgs a)
{
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    const BatchDev &b = a.b;
    if (r >= live_reads(b)) return;
    ReadStateDev &s = a.s;

    const uint32_t L = b.read_len[r];
    const uint32_t rs = b.read_start[r];
    const uint32_t qs = b.qual_start[r];
    const uint32_t *ops = b.cigars + b.cigar_start[r];
    const uint32_t nops = b.cigar_len[r];
    const uint16_t flags = b.flags[r];
    const uint32_t astart = b.alignment_start[r];
    const uint32_t padded = (L + 7) & ~7u;

    // defined initial state (per read)
    s.rwc[r] = 0; s.noins[r] = 0; s.refw[r] = 0; s.num_errors[r] = 0;
    s.hmmw[r] = 0; s.bandwidth[r] = 0; s.hqw[r] = 0;
    for (uint32_t i = 0; i < padded; i++) s.base_flags[rs + i] = 0;
    if (L > 0) atomicMax(&a.counters[3], L);

    // ---- filter_invalid_reads
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
    if (!ok) {
        s.status[r] = RS_FILTERED;
        atomicAdd(&a.counters[1], 1u);
        return;
    }

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
    uint16_t nix = 0, niy = 0;
    if (cur.first(ops, nops)) {
        for (;;) {
            const uint16_t ro = cur.rc();
            if (!(ro == NONE16 || ro < rwx || cur.type() == EV_I)) { nix = ro; break; }
            if (!cur.next()) break;
        }
    }
    if (cur.last(ops, nops)) {
        for (;;) {
            const uint16_t ro = cur.rc();
            if (!(ro == NONE16 || ro > rwy || cur.type() == EV_I)) { niy = ro; break; }
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
    s.rwc[r] = pack2(rwx, rwy);
    s.noins[r] = pack2(nix, niy);
    s.refw[r] = pack2(fwx, fwy);

    // ---- error vectors (cigar.cu:612-733)
    const uint32_t chrom = b.chromosome[r];
    const bool neg = (flags & F_REVERSE) != 0;
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
                        s.base_flags[rs + bp] |= BF_SNP;
                        if (bp >= rwx && bp <= rwy) nerr++;
                    }
                }
            } else if (op == OP_I) {
                if (l) {
                    bp = ro;
                    if (bp >= rwx && bp <= rwy) {
                        const int off = neg ? (int)bp + (int)l : (int)bp - 1;
                        if (off >= 0 && off <= rwy) { s.base_flags[rs + off] |= BF_INS; nerr++; }
                    }
                }
            } else if (op == OP_D) {
                if (l) {
                    const uint16_t cri = fo;
                    if (cri >= fwx && cri <= fwy) {
                        if (!neg) { s.base_flags[rs + bp] |= BF_DEL; nerr++; }
                        else {
                            const uint16_t off = (uint16_t)(bp + 1);
                            if (off < L) { s.base_flags[rs + off] |= BF_DEL; nerr++; }
                        }
                    }
                }
            }
            if (op_reads(op)) ro = (uint16_t)(ro + l);
            if (op_refs(op)) fo = (uint16_t)(fo + l);
        }
    }
    s.num_errors[r] = nerr;

    // ---- per-base filters (read_filters.cu:376-438)
    for (uint32_t i = 0; i < L; i++) {
        const uint8_t bpv = read_base(b.reads, (uint64_t)rs + i);
        const bool regular = bpv == B_A || bpv == B_C || bpv == B_G || bpv == B_T;
        if (!regular || b.quals[qs + i] < 6) s.base_flags[rs + i] |= BF_MASKED;
    }

    // ---- known sites (snp_filter.cu:162-395)
    if (chrom < a.g.num_variant_contigs) {
        const uint64_t f0 = a.g.feature_offset[chrom];
        const uint32_t nf = (uint32_t)(a.g.feature_offset[chrom + 1] - f0);
        const uint32_t *fstart = a.g.feature_start + f0;
        const uint32_t *fstop = a.g.feature_stop + f0;
        const uint32_t *mel = a.g.max_end_left + f0;
        const uint32_t qx = astart + fwx, qy = astart + fwy;
        // lower_bound over max_end_point_left (primitives/algorithms.h:39-70)
        uint32_t vx;
        if (nf == 0 || qx < mel[0]) vx = 0;
        else if (mel[nf - 1] < qx) vx = nf;
        else {
            uint32_t lo = 0, cnt = nf;
            while (cnt > 0) {
                const uint32_t h = cnt / 2;
                if (mel[lo + h] < qx) { lo = lo + h + 1; cnt -= h + 1; }
                else cnt = h;
            }
            vx = lo;
        }
        if (vx < nf) {
            // upper_bound over feature_start (primitives/algorithms.h:77-100)
            uint32_t lo = 0, cnt = nf;
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
                    if (!ev.first(ops, nops)) continue;
                    bool found = false;
                    uint16_t refc = NONE16;
                    for (;;) {
                        refc = ev.fc();
                        if (refc != NONE16 && (int)refc >= fs) { found = true; break; }
                        if (!ev.next()) break;
                    }
                    if (!found) continue;
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
                    for (uint32_t d = r0; d <= r1; d++) s.base_flags[rs + d] |= BF_MASKED;
                }
            }
        }
    }

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
            s.status[r] = RS_BAQ_DROPPED;
            return;
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
    s.status[r] = RS_ACTIVE;
    atomicAdd(&a.counters[0], 1u);
    if (nerr != 0) atomicAdd(&a.counters[2], 1u);
}

} // namespace fpd

// host launcher
extern "C" hipError_t fpd_launch_prep(const fpd::BatchDev *b, const fpd::GenomeDev *g,
                                      const fpd::ReadStateDev *s, uint32_t *counters, hipStream_t st)
{
    fpd::PrepArgs a{};
    a.b = *b; a.g = *g; a.s = *s; a.counters = counters;
    const uint32_t n = b->num_reads;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(fpd::k_prep, dim3((n + 255) / 256), dim3(256), 0, st, a);
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

extern "C" hipError_t fpd_rebase(uint32_t n, uint32_t *cigar_start, uint32_t *read_start, uint32_t *qual_start,
                                 uint32_t *read_group, uint32_t cig_off, uint32_t base_off, uint32_t qual_off,
                                 uint32_t rg_off, uint32_t rg_limit, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(fpd::k_rebase, dim3((n + 255) / 256), dim3(256), 0, st, n, cigar_start, read_start, qual_start,
                       read_group, cig_off, base_off, qual_off, rg_off, rg_limit);
    return hipGetLastError();
}
