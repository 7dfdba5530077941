/*
 * K_cov: fractional errors (device/fractional_errors.cu:39-141), the high
 * quality window (device/covariates.cu:215-246) and covariate key packing
 * (covariates.cu:50-103, covariates/packer_quality_score.h and siblings,
 * covariates/bit_packers/cycle_illumina.h, context.h), accumulated straight
 * into the device-resident covariate tables.
 *
 * The tables are direct-indexed by covariate key: the read group is the most
 * significant field of every key layout, so the dense array index of a key is
 * the key itself and the rows of a table come out in key order -- the order
 * the reference obtains by radix-sorting (covariate_table.cu:43-58).  Counts
 * are u64; mismatch sums are exact fixed point kept as three 27-bit limb sums
 * (fp_device.h), so the accumulated table does not depend on batch size,
 * launch order or the order atomics land in, and no update waits for a
 * returned value.
 *
 * Work layout: reads are grouped by read group; a workgroup of 12 waves takes a
 * tile of up to 1024 reads of one group and keeps that group's table entries in
 * LDS: every count, and the exact sums of the quality, context and cycle I/D
 * entries.  Only the sums of the cycle table's mismatch entries (few nonzero,
 * thousands of distinct keys) go to the resident tables during the tile.
 * Global atomics on a handful of hot keys serialise at the coherence point and
 * stall the whole memory system, so the per-base counts never leave LDS
 * until the tile's flush.
 * Each wavefront handles one read at a time with its 64 lanes spread over the
 * read's bases (positions lane, lane+64, lane+128, lane+192 of the clipped
 * window), so every per-base load is coalesced.  The GATK block-averaged
 * fractional errors come from 256-bit ballot masks of the BAQ blocks and error
 * bits: one wave-uniform walk over a read's blocks gives each lane the error
 * counts of the block range over its positions.  The context keys read neighbouring bases from a
 * per-wave LDS copy of the window.  Windows longer than 256 bases take a
 * serial path on one lane.
 *
 * A wave loads the per-read fields (list entry, windows, starts, lengths, flags,
 * CIGAR position) for 64 reads at once, one read per lane, and takes read k's
 * with readlane; a read's CIGAR ops arrive one per lane beside its bases.  In
 * groups of reads of at most 128 window bases the next read's inputs are loaded
 * while the current read is processed.
 */
#include "fp_device.h"

namespace fpd {

struct CovArgs {
    BatchDev b;
    ReadStateDev s;
    TablesDev t;
    uint32_t frac_f64;
    uint32_t keep_frac;          // also store the per-base fractional errors (inspection output)
    uint32_t ncodes;             // cycle codes whose I/D entries are kept in LDS
    uint32_t mcodes;             // cycle codes whose M counts are kept in LDS
    const uint32_t *rg_list;     // active reads grouped by read group
    const uint32_t *rg_offset;   // num_rg + 1
    const uint32_t *tile_offset; // num_rg + 1 (tiles of `tile` reads)
    uint32_t tile;               // reads per workgroup tile
    // pair mode (sort aggregation): one (key, value) pair per observation and table
    uint32_t *pkey[3];
    double *pval[3];
    uint32_t *pcursor;           // [3] append cursors (elements)
    uint32_t *perr;              // set when the pairs of a batch would overflow pcap
    uint64_t pcap;               // pair capacity per table
};

// calculateFractionalErrorArray for one error class over the clipped window (serial form)
__device__ void fractional_errors(const uint8_t *baq, const uint8_t *flags, int len, uint8_t bit,
                                  double *out /* stride 3 */)
{
    bool inBlock = false;
    int bs = -1;
    int iii;
    for (iii = 0; iii < len; iii++) {
        if (baq[iii] == 64) {
            if (!inBlock) {
                out[3 * iii] = (double)((flags[iii] & bit) ? 1 : 0);
            } else {
                const int s0 = bs - 1 > 0 ? bs - 1 : 0;
                int tot = 0;
                for (int j = s0; j <= iii; j++) tot += (flags[j] & bit) ? 1 : 0;
                const double v = ((double)tot) / ((double)(iii - s0 + 1));
                for (int j = s0; j <= iii; j++) out[3 * j] = v;
                inBlock = false;
                bs = -1;
            }
        } else {
            inBlock = true;
            if (bs == -1) bs = iii;
        }
    }
    if (inBlock) {
        const int end = iii - 1;
        const int s0 = bs - 1 > 0 ? bs - 1 : 0;
        int tot = 0;
        for (int j = s0; j <= end; j++) tot += (flags[j] & bit) ? 1 : 0;
        const double v = ((double)tot) / ((double)(end - s0 + 1));
        for (int j = s0; j <= end; j++) out[3 * j] = v;
    }
}

// covariate_Cycle_Illumina (bit_packers/cycle_illumina.h:46-95)
__device__ inline void cycle_keys(uint16_t flags, uint16_t wx, uint16_t wy, uint16_t bp, uint32_t &km, uint32_t &kid)
{
    const bool paired = flags & F_PAIRED, second = flags & F_READ2, neg = flags & F_REVERSE;
    const int readLength = (int)wy - (int)wx + 1;
    const int rof = (paired && second) ? -1 : 1;
    const int i = (int)bp - (int)wx;
    int cycle, inc;
    if (neg) { cycle = readLength * rof; inc = -1 * rof; }
    else { cycle = rof; inc = rof; }
    cycle += i * inc;
    uint32_t key = (uint32_t)(cycle < 0 ? -cycle : cycle);
    key = key << 1;
    if (cycle < 0) key++;
    km = key;
    kid = (i < 4 || i > readLength - 4 - 1) ? 1023u : key;
}

// covariate_Context with 2 mismatch / 3 indel bases (bit_packers/context.h:104-174);
// base(i) returns the read base at read position i
template <class BASE>
__device__ inline void context_keys(BASE base, uint16_t hx, uint16_t hy, bool neg, uint16_t bp, uint32_t &km,
                                    uint32_t &kid)
{
    km = 2047u; kid = 2047u;
    const int dir = neg ? 1 : -1;
    int avail = neg ? (int)hy - (int)bp + 1 : (int)bp - (int)hx + 1;
    const int cb = avail < 3 ? avail : 3;
    if (cb < 2) return;
    uint32_t ctx = 0;
    int n = 0;
    for (int i = bp, stop = (int)bp + cb * dir; i != stop; i += dir) {
        const uint8_t bb = base(i);
        if (bb != B_A && bb != B_C && bb != B_T && bb != B_G) break;
        ctx <<= 2;
        ctx |= 31 - __builtin_clz((uint32_t)bb);
        n++;
    }
    if (neg) ctx = ~ctx & ((1u << (n * 2)) - 1);
    if (n >= 2) km = ((ctx >> ((n - 2) * 2)) << 4) | 2;
    if (n >= 3) kid = ((ctx >> ((n - 3) * 2)) << 4) | 3;
}

// context_keys without the loop: the (up to) three bases are read at once, positions past the
// high-quality window clamped to the base itself and masked out; the same keys
template <class BASE>
__device__ inline void context_keys3(BASE base, uint16_t hx, uint16_t hy, bool neg, uint16_t bp, uint32_t &km,
                                     uint32_t &kid)
{
    const int avail = neg ? (int)hy - (int)bp + 1 : (int)bp - (int)hx + 1;
    const int dir = neg ? 1 : -1;
    const int p1 = avail >= 2 ? (int)bp + dir : (int)bp, p2 = avail >= 3 ? (int)bp + 2 * dir : (int)bp;
    const uint32_t b0 = base((int)bp), b1 = base(p1), b2 = base(p2);
    auto acgt = [](uint32_t b) { return b == B_A || b == B_C || b == B_G || b == B_T; };
    auto code = [neg](uint32_t b) {   // A C G T -> 0 1 2 3, complemented on the reverse strand
        const uint32_t c = 31u - (uint32_t)__builtin_clz(b | 1u);
        return neg ? 3u - c : c;
    };
    // bitwise ands: all three tests are cheap, and a short-circuit would branch on the exec mask
    const bool v2 = acgt(b0) & (avail >= 2) & acgt(b1);
    const bool v3 = v2 & (avail >= 3) & acgt(b2);
    const uint32_t c0 = code(b0), c1 = code(b1), c2 = code(b2);
    km = v2 ? ((((c0 << 2) | c1) << 4) | 2u) : 2047u;
    kid = v3 ? ((((c0 << 4) | (c1 << 2) | c2) << 4) | 3u) : 2047u;
}

// an exact per-base value as limbs
struct Limbs {
    uint64_t p0, p1, p2;
};

__device__ inline Limbs to_limbs(double v, uint32_t f64)
{
    Limbs l;
    // most per-base errors are exactly 0 or 1 (outside BAQ blocks); 2^80 has limbs (0, 0, 2^26)
    if (v == 0.0) { l.p0 = 0; l.p1 = 0; l.p2 = 0; return l; }
    if (v == 1.0) { l.p0 = 0; l.p1 = 0; l.p2 = 1ull << 26; return l; }
    fixed_split(fixed_from_double(f64 ? v : (double)(float)v), l.p0, l.p1, l.p2);
    return l;
}

__device__ inline bool nonzero(const Limbs &l) { return (l.p0 | l.p1 | l.p2) != 0; }

// LDS-privatised accumulation: an entry holds a count and three limb sums; the
// cycle table's M entries keep only a 16-bit count, two per word, quality-major
// (word index (q * mcodes + code) / 2) so lanes at consecutive cycles hit
// consecutive banks.  A tile adds at most one count per read to such an entry.
struct LdsAcc {
    unsigned long long *p0, *p1, *p2;
    uint32_t *cnt;
    uint32_t *mcnt;
    uint32_t mcodes;
    uint32_t dump, mdump;   // first of 64 per-lane count words (and packed-count words) never flushed
};

__device__ inline bool lds_cycle_m(const LdsAcc &L, uint32_t code, uint32_t q)
{
    if (code >= L.mcodes || q >= 64) return false;
    const uint32_t i = q * L.mcodes + code;
    atomicAdd(&L.mcnt[i >> 1], 1u << (16 * (i & 1)));
    return true;
}

__device__ inline void lds_add_n(const LdsAcc &L, uint32_t e, uint32_t n, const Limbs &v)
{
    atomicAdd(&L.cnt[e], n);
    if (nonzero(v)) {
        atomicAdd(&L.p0[e], (unsigned long long)v.p0);
        atomicAdd(&L.p1[e], (unsigned long long)v.p1);
        atomicAdd(&L.p2[e], (unsigned long long)v.p2);
    }
}

__device__ inline void lds_add(const LdsAcc &L, uint32_t e, const Limbs &v) { lds_add_n(L, e, 1, v); }

// the sums of an entry whose count is kept by its partner (an I/D key pair counts every base in both;
// the D entry's count is read from the I entry's when the tile is flushed)
__device__ inline void lds_add_sums(const LdsAcc &L, uint32_t e, const Limbs &v)
{
    if (nonzero(v)) {
        atomicAdd(&L.p0[e], (unsigned long long)v.p0);
        atomicAdd(&L.p1[e], (unsigned long long)v.p1);
        atomicAdd(&L.p2[e], (unsigned long long)v.p2);
    }
}

__device__ inline void global_add(const TablesDev &t, int tab, uint32_t key, uint64_t count, const Limbs &v)
{
    atomicAdd((unsigned long long *)&t.count[tab][key], (unsigned long long)count);
    if (nonzero(v)) {
        atomicAdd((unsigned long long *)&t.part[tab][0][key], (unsigned long long)v.p0);
        atomicAdd((unsigned long long *)&t.part[tab][1][key], (unsigned long long)v.p1);
        atomicAdd((unsigned long long *)&t.part[tab][2][key], (unsigned long long)v.p2);
    }
}

// LDS entry map of one read-group tile
//   [0, 64)                quality table, event M, quality q < 64
//   64, 65                 quality table, events I and D (quality 45)
//   [66, 66+1024)          context M: q < 64 x 16 two-base contexts
//   [1090, 1090+128)       context I/D: 2 x 64 three-base contexts
//   [1218, 1218+2*ncodes)  cycle I/D: 2 x cycle codes
static constexpr uint32_t E_CM = 66, E_CID = 66 + 1024, E_YID = 66 + 1024 + 128;
static constexpr uint32_t LDS_ENTRY_BYTES = 3 * 8 + 4;
static constexpr uint32_t COV_WAVES = 12;
static constexpr uint32_t TILE = 1024;                     // reads per workgroup tile (large batches)
static constexpr uint32_t TILE_MIN = 128;                  // small batches: more, smaller tiles
static constexpr uint32_t STAGE_BYTES = COV_WAVES * 256;   // per-wave copies of the window's read bases

// the covariate keys of one base and their accumulation
template <class BASE>
__device__ inline void emit_base(const CovArgs &a, const LdsAcc &L, uint32_t rg, uint32_t ncodes, uint16_t flags,
                                 bool neg, BASE base, uint16_t rwx, uint16_t rwy, uint16_t hx, uint16_t hy,
                                 uint16_t bp, uint32_t q, const Limbs &f0, const Limbs &f1, const Limbs &f2)
{
    // RecalTable1: EventTracker(2) | QualityScore(8) | ReadGroup(8)
    if (q < 64) lds_add(L, q, f0);
    else global_add(a.t, 0, 0u | (q << 2) | (rg << 10), 1, f0);
    lds_add(L, 64, f1);
    lds_add_sums(L, 65, f2);
    // cycle: EventTracker(2) | Cycle(10) | QualityScore(8) | ReadGroup(8)
    uint32_t cm, cid;
    cycle_keys(flags, rwx, rwy, bp, cm, cid);
    {
        const uint32_t key = 0u | (cm << 2) | (q << 12) | (rg << 20);
        if (lds_cycle_m(L, cm, q)) {
            if (nonzero(f0)) {
                atomicAdd((unsigned long long *)&a.t.part[1][0][key], (unsigned long long)f0.p0);
                atomicAdd((unsigned long long *)&a.t.part[1][1][key], (unsigned long long)f0.p1);
                atomicAdd((unsigned long long *)&a.t.part[1][2][key], (unsigned long long)f0.p2);
            }
        } else {
            global_add(a.t, 1, key, 1, f0);
        }
    }
    if (cid != 1023u) {
        if (cid < ncodes) {
            lds_add(L, E_YID + cid, f1);
            lds_add_sums(L, E_YID + ncodes + cid, f2);
        } else {
            global_add(a.t, 1, 1u | (cid << 2) | (45u << 12) | (rg << 20), 1, f1);
            global_add(a.t, 1, 2u | (cid << 2) | (45u << 12) | (rg << 20), 1, f2);
        }
    }
    // context: EventTracker(2) | Context(11) | QualityScore(8) | ReadGroup(8)
    uint32_t xm, xid;
    context_keys(base, hx, hy, neg, bp, xm, xid);
    if (xm != 2047u) {
        if (q < 64) lds_add(L, E_CM + q * 16 + (xm >> 4), f0);
        else global_add(a.t, 2, 0u | (xm << 2) | (q << 13) | (rg << 21), 1, f0);
    }
    if (xid != 2047u) {
        lds_add(L, E_CID + (xid >> 4), f1);
        lds_add_sums(L, E_CID + 64 + (xid >> 4), f2);
    }
}

// The same keys and updates as emit_base for the wave path, with the five count updates of a base
// issued unconditionally: a base whose key has no LDS entry (no key at all, or one kept in the
// resident tables) counts into its lane's dump word, which the flush never reads.  The sums (nonzero
// errors only) and the resident-table fallbacks (qualities >= 64, cycle codes beyond the LDS entries)
// run under one wave-uniform test; most bases have neither.
template <class BASE>
__device__ inline void emit_base_wave(const CovArgs &a, const LdsAcc &L, uint32_t rg, uint32_t ncodes,
                                      uint16_t flags, bool neg, BASE base, uint16_t rwx, uint16_t rwy,
                                      uint16_t hx, uint16_t hy, uint16_t bp, uint32_t q, const Limbs &f0,
                                      const Limbs &f1, const Limbs &f2)
{
    const uint32_t lane = threadIdx.x & 63;
    uint32_t cm, cid, xm, xid;
    cycle_keys(flags, rwx, rwy, bp, cm, cid);
    context_keys3(base, hx, hy, neg, bp, xm, xid);
    const bool qok = q < 64;                            // quality and context M entries in LDS
    const bool mok = qok && cm < L.mcodes;              // cycle M count in LDS
    const bool yid = cid != 1023u, yok = yid && cid < ncodes;
    const bool xmk = xm != 2047u, xik = xid != 2047u;
    const uint32_t dump = L.dump + lane;
    const uint32_t mi = q * L.mcodes + cm;
    atomicAdd(&L.cnt[qok ? q : dump], 1u);
    atomicAdd(&L.mcnt[mok ? (mi >> 1) : L.mdump + lane], mok ? 1u << (16 * (mi & 1)) : 0u);
    atomicAdd(&L.cnt[yok ? E_YID + cid : dump], 1u);
    atomicAdd(&L.cnt[xmk && qok ? E_CM + q * 16 + (xm >> 4) : dump], 1u);
    atomicAdd(&L.cnt[xik ? E_CID + (xid >> 4) : dump], 1u);
    const bool sums = nonzero(f0) || nonzero(f1) || nonzero(f2);
    const bool rare = !qok || !mok || (yid && !yok);
    if (!wave_any(sums || rare)) return;
    // RecalTable1: EventTracker(2) | QualityScore(8) | ReadGroup(8); the I and D keys are the same
    // for every base of the read (their counts are added once per read)
    lds_add_sums(L, 64, f1);
    lds_add_sums(L, 65, f2);
    if (qok) lds_add_sums(L, q, f0);
    else global_add(a.t, 0, 0u | (q << 2) | (rg << 10), 1, f0);
    // cycle: EventTracker(2) | Cycle(10) | QualityScore(8) | ReadGroup(8)
    const uint32_t key = 0u | (cm << 2) | (q << 12) | (rg << 20);
    if (mok) {
        if (nonzero(f0)) {
            atomicAdd((unsigned long long *)&a.t.part[1][0][key], (unsigned long long)f0.p0);
            atomicAdd((unsigned long long *)&a.t.part[1][1][key], (unsigned long long)f0.p1);
            atomicAdd((unsigned long long *)&a.t.part[1][2][key], (unsigned long long)f0.p2);
        }
    } else {
        global_add(a.t, 1, key, 1, f0);
    }
    if (yok) {
        lds_add_sums(L, E_YID + cid, f1);
        lds_add_sums(L, E_YID + ncodes + cid, f2);
    } else if (yid) {
        global_add(a.t, 1, 1u | (cid << 2) | (45u << 12) | (rg << 20), 1, f1);
        global_add(a.t, 1, 2u | (cid << 2) | (45u << 12) | (rg << 20), 1, f2);
    }
    // context: EventTracker(2) | Context(11) | QualityScore(8) | ReadGroup(8)
    if (xmk) {
        if (qok) lds_add_sums(L, E_CM + q * 16 + (xm >> 4), f0);
        else global_add(a.t, 2, 0u | (xm << 2) | (q << 13) | (rg << 21), 1, f0);
    }
    if (xik) {
        lds_add_sums(L, E_CID + (xid >> 4), f1);
        lds_add_sums(L, E_CID + 64 + (xid >> 4), f2);
    }
}

// ---- pair mode -------------------------------------------------------------

static constexpr uint32_t PAIR_SENTINEL = 0xffffffffu;

// a workgroup's slice of the pair arrays: base[t] and the slots used so far (LDS)
struct PairRegion {
    uint32_t *used;
    const uint64_t *base;
};

// append n (0..3) pairs per lane to table t; a collective of the active lanes of the wavefront
__device__ inline void append_pairs(const CovArgs &a, const PairRegion &pr, int t, uint32_t n, uint32_t k0,
                                    uint32_t k1, uint32_t k2, double v0, double v1, double v2)
{
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t lt = (1ull << lane) - 1;
    const uint64_t b0 = __ballot(n & 1), b1 = __ballot((n >> 1) & 1);
    const uint32_t total = (uint32_t)__popcll(b0) + 2 * (uint32_t)__popcll(b1);
    if (total == 0) return;
    const int leader = __builtin_ctzll(__ballot(1));
    uint32_t off = 0;
    if ((int)lane == leader) off = atomicAdd(&pr.used[t], total);
    off = (uint32_t)__shfl((int)off, leader);
    const uint64_t pos = pr.base[t] + off + (uint32_t)__popcll(b0 & lt) + 2 * (uint32_t)__popcll(b1 & lt);
    uint32_t *kk = a.pkey[t] + pos;
    double *vv = a.pval[t] + pos;
    if (n > 0) { kk[0] = k0; vv[0] = v0; }
    if (n > 1) { kk[1] = k1; vv[1] = v1; }
    if (n > 2) { kk[2] = k2; vv[2] = v2; }
}

// the covariate keys of one base as pairs (the same keys emit_base accumulates); emit false:
// the lane takes part in the collective appends with no pairs
template <class BASE>
__device__ inline void emit_pairs(const CovArgs &a, const PairRegion &pr, uint32_t rg, uint16_t flags, bool neg,
                                  BASE base, uint16_t rwx, uint16_t rwy, uint16_t hx, uint16_t hy, uint16_t bp,
                                  uint32_t q, double vs, double vi, double vd, bool emit)
{
    if (!a.frac_f64) { vs = (double)(float)vs; vi = (double)(float)vi; vd = (double)(float)vd; }
    uint32_t nq = 0, ny = 0, nx = 0;
    uint32_t y0 = 0, y1 = 0, y2 = 0, x0 = 0, x1 = 0, x2 = 0;
    double xv0 = 0.0, xv1 = 0.0, xv2 = 0.0;
    if (emit) {
        nq = 3;
        uint32_t cm, cid;
        cycle_keys(flags, rwx, rwy, bp, cm, cid);
        y0 = 0u | (cm << 2) | (q << 12) | (rg << 20);
        ny = 1;
        if (cid != 1023u) {
            y1 = 1u | (cid << 2) | (45u << 12) | (rg << 20);
            y2 = 2u | (cid << 2) | (45u << 12) | (rg << 20);
            ny = 3;
        }
        uint32_t xm, xid;
        context_keys(base, hx, hy, neg, bp, xm, xid);
        const uint32_t km = 0u | (xm << 2) | (q << 13) | (rg << 21);
        const uint32_t ki = 1u | (xid << 2) | (45u << 13) | (rg << 21);
        const uint32_t kd = 2u | (xid << 2) | (45u << 13) | (rg << 21);
        if (xm != 2047u && xid != 2047u) { x0 = km; xv0 = vs; x1 = ki; xv1 = vi; x2 = kd; xv2 = vd; nx = 3; }
        else if (xm != 2047u) { x0 = km; xv0 = vs; nx = 1; }
        else if (xid != 2047u) { x0 = ki; xv0 = vi; x1 = kd; xv1 = vd; nx = 2; }
    }
    // RecalTable1: EventTracker(2) | QualityScore(8) | ReadGroup(8)
    append_pairs(a, pr, 0, nq, 0u | (q << 2) | (rg << 10), 1u | (45u << 2) | (rg << 10), 2u | (45u << 2) | (rg << 10),
                 vs, vi, vd);
    // cycle: EventTracker(2) | Cycle(10) | QualityScore(8) | ReadGroup(8)
    append_pairs(a, pr, 1, ny, y0, y1, y2, vs, vi, vd);
    // context: EventTracker(2) | Context(11) | QualityScore(8) | ReadGroup(8)
    append_pairs(a, pr, 2, nx, x0, x1, x2, xv0, xv1, xv2);
}

// one read on one lane (windows longer than 256 bases)
template <bool PAIRS>
__device__ __forceinline__ void cov_read_serial(const CovArgs &a, const LdsAcc &L, const PairRegion &pr, uint32_t r, uint32_t rg,
                                uint32_t ncodes)
{
    const BatchDev &b = a.b;
    const uint32_t rw = a.s.rwc[r];
    const uint16_t rwx = lo16(rw), rwy = hi16(rw);
    const uint32_t rs = b.read_start[r];
    const uint32_t qs = b.qual_start[r];
    const int len = (int)rwy - (int)rwx + 1;
    const uint8_t *baq = a.s.baq + qs + rwx;
    const uint8_t *bfl = a.s.base_flags + rs + rwx;
    // the serial path works in the fractional-error array itself
    double *frac = a.s.frac + 3ull * (qs + rwx);
    for (uint32_t k = 0; k < b.read_len[r]; k++)
        if (k < rwx || k > rwy) { double *z = a.s.frac + 3ull * (qs + k); z[0] = 0.0; z[1] = 0.0; z[2] = 0.0; }
    fractional_errors(baq, bfl, len, BF_SNP, frac + 0);
    fractional_errors(baq, bfl, len, BF_INS, frac + 1);
    fractional_errors(baq, bfl, len, BF_DEL, frac + 2);

    uint16_t hx = rwx, hy = rwy;
    while (b.quals[qs + hx] <= 2 && hx < hy) hx++;
    while (b.quals[qs + hy] <= 2 && hy > hx) hy--;
    a.s.hqw[r] = pack2(hx, hy);

    const uint16_t flags = b.flags[r];
    const bool neg = (flags & F_REVERSE) != 0;
    const uint32_t *ops = b.cigars + b.cigar_start[r];
    const uint32_t nops = b.cigar_len[r];
    auto base = [&](int i) { return read_base(b.reads, (uint64_t)rs + (uint32_t)i); };
    uint16_t ro = 0;
    for (uint32_t oi = 0; oi < nops; oi++) {
        const uint32_t op = op_code(ops[oi]), l = op_len(ops[oi]);
        if (!op_reads(op)) continue;
        if (op_event(op) != EV_S) {
            for (uint32_t k = 0; k < l; k++) {
                const uint16_t bp = (uint16_t)(ro + k);
                if (bp < rwx || bp > rwy) continue;
                if (a.s.base_flags[rs + bp] & BF_MASKED) continue;
                const double *fe = a.s.frac + 3ull * (qs + bp);
                if constexpr (PAIRS)
                    emit_pairs(a, pr, rg, flags, neg, base, rwx, rwy, hx, hy, bp, b.quals[qs + bp], fe[0], fe[1],
                               fe[2], true);
                else
                    emit_base(a, L, rg, ncodes, flags, neg, base, rwx, rwy, hx, hy, bp, b.quals[qs + bp],
                              to_limbs(fe[0], a.frac_f64), to_limbs(fe[1], a.frac_f64),
                              to_limbs(fe[2], a.frac_f64));
            }
        }
        ro = (uint16_t)(ro + l);
    }
}

// ---- 256-bit masks (4 x u64, bit j = window position j) -------------------

// The helpers below visit the four words with compile-time indices, so the masks stay in
// registers (a word picked by a per-lane index would put the arrays in scratch memory).

// highest index < j whose bit is clear, or -1
__device__ inline int prev_zero(const uint64_t *m, int j)
{
    int res = -1;
#pragma unroll
    for (int w = 0; w < 4; w++) {
        const int top = j - 64 * w;           // bits of word w below j
        const uint64_t z0 = ~m[w];
        const uint64_t z = top <= 0 ? 0ull : (top >= 64 ? z0 : (z0 & ((1ull << (top & 63)) - 1)));
        if (z) res = w * 64 + 63 - __builtin_clzll(z);
    }
    return res;
}

// lowest index >= j whose bit is clear, or 256
__device__ inline int next_zero(const uint64_t *m, int j)
{
    int res = 256;
#pragma unroll
    for (int w = 3; w >= 0; w--) {
        const int lo = j - 64 * w;            // first bit of word w at or above j
        const uint64_t z0 = ~m[w];
        const uint64_t z = lo >= 64 ? 0ull : (lo <= 0 ? z0 : (z0 & (~0ull << (lo & 63))));
        if (z) res = w * 64 + __builtin_ctzll(z);
    }
    return res;
}

// lowest index >= j whose bit is set, or 256
__device__ inline int next_one(const uint64_t *m, int j)
{
    int res = 256;
#pragma unroll
    for (int w = 3; w >= 0; w--) {
        const int lo = j - 64 * w;
        const uint64_t z = lo >= 64 ? 0ull : (lo <= 0 ? m[w] : (m[w] & (~0ull << (lo & 63))));
        if (z) res = w * 64 + __builtin_ctzll(z);
    }
    return res;
}

// number of set bits in [a, b]
__device__ inline int pop_range(const uint64_t *m, int a, int b)
{
    int n = 0;
#pragma unroll
    for (int w = 0; w < 4; w++) {
        const int lo = a - 64 * w, hi = b - 64 * w;
        if (hi < 0 || lo > 63) continue;
        const uint64_t lm = lo <= 0 ? ~0ull : (~0ull << (lo & 63));
        const uint64_t hm = hi >= 63 ? ~0ull : ((1ull << ((hi + 1) & 63)) - 1);
        n += __popcll(m[w] & lm & hm);
    }
    return n;
}

// set bits [lo, hi) of a 256-bit mask (wave-uniform values)
__device__ inline void mset_range(uint64_t *m, int lo, int hi)
{
    lo = lo < 0 ? 0 : lo;
    hi = hi > 256 ? 256 : hi;
    for (int w = 0; w < 4; w++) {
        const int a0 = lo - 64 * w, a1 = hi - 64 * w;
        if (a1 <= 0 || a0 >= 64) continue;
        const int x0 = a0 < 0 ? 0 : a0, x1 = a1 > 64 ? 64 : a1;
        const uint64_t hm = x1 >= 64 ? ~0ull : ((1ull << x1) - 1);
        const uint64_t lm = (1ull << x0) - 1;
        m[w] |= hm & ~lm;
    }
}

__device__ inline uint64_t wave_sum(uint64_t v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// the per-read fields a wave needs before it touches a read's bases.  A wave loads them for up
// to 64 reads at once, one read per lane, and takes read k's with readlane: one round trip per
// 64 reads instead of a chain of dependent loads in front of every read.
struct ReadMeta {
    uint32_t r, rw, rs, qs, len, cs, cl, flags;
};

__device__ inline ReadMeta meta_lane(const CovArgs &a, uint32_t li, bool v)
{
    ReadMeta m{};
    if (v) {
        m.r = a.rg_list[li];
        m.rw = a.s.rwc[m.r];
        m.rs = a.b.read_start[m.r];
        m.qs = a.b.qual_start[m.r];
        m.len = a.b.read_len[m.r];
        m.cs = a.b.cigar_start[m.r];
        m.cl = a.b.cigar_len[m.r];
        m.flags = a.b.flags[m.r];
    }
    return m;
}

__device__ inline ReadMeta meta_of(const ReadMeta &m, uint32_t k)
{
    ReadMeta u;
    u.r = __builtin_amdgcn_readlane(m.r, k);
    u.rw = __builtin_amdgcn_readlane(m.rw, k);
    u.rs = __builtin_amdgcn_readlane(m.rs, k);
    u.qs = __builtin_amdgcn_readlane(m.qs, k);
    u.len = __builtin_amdgcn_readlane(m.len, k);
    u.cs = __builtin_amdgcn_readlane(m.cs, k);
    u.cl = __builtin_amdgcn_readlane(m.cl, k);
    u.flags = __builtin_amdgcn_readlane(m.flags, k);
    return u;
}

// the per-base inputs of one read's window positions lane + 64c (c < NC) and its CIGAR ops, one
// per lane: issued for the next read while the current one is processed
template <int NC>
struct CovLoads {
    uint32_t tag[NC], fl[NC], qv[NC], rb[NC];
    uint32_t op;
};

template <int NC>
__device__ inline void cov_load(const CovArgs &a, const ReadMeta &m, CovLoads<NC> &x)
{
    const BatchDev &b = a.b;
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t rwx = lo16(m.rw);
    const int len = (int)hi16(m.rw) - (int)rwx + 1;
    x.op = lane < m.cl ? b.cigars[m.cs + lane] : 0u;
#pragma unroll
    for (int c = 0; c < NC; c++) {
        const int j = (int)lane + 64 * c;
        const bool v = j < len;
        const uint32_t p = rwx + j;
        x.tag[c] = v ? a.s.baq[m.qs + p] : 64u;
        x.fl[c] = v ? a.s.base_flags[m.rs + p] : 0u;
        x.qv[c] = v ? b.quals[m.qs + p] : 255u;
        x.rb[c] = v ? read_base(b.reads, (uint64_t)m.rs + p) : 0u;
    }
}

// one read on one wavefront (clipped window of at most 64 * NC bases); m is wave-uniform
// (x: its loads; with PF, once its masks are built, the loads of read mn start into nx)
template <bool PAIRS, int NC, bool PF>
__device__ __forceinline__ void cov_read_wave(const CovArgs &a, const LdsAcc &L, const PairRegion &pr, uint8_t *stage,
                                              const ReadMeta &m, const CovLoads<NC> &x, uint32_t rg, uint32_t ncodes,
                                              const ReadMeta &mn, CovLoads<NC> &nx)
{
    const BatchDev &b = a.b;
    const int lane = threadIdx.x & 63;
    const uint32_t r = m.r;
    const uint32_t rw = m.rw;
    const uint16_t rwx = lo16(rw), rwy = hi16(rw);
    const uint32_t qs = m.qs;
    const uint32_t L_read = m.len;
    const int len = (int)rwy - (int)rwx + 1;
    // the read's CIGAR ops, one per lane, loaded beside its bases (longer CIGARs are walked from memory)
    const uint32_t *ops = b.cigars + m.cs;
    const uint32_t nops = m.cl;
    const uint32_t op_l = x.op;

    // zero the fractional errors outside the window (defined output for every quality of the read)
    if (a.keep_frac)
        for (uint32_t k = lane; k < L_read; k += 64)
            if (k < rwx || k > rwy) { double *z = a.s.frac + 3ull * (qs + k); z[0] = 0.0; z[1] = 0.0; z[2] = 0.0; }

    uint8_t tag[NC], fl[NC], qv[NC];
    uint64_t IB[4] = {0, 0, 0, 0}, ES[4] = {0, 0, 0, 0}, EI_[4] = {0, 0, 0, 0}, ED[4] = {0, 0, 0, 0},
             LQ[4] = {0, 0, 0, 0};
#pragma unroll
    for (int c = 0; c < NC; c++) {
        const int j = lane + 64 * c;
        tag[c] = (uint8_t)x.tag[c];
        fl[c] = (uint8_t)x.fl[c];
        qv[c] = (uint8_t)x.qv[c];
        if (j < len) stage[j] = (uint8_t)x.rb[c];
    }
#pragma unroll
    for (int c = 0; c < NC; c++) {
        const bool v = lane + 64 * c < len;
        IB[c] = __ballot(v && tag[c] != 64);
        ES[c] = __ballot(v && (fl[c] & BF_SNP));
        EI_[c] = __ballot(v && (fl[c] & BF_INS));
        ED[c] = __ballot(v && (fl[c] & BF_DEL));
        LQ[c] = __ballot(v && qv[c] <= 2);
    }
    // high quality window (covariates.cu:215-246)
    // (no base of quality <= 2: the whole window)
    const bool lowq = (LQ[0] | LQ[1] | LQ[2] | LQ[3]) != 0;
    int hxr = lowq ? next_zero(LQ, 0) : 0;
    if (hxr > len - 1) hxr = len - 1;
    int hyr = lowq ? prev_zero(LQ, len) : len - 1;
    if (hyr < hxr) hyr = hxr;
    const uint16_t hx = (uint16_t)(rwx + hxr), hy = (uint16_t)(rwx + hyr);
    if (lane == 0) a.s.hqw[r] = pack2(hx, hy);
    if constexpr (PF) cov_load<NC>(a, mn, nx);

    const uint16_t flags = (uint16_t)m.flags;
    const bool neg = (flags & F_REVERSE) != 0;
    // soft-clipped window positions (S ops emit nothing), from a wave-uniform walk of the CIGAR
    uint64_t SM[4] = {0, 0, 0, 0};
    {
        int ro = 0;
        for (uint32_t oi = 0; oi < nops; oi++) {
            const uint32_t w = oi < 64 ? __builtin_amdgcn_readlane(op_l, oi) : ops[oi];
            const uint32_t op = op_code(w), l = op_len(w);
            if (!op_reads(op)) continue;
            if (op_event(op) == EV_S) mset_range(SM, ro - (int)rwx, ro + (int)l - (int)rwx);
            ro += (int)l;
        }
    }
    // the stage holds the window's bases; context positions stay inside the high-quality window
    __builtin_amdgcn_wave_barrier();
    auto base = [&](int i) -> uint8_t { return stage[i - (int)rwx]; };

    // bit of window position j = lane + 64c, with c known at compile time after unrolling (one
    // shift of a wave-uniform word instead of a lookup over four words)
    auto bit_at = [&](const uint64_t *M, int c) -> bool { return (M[c] >> lane) & 1; };

    // GATK block-averaged fractional errors (calculateFractionalErrorArray): a BAQ block [bs, be]
    // (positions whose tag is not 64) averages the error bits over [max(bs - 1, 0), min(be + 1,
    // len - 1)] and every position of that range takes the average; a position in two ranges takes
    // the later block's, as the reference assigns the ranges in order.  The blocks are walked in
    // order with wave-uniform arithmetic on the masks; each lane keeps the size and error counts of
    // the last range over its positions j = lane + 64c.
    uint32_t rn[NC], rc[NC];   // range size (0: none) and its SNP | INS << 10 | DEL << 20 error counts
#pragma unroll
    for (int c = 0; c < NC; c++) { rn[c] = 0; rc[c] = 0; }
    for (int pos = next_one(IB, 0); pos < len;) {
        const int be = next_zero(IB, pos) - 1;
        const int s0 = pos > 0 ? pos - 1 : 0, e = be + 1 < len ? be + 1 : be;
        const uint32_t n = (uint32_t)(e - s0 + 1);
        const uint32_t cnt = (uint32_t)pop_range(ES, s0, e) | ((uint32_t)pop_range(EI_, s0, e) << 10) |
                             ((uint32_t)pop_range(ED, s0, e) << 20);
#pragma unroll
        for (int c = 0; c < NC; c++) {
            const int j = lane + 64 * c;
            const bool in = j >= s0 && j <= e;
            rn[c] = in ? n : rn[c];
            rc[c] = in ? cnt : rc[c];
        }
        pos = next_one(IB, be + 2);   // be + 1 is outside the block
    }
    auto frac_at = [&](int j, int c, double &vs, double &vi, double &vd) {
        if (rn[c]) {
            const double n = (double)rn[c];
            vs = (double)(rc[c] & 1023u) / n;
            vi = (double)((rc[c] >> 10) & 1023u) / n;
            vd = (double)(rc[c] >> 20) / n;
        } else {
            vs = bit_at(ES, c) ? 1.0 : 0.0;
            vi = bit_at(EI_, c) ? 1.0 : 0.0;
            vd = bit_at(ED, c) ? 1.0 : 0.0;
        }
        if (a.keep_frac) {
            double *fo = a.s.frac + 3ull * (qs + rwx + j);
            fo[0] = vs; fo[1] = vi; fo[2] = vd;
        }
    };

    if constexpr (PAIRS) {
        // every lane stays in the loop: the appends are wavefront collectives
#pragma unroll
        for (int c = 0; c < NC; c++) {
            if (64 * c >= len) break;
            const int j = lane + 64 * c;
            double vs = 0.0, vi = 0.0, vd = 0.0;
            const bool in = j < len;
            if (in) frac_at(j, c, vs, vi, vd);
            const bool emit = in && !(fl[c] & BF_MASKED) && !bit_at(SM, c);
            emit_pairs(a, pr, rg, flags, neg, base, rwx, rwy, hx, hy, (uint16_t)(rwx + (in ? j : 0)), qv[c], vs, vi,
                       vd, emit);
        }
        __builtin_amdgcn_wave_barrier();
        return;
    }

    uint32_t nid = 0;
#pragma unroll
    for (int c = 0; c < NC; c++) {
        const int j = lane + 64 * c;
        bool em = false;
        if (j < len) {
            double vs, vi, vd;
            frac_at(j, c, vs, vi, vd);
            const uint16_t bp = (uint16_t)(rwx + j);
            if (!(fl[c] & BF_MASKED) && !bit_at(SM, c)) {
                // outside BAQ blocks every value is exactly 0 or 1 (limbs 0, 0 and 0 or 2^26)
                Limbs f0, f1, f2;
                auto unit = [](double v) { return v == 0.0 || v == 1.0; };
                if (wave_all(unit(vs) && unit(vi) && unit(vd))) {
                    f0 = {0, 0, vs == 1.0 ? 1ull << 26 : 0ull};
                    f1 = {0, 0, vi == 1.0 ? 1ull << 26 : 0ull};
                    f2 = {0, 0, vd == 1.0 ? 1ull << 26 : 0ull};
                } else {
                    f0 = to_limbs(vs, a.frac_f64); f1 = to_limbs(vi, a.frac_f64); f2 = to_limbs(vd, a.frac_f64);
                }
                em = true;
                emit_base_wave(a, L, rg, ncodes, flags, neg, base, rwx, rwy, hx, hy, bp, qv[c], f0, f1, f2);
            }
        }
        nid += (uint32_t)__popcll(__ballot(em));
    }
    // the quality-table I and D keys are the same for every base of the read: one count update per read
    if (lane == 0 && nid) atomicAdd(&L.cnt[64], nid);
    // the stage is reused by the wave's next read
    __builtin_amdgcn_wave_barrier();
}

// one workgroup per (read group, tile of a.tile reads of that group).  Two 12-wave workgroups share a
// CU (their LDS fits twice in the CU's 160 KB), six waves per SIMD, for which the registers are held
// to 80 per lane; eight waves per workgroup left four waves per SIMD, sixteen spilled.
template <bool PAIRS>
__global__ void __launch_bounds__(64 * COV_WAVES) __attribute__((amdgpu_waves_per_eu(6, 8))) k_cov(CovArgs a)
{
    extern __shared__ unsigned long long lds_raw[];
    __shared__ uint32_t pr_used[3], pr_bound, pr_over;
    __shared__ uint64_t pr_base[3];
    const uint32_t ncodes = a.ncodes;
    const uint32_t E = E_YID + 2 * ncodes;
    const uint32_t MW = (a.mcodes * 64 + 1) / 2;    // words of packed cycle-M counts
    LdsAcc L;
    L.p0 = lds_raw;
    L.p1 = lds_raw + E;
    L.p2 = lds_raw + 2 * E;
    L.cnt = (uint32_t *)(lds_raw + 3 * E);          // E entries, then 64 dump words
    L.mcnt = L.cnt + E + 64;                        // MW words, then 64 dump words
    L.mcodes = a.mcodes;
    L.dump = E;
    L.mdump = MW;
    uint8_t *stage = PAIRS ? (uint8_t *)lds_raw + (threadIdx.x >> 6) * 256
                           : (uint8_t *)(L.mcnt + MW + 64) + (threadIdx.x >> 6) * 256;
    const PairRegion pr{pr_used, pr_base};

    // which read group / tile this block covers
    const uint32_t nrg = a.t.num_rg;
    const uint32_t blk = blockIdx.x;
    if (blk >= a.tile_offset[nrg]) return;
    uint32_t lo_rg = 0, hi_rg = nrg - 1;
    while (lo_rg < hi_rg) {      // last rg with tile_offset[rg] <= blk
        const uint32_t mid = (lo_rg + hi_rg + 1) >> 1;
        if (a.tile_offset[mid] <= blk) lo_rg = mid; else hi_rg = mid - 1;
    }
    const uint32_t rg = lo_rg;
    const uint32_t tile = blk - a.tile_offset[rg];
    const uint32_t first = a.rg_offset[rg] + tile * a.tile;
    const uint32_t last = min(a.rg_offset[rg + 1], first + a.tile);

    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    if constexpr (PAIRS) {
        // this tile's slice of each table's pair array: 3 slots per window base bound its pairs
        if (threadIdx.x == 0) { pr_bound = 0; pr_over = 0; }
        if (threadIdx.x < 3) pr_used[threadIdx.x] = 0;
        __syncthreads();
        uint32_t part = 0;
        for (uint32_t li = first + threadIdx.x; li < last; li += blockDim.x) {
            const uint32_t rw = a.s.rwc[a.rg_list[li]];
            part += (uint32_t)((int)hi16(rw) - (int)lo16(rw) + 1);
        }
        part = (uint32_t)wave_sum(part);
        if (lane == 0 && part) atomicAdd(&pr_bound, part);
        __syncthreads();
        if (threadIdx.x < 3) {
            const uint64_t want = 3ull * pr_bound;
            const uint32_t b = atomicAdd(&a.pcursor[threadIdx.x], (uint32_t)want);
            pr_base[threadIdx.x] = b;
            if ((uint64_t)b + want > a.pcap) {
                atomicOr(a.perr, 1u);
                pr_over = 1;
            }
        }
        __syncthreads();
        // a slice past the capacity is never written; the engine reports the overflow
        if (pr_over) return;
    } else {
        for (uint32_t e = threadIdx.x; e < E; e += blockDim.x) { L.p0[e] = 0; L.p1[e] = 0; L.p2[e] = 0; L.cnt[e] = 0; }
        for (uint32_t w = threadIdx.x; w < MW; w += blockDim.x) L.mcnt[w] = 0;
        __syncthreads();
    }

    // each wave takes a contiguous run of the tile's reads, 64 reads' fields at a time
    const uint32_t per = (last - first + COV_WAVES - 1) / COV_WAVES;
    const uint32_t w0 = first + wave * per, w1 = min(last, w0 + per);
    for (uint32_t g = w0; g < w1; g += 64) {
        const uint32_t cnt = min(64u, w1 - g);
        const ReadMeta ml = meta_lane(a, g + lane, lane < cnt);
        const int len_l = (int)hi16(ml.rw) - (int)lo16(ml.rw) + 1;
        if (!wave_any(lane < cnt && len_l > 128)) {
            // reads of up to 128 window bases: software-pipelined, the next read's inputs are
            // in flight while this one is processed
            ReadMeta m = meta_of(ml, 0);
            CovLoads<2> cur, nxt;
            cov_load<2>(a, m, cur);
            for (uint32_t k = 0; k < cnt; k++) {
                const ReadMeta mn = meta_of(ml, k + 1 < cnt ? k + 1 : k);
                cov_read_wave<PAIRS, 2, true>(a, L, pr, stage, m, cur, rg, ncodes, mn, nxt);
                m = mn;
                cur = nxt;
            }
            continue;
        }
        for (uint32_t k = 0; k < cnt; k++) {
            const ReadMeta m = meta_of(ml, k);
            const int len = (int)hi16(m.rw) - (int)lo16(m.rw) + 1;
            if (len <= 128) {
                CovLoads<2> x;
                cov_load<2>(a, m, x);
                cov_read_wave<PAIRS, 2, false>(a, L, pr, stage, m, x, rg, ncodes, m, x);
            } else if (len <= 256) {
                CovLoads<4> x;
                cov_load<4>(a, m, x);
                cov_read_wave<PAIRS, 4, false>(a, L, pr, stage, m, x, rg, ncodes, m, x);
            } else if (lane == 0) {
                cov_read_serial<PAIRS>(a, L, pr, m.r, rg, ncodes);
            }
        }
    }
    __syncthreads();
    if constexpr (PAIRS) {
        // unused slots of the slice sort last and are dropped by the first radix pass
        const uint64_t want = 3ull * pr_bound;
        for (int t = 0; t < 3; t++)
            for (uint64_t i = pr_used[t] + threadIdx.x; i < want; i += blockDim.x) a.pkey[t][pr_base[t] + i] = PAIR_SENTINEL;
        return;
    }

    // flush the tile's entries into the resident tables (a D entry's count is its I entry's)
    for (uint32_t e = threadIdx.x; e < E; e += blockDim.x) {
        const uint32_t ce = e == 65 ? 64u
                          : (e >= E_CID + 64 && e < E_YID) ? e - 64
                          : (e >= E_YID + ncodes) ? e - ncodes : e;
        const uint32_t c = L.cnt[ce];
        if (!c) continue;
        int tab;
        uint32_t key;
        if (e < 64) { tab = 0; key = 0u | (e << 2) | (rg << 10); }
        else if (e < 66) { tab = 0; key = (e - 63) | (45u << 2) | (rg << 10); }
        else if (e < E_CID) {
            const uint32_t q = (e - E_CM) >> 4, bits = (e - E_CM) & 15;
            tab = 2; key = 0u | (((bits << 4) | 2) << 2) | (q << 13) | (rg << 21);
        } else if (e < E_YID) {
            const uint32_t ev = 1 + (e - E_CID) / 64, bits = (e - E_CID) & 63;
            tab = 2; key = ev | (((bits << 4) | 3) << 2) | (45u << 13) | (rg << 21);
        } else {
            const uint32_t ev = 1 + (e - E_YID) / ncodes, code = (e - E_YID) % ncodes;
            tab = 1; key = ev | (code << 2) | (45u << 12) | (rg << 20);
        }
        const Limbs v = {L.p0[e], L.p1[e], L.p2[e]};
        global_add(a.t, tab, key, c, v);
    }
    for (uint32_t i = threadIdx.x; i < a.mcodes * 64; i += blockDim.x) {
        const uint32_t c = (L.mcnt[i >> 1] >> (16 * (i & 1))) & 0xffff;
        if (!c) continue;
        const uint32_t q = i / a.mcodes, code = i % a.mcodes;
        atomicAdd((unsigned long long *)&a.t.count[1][0u | (code << 2) | (q << 12) | (rg << 20)],
                  (unsigned long long)c);
    }
}

// wave-aggregated append: one atomic per distinct key in the wavefront.
// Returns this lane's slot (counter value before its own increment).
__device__ inline uint32_t wave_append(uint32_t *counters, uint32_t key, bool active)
{
    const uint32_t lane = threadIdx.x & 63;
    uint32_t slot = 0;
    uint64_t pending = __ballot(active);
    while (pending) {
        const int leader = __builtin_ctzll(pending);
        const uint32_t k = __shfl(key, leader);
        const uint64_t same = __ballot(active && key == k);
        uint32_t base = 0;
        if ((int)lane == leader) base = atomicAdd(&counters[k], (uint32_t)__popcll(same));
        base = __shfl(base, leader);
        if (active && key == k) slot = base + (uint32_t)__popcll(same & ((1ull << lane) - 1));
        pending &= ~same;
    }
    return slot;
}

// graph-safe clear of a small counter array
__global__ void __launch_bounds__(256) k_fill_u32c(uint32_t *p, uint32_t n, uint32_t v)
{
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) p[i] = v;
}

// per-read-group read lists: counts -> offsets -> scatter.  A fixed grid of RG_BLOCKS workgroups
// strides over the reads; each keeps its per-read-group counts in LDS and touches a global
// counter once per read group.
static constexpr uint32_t RG_BLOCKS = 512;

__global__ void __launch_bounds__(256) k_rg_count(uint32_t n, const uint32_t *live, const uint8_t *status,
                                                  const uint32_t *read_group, uint32_t num_rg, uint32_t *rg_count,
                                                  uint32_t *errors)
{
    __shared__ uint32_t cnt[256];
    __shared__ uint32_t nbad;
    for (uint32_t g = threadIdx.x; g < 256; g += blockDim.x) cnt[g] = 0;
    if (threadIdx.x == 0) nbad = 0;
    __syncthreads();
    if (live) n = live[0];
    for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
        const uint32_t r = base + threadIdx.x;
        const bool act = r < n && status[r] == RS_ACTIVE;
        const uint32_t rg = act ? read_group[r] : 0;
        const bool bad = act && rg >= num_rg;
        if (bad) atomicAdd(&nbad, 1u);
        wave_append(cnt, rg, act && !bad);
    }
    __syncthreads();
    for (uint32_t g = threadIdx.x; g < num_rg; g += blockDim.x)
        if (cnt[g]) atomicAdd(&rg_count[g], cnt[g]);
    if (threadIdx.x == 0 && nbad && errors) atomicAdd(errors, nbad);
}

__global__ void __launch_bounds__(256) k_rg_offsets(uint32_t num_rg, const uint32_t *rg_count, uint32_t *rg_offset,
                                                    uint32_t *tile_offset, uint32_t *rg_cursor, uint32_t tile)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    uint32_t o = 0, t = 0;
    for (uint32_t g = 0; g < num_rg; g++) {
        rg_offset[g] = o;
        tile_offset[g] = t;
        rg_cursor[g] = 0;
        o += rg_count[g];
        t += (rg_count[g] + tile - 1) / tile;
    }
    rg_offset[num_rg] = o;
    tile_offset[num_rg] = t;
}

__global__ void __launch_bounds__(256) k_rg_scatter(uint32_t n, const uint32_t *live, const uint8_t *status,
                                                    const uint32_t *read_group, uint32_t num_rg,
                                                    const uint32_t *rg_offset, uint32_t *rg_cursor, uint32_t *rg_list)
{
    __shared__ uint32_t cnt[256], first[256];
    for (uint32_t g = threadIdx.x; g < 256; g += blockDim.x) cnt[g] = 0;
    __syncthreads();
    if (live) n = live[0];
    const uint32_t stride = gridDim.x * blockDim.x;
    // this workgroup's reads per read group, then one reservation per read group
    for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += stride) {
        const uint32_t r = base + threadIdx.x;
        const bool act = r < n && status[r] == RS_ACTIVE;
        const uint32_t rg = act ? read_group[r] : 0;
        wave_append(cnt, rg, act && rg < num_rg);
    }
    __syncthreads();
    for (uint32_t g = threadIdx.x; g < num_rg; g += blockDim.x) {
        first[g] = cnt[g] ? rg_offset[g] + atomicAdd(&rg_cursor[g], cnt[g]) : 0;
        cnt[g] = 0;
    }
    __syncthreads();
    for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += stride) {
        const uint32_t r = base + threadIdx.x;
        const bool act = r < n && status[r] == RS_ACTIVE;
        const uint32_t rg = act ? read_group[r] : 0;
        const bool ok = act && rg < num_rg;
        const uint32_t slot = wave_append(cnt, rg, ok);
        if (ok) rg_list[first[rg] + slot] = r;
    }
}

} // namespace fpd

// LDS layout of one covariate workgroup for reads up to max_read_len within the device's
// per-workgroup limit: cycle codes beyond what fits go to the resident tables directly
static int g_cov_lds_limit = -1;

static void cov_layout(uint32_t max_read_len, uint32_t *ncodes, uint32_t *mcodes, size_t *bytes)
{
    const int limit = g_cov_lds_limit > 0 ? g_cov_lds_limit : 64 * 1024;
    const uint32_t budget = (uint32_t)limit - fpd::STAGE_BYTES - 256;   // 256: the kernel's static LDS
    uint32_t nc = 2 * max_read_len + 2, mc = 2 * max_read_len + 2;
    auto need = [&](uint32_t n, uint32_t m) {   // entries, packed counts, 2 x 64 dump words
        return (fpd::E_YID + 2 * n) * fpd::LDS_ENTRY_BYTES + 4 * ((m * 64 + 1) / 2) + 512;
    };
    while (need(nc, mc) > budget && mc > 0) mc = mc > 16 ? mc - 16 : 0;
    while (need(nc, mc) > budget && nc > 0) nc = nc > 16 ? nc - 16 : 0;
    *ncodes = nc;
    *mcodes = mc;
    *bytes = need(nc, mc) + fpd::STAGE_BYTES;
}

// once per device, before any launch (and outside stream capture): the per-workgroup LDS limit and the
// matching dynamic-LDS attribute of the covariate kernel
extern "C" hipError_t fpd_cov_prepare()
{
    int dev = 0, v = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    e = hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
    if (e != hipSuccess) return e;
    g_cov_lds_limit = v > 0 ? v : 64 * 1024;
    (void)hipFuncSetAttribute((const void *)fpd::k_cov<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              g_cov_lds_limit);
    return hipSuccess;
}

extern "C" hipError_t fpd_launch_cov(const fpd::BatchDev *b, const fpd::ReadStateDev *s, const fpd::TablesDev *t,
                                     uint32_t frac_f64, uint32_t keep_frac, uint32_t *batch, uint32_t *rg_buf,
                                     uint32_t max_read_len, uint32_t *const *pkey, double *const *pval,
                                     uint32_t *pcursor, uint64_t pcap, hipStream_t st)
{
    const uint32_t n = b->num_reads;
    if (n == 0) return hipSuccess;
    const uint32_t nrg = t->num_rg;
    // rg_buf layout: count[nrg] | cursor[nrg] | offset[nrg+1] | tile_offset[nrg+1] | list[n]
    uint32_t *cnt = rg_buf, *cur = rg_buf + nrg, *off = rg_buf + 2 * nrg, *toff = rg_buf + 3 * nrg + 1;
    uint32_t *list = rg_buf + 4 * nrg + 2;
    hipLaunchKernelGGL(fpd::k_fill_u32c, dim3(1), dim3(256), 0, st, cnt, nrg, 0u);
    const uint32_t rg_blocks = std::min<uint32_t>((n + 255) / 256, fpd::RG_BLOCKS);
    hipLaunchKernelGGL(fpd::k_rg_count, dim3(rg_blocks), dim3(256), 0, st, n, b->live, s->status, b->read_group,
                       nrg, cnt, (uint32_t *)nullptr);   // the prep kernel counts such reads (BW_BAD_RG)
    // tiles of 1024 reads for large batches; smaller ones (down to 128) keep a small batch spread
    // over the whole device
    const uint32_t tile = std::max<uint32_t>(fpd::TILE_MIN, std::min<uint32_t>(fpd::TILE, (n / 1024 + 63) & ~63u));
    hipLaunchKernelGGL(fpd::k_rg_offsets, dim3(1), dim3(64), 0, st, nrg, cnt, off, toff, cur, tile);
    hipLaunchKernelGGL(fpd::k_rg_scatter, dim3(rg_blocks), dim3(256), 0, st, n, b->live, s->status,
                       b->read_group, nrg, off, cur, list);
    fpd::CovArgs a{};
    a.b = *b; a.s = *s; a.t = *t; a.frac_f64 = frac_f64; a.keep_frac = keep_frac;
    a.rg_list = list; a.rg_offset = off; a.tile_offset = toff; a.tile = tile;
    const uint32_t blocks = (n + tile - 1) / tile + nrg;
    if (pkey) {
        // pair mode (sort aggregation): pairs out instead of table updates; LDS holds the base stages
        for (int i = 0; i < 3; i++) { a.pkey[i] = pkey[i]; a.pval[i] = pval[i]; }
        a.pcursor = pcursor;
        a.perr = batch + fpd::BW_OVERFLOW;
        a.pcap = pcap;
        hipLaunchKernelGGL(fpd::k_cov<true>, dim3(blocks), dim3(64 * fpd::COV_WAVES), fpd::STAGE_BYTES, st, a);
        return hipGetLastError();
    }
    size_t lds = 0;
    cov_layout(max_read_len, &a.ncodes, &a.mcodes, &lds);
    hipLaunchKernelGGL(fpd::k_cov<false>, dim3(blocks), dim3(64 * fpd::COV_WAVES), lds, st, a);
    return hipGetLastError();
}
