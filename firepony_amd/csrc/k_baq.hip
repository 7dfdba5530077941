/*
 * K_baq: the BAQ glocal banded HMM (device/baq.cu:214-697) plus MAP quality,
 * capping and BQ-tag recoding (baq.cu:842-1081).
 *
 * One lane per BAQ read.  The band of the current row lives in registers in
 * band-relative coordinates r = k - i + bw, which makes every dependency a
 * fixed offset: M[i][r] <- row i-1 at r, I[i][r] <- row i-1 at r+1,
 * D[i][r] <- row i at r-1 (forward), mirrored for backward.  The forward pass
 * spills each scaled row's M and I states to a per-wave scratch slot,
 * interleaved by lane so every store/load moves 64 consecutive doubles; the
 * backward pass recomputes nothing and consumes those rows in reverse for the
 * MAP step, so the backward matrix is never stored.
 *
 * Arithmetic is FP64 with the reference's operation order and association
 * (this file must be compiled with -ffp-contract=off).  qual2prob and the gap
 * open probability come from host tables computed with the host libm; the MAP
 * quality (int)(-4.343*log(1-max)+.499) is evaluated through host-computed
 * thresholds on 1-max, which reproduces the host libm bit for bit.
 */
#include "fp_device.h"

#include <algorithm>

namespace fpd {

__constant__ double c_qual2prob[256];
__constant__ double c_map_thr[128];     // thresholds T[v], v = 5..101
__constant__ double c_gap_open;

static constexpr double EM = 0.33333333333;
static constexpr double EI = 0.25;
static constexpr double GEP = 0.1;

// compile-time form of a row: Steady<true> rows carry no band or N masks
template <bool B> struct Steady { static constexpr bool value = B; };

// Starting guess in [4, 100] for the MAP quality of y > 0: -4.343 ln(y) + 0.499 in single precision
// (one hardware log2, -4.343 ln 2 = -3.0103; y below the float range gives +inf, clamped to 100)
__device__ inline int map_guess(double y)
{
    const float g = -3.0103f * __builtin_amdgcn_logf((float)y) + 0.499f;
    return (int)fminf(fmaxf(g, 4.0f), 100.0f);
}

// MAP quality from y = 1 - max/sum (baq.cu:673-674).  The exact value is the
// largest v in [5, 101] with y <= T[v] (T decreasing, computed with the host
// libm); the guess only provides the starting point, which the threshold
// comparisons then correct (from any start), so the result never depends on
// the device's logarithm.
__device__ inline uint8_t map_quality(double y)
{
    if (y == 0.0) return 4;                 // log(0): (int)+inf is INT_MIN on the host
    int k = map_guess(y);
    while (k >= 5 && !(y <= c_map_thr[k])) k--;
    while (k + 1 <= 101 && y <= c_map_thr[k + 1]) k++;
    if (k <= 4) return 4;
    if (k >= 101) return 99;
    return (uint8_t)k;
}

struct BaqArgs {
    BatchDev b;
    GenomeDev g;
    ReadStateDev s;
    const uint32_t *list;     // BAQ read ids of this class
    const uint32_t *count;    // number of ids
    double *scratch;          // per-wave slots
    uint64_t slot_doubles;    // doubles per wave slot
    uint32_t rows_cap;        // >= max query length
    uint32_t wmax_runtime;    // band cells for the generic kernel
    unsigned long long *cells;  // total band cells evaluated (algorithmic work)
    uint32_t *next;           // band kernels: chunk counter for dynamic chunk order (zero at launch), or null
};

// WPE: waves per SIMD the register budget is sized for (1: band fully in
// registers; 2: twice the latency hiding at the cost of a few spills)
template <int WMAX, int WPE>
__global__ void __launch_bounds__(64, WPE) k_baq_hmm(BaqArgs a)
{
    constexpr bool REG = WMAX > 0;
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t wave = blockIdx.x;
    const uint32_t nwaves = gridDim.x;
    const uint32_t total = *a.count;
    const int W = REG ? WMAX : (int)a.wmax_runtime;

    double *slot = a.scratch + (uint64_t)wave * a.slot_doubles;
    double *Fst = slot;                                              // rows x W x 2 x 64
    double *Sst = slot + (uint64_t)(a.rows_cap + 1) * W * 2 * 64;    // (rows+2) x 64
    double *Gst = Sst + (uint64_t)(a.rows_cap + 2) * 64;             // generic rows: 3 x W x 64

    double Mr[REG ? WMAX : 1], Ir[REG ? WMAX : 1], Dr[REG ? WMAX : 1];

    auto M = [&](int r) -> double & { if constexpr (REG) return Mr[r]; else return Gst[(0 * W + r) * 64 + lane]; };
    auto I = [&](int r) -> double & { if constexpr (REG) return Ir[r]; else return Gst[(1 * W + r) * 64 + lane]; };
    auto D = [&](int r) -> double & { if constexpr (REG) return Dr[r]; else return Gst[(2 * W + r) * 64 + lane]; };
    auto F = [&](int i, int r, int st) -> double & { return Fst[(((uint64_t)i * W + r) * 2 + st) * 64 + lane]; };
    auto S = [&](int i) -> double & { return Sst[(uint64_t)i * 64 + lane]; };

    for (uint32_t chunk = wave; (uint64_t)chunk * 64 < total; chunk += nwaves) {
        const uint32_t idx = chunk * 64 + lane;
        if (idx >= total) continue;
        const uint32_t rd = a.list[idx];
        const BatchDev &b = a.b;

        // ---- setup (baq.cu:260-354)
        const uint32_t hw = a.s.hmmw[rd];
        const int hx = (int16_t)lo16(hw), hy = (int16_t)hi16(hw);
        const uint32_t rwv = a.s.rwc[rd];
        const int qStart = lo16(rwv);
        const int qLen = (int)hi16(rwv) - qStart + 1;
        const int bw = a.s.bandwidth[rd];
        const uint32_t astart = b.alignment_start[rd];
        const uint32_t chrom = b.chromosome[rd];
        const int refLenW = hy - hx + 1;
        const uint32_t ahx = hx < 0 ? (uint32_t)(-hx) : (uint32_t)hx;
        const bool clamp = hx < 0 && astart < ahx;
        const int refLen = clamp ? refLenW - 1 : refLenW;
        const uint32_t hoff = clamp ? 0u : (uint32_t)(astart + hx);

        const double GOP = c_gap_open;
        const double sM = 1.0 / (2 * qLen + 2);
        const double sI = sM;
        const double bM = (1 - GOP) / refLenW;
        const double bI = GOP / refLenW;
        const double m0 = (1 - GOP - GOP) * (1 - sM);
        const double m1 = GOP * (1 - sM);
        const double m2 = m1;
        const double m3 = (1 - GEP) * (1 - sI);
        const double m4 = GEP * (1 - sI);
        const double m6 = 1 - GEP;
        const double m8 = GEP;
        const double EIm1 = EI * m1;
        const double EIm4 = EI * m4;

        const uint64_t qbase = (uint64_t)b.read_start[rd] + qStart;
        const uint32_t qqual = b.qual_start[rd] + qStart;
        const uint32_t *cw = a.g.words + a.g.word_offset[chrom];
        const uint32_t clen = a.g.length[chrom];
        auto refbase = [&](int idx0) -> uint8_t {   // 0-based index into the HMM window
            if (idx0 < 0 || idx0 >= refLen) return B_N;
            const uint32_t p = hoff + (uint32_t)idx0;
            if (p >= clen) return B_N;
            return (uint8_t)((cw[p >> 3] >> ((p & 7) * 4)) & 0xf);
        };
        // 15-cell bands keep the band's reference bases as 16 packed nibbles
        constexpr bool WIN = REG && WMAX <= 16;
        uint64_t win = 0;
        auto win_fill = [&](int idx0) {
            win = 0;
            for (int j = 0; j < 16; j++) win |= (uint64_t)refbase(idx0 + j) << (4 * j);
        };

        uint64_t ncell = 0;
        // ---- forward row 1 (baq.cu:396-442)
        {
            const uint8_t qb = read_base(b.reads, qbase);
            const uint8_t qq = b.quals[qqual];
            const double p = c_qual2prob[qq < 4 ? 4 : qq];
            const double em_eq = 1 - p, em_ne = p * EM;
            const int khi = refLen < bw + 1 ? refLen : bw + 1;
            ncell += khi >= 1 ? (uint64_t)khi : 0;
            if constexpr (WIN) win_fill(1 - bw - 1);   // nibble r = base at k - 1, k = 1 - bw + r
            double sum = 0.0;
#pragma unroll
            for (int r = 0; r < (REG ? WMAX : 1) * 0 + W; r++) {
                const int k = 1 - bw + r;
                const bool in = k >= 1 && k <= khi;
                double mv = 0.0, iv = 0.0;
                if (in) {
                    const uint8_t rb = WIN ? (uint8_t)((win >> (4 * r)) & 15) : refbase(k - 1);
                    const double e = (rb == B_N || qb == B_N) ? 1.0 : (rb == qb ? em_eq : em_ne);
                    mv = e * bM;
                    iv = EI * bI;
                    sum += mv + iv;
                }
                M(r) = mv; I(r) = iv; D(r) = 0.0;
            }
            S(0) = 1.0;
            S(1) = sum;
#pragma unroll
            for (int r = 0; r < W; r++) {
                M(r) = M(r) / sum; I(r) = I(r) / sum; D(r) = D(r) / sum;
                F(1, r, 0) = M(r); F(1, r, 1) = I(r);
            }
        }
        // ---- forward rows 2..qLen (baq.cu:444-512)
        // the next row's query base and quality are loaded one row ahead
        uint8_t nqb = qLen >= 2 ? read_base(b.reads, qbase + 1) : 0;
        uint8_t nqq = qLen >= 2 ? b.quals[qqual + 1] : 0;
        for (int i = 2; i <= qLen; i++) {
            const uint8_t qb = nqb;
            const uint8_t qq = nqq;
            if (i < qLen) { nqb = read_base(b.reads, qbase + i); nqq = b.quals[qqual + i]; }
            const double p = c_qual2prob[qq < 4 ? 4 : qq];
            const double em_eq = 1 - p, em_ne = p * EM;
            const int klo = 1 > i - bw ? 1 : i - bw;
            const int khi = refLen < i + bw ? refLen : i + bw;
            if (khi >= klo) ncell += (uint64_t)(khi - klo + 1);
            if constexpr (WIN) win = (win >> 4) | ((uint64_t)refbase(i - bw - 1 + 15) << 60);
            double sum = 0.0, cM = 0.0, cD = 0.0;
#pragma unroll
            for (int r = 0; r < W; r++) {
                const int k = i - bw + r;
                const bool in = k >= klo && k <= khi;
                const double pM = M(r), pI = I(r), pD = D(r);
                const double nM = (r + 1 < W) ? M(r + 1) : 0.0;
                const double nI = (r + 1 < W) ? I(r + 1) : 0.0;
                double mv = 0.0, iv = 0.0, dv = 0.0;
                if (in) {
                    const uint8_t rb = WIN ? (uint8_t)((win >> (4 * r)) & 15) : refbase(k - 1);
                    const double e = (rb == B_N || qb == B_N) ? 1.0 : (rb == qb ? em_eq : em_ne);
                    mv = e * (m0 * pM + m3 * pI + m6 * pD);
                    iv = EI * (m1 * nM + m4 * nI);
                    dv = m2 * cM + m8 * cD;
                    sum += mv + iv + dv;
                }
                M(r) = mv; I(r) = iv; D(r) = dv;
                cM = mv; cD = dv;
            }
            S(i) = sum;
            const double inv = 1. / sum;
#pragma unroll
            for (int r = 0; r < W; r++) {
                M(r) = M(r) * inv; I(r) = I(r) * inv; D(r) = D(r) * inv;
                F(i, r, 0) = M(r); F(i, r, 1) = I(r);
            }
        }
        // ---- s[qLen+1] (baq.cu:514-525)
        {
            double sum = 0.0;
            const int klo = 1 > qLen - bw ? 1 : qLen - bw;
            const int khi = refLen < qLen + bw ? refLen : qLen + bw;
#pragma unroll
            for (int r = 0; r < W; r++) {
                const int k = qLen - bw + r;
                if (k >= klo && k <= khi) sum += M(r) * sM + I(r) * sI;
            }
            S(qLen + 1) = sum;
        }
        atomicAdd(a.cells, (unsigned long long)ncell);

        // ---- backward + MAP (baq.cu:528-681)
        const uint32_t qout = qqual;
        auto map_row = [&](int i) {
            double sum = 0.0, mx = 0.0;
            int32_t st = -1;
#pragma unroll
            for (int r = 0; r < W; r++) {
                const int k = i - bw + r;
                double z = F(i, r, 0) * M(r);
                sum += z;
                if (z > mx) { mx = z; st = (k - 1) << 2 | 0; }
                z = F(i, r, 1) * I(r);
                sum += z;
                if (z > mx) { mx = z; st = (k - 1) << 2 | 1; }
            }
            mx /= sum;
            a.s.baq[qout + i - 1] = map_quality(1.0 - mx);
            a.s.baq_state[qout + i - 1] = (uint32_t)st;
        };
        {
            const double sq = S(qLen), sq1 = S(qLen + 1);
            const int klo = 1 > qLen - bw ? 1 : qLen - bw;
            const int khi = refLen < qLen + bw ? refLen : qLen + bw;
            const double vM = sM / sq / sq1, vI = sI / sq / sq1;
#pragma unroll
            for (int r = 0; r < W; r++) {
                const int k = qLen - bw + r;
                const bool in = k >= klo && k <= khi;
                M(r) = in ? vM : 0.0; I(r) = in ? vI : 0.0;
            }
            map_row(qLen);
        }
        if constexpr (WIN) win_fill(qLen - bw);
        uint8_t bqb = qLen >= 2 ? read_base(b.reads, qbase + qLen - 1) : 0;
        uint8_t bqq = qLen >= 2 ? b.quals[qqual + qLen - 1] : 0;
        for (int i = qLen - 1; i >= 1; i--) {
            // nibble r = base at k = i - bw + r
            if constexpr (WIN) win = (win << 4) | (uint64_t)refbase(i - bw);
            const double y = (i > 1) ? 1. : 0.;
            const uint8_t qb = bqb;
            const uint8_t qq = bqq;
            if (i > 1) { bqb = read_base(b.reads, qbase + i - 1); bqq = b.quals[qqual + i - 1]; }
            const double p = c_qual2prob[qq < 4 ? 4 : qq];
            const double em_eq = 1 - p, em_ne = p * EM;
            const int klo = 1 > i - bw ? 1 : i - bw;
            const int khi = refLen < i + bw ? refLen : i + bw;
            double cD = 0.0;
#pragma unroll
            for (int r = W - 1; r >= 0; r--) {
                const int k = i - bw + r;
                const bool in = k >= klo && k <= khi;
                const double pM = M(r);
                const double pIl = r >= 1 ? I(r - 1) : 0.0;
                double mv = 0.0, iv = 0.0, dv = 0.0;
                if (in) {
                    double e;
                    if (k >= refLen) e = 0;
                    else {
                        const uint8_t rb = WIN ? (uint8_t)((win >> (4 * r)) & 15) : refbase(k);
                        e = ((rb == B_N || qb == B_N) ? 1.0 : (rb == qb ? em_eq : em_ne)) * pM;
                    }
                    mv = e * m0 + EIm1 * pIl + m2 * cD;
                    iv = e * m3 + EIm4 * pIl;
                    dv = (e * m6 + m8 * cD) * y;
                }
                // the previous row's D is never read by the backward pass: only the carry
                M(r) = mv; I(r) = iv;
                cD = dv;
            }
            const double yy = 1.0 / S(i);
#pragma unroll
            for (int r = 0; r < W; r++) { M(r) = M(r) * yy; I(r) = I(r) * yy; }
            map_row(i);
        }
    }
}

// ---------------------------------------------------------------------------
// The event walk of cap_baq_qualities (baq.cu:870-1056).
//
// Visits the read positions of the BAQ window in order, from the first event
// inside the window (baq_start) to the end of the window, calling
// visit(rb, kind, expectedPos) with kind CAP_M (match/mismatch: the MAP state
// must be a match state at expectedPos), CAP_I (insertion: the original
// quality is kept) or CAP_S (soft clip: the MAP quality is kept).  Returns the
// reference's check: the window holds exactly its M and I events.  The
// reference applies its update pass only when the check passes, and then that
// pass visits every window position exactly once; a failed check keeps the
// original qualities across the window.
enum : uint32_t { CAP_M = 0, CAP_I = 1, CAP_S = 3 };

template <class V>
__device__ inline bool cap_walk(const BatchDev &b, const ReadStateDev &s, uint32_t r, V visit)
{
    const uint32_t rw = s.rwc[r];
    const uint16_t rwx = lo16(rw), rwy = hi16(rw);
    const uint32_t *ops = b.cigars + b.cigar_start[r];
    const uint32_t nops = b.cigar_len[r];
    const int hx = (int16_t)lo16(s.hmmw[r]);
    const uint16_t fwx = lo16(s.refw[r]);
    const uint32_t astart = b.alignment_start[r];

    int16_t reference_shift, read_shift;
    const uint32_t ahx = hx < 0 ? (uint32_t)(-hx) : (uint32_t)hx;
    if (hx < 0 && astart < ahx) { reference_shift = (int16_t)ahx; read_shift = (int16_t)astart; }
    else { reference_shift = 0; read_shift = 0; }
    const int16_t refOffset = (int16_t)(hx - (int)fwx + reference_shift);

    // baq_start: event index of the first read position inside the window
    uint32_t baq_start = 0;
    {
        uint32_t e = 0;
        uint16_t ro = 0;
        for (uint32_t i = 0; i < nops; i++) {
            const uint32_t op = op_code(ops[i]), l = op_len(ops[i]);
            if (op_reads(op)) {
                if ((uint32_t)ro + l > rwx) { baq_start = e + (rwx > ro ? rwx - ro : 0); break; }
                ro = (uint16_t)(ro + l);
            }
            e += l;
        }
    }
    uint32_t readI = 0, refI = 0, numD = 0;
    uint32_t e = 0;
    uint16_t ro = 0;
    bool stop = false;
    for (uint32_t oi = 0; oi < nops && !stop; oi++) {
        const uint32_t op = op_code(ops[oi]), l = op_len(ops[oi]);
        const uint32_t t = op_event(op);
        const bool reads = op_reads(op);
        for (uint32_t k = 0; k < l; k++) {
            const uint32_t i = e + k;
            if (i < baq_start) continue;
            const uint16_t rb = reads ? (uint16_t)(ro + k) : NONE16;
            if (rb != NONE16) {
                if (rb < rwx) continue;
                if (rb > rwy) { stop = true; break; }
            }
            if (t == EV_S) {
                visit(rb, CAP_S, 0u);
                refI++;
            } else if (t == EV_I) {
                visit(rb, CAP_I, 0u);
                readI++;
            } else if (t == EV_D) {
                refI++; numD++;
            } else {
                visit(rb, CAP_M, refI - refOffset + (i - numD - baq_start - readI) + read_shift);
                readI++; refI++;
            }
        }
        if (reads) ro = (uint16_t)(ro + l);
        e += l;
    }
    return readI == (uint32_t)(rwy - rwx + 1);
}

// a window position's cap-walk code in 32 bits: kind in the top two bits, the expected MAP state
// below.  An expected state that does not fit is stored as CAP_EP_MASK, which no MAP state with
// state & 3 == 0 reaches (band positions are far below 2^30), so the comparison in cap_tag keeps
// its outcome.
static constexpr uint32_t CAP_EP_MASK = 0x3fffffffu;
__device__ inline uint32_t cap_code(uint32_t kind, uint32_t ep)
{
    return (kind << 30) | (ep < CAP_EP_MASK ? ep : CAP_EP_MASK);
}

// the capped quality of one window position (cap_baq_qualities) recoded into its BQ tag (recode, baq.cu:1058-1081)
__device__ inline uint8_t cap_tag(uint32_t kind, uint32_t expectedPos, uint8_t oq, uint8_t mapq, uint32_t state)
{
    uint8_t v;
    if (kind == CAP_I) v = oq;
    else if (kind == CAP_S) v = mapq;
    else if ((state & 3) != 0 || (state >> 2) != expectedPos) v = 4;
    else v = mapq < oq ? mapq : oq;
    return (uint8_t)((uint8_t)(oq + 64) - v);
}

// ---------------------------------------------------------------------------
// Band kernels for the common band widths (15, 23 and 33 cells).
//
// Same arithmetic as above, laid out for occupancy: 256-lane workgroups with
// the qual2prob and MAP threshold tables in LDS; each forward row is stored as
// W (M, I) pairs through one row pointer (16-byte lane-interleaved stores);
// the next row's query base, quality, reference base and scale factor and,
// in the backward pass, the forward row that the MAP step consumes are all
// loaded one row ahead so their latency overlaps the current row's FP64 work.

// STAGE: each lane first copies its read's query bases and qualities (one u16 per window position)
// and its reference window (one byte per cell) into LDS, laid out [index][lane]; the row loops then
// read nothing from global memory but the spilled forward rows.  Per-lane global reads of a few
// bytes per row are otherwise fetched again from HBM on every row, because the forward-row stream
// evicts them from the caches in between.
template <int W, int WPE, bool STAGE, bool TAGL>
__global__ void __launch_bounds__(256, WPE) k_baq_band(BaqArgs a)
{
    constexpr int NWIN = (W + 15) / 16;       // u64 words of band reference nibbles
    constexpr int NREF = 2 * NWIN + 1;        // u32 reference words a band window can touch
    constexpr int TOP = (W - 1) - 16 * (NWIN - 1);   // last used nibble of the last window word
    constexpr uint64_t TOPMASK = TOP == 15 ? ~0ull : ((1ull << (4 * (TOP + 1))) - 1);

    __shared__ double s_q2p[256];
    __shared__ double s_thr[128];
    for (uint32_t t = threadIdx.x; t < 256; t += blockDim.x) s_q2p[t] = c_qual2prob[t];
    for (uint32_t t = threadIdx.x; t < 128; t += blockDim.x) s_thr[t] = c_map_thr[t];
    __syncthreads();

    const uint32_t lane = threadIdx.x & 63;
    const uint32_t gwave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const uint32_t nwaves = gridDim.x * (blockDim.x >> 6);
    const uint32_t total = *a.count;
    double *slot = a.scratch + (uint64_t)gwave * a.slot_doubles;
    double2 *Fst = (double2 *)slot;                                  // [row][r][lane] (M, I)
    double *Sst = slot + (uint64_t)(a.rows_cap + 1) * W * 2 * 64;    // [row][lane]
    // [row][lane] the cap walk's code (kind << 30 | expected state, capped) and then, unless TAGL, the tag
    uint32_t *Ost = (uint32_t *)(Sst + (uint64_t)(a.rows_cap + 2) * 64);
    const BatchDev &b = a.b;
    uint64_t ncell = 0;
    // per-wave staging area (STAGE), [index][lane]: rows_cap query qualities (bytes), the query
    // bases and the rows_cap + 2W reference cells as nibbles, two per byte, and (TAGL) the read's
    // BQ tags as the MAP steps produce them
    extern __shared__ unsigned char baq_stage[];
    const uint32_t base_bytes = (a.rows_cap + 1) / 2, ref_bytes = (a.rows_cap + 2 * W + 1) / 2;
    const uint32_t wave_bytes = (a.rows_cap + base_bytes + ref_bytes + (TAGL ? a.rows_cap : 0u)) * 64;
    uint8_t *rowq = baq_stage + (threadIdx.x >> 6) * wave_bytes;
    uint8_t *rowb = rowq + (uint64_t)a.rows_cap * 64;
    uint8_t *refd = rowb + (uint64_t)base_bytes * 64;
    uint8_t *tagl = refd + (uint64_t)ref_bytes * 64;
    auto refn = [&](int i) -> uint32_t { return (refd[(i >> 1) * 64 + (threadIdx.x & 63)] >> ((i & 1) * 4)) & 15; };
    // query position q as quality | base << 8
    auto rowv = [&](int q) -> uint32_t {
        const uint32_t l = threadIdx.x & 63;
        return rowq[q * 64 + l] | (((rowb[(q >> 1) * 64 + l] >> ((q & 1) * 4)) & 15) << 8);
    };

    auto qprob = [&](uint8_t q) -> double { return s_q2p[q < 4 ? 4 : q]; };
    // Compiler fence: makes a value materialise here (emits no instruction).  The row loops decode
    // the next row's prefetched inputs before their own stores; left alone the compiler sinks the
    // decode into the next row, behind those stores, where it waits for them.
    auto pin = [](auto &v) { asm volatile("" : "+v"(v)); };
    // MAP quality through the thresholds (largest v in [5, 101] with y <= T[v])
    auto mapq = [&](double y) -> uint8_t {
        if (y == 0.0) return 4;
        int k = map_guess(y);
        const double t0 = s_thr[k], t1 = s_thr[k + 1];
        if (k >= 5 && !(y <= t0)) {
            k--;
            while (k >= 5 && !(y <= s_thr[k])) k--;
        } else if (y <= t1) {
            k++;
            while (k + 1 <= 101 && y <= s_thr[k + 1]) k++;
        }
        if (k <= 4) return 4;
        if (k >= 101) return 99;
        return (uint8_t)k;
    };

    // one chunk of 64 list entries, one per lane (a lambda: lanes past the list leave it early and
    // rejoin the wave when it returns)
    auto run_chunk = [&](uint32_t chunk) {
        const uint32_t idx = chunk * 64 + lane;
        const bool act = idx < total;
        const uint32_t rd = act ? a.list[idx] : 0u;
        const uint32_t rwv = act ? a.s.rwc[rd] : 0u;
        const int qStart = lo16(rwv);
        const int qLen = act ? (int)hi16(rwv) - qStart + 1 : 0;
        // the backward loop runs on the wave's longest window, so all lanes touch the same row at once
        int qmax = qLen;
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            const int o = __shfl_xor(qmax, off);
            qmax = o > qmax ? o : qmax;
        }
        if (!act) return;

        // ---- setup (baq.cu:260-354)
        const uint32_t hw = a.s.hmmw[rd];
        const int hx = (int16_t)lo16(hw), hy = (int16_t)hi16(hw);
        const int bw = a.s.bandwidth[rd];
        const uint32_t astart = b.alignment_start[rd];
        const uint32_t chrom = b.chromosome[rd];
        const int refLenW = hy - hx + 1;
        const uint32_t ahx = hx < 0 ? (uint32_t)(-hx) : (uint32_t)hx;
        const bool clamp = hx < 0 && astart < ahx;
        const int refLen = clamp ? refLenW - 1 : refLenW;
        const int64_t hoff = clamp ? 0 : (int64_t)astart + hx;

        const double GOP = c_gap_open;
        const double sM = 1.0 / (2 * qLen + 2);
        const double sI = sM;
        const double bM = (1 - GOP) / refLenW;
        const double bI = GOP / refLenW;
        const double m0 = (1 - GOP - GOP) * (1 - sM);
        const double m1 = GOP * (1 - sM);
        const double m2 = m1;
        const double m3 = (1 - GEP) * (1 - sI);
        const double m4 = GEP * (1 - sI);
        const double m6 = 1 - GEP;
        const double m8 = GEP;
        const double EIm1 = EI * m1;
        const double EIm4 = EI * m4;

        const uint64_t qbase = (uint64_t)b.read_start[rd] + qStart;
        const uint32_t qqual = b.qual_start[rd] + qStart;
        const uint32_t *cw = a.g.words + a.g.word_offset[chrom];
        const int64_t clen = a.g.length[chrom];
        const int64_t wlast = clen > 0 ? ((clen + 7) >> 3) - 1 : 0;
        // reference word holding contig position p, index clamped into the contig (loads never branch)
        auto ref_word = [&](int64_t p) -> uint32_t {
            int64_t wi = p >> 3;
            wi = wi < 0 ? 0 : (wi > wlast ? wlast : wi);
            return cw[wi];
        };
        // cells of the HMM window [0, refLen) that lie inside the contig carry a base; others are N
        // (bitwise tests and an arithmetic select: no branch, so the row loops stay straight-line)
        auto nibble = [&](uint32_t w, int idx0) -> uint8_t {
            const uint32_t ok = (uint32_t)(idx0 >= 0) & (uint32_t)(idx0 < refLen) & (uint32_t)(hoff + idx0 < clen);
            const uint32_t nb = (w >> (((uint32_t)(hoff + idx0) & 7) * 4)) & 15;
            return (uint8_t)((nb & (0u - ok)) | (15u & (ok - 1u)));
        };

        // band reference bases: nibble r of the window = HMM window index idx0 + r
        uint64_t win[NWIN];
        auto rb_of = [&](int r) -> uint8_t { return (uint8_t)((win[r >> 4] >> (4 * (r & 15))) & 15); };
        auto win_fill = [&](int idx0) {
            if constexpr (STAGE) {
#pragma unroll
                for (int w = 0; w < NWIN; w++) {
                    uint64_t v = 0;
#pragma unroll
                    for (int j = 0; j < 16; j++)
                        if (w * 16 + j < W) v |= (uint64_t)refn(idx0 + w * 16 + j + W) << (4 * j);
                    win[w] = v;
                }
                return;
            }
            const int64_t p0 = hoff + idx0;
            const int64_t w0 = p0 >> 3;               // floor
            const int sh = (int)(p0 & 7) * 4;
            uint32_t rw[NREF];
#pragma unroll
            for (int j = 0; j < NREF; j++) rw[j] = ref_word((w0 + j) * 8);
            // valid cells [jlo, jhi)
            int64_t jlo = -idx0, jhi = W;
            if (jlo < 0) jlo = 0;
            if (refLen - idx0 < jhi) jhi = refLen - idx0;
            if (clen - p0 < jhi) jhi = clen - p0;
#pragma unroll
            for (int w = 0; w < NWIN; w++) {
                const uint64_t lo = (uint64_t)rw[2 * w] | ((uint64_t)rw[2 * w + 1] << 32);
                uint64_t v = lo >> sh;
                if (sh) v |= (uint64_t)rw[2 * w + 2] << (64 - sh);
                int64_t a0 = jlo - 16 * w, a1 = jhi - 16 * w;
                a0 = a0 < 0 ? 0 : (a0 > 16 ? 16 : a0);
                a1 = a1 < 0 ? 0 : (a1 > 16 ? 16 : a1);
                const uint64_t below1 = a1 >= 16 ? ~0ull : ((1ull << (4 * a1)) - 1);
                const uint64_t below0 = a0 >= 16 ? ~0ull : ((1ull << (4 * a0)) - 1);
                const uint64_t m = a1 > a0 ? (below1 & ~below0) : 0ull;
                v = (v & m) | (~m & 0xffffffffffffffffull);   // N = 15 in every invalid cell
                win[w] = w == NWIN - 1 ? (v & TOPMASK) : v;
            }
        };
        auto win_shr = [&](uint8_t nb) {     // band moves right: drop cell 0, append cell W-1
#pragma unroll
            for (int w = 0; w < NWIN; w++) {
                const uint64_t in = (w + 1 < NWIN) ? (win[w + 1 < NWIN ? w + 1 : w] & 15) : (uint64_t)nb;
                const int top = (w + 1 < NWIN) ? 15 : TOP;
                win[w] = (win[w] >> 4) | (in << (4 * top));
            }
        };
        auto win_shl = [&](uint8_t nb) {     // band moves left: drop cell W-1, insert cell 0
#pragma unroll
            for (int w = NWIN - 1; w >= 0; w--) {
                const uint64_t in = w > 0 ? (win[w > 0 ? w - 1 : 0] >> 60) & 15 : (uint64_t)nb;
                const uint64_t keep = (w + 1 < NWIN) ? ~0ull : TOPMASK;
                win[w] = ((win[w] << 4) | in) & keep;
            }
        };

        // expected MAP state of every window position (the cap_baq_qualities walk), staged per row;
        // the MAP step turns it into the position's final tag
        const bool good = cap_walk(b, a.s, rd, [&](uint16_t p, uint32_t kind, uint32_t ep) {
            Ost[(uint64_t)((int)p - qStart + 1) * 64 + lane] = cap_code(kind, ep);
        });

        if constexpr (STAGE) {
            // this read's inputs into the wave's LDS area (each lane writes and later reads only its own column)
#pragma unroll 4
            for (int q = 0; q < qLen; q++) rowq[q * 64 + lane] = b.quals[qqual + q];
#pragma unroll 4
            for (int q = 0; q < qLen; q += 2) {
                const uint32_t w = b.reads[(qbase + q) >> 3];
                const uint32_t b0 = (w >> (((qbase + q) & 7) * 4)) & 15;
                const uint64_t i1 = qbase + (q + 1 < qLen ? q + 1 : q);   // never past the window
                const uint32_t b1 = q + 1 < qLen ? (b.reads[i1 >> 3] >> ((i1 & 7) * 4)) & 15 : 15u;
                rowb[(q >> 1) * 64 + lane] = (uint8_t)(b0 | (b1 << 4));
            }
#pragma unroll 4
            // cells k = -bw .. qLen + W - 1 - bw at index k + W, two per byte (an odd first index also
            // fills the cell below the window, never read)
            for (int bi = (W - bw) >> 1; bi <= (qLen + 2 * W - 1 - bw) >> 1; bi++) {
                const int k0 = 2 * bi - W, k1 = k0 + 1;
                refd[bi * 64 + lane] = (uint8_t)(nibble(ref_word(hoff + k0), k0) | (nibble(ref_word(hoff + k1), k1) << 4));
            }
        }

        double M[W], I[W], D[W];

        // ---- forward row 1 (baq.cu:396-442)
        {
            uint8_t qb;
            double p;
            if constexpr (STAGE) {
                const uint32_t v = rowv(0);
                qb = (uint8_t)(v >> 8);
                p = qprob((uint8_t)(v & 255));
            } else {
                qb = read_base(b.reads, qbase);
                p = qprob(b.quals[qqual]);
            }
            const double em_eq = 1 - p, em_ne = p * EM;
            const int khi = refLen < bw + 1 ? refLen : bw + 1;
            ncell += khi >= 1 ? (uint64_t)khi : 0;
            win_fill(-bw);                       // cell r: k = 1 - bw + r, window index k - 1
            double sum = 0.0;
#pragma unroll
            for (int r = 0; r < W; r++) {
                const int k = 1 - bw + r;
                const bool in = k >= 1 && k <= khi;
                const uint8_t rb = rb_of(r);
                const double e = (rb == B_N || qb == B_N) ? 1.0 : (rb == qb ? em_eq : em_ne);
                const double mv = in ? e * bM : 0.0;
                const double iv = in ? EI * bI : 0.0;
                sum += in ? mv + iv : 0.0;
                M[r] = mv; I[r] = iv; D[r] = 0.0;
            }
            Sst[64 + lane] = sum;
            double2 *fr = Fst + (uint64_t)1 * W * 64 + lane;
#pragma unroll
            for (int r = 0; r < W; r++) {
                M[r] = M[r] / sum; I[r] = I[r] / sum; D[r] = D[r] / sum;
                fr[r * 64] = make_double2(M[r], I[r]);
            }
        }
        // ---- forward rows 2..qLen (baq.cu:444-512)
        // Each row issues the loads of the next row's query word, quality and entering reference word
        // first and decodes them after its FP64 work but before its own row stores.  Loads and stores
        // retire in order, so a decode placed behind the stores would wait for them.
        const int qlast = qLen - 1;
        const int q1 = 1 < qlast ? 1 : qlast;
        uint32_t cqb, crb;
        double cp;
        if constexpr (STAGE) {
            const uint32_t v = rowv(q1);
            cqb = v >> 8;
            cp = qprob((uint8_t)(v & 255));
            crb = refn(2 - bw - 1 + (W - 1) + W);
        } else {
            cqb = (read_word(b.reads, qbase + q1) >> (((qbase + q1) & 7) * 4)) & 15;
            cp = qprob(b.quals[qqual + q1]);
            crb = nibble(ref_word(hoff + (2 - bw - 1 + (W - 1))), 2 - bw - 1 + (W - 1));
        }
        for (int i = 2; i <= qLen; i++) {
            const int in_ = i < qlast ? i : qlast;
            const int nidx = i + 1 - bw - 1 + (W - 1);
            uint32_t nqw = 0, nrw = 0, nv = 0;
            uint8_t nqq = 0;
            if constexpr (STAGE) {
                nv = rowv(in_);
                nrw = refn(nidx + W);
            } else {
                nqw = read_word(b.reads, qbase + in_);
                nqq = b.quals[qqual + in_];
                nrw = ref_word(hoff + nidx);
            }
            const uint8_t qb = (uint8_t)cqb;
            win_shr((uint8_t)crb);
            const double em_eq = 1 - cp, em_ne = cp * EM;
            const int klo = 1 > i - bw ? 1 : i - bw;
            const int khi = refLen < i + bw ? refLen : i + bw;
            if (khi >= klo) ncell += (uint64_t)(khi - klo + 1);
            double sum = 0.0, cM = 0.0, cD = 0.0;
#pragma unroll
            for (int r = 0; r < W; r++) {
                const int k = i - bw + r;
                const bool in = k >= klo && k <= khi;
                const double pM = M[r], pI = I[r], pD = D[r];
                const double nM = (r + 1 < W) ? M[r + 1 < W ? r + 1 : r] : 0.0;
                const double nI = (r + 1 < W) ? I[r + 1 < W ? r + 1 : r] : 0.0;
                const uint8_t rb = rb_of(r);
                const double e = (rb == B_N || qb == B_N) ? 1.0 : (rb == qb ? em_eq : em_ne);
                const double mv = in ? e * (m0 * pM + m3 * pI + m6 * pD) : 0.0;
                const double iv = in ? EI * (m1 * nM + m4 * nI) : 0.0;
                const double dv = in ? m2 * cM + m8 * cD : 0.0;
                sum += in ? mv + iv + dv : 0.0;
                M[r] = mv; I[r] = iv; D[r] = dv;
                cM = mv; cD = dv;
            }
            if constexpr (STAGE) {
                cqb = nv >> 8;
                cp = qprob((uint8_t)(nv & 255));
                crb = nrw;
            } else {
                cqb = (nqw >> (((qbase + in_) & 7) * 4)) & 15;
                cp = qprob(nqq);
                crb = nibble(nrw, nidx);
            }
            pin(cqb); pin(cp); pin(crb);
            Sst[(uint64_t)i * 64 + lane] = sum;
            const double inv = 1. / sum;
            double2 *fr = Fst + (uint64_t)i * W * 64 + lane;
#pragma unroll
            for (int r = 0; r < W; r++) {
                M[r] = M[r] * inv; I[r] = I[r] * inv; D[r] = D[r] * inv;
                fr[r * 64] = make_double2(M[r], I[r]);
            }
        }
        // ---- s[qLen+1] (baq.cu:514-525)
        double sq1 = 0.0;
        {
            const int klo = 1 > qLen - bw ? 1 : qLen - bw;
            const int khi = refLen < qLen + bw ? refLen : qLen + bw;
#pragma unroll
            for (int r = 0; r < W; r++) {
                const int k = qLen - bw + r;
                const double t = M[r] * sM + I[r] * sI;
                sq1 += (k >= klo && k <= khi) ? t : 0.0;
            }
        }

        // ---- backward + MAP (baq.cu:528-681)
        // with the register budget of one wave per SIMD the forward row that the next MAP step
        // consumes is loaded right after the current one is used, a full backward row ahead
        constexpr bool PREF = WPE == 1;
        double2 Fp[PREF ? W : 1];
        uint32_t Cp = 0;
        auto load_f = [&](int i) {
            if constexpr (PREF) {
                const double2 *fr = Fst + (uint64_t)i * W * 64 + lane;
#pragma unroll
                for (int r = 0; r < W; r++) Fp[r] = fr[r * 64];
                Cp = Ost[(uint64_t)i * 64 + lane];
            }
        };
        auto map_row = [&](int i, uint8_t oq) {
            const double2 *fr = Fst + (uint64_t)i * W * 64 + lane;
            double sum = 0.0, mx = 0.0;
            int32_t stt = -1;
#pragma unroll
            for (int r = 0; r < W; r++) {
                const int k = i - bw + r;
                double2 f;
                if constexpr (PREF) f = Fp[r < W ? r : 0];
                else f = fr[r * 64];
                double z = f.x * M[r];
                sum += z;
                if (z > mx) { mx = z; stt = (k - 1) << 2 | 0; }
                z = f.y * I[r];
                sum += z;
                if (z > mx) { mx = z; stt = (k - 1) << 2 | 1; }
            }
            mx /= sum;
            const uint32_t code = PREF ? Cp : Ost[(uint64_t)i * 64 + lane];
            const uint8_t tag = cap_tag(code >> 30, code & CAP_EP_MASK, oq, mapq(1.0 - mx), (uint32_t)stt);
            // the final tag goes to LDS (TAGL) or replaces the code in the slot (one coalesced row store)
            if constexpr (TAGL) tagl[(i - 1) * 64 + lane] = tag;
            else Ost[(uint64_t)i * 64 + lane] = tag;
        };
        load_f(qLen);
        // backward row i emits query index i (the row below it); map_row(i) needs quality index i - 1
        const uint8_t lastq = STAGE ? (uint8_t)(rowv(qlast) & 255) : b.quals[qqual + qlast];
        {
            const double sq = Sst[(uint64_t)qLen * 64 + lane];
            const int klo = 1 > qLen - bw ? 1 : qLen - bw;
            const int khi = refLen < qLen + bw ? refLen : qLen + bw;
            const double vM = sM / sq / sq1, vI = sI / sq / sq1;
#pragma unroll
            for (int r = 0; r < W; r++) {
                const int k = qLen - bw + r;
                const bool in = k >= klo && k <= khi;
                M[r] = in ? vM : 0.0; I[r] = in ? vI : 0.0;
            }
            map_row(qLen, lastq);
        }
        if (qLen >= 2) load_f(qLen - 1);
        win_fill(qLen - bw);                     // backward cell r of row i: window index i - bw + r
        // inputs of backward row qLen - 1, then loaded a row ahead and decoded as in the forward pass
        if constexpr (STAGE) {
            cqb = rowv(qlast) >> 8;
            crb = refn(qLen - 1 - bw + W);
        } else {
            cqb = (read_word(b.reads, qbase + qlast) >> (((qbase + qlast) & 7) * 4)) & 15;
            crb = nibble(ref_word(hoff + (qLen - 1 - bw)), qLen - 1 - bw);
        }
        cp = qprob(lastq);
        double cyy = 1.0 / Sst[(uint64_t)(qLen >= 2 ? qLen - 1 : qLen) * 64 + lane];   // 1 / s[i]
        for (int i = qmax - 1; i >= 1; i--) {
            if (i >= qLen) continue;            // rows above this lane's window
            const int in_ = i - 1;
            const int nidx = i - 1 - bw;
            uint32_t nqw = 0, nrw = 0, nv = 0;
            uint8_t nqq = 0;
            if constexpr (STAGE) {
                nv = rowv(in_);
                nrw = refn(nidx + W);
                nqq = (uint8_t)(nv & 255);
            } else {
                nqw = read_word(b.reads, qbase + in_);
                nqq = b.quals[qqual + in_];
                nrw = ref_word(hoff + nidx);
            }
            const double nS = Sst[(uint64_t)(i >= 2 ? i - 1 : i) * 64 + lane];
            win_shl((uint8_t)crb);
            const double y = (i > 1) ? 1. : 0.;
            const uint8_t qb = (uint8_t)cqb;
            const double em_eq = 1 - cp, em_ne = cp * EM;
            const int klo = 1 > i - bw ? 1 : i - bw;
            const int khi = refLen < i + bw ? refLen : i + bw;
            double cD = 0.0;
#pragma unroll
            for (int r = W - 1; r >= 0; r--) {
                const int k = i - bw + r;
                const bool in = k >= klo && k <= khi;
                const double pM = M[r];
                const double pIl = r >= 1 ? I[r >= 1 ? r - 1 : 0] : 0.0;
                const uint8_t rb = rb_of(r);
                const double e0 = ((rb == B_N || qb == B_N) ? 1.0 : (rb == qb ? em_eq : em_ne)) * pM;
                const double e = k >= refLen ? 0.0 : e0;
                const double mv = in ? e * m0 + EIm1 * pIl + m2 * cD : 0.0;
                const double iv = in ? e * m3 + EIm4 * pIl : 0.0;
                const double dv = in ? (e * m6 + m8 * cD) * y : 0.0;
                M[r] = mv; I[r] = iv;
                cD = dv;
            }
            const double yy = cyy;
#pragma unroll
            for (int r = 0; r < W; r++) { M[r] = M[r] * yy; I[r] = I[r] * yy; }
            if constexpr (STAGE) {
                cqb = nv >> 8;
                crb = nrw;
            } else {
                cqb = (nqw >> (((qbase + in_) & 7) * 4)) & 15;
                crb = nibble(nrw, nidx);
            }
            cp = qprob(nqq);
            cyy = 1.0 / nS;
            pin(cqb); pin(cp); pin(crb); pin(cyy);
            map_row(i, nqq);
            if (i > 1) load_f(i - 1);
        }
        // the read's tags, row by row in lockstep across the wave: coalesced slot loads, and each lane
        // writes its read's window as one burst of consecutive bytes
        uint8_t *out = a.s.baq + b.qual_start[rd] + qStart;
        if constexpr (TAGL) {
            const uint8_t *trow = tagl + lane;                  // index j holds position qStart + j
            for (int j = 0; j < qLen; j++) out[j] = good ? trow[j * 64] : (uint8_t)64;
        } else {
            const uint32_t *orow = Ost + 64 + lane;             // row j + 1 holds position qStart + j
            int j = 0;
            for (; j + 8 <= qLen; j += 8) {
                uint32_t t[8];
#pragma unroll
                for (int u = 0; u < 8; u++) t[u] = orow[(uint64_t)(j + u) * 64];
#pragma unroll
                for (int u = 0; u < 8; u++) out[j + u] = good ? (uint8_t)t[u] : (uint8_t)64;
            }
            for (; j < qLen; j++) out[j] = good ? (uint8_t)orow[(uint64_t)j * 64] : (uint8_t)64;
        }
    };
    // Chunks in a fixed round-robin over the resident waves, or, with a.next, taken one at a time
    // from a counter (zero at launch) in descending list order, longest windows first: waves that
    // start late, beside another kernel, then take fewer chunks and all finish together.
    const uint32_t nchunks = (uint32_t)((total + 63) / 64);
    auto take = [&]() -> uint32_t {
        uint32_t c = 0;
        if (lane == 0) c = atomicAdd(a.next, 1u);
        return __builtin_amdgcn_readfirstlane(c);
    };
    for (uint32_t k = a.next ? take() : gwave; k < nchunks; k = a.next ? take() : k + nwaves)
        run_chunk(a.next ? nchunks - 1 - k : k);
    // one counter update per wave
    for (int off = 32; off >= 1; off >>= 1) ncell += __shfl_xor(ncell, off);
    if (lane == 0 && ncell) atomicAdd(a.cells, (unsigned long long)ncell);
}

// ---------------------------------------------------------------------------
// Checkpointed band kernel (inputs staged in LDS, one wave per SIMD).
//
// The forward pass keeps every row's scale factor but the state of only every K-th row, from row 1.
// The backward pass then walks K-row segments from the top: it reloads a segment's checkpoint,
// recomputes the segment's other forward rows into registers with the forward pass's own row code
// (so they are the same bits), and consumes them in reverse for the MAP steps.
//
// A checkpoint row i >= 2 is kept as its unscaled M and I (16 B per cell): the row's D follows from
// its own M along the band (D[k] = m2 M[k-1] + m8 D[k-1], the recurrence of the forward pass), and
// scaling M, I and D by the stored 1 / s[i] gives the forward pass's scaled row bit for bit.  Row 1
// (scaled by division, D = 0) is kept scaled.  Forward-row traffic falls from 32 B per cell (M and I
// of every row written and read back) to 32/K B; the recompute repeats the forward arithmetic once.
//
// Registers: the K recomputed rows as (M, I) pairs, the forward D of the recompute and the
// backward row's M and I; the forward pass's own state dies before them.  The next segment's
// checkpoint (M, I) and its scale factor are loaded while this segment's backward rows run, into the
// top row's slot once that row's MAP step is done.  The segment loop runs on a wave-uniform counter,
// so all lanes load the same checkpoint row at once.
template <int W, int K, bool TAGL>
__global__ void __launch_bounds__(256, 1) k_baq_ckpt(BaqArgs a)
{
    constexpr int NWIN = (W + 15) / 16;
    constexpr int TOP = (W - 1) - 16 * (NWIN - 1);
    constexpr uint64_t TOPMASK = TOP == 15 ? ~0ull : ((1ull << (4 * (TOP + 1))) - 1);

    __shared__ double s_q2p[256];
    __shared__ double s_thr[128];
    for (uint32_t t = threadIdx.x; t < 256; t += blockDim.x) s_q2p[t] = c_qual2prob[t];
    for (uint32_t t = threadIdx.x; t < 128; t += blockDim.x) s_thr[t] = c_map_thr[t];
    __syncthreads();

    const uint32_t lane = threadIdx.x & 63;
    const uint32_t gwave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const uint32_t nwaves = gridDim.x * (blockDim.x >> 6);
    const uint32_t total = *a.count;
    double *slot = a.scratch + (uint64_t)gwave * a.slot_doubles;
    // checkpoints of rows 1, 1 + K, 1 + 2K, ...: (M, I) [j][r][lane]; they fit in the forward-row area
    // of the every-row kernel, and the scale factors and cap codes keep its offsets
    double2 *Cmi = (double2 *)slot;
    double *Sst = slot + (uint64_t)(a.rows_cap + 1) * W * 2 * 64;         // [row][lane]
    uint32_t *Ost = (uint32_t *)(Sst + (uint64_t)(a.rows_cap + 2) * 64);  // [row][lane] cap code, then tag
    const BatchDev &b = a.b;
    uint64_t ncell = 0;
    extern __shared__ unsigned char baq_stage[];
    const uint32_t base_bytes = (a.rows_cap + 1) / 2, ref_bytes = (a.rows_cap + 2 * W + 1) / 2;
    const uint32_t wave_bytes = (a.rows_cap + base_bytes + ref_bytes + (TAGL ? a.rows_cap : 0u)) * 64;
    uint8_t *rowq = baq_stage + (threadIdx.x >> 6) * wave_bytes;
    uint8_t *rowb = rowq + (uint64_t)a.rows_cap * 64;
    uint8_t *refd = rowb + (uint64_t)base_bytes * 64;
    uint8_t *tagl = refd + (uint64_t)ref_bytes * 64;
    auto refn = [&](int i) -> uint32_t { return (refd[(i >> 1) * 64 + lane] >> ((i & 1) * 4)) & 15; };
    auto rowv = [&](int q) -> uint32_t {
        return rowq[q * 64 + lane] | (((rowb[(q >> 1) * 64 + lane] >> ((q & 1) * 4)) & 15) << 8);
    };
    auto qprob = [&](uint8_t q) -> double { return s_q2p[q < 4 ? 4 : q]; };
    auto pin = [](auto &v) { asm volatile("" : "+v"(v)); };
    auto mapq = [&](double y) -> uint8_t {
        if (y == 0.0) return 4;
        // the double-precision guess here: with the single-precision one (map_guess) this kernel's
        // register allocation goes past 512 into spills
        int k = (int)(-4.343 * log(y) + 0.499);
        k = k < 4 ? 4 : (k > 100 ? 100 : k);
        const double t0 = s_thr[k], t1 = s_thr[k + 1];
        if (k >= 5 && !(y <= t0)) {
            k--;
            while (k >= 5 && !(y <= s_thr[k])) k--;
        } else if (y <= t1) {
            k++;
            while (k + 1 <= 101 && y <= s_thr[k + 1]) k++;
        }
        if (k <= 4) return 4;
        if (k >= 101) return 99;
        return (uint8_t)k;
    };
    // band reference windows: nibble r = HMM window index idx0 + r
    auto win_fill = [&](uint64_t (&win)[NWIN], int idx0) {
#pragma unroll
        for (int w = 0; w < NWIN; w++) {
            uint64_t v = 0;
#pragma unroll
            for (int j = 0; j < 16; j++)
                if (w * 16 + j < W) v |= (uint64_t)refn(idx0 + w * 16 + j + W) << (4 * j);
            win[w] = v;
        }
    };
    auto win_shr = [&](uint64_t (&win)[NWIN], uint32_t nb) {     // band moves right: append cell W-1
#pragma unroll
        for (int w = 0; w < NWIN; w++) {
            const uint64_t in = (w + 1 < NWIN) ? (win[w + 1 < NWIN ? w + 1 : w] & 15) : (uint64_t)nb;
            const int top = (w + 1 < NWIN) ? 15 : TOP;
            win[w] = (win[w] >> 4) | (in << (4 * top));
        }
    };
    auto win_shl = [&](uint64_t (&win)[NWIN], uint32_t nb) {     // band moves left: insert cell 0
#pragma unroll
        for (int w = NWIN - 1; w >= 0; w--) {
            const uint64_t in = w > 0 ? (win[w > 0 ? w - 1 : 0] >> 60) & 15 : (uint64_t)nb;
            const uint64_t keep = (w + 1 < NWIN) ? ~0ull : TOPMASK;
            win[w] = ((win[w] << 4) | in) & keep;
        }
    };

    auto run_chunk = [&](uint32_t chunk) {
        const uint32_t idx = chunk * 64 + lane;
        const bool act = idx < total;
        const uint32_t rd = act ? a.list[idx] : 0u;
        const uint32_t rwv = act ? a.s.rwc[rd] : 0u;
        const int qStart = lo16(rwv);
        const int qLen = act ? (int)hi16(rwv) - qStart + 1 : 0;
        int qmax = qLen;
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            const int o = __shfl_xor(qmax, off);
            qmax = o > qmax ? o : qmax;
        }
        if (!act) return;

        // ---- setup (baq.cu:260-354)
        const uint32_t hw = a.s.hmmw[rd];
        const int hx = (int16_t)lo16(hw), hy = (int16_t)hi16(hw);
        const int bw = a.s.bandwidth[rd];
        const uint32_t astart = b.alignment_start[rd];
        const uint32_t chrom = b.chromosome[rd];
        const int refLenW = hy - hx + 1;
        const uint32_t ahx = hx < 0 ? (uint32_t)(-hx) : (uint32_t)hx;
        const bool clamp = hx < 0 && astart < ahx;
        const int refLen = clamp ? refLenW - 1 : refLenW;
        const int64_t hoff = clamp ? 0 : (int64_t)astart + hx;

        const double GOP = c_gap_open;
        const double sM = 1.0 / (2 * qLen + 2);
        const double sI = sM;
        const double bM = (1 - GOP) / refLenW;
        const double bI = GOP / refLenW;
        const double m0 = (1 - GOP - GOP) * (1 - sM);
        const double m1 = GOP * (1 - sM);
        const double m2 = m1;
        const double m3 = (1 - GEP) * (1 - sI);
        const double m4 = GEP * (1 - sI);
        const double m6 = 1 - GEP;
        const double m8 = GEP;
        const double EIm1 = EI * m1;
        const double EIm4 = EI * m4;

        const uint64_t qbase = (uint64_t)b.read_start[rd] + qStart;
        const uint32_t qqual = b.qual_start[rd] + qStart;
        const uint32_t *cw = a.g.words + a.g.word_offset[chrom];
        const int64_t clen = a.g.length[chrom];
        const int64_t wlast = clen > 0 ? ((clen + 7) >> 3) - 1 : 0;
        auto ref_word = [&](int64_t p) -> uint32_t {
            int64_t wi = p >> 3;
            wi = wi < 0 ? 0 : (wi > wlast ? wlast : wi);
            return cw[wi];
        };
        auto nibble = [&](uint32_t w, int idx0) -> uint8_t {
            const uint32_t ok = (uint32_t)(idx0 >= 0) & (uint32_t)(idx0 < refLen) & (uint32_t)(hoff + idx0 < clen);
            const uint32_t nb = (w >> (((uint32_t)(hoff + idx0) & 7) * 4)) & 15;
            return (uint8_t)((nb & (0u - ok)) | (15u & (ok - 1u)));
        };

        const bool good = cap_walk(b, a.s, rd, [&](uint16_t p, uint32_t kind, uint32_t ep) {
            Ost[(uint64_t)((int)p - qStart + 1) * 64 + lane] = cap_code(kind, ep);
        });

        // this read's inputs into the wave's LDS area (each lane writes and later reads only its own column)
#pragma unroll 4
        for (int q = 0; q < qLen; q++) rowq[q * 64 + lane] = b.quals[qqual + q];
#pragma unroll 4
        for (int q = 0; q < qLen; q += 2) {
            const uint32_t w = b.reads[(qbase + q) >> 3];
            const uint32_t b0 = (w >> (((qbase + q) & 7) * 4)) & 15;
            const uint64_t i1 = qbase + (q + 1 < qLen ? q + 1 : q);
            const uint32_t b1 = q + 1 < qLen ? (b.reads[i1 >> 3] >> ((i1 & 7) * 4)) & 15 : 15u;
            rowb[(q >> 1) * 64 + lane] = (uint8_t)(b0 | (b1 << 4));
        }
#pragma unroll 4
        for (int bi = (W - bw) >> 1; bi <= (qLen + 2 * W - 1 - bw) >> 1; bi++) {
            const int k0 = 2 * bi - W, k1 = k0 + 1;
            refd[bi * 64 + lane] = (uint8_t)(nibble(ref_word(hoff + k0), k0) | (nibble(ref_word(hoff + k1), k1) << 4));
        }

        // A steady row has every band cell inside the band and the HMM window, no N among the band's
        // reference cells and a query base that is not N: its cells need no masks and its emissions
        // only the match test.  The kernels take the mask-free code for a row when every lane still
        // running is steady there (a wave-uniform branch); both forms compute the same expressions.
        const bool full_band = 2 * bw + 1 == W;
        auto no_n = [&](const uint64_t (&win)[NWIN], uint32_t qb) -> bool {
            uint64_t t = 0;
#pragma unroll
            for (int w = 0; w < NWIN; w++) t |= win[w] & (win[w] >> 1) & (win[w] >> 2) & (win[w] >> 3);
            return (t & 0x1111111111111111ull) == 0 && qb != B_N;
        };
        // forward row i: cells k = i - bw .. i - bw + W - 1 all in [1, refLen]
        auto fwd_steady = [&](int i, const uint64_t (&win)[NWIN], uint32_t qb) -> bool {
            return full_band && i - bw >= 1 && i - bw + W - 1 <= refLen && no_n(win, qb);
        };
        // backward row i: the same, and no cell at k = refLen (whose emission term is zero)
        auto bwd_steady = [&](int i, const uint64_t (&win)[NWIN], uint32_t qb) -> bool {
            return full_band && i - bw >= 1 && i - bw + W - 1 < refLen && no_n(win, qb);
        };

        // forward row i >= 2 from row i - 1 (baq.cu:444-512) into O (may be P itself), unscaled, and its
        // sum.  The forward pass and the segment recompute both run this code: the same bits.
        auto fwd_row = [&](auto steady, int i, const double2 (&P)[W], double2 (&O)[W], double (&D)[W],
                           const uint64_t (&win)[NWIN], uint32_t qb, double p) -> double {
            constexpr bool S = decltype(steady)::value;
            const double em_eq = 1 - p, em_ne = p * EM;
            const int klo = 1 > i - bw ? 1 : i - bw;
            const int khi = refLen < i + bw ? refLen : i + bw;
            double sum = 0.0, cM = 0.0, cD = 0.0;
#pragma unroll
            for (int r = 0; r < W; r++) {
                const int k = i - bw + r;
                const bool in = S || (k >= klo && k <= khi);
                const double pM = P[r].x, pI = P[r].y, pD = D[r];
                const double nM = (r + 1 < W) ? P[r + 1 < W ? r + 1 : r].x : 0.0;
                const double nI = (r + 1 < W) ? P[r + 1 < W ? r + 1 : r].y : 0.0;
                const uint32_t rb = (uint32_t)((win[r >> 4] >> (4 * (r & 15))) & 15);
                const double e = (!S && (rb == B_N || qb == B_N)) ? 1.0 : (rb == qb ? em_eq : em_ne);
                const double mv = in ? e * (m0 * pM + m3 * pI + m6 * pD) : 0.0;
                const double iv = in ? EI * (m1 * nM + m4 * nI) : 0.0;
                const double dv = in ? m2 * cM + m8 * cD : 0.0;
                sum += in ? mv + iv + dv : 0.0;
                O[r] = make_double2(mv, iv);
                D[r] = dv;
                cM = mv; cD = dv;
            }
            return sum;
        };
        auto fwd_any = [&](int i, const double2 (&P)[W], double2 (&O)[W], double (&D)[W],
                           const uint64_t (&win)[NWIN], uint32_t qb, double p) -> double {
            if (wave_all(fwd_steady(i, win, qb))) return fwd_row(Steady<true>{}, i, P, O, D, win, qb, p);
            return fwd_row(Steady<false>{}, i, P, O, D, win, qb, p);
        };
        // rows >= 2 are scaled by the reciprocal of their sum (baq.cu:505-510)
        auto fwd_scale = [&](double2 (&O)[W], double (&D)[W], double inv) {
#pragma unroll
            for (int r = 0; r < W; r++) { O[r] = make_double2(O[r].x * inv, O[r].y * inv); D[r] = D[r] * inv; }
        };
        auto store_ckpt = [&](int j, const double2 (&O)[W]) {
            double2 *cmi = Cmi + (uint64_t)j * W * 64 + lane;
#pragma unroll
            for (int r = 0; r < W; r++) cmi[r * 64] = O[r];
        };
        // a checkpoint row i >= 2 from its unscaled (M, I) in O: its D along the band as the forward
        // row code computes it (cells outside the band hold M = 0 and get D = 0), then the row scaled
        // by inv = 1 / s[i]
        auto restore_ckpt = [&](int i, double2 (&O)[W], double (&D)[W], double inv) {
            const int klo = 1 > i - bw ? 1 : i - bw;
            const int khi = refLen < i + bw ? refLen : i + bw;
            double cM = 0.0, cD = 0.0;
#pragma unroll
            for (int r = 0; r < W; r++) {
                const int k = i - bw + r;
                const bool in = k >= klo && k <= khi;
                const double dv = in ? m2 * cM + m8 * cD : 0.0;
                D[r] = dv;
                cM = O[r].x; cD = dv;
            }
#pragma unroll
            for (int r = 0; r < W; r++) { O[r] = make_double2(O[r].x * inv, O[r].y * inv); D[r] = D[r] * inv; }
        };

        // ---- forward pass.  The scale factor area keeps 1 / s[i], the factor the backward rows
        // multiply by (the same division as the forward's own), and s[qLen] stays in a register.
        double sq1 = 0.0, sq = 0.0;
        {
            double2 MI[W];
            double D[W];
            uint64_t win[NWIN];
            // row 1 (baq.cu:396-442)
            {
                const uint32_t v = rowv(0);
                const uint8_t qb = (uint8_t)(v >> 8);
                const double p = qprob((uint8_t)(v & 255));
                const double em_eq = 1 - p, em_ne = p * EM;
                const int khi = refLen < bw + 1 ? refLen : bw + 1;
                ncell += khi >= 1 ? (uint64_t)khi : 0;
                win_fill(win, -bw);
                double sum = 0.0;
#pragma unroll
                for (int r = 0; r < W; r++) {
                    const int k = 1 - bw + r;
                    const bool in = k >= 1 && k <= khi;
                    const uint8_t rb = (uint8_t)((win[r >> 4] >> (4 * (r & 15))) & 15);
                    const double e = (rb == B_N || qb == B_N) ? 1.0 : (rb == qb ? em_eq : em_ne);
                    const double mv = in ? e * bM : 0.0;
                    const double iv = in ? EI * bI : 0.0;
                    sum += in ? mv + iv : 0.0;
                    MI[r] = make_double2(mv, iv); D[r] = 0.0;
                }
                Sst[64 + lane] = 1. / sum;
                sq = sum;
#pragma unroll
                for (int r = 0; r < W; r++) {
                    MI[r] = make_double2(MI[r].x / sum, MI[r].y / sum); D[r] = D[r] / sum;
                }
                store_ckpt(0, MI);
            }
            // rows 2..qLen: the next row's LDS inputs read a row ahead
            const int qlast = qLen - 1;
            const int q1 = 1 < qlast ? 1 : qlast;
            uint32_t cv = rowv(q1);
            uint32_t crb = refn(2 - bw - 1 + (W - 1) + W);
            for (int i = 2; i <= qLen; i++) {
                const int in_ = i < qlast ? i : qlast;
                const uint32_t nv = rowv(in_);
                const uint32_t nrw = refn(i + 1 - bw - 1 + (W - 1) + W);
                win_shr(win, crb);
                const int klo = 1 > i - bw ? 1 : i - bw;
                const int khi = refLen < i + bw ? refLen : i + bw;
                if (khi >= klo) ncell += (uint64_t)(khi - klo + 1);
                const double sum = fwd_any(i, MI, MI, D, win, cv >> 8, qprob((uint8_t)(cv & 255)));
                cv = nv;
                crb = nrw;
                pin(cv); pin(crb);
                const double inv = 1. / sum;
                Sst[(uint64_t)i * 64 + lane] = inv;
                sq = sum;
                if ((i - 1) % K == 0) store_ckpt((i - 1) / K, MI);   // unscaled
                fwd_scale(MI, D, inv);
            }
            // s[qLen+1] (baq.cu:514-525)
            const int klo = 1 > qLen - bw ? 1 : qLen - bw;
            const int khi = refLen < qLen + bw ? refLen : qLen + bw;
#pragma unroll
            for (int r = 0; r < W; r++) {
                const int k = qLen - bw + r;
                const double t = MI[r].x * sM + MI[r].y * sI;
                sq1 += (k >= klo && k <= khi) ? t : 0.0;
            }
        }

        // ---- backward + MAP (baq.cu:528-681), K rows at a time from the top
        const int nseg_w = (qmax - 1) / K + 1;
        const int last_seg = (qLen - 1) / K;
        double2 F[K][W];
        double Dr[W];
        double bMv[W], bIv[W];
        uint64_t bwin[NWIN];
        // backward row i < qLen from row i + 1 (baq.cu:547-604) in place in (bMv, bIv), unscaled
        auto bwd_row = [&](auto steady, int i, uint32_t qb, double p) {
            constexpr bool S = decltype(steady)::value;
            const double y = (i > 1) ? 1. : 0.;             // a steady row has i > bw >= 1
            const double em_eq = 1 - p, em_ne = p * EM;
            const int klo = 1 > i - bw ? 1 : i - bw;
            const int khi = refLen < i + bw ? refLen : i + bw;
            double cD = 0.0;
#pragma unroll
            for (int r = W - 1; r >= 0; r--) {
                const int k = i - bw + r;
                const bool in = S || (k >= klo && k <= khi);
                const double pM = bMv[r];
                const double pIl = r >= 1 ? bIv[r >= 1 ? r - 1 : 0] : 0.0;
                const uint8_t rb = (uint8_t)((bwin[r >> 4] >> (4 * (r & 15))) & 15);
                const double e0 = ((!S && (rb == B_N || qb == B_N)) ? 1.0 : (rb == qb ? em_eq : em_ne)) * pM;
                const double e = (!S && k >= refLen) ? 0.0 : e0;
                const double mv = in ? e * m0 + EIm1 * pIl + m2 * cD : 0.0;
                const double iv = in ? e * m3 + EIm4 * pIl : 0.0;
                // (x * 1.0 is x: a steady row leaves the factor out)
                const double dv = in ? (S ? e * m6 + m8 * cD : (e * m6 + m8 * cD) * y) : 0.0;
                bMv[r] = mv; bIv[r] = iv;
                cD = dv;
            }
        };
        {
            // backward row qLen (baq.cu:528-545)
            const int klo = 1 > qLen - bw ? 1 : qLen - bw;
            const int khi = refLen < qLen + bw ? refLen : qLen + bw;
            const double vM = sM / sq / sq1, vI = sI / sq / sq1;
#pragma unroll
            for (int r = 0; r < W; r++) {
                const int k = qLen - bw + r;
                const bool in = k >= klo && k <= khi;
                bMv[r] = in ? vM : 0.0; bIv[r] = in ? vI : 0.0;
            }
            win_fill(bwin, qLen - bw);
        }
        // the first segment's checkpoint (a lane whose window ends below the wave's top segment loads
        // its own last one, which the walk reaches when its rows begin)
        double ck_inv;                           // 1 / s of the checkpoint row in the top slot
        {
            const int j0 = (nseg_w - 1) < last_seg ? (nseg_w - 1) : last_seg;
            const double2 *cmi = Cmi + (uint64_t)j0 * W * 64 + lane;
#pragma unroll
            for (int r = 0; r < W; r++) F[K - 1][r] = cmi[r * 64];
            ck_inv = Sst[(uint64_t)(1 + j0 * K) * 64 + lane];
        }
        for (int j = nseg_w - 1; j >= 0; j--) {
            const int s0 = 1 + j * K;
            if (j > last_seg) continue;          // rows above this lane's window
            // The segment's scale factors and cap codes are loaded first: loads retire in issue order,
            // so a load issued after the next checkpoint's prefetch would wait for the prefetch.
            double rinv[K];
            uint32_t rcode[K];
#pragma unroll
            for (int t = 0; t < K; t++) {
                const int i = s0 + t <= qLen ? s0 + t : qLen;
                rinv[t] = Sst[(uint64_t)i * 64 + lane];
                rcode[t] = Ost[(uint64_t)i * 64 + lane];
            }
            // the checkpoint row, loaded into the top slot during the previous segment: row 1 as it is
            // (D = 0), a later row restored from its unscaled (M, I)
#pragma unroll
            for (int r = 0; r < W; r++) F[0][r] = F[K - 1][r];
            if (j > 0) {
                restore_ckpt(s0, F[0], Dr, ck_inv);
            } else {
#pragma unroll
                for (int r = 0; r < W; r++) Dr[r] = 0.0;
            }
            uint64_t fwin[NWIN];
            win_fill(fwin, s0 - bw);             // forward row s0 + 1: cell r at window index s0 - bw + r
#pragma unroll
            for (int t = 1; t < K; t++) {
                const int i = s0 + t;
                if (i <= qLen) {
                    if (t > 1) win_shr(fwin, refn(i - bw + W - 2 + W));
                    const uint32_t v = rowv(i - 1);
                    const double sum = fwd_any(i, F[t - 1], F[t], Dr, fwin, v >> 8, qprob((uint8_t)(v & 255)));
                    fwd_scale(F[t], Dr, 1. / sum);
                }
            }
            const int jn = j > 0 ? j - 1 : 0;
#pragma unroll
            for (int t = K - 1; t >= 0; t--) {
                const int i = s0 + t;
                if (i > qLen) continue;
                const uint32_t code = rcode[t];
                if (i < qLen) {
                    win_shl(bwin, refn(i - bw + W));
                    const uint32_t v = rowv(i);
                    const uint8_t qb = (uint8_t)(v >> 8);
                    const double p = qprob((uint8_t)(v & 255));
                    if (wave_all(bwd_steady(i, bwin, qb))) bwd_row(Steady<true>{}, i, qb, p);
                    else bwd_row(Steady<false>{}, i, qb, p);
                    const double yy = rinv[t];
#pragma unroll
                    for (int r = 0; r < W; r++) { bMv[r] = bMv[r] * yy; bIv[r] = bIv[r] * yy; }
                }
                // MAP step of row i (quality of query position i - 1)
                double sum = 0.0, mx = 0.0;
                int32_t stt = -1;
#pragma unroll
                for (int r = 0; r < W; r++) {
                    const int k = i - bw + r;
                    const double2 f = F[t][r];
                    double z = f.x * bMv[r];
                    sum += z;
                    // the first strictly larger value wins (values are finite and >= 0, so fmax is z > mx ? z : mx)
                    if (z > mx) { mx = z; stt = (k - 1) << 2 | 0; }
                    z = f.y * bIv[r];
                    sum += z;
                    if (z > mx) { mx = z; stt = (k - 1) << 2 | 1; }
                }
                mx /= sum;
                const uint8_t oq = (uint8_t)(rowv(i - 1) & 255);
                const uint8_t tag = cap_tag(code >> 30, code & CAP_EP_MASK, oq, mapq(1.0 - mx), (uint32_t)stt);
                if constexpr (TAGL) tagl[(i - 1) * 64 + lane] = tag;
                else Ost[(uint64_t)i * 64 + lane] = tag;
                if (t == K - 1) {
                    // the top row is consumed: its slot takes the next checkpoint's (M, I)
                    const double2 *cmi = Cmi + (uint64_t)jn * W * 64 + lane;
#pragma unroll
                    for (int r = 0; r < W; r++) F[K - 1][r] = cmi[r * 64];
                    ck_inv = Sst[(uint64_t)(1 + jn * K) * 64 + lane];
                }
            }
            if (s0 + K - 1 > qLen) {
                // the top row was not this lane's (its window ends inside the segment): load here
                const double2 *cmi = Cmi + (uint64_t)jn * W * 64 + lane;
#pragma unroll
                for (int r = 0; r < W; r++) F[K - 1][r] = cmi[r * 64];
                ck_inv = Sst[(uint64_t)(1 + jn * K) * 64 + lane];
            }
        }
        // the read's tags, row by row in lockstep across the wave
        uint8_t *out = a.s.baq + b.qual_start[rd] + qStart;
        if constexpr (TAGL) {
            const uint8_t *trow = tagl + lane;
            for (int j = 0; j < qLen; j++) out[j] = good ? trow[j * 64] : (uint8_t)64;
        } else {
            const uint32_t *orow = Ost + 64 + lane;
            for (int j = 0; j < qLen; j++) out[j] = good ? (uint8_t)orow[(uint64_t)j * 64] : (uint8_t)64;
        }
    };
    // chunks from the counter (zero at launch), longest windows first, or round-robin without it
    const uint32_t nchunks = (uint32_t)((total + 63) / 64);
    auto take = [&]() -> uint32_t {
        uint32_t c = 0;
        if (lane == 0) c = atomicAdd(a.next, 1u);
        return __builtin_amdgcn_readfirstlane(c);
    };
    for (uint32_t k = a.next ? take() : gwave; k < nchunks; k = a.next ? take() : k + nwaves)
        run_chunk(a.next ? nchunks - 1 - k : k);
    for (int off = 32; off >= 1; off >>= 1) ncell += __shfl_xor(ncell, off);
    if (lane == 0 && ncell) atomicAdd(a.cells, (unsigned long long)ncell);
}

// ---------------------------------------------------------------------------
// flat BAQ, capping and BQ-tag recoding, one lane per read
// (read_flat_baq baq.cu:842-868, cap_baq_qualities :870-1056, recode :1058-1081)

struct PostArgs {
    BatchDev b;
    ReadStateDev s;
    uint32_t band_done;   // the band kernels already wrote the tags of reads with bands of <= 33 cells
};

__global__ void __launch_bounds__(256) k_baq_post(PostArgs a)
{
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    const BatchDev &b = a.b;
    if (r >= live_reads(b)) return;
    if (a.s.status[r] != RS_ACTIVE) return;
    const uint32_t rw = a.s.rwc[r];
    const uint16_t rwx = lo16(rw), rwy = hi16(rw);
    const uint32_t qs = b.qual_start[r];
    uint8_t *baq = a.s.baq;
    if (a.s.num_errors[r] == 0) {
        for (uint32_t i = rwx; i <= rwy; i++) baq[qs + i] = 64;
        return;
    }
    if (a.band_done && 2u * a.s.bandwidth[r] + 1 <= 33) return;
    const uint32_t *st = a.s.baq_state;
    if (cap_walk(b, a.s, r, [](uint16_t, uint32_t, uint32_t) {})) {
        cap_walk(b, a.s, r, [&](uint16_t i, uint32_t kind, uint32_t ep) {
            baq[qs + i] = cap_tag(kind, ep, b.quals[qs + i], baq[qs + i], st[qs + i]);
        });
    } else {
        for (uint32_t i = rwx; i <= rwy; i++) baq[qs + i] = 64;
    }
}

// BAQ read lists per band class
struct ClassifyArgs {
    uint32_t num_reads;
    const uint32_t *live;  // staged batch: live read count (else num_reads)
    const uint8_t *status;
    const uint16_t *num_errors;
    const uint16_t *bandwidth;
    const uint32_t *rwc;
    uint32_t *lists;       // 4 lists of capacity num_reads
    uint32_t *counts;      // 4 class counters, then the widest band of class 3
    uint32_t capacity;
};

// BAQ reads grouped by band class (which band kernel runs them) and, inside a class, by window
// length in 16-base buckets, so the 64 reads of a wavefront have similar row counts.  counts layout:
// [0..3] class sizes, [4] widest class-3 band, [16 + 32c + b] (class, bucket) sizes,
// [144 + 32c + b] their offsets in the class list, [272 + 32c + b] scatter cursors.  A fixed grid
// strides over the reads; counts are kept in LDS and reach global memory once per workgroup.
static constexpr uint32_t CLS_BLOCKS = 512, CLS_BUCKETS = 32;
static constexpr uint32_t CB_COUNT = 16, CB_OFFSET = 144, CB_CURSOR = 272;

__device__ inline bool classify_read(const ClassifyArgs &a, uint32_t r, uint32_t n, uint32_t &key, uint32_t &w)
{
    if (r >= n || a.status[r] != RS_ACTIVE || a.num_errors[r] == 0) return false;
    w = 2u * a.bandwidth[r] + 1;
    const uint32_t cls = w <= 15 ? 0 : w <= 23 ? 1 : w <= 33 ? 2 : 3;
    const uint32_t rw = a.rwc[r];
    const uint32_t qlen = (uint32_t)((int)hi16(rw) - (int)lo16(rw) + 1);
    const uint32_t bucket = qlen / 16 < CLS_BUCKETS ? qlen / 16 : CLS_BUCKETS - 1;
    key = cls * CLS_BUCKETS + bucket;
    return true;
}

// one LDS increment per distinct key in the wavefront; returns this lane's slot
__device__ inline uint32_t lds_append(uint32_t *cnt, uint32_t key, bool active)
{
    const uint32_t lane = threadIdx.x & 63;
    uint32_t slot = 0;
    uint64_t pending = __ballot(active);
    while (pending) {
        const int leader = __builtin_ctzll(pending);
        const uint32_t k = (uint32_t)__shfl((int)key, leader);
        const uint64_t same = __ballot(active && key == k);
        uint32_t base = 0;
        if ((int)lane == leader) base = atomicAdd(&cnt[k], (uint32_t)__popcll(same));
        base = (uint32_t)__shfl((int)base, leader);
        if (active && key == k) slot = base + (uint32_t)__popcll(same & ((1ull << lane) - 1));
        pending &= ~same;
    }
    return slot;
}

__global__ void __launch_bounds__(256) k_baq_class_count(ClassifyArgs a)
{
    __shared__ uint32_t cnt[4 * CLS_BUCKETS];
    __shared__ uint32_t wmax;
    for (uint32_t i = threadIdx.x; i < 4 * CLS_BUCKETS; i += blockDim.x) cnt[i] = 0;
    if (threadIdx.x == 0) wmax = 0;
    __syncthreads();
    const uint32_t n = a.live ? a.live[0] : a.num_reads;
    for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
        uint32_t key = 0, w = 0;
        const bool act = classify_read(a, base + threadIdx.x, n, key, w);
        if (act && key >= 3 * CLS_BUCKETS) atomicMax(&wmax, w);
        lds_append(cnt, key, act);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < 4 * CLS_BUCKETS; i += blockDim.x)
        if (cnt[i]) atomicAdd(&a.counts[CB_COUNT + i], cnt[i]);
    if (threadIdx.x == 0 && wmax) atomicMax(&a.counts[4], wmax);
}

// class sizes and bucket offsets: one thread per class
__global__ void __launch_bounds__(64) k_baq_class_offsets(ClassifyArgs a)
{
    const uint32_t c = threadIdx.x;
    if (c >= 4) return;
    uint32_t o = 0;
    for (uint32_t b = 0; b < CLS_BUCKETS; b++) {
        a.counts[CB_OFFSET + c * CLS_BUCKETS + b] = o;
        a.counts[CB_CURSOR + c * CLS_BUCKETS + b] = 0;
        o += a.counts[CB_COUNT + c * CLS_BUCKETS + b];
    }
    a.counts[c] = o;
}

__global__ void __launch_bounds__(256) k_baq_class_scatter(ClassifyArgs a)
{
    __shared__ uint32_t cnt[4 * CLS_BUCKETS], first[4 * CLS_BUCKETS];
    for (uint32_t i = threadIdx.x; i < 4 * CLS_BUCKETS; i += blockDim.x) cnt[i] = 0;
    __syncthreads();
    const uint32_t n = a.live ? a.live[0] : a.num_reads;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += stride) {
        uint32_t key = 0, w = 0;
        const bool act = classify_read(a, base + threadIdx.x, n, key, w);
        lds_append(cnt, key, act);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < 4 * CLS_BUCKETS; i += blockDim.x) {
        first[i] = cnt[i] ? a.counts[CB_OFFSET + i] + atomicAdd(&a.counts[CB_CURSOR + i], cnt[i]) : 0;
        cnt[i] = 0;
    }
    __syncthreads();
    for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += stride) {
        const uint32_t r = base + threadIdx.x;
        uint32_t key = 0, w = 0;
        const bool act = classify_read(a, r, n, key, w);
        const uint32_t slot = lds_append(cnt, key, act);
        if (act) a.lists[(uint64_t)(key / CLS_BUCKETS) * a.capacity + first[key] + slot] = r;
    }
}

__global__ void __launch_bounds__(256) k_fill_u8(uint8_t *p, uint64_t n, const uint32_t *live, uint8_t v)
{
    const uint64_t m = live ? ((uint64_t)live[1] | ((uint64_t)live[2] << 32)) : n;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void __launch_bounds__(256) k_fill_u32(uint32_t *p, uint64_t n, uint32_t v)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

} // namespace fpd

extern "C" hipError_t fpd_baq_set_constants(const double *qual2prob, const double *thr, double gap_open)
{
    hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(fpd::c_qual2prob), qual2prob, 256 * sizeof(double));
    if (e != hipSuccess) return e;
    e = hipMemcpyToSymbol(HIP_SYMBOL(fpd::c_map_thr), thr, 128 * sizeof(double));
    if (e != hipSuccess) return e;
    return hipMemcpyToSymbol(HIP_SYMBOL(fpd::c_gap_open), &gap_open, sizeof(double));
}

extern "C" hipError_t fpd_launch_baq_classify(uint32_t num_reads, const uint32_t *live, const fpd::ReadStateDev *s,
                                              uint32_t *lists, uint32_t *counts, uint32_t capacity, hipStream_t st)
{
    if (num_reads == 0) return hipSuccess;
    fpd::ClassifyArgs a{};
    a.num_reads = num_reads; a.live = live; a.status = s->status; a.num_errors = s->num_errors; a.bandwidth = s->bandwidth;
    a.rwc = s->rwc; a.lists = lists; a.counts = counts; a.capacity = capacity;
    const uint32_t blocks = std::min<uint32_t>((num_reads + 255) / 256, fpd::CLS_BLOCKS);
    hipLaunchKernelGGL(fpd::k_baq_class_count, dim3(blocks), dim3(256), 0, st, a);
    hipLaunchKernelGGL(fpd::k_baq_class_offsets, dim3(1), dim3(64), 0, st, a);
    hipLaunchKernelGGL(fpd::k_baq_class_scatter, dim3(blocks), dim3(256), 0, st, a);
    return hipGetLastError();
}

static int g_baq_lds_limit = 64 * 1024;
static int g_baq_cu_lds = 64 * 1024;     // LDS of one CU

// LDS bytes of a STAGE band workgroup (4 waves): per wave rows_cap quality bytes, rows_cap query
// base nibbles and rows_cap + 2W reference nibbles, and with tags rows_cap tag bytes
static size_t baq_stage_bytes(int W, uint32_t rows_cap, bool tags = false)
{
    return 4 * ((size_t)rows_cap + ((size_t)rows_cap + 1) / 2 + ((size_t)rows_cap + 2 * W + 1) / 2 +
                (tags ? (size_t)rows_cap : 0)) * 64;
}

extern "C" hipError_t fpd_baq_prepare()
{
    int dev = 0, v = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    e = hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
    if (e != hipSuccess) return e;
    // the static tables (qualities and thresholds, 3 KiB) share the limit
    g_baq_lds_limit = (v > 0 ? v : 64 * 1024) - 4096;
    int cu = 0;
    if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) == hipSuccess && cu > 0)
        g_baq_cu_lds = cu;
    const int lim = g_baq_lds_limit;
    (void)hipFuncSetAttribute((const void *)fpd::k_baq_band<15, 1, true, false>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
    (void)hipFuncSetAttribute((const void *)fpd::k_baq_band<15, 2, true, false>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
    (void)hipFuncSetAttribute((const void *)fpd::k_baq_band<23, 1, true, false>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
    (void)hipFuncSetAttribute((const void *)fpd::k_baq_band<33, 1, true, false>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
    (void)hipFuncSetAttribute((const void *)fpd::k_baq_band<15, 1, true, true>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
    (void)hipFuncSetAttribute((const void *)fpd::k_baq_band<23, 1, true, true>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
    (void)hipFuncSetAttribute((const void *)fpd::k_baq_band<33, 1, true, true>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
    (void)hipFuncSetAttribute((const void *)fpd::k_baq_ckpt<15, 3, true>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
    (void)hipFuncSetAttribute((const void *)fpd::k_baq_ckpt<15, 3, false>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
    (void)hipFuncSetAttribute((const void *)fpd::k_baq_ckpt<15, 4, true>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
    (void)hipFuncSetAttribute((const void *)fpd::k_baq_ckpt<15, 4, false>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
    (void)hipFuncSetAttribute((const void *)fpd::k_baq_ckpt<15, 5, true>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
    (void)hipFuncSetAttribute((const void *)fpd::k_baq_ckpt<15, 5, false>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
    return hipSuccess;
}

// whether a band class runs its STAGE kernel for this row capacity (class 0 at one wave per SIMD, or at
// two when two workgroups' staging areas fit one CU's LDS)
extern "C" int fpd_baq_staged(int cls, uint32_t rows_cap, int wpe, int allow)
{
    if (!allow || cls > 2 || (cls == 0 && wpe > 2)) return 0;
    const int W = cls == 0 ? 15 : (cls == 1 ? 23 : 33);
    if (cls == 0 && wpe == 2) return 2 * (baq_stage_bytes(W, rows_cap) + 4096) <= (size_t)g_baq_cu_lds;
    return baq_stage_bytes(W, rows_cap) <= (size_t)g_baq_lds_limit;
}

// class: 0 -> 15 cells, 1 -> 23, 2 -> 33, 3 -> generic (wmax_runtime)
extern "C" hipError_t fpd_launch_baq_hmm(int cls, const fpd::BatchDev *b, const fpd::GenomeDev *g,
                                         const fpd::ReadStateDev *s, const uint32_t *list,
                                         const uint32_t *count, double *scratch, uint64_t slot_doubles,
                                         uint32_t rows_cap, uint32_t wmax_runtime, uint32_t waves,
                                         unsigned long long *cells, int wpe, int allow_stage, int recompute,
                                         uint32_t *next_chunk, hipStream_t st)
{
    fpd::BaqArgs a{};
    a.b = *b; a.g = *g; a.s = *s; a.list = list; a.count = count; a.scratch = scratch;
    a.slot_doubles = slot_doubles; a.rows_cap = rows_cap; a.wmax_runtime = wmax_runtime; a.cells = cells;
    a.next = next_chunk;
    if (waves == 0) return hipSuccess;
    // band kernels: 4 waves per workgroup (the caller sizes waves as a multiple of 4)
    const dim3 g4((waves + 3) / 4), b4(256);
    const bool stage = fpd_baq_staged(cls, rows_cap, wpe, allow_stage) != 0;
    const int W = cls == 0 ? 15 : (cls == 1 ? 23 : 33);
    // the tags stay in LDS too when the budget holds them (short reads)
    const bool tags = stage && cls <= 2 && !(cls == 0 && wpe == 2) &&
                      baq_stage_bytes(W, rows_cap, true) <= (size_t)g_baq_lds_limit;
    const size_t lds = stage ? baq_stage_bytes(W, rows_cap, tags) : 0;
    switch (cls) {
    case 0:
        if (stage && recompute >= 2) {
            // checkpoint every `recompute` rows; the tags stay in LDS when the budget holds them
            switch (recompute) {
            case 3: if (tags) hipLaunchKernelGGL((fpd::k_baq_ckpt<15, 3, true>), g4, b4, lds, st, a);
                    else hipLaunchKernelGGL((fpd::k_baq_ckpt<15, 3, false>), g4, b4, lds, st, a); break;
            case 5: if (tags) hipLaunchKernelGGL((fpd::k_baq_ckpt<15, 5, true>), g4, b4, lds, st, a);
                    else hipLaunchKernelGGL((fpd::k_baq_ckpt<15, 5, false>), g4, b4, lds, st, a); break;
            default: if (tags) hipLaunchKernelGGL((fpd::k_baq_ckpt<15, 4, true>), g4, b4, lds, st, a);
                     else hipLaunchKernelGGL((fpd::k_baq_ckpt<15, 4, false>), g4, b4, lds, st, a); break;
            }
            break;
        }
        if (stage && wpe == 2) { hipLaunchKernelGGL((fpd::k_baq_band<15, 2, true, false>), g4, b4, lds, st, a); break; }
        if (tags) { hipLaunchKernelGGL((fpd::k_baq_band<15, 1, true, true>), g4, b4, lds, st, a); break; }
        if (stage) { hipLaunchKernelGGL((fpd::k_baq_band<15, 1, true, false>), g4, b4, lds, st, a); break; }
        switch (wpe) {
        case 1: hipLaunchKernelGGL((fpd::k_baq_band<15, 1, false, false>), g4, b4, 0, st, a); break;
        case 3: hipLaunchKernelGGL((fpd::k_baq_band<15, 3, false, false>), g4, b4, 0, st, a); break;
        case 4: hipLaunchKernelGGL((fpd::k_baq_band<15, 4, false, false>), g4, b4, 0, st, a); break;
        default: hipLaunchKernelGGL((fpd::k_baq_band<15, 2, false, false>), g4, b4, 0, st, a); break;
        }
        break;
    case 1:
        if (tags) hipLaunchKernelGGL((fpd::k_baq_band<23, 1, true, true>), g4, b4, lds, st, a);
        else if (stage) hipLaunchKernelGGL((fpd::k_baq_band<23, 1, true, false>), g4, b4, lds, st, a);
        else hipLaunchKernelGGL((fpd::k_baq_band<23, 1, false, false>), g4, b4, 0, st, a);
        break;
    case 2:
        if (tags) hipLaunchKernelGGL((fpd::k_baq_band<33, 1, true, true>), g4, b4, lds, st, a);
        else if (stage) hipLaunchKernelGGL((fpd::k_baq_band<33, 1, true, false>), g4, b4, lds, st, a);
        else hipLaunchKernelGGL((fpd::k_baq_band<33, 1, false, false>), g4, b4, 0, st, a);
        break;
    default: hipLaunchKernelGGL((fpd::k_baq_hmm<0, 2>), dim3(waves), dim3(64), 0, st, a); break;
    }
    return hipGetLastError();
}

extern "C" hipError_t fpd_launch_baq_post(const fpd::BatchDev *b, const fpd::ReadStateDev *s, uint32_t band_done,
                                          hipStream_t st)
{
    if (b->num_reads == 0) return hipSuccess;
    fpd::PostArgs a{};
    a.b = *b; a.s = *s; a.band_done = band_done;
    hipLaunchKernelGGL(fpd::k_baq_post, dim3((b->num_reads + 255) / 256), dim3(256), 0, st, a);
    return hipGetLastError();
}

// fill n bytes, or the staged batch's quality count when live is set (n is then the capacity)
extern "C" hipError_t fpd_fill_u8(uint8_t *p, uint64_t n, const uint32_t *live, uint8_t v, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    const uint64_t blocks = std::min<uint64_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(fpd::k_fill_u8, dim3((unsigned)blocks), dim3(256), 0, st, p, n, live, v);
    return hipGetLastError();
}

// device-side fill (graph-safe replacement for memset)
extern "C" hipError_t fpd_fill_u32(uint32_t *p, uint64_t n, uint32_t v, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    const uint64_t blocks = std::min<uint64_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(fpd::k_fill_u32, dim3((unsigned)blocks), dim3(256), 0, st, p, n, v);
    return hipGetLastError();
}
