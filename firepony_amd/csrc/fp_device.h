/*
 * Shared device-side definitions of the firepony_amd engine.
 *
 * Layouts in HBM (per batch, all SoA):
 *   batch inputs        exactly the fp_batch_view arrays (include/firepony_amd.h)
 *   per read            status u8, windows (ushort2 as u32), num_errors u16,
 *                       hmm window (short2 as u32), bandwidth u16
 *   per padded base     flags u8: bit0 snp, bit1 insertion, bit2 deletion,
 *                       bit3 masked (non-ACGT, Q<6 or known site)
 *   per quality         BAQ tag u8, MAP state u32, fractional errors 3 x f64
 *   tables (resident)   direct-indexed by covariate key: count u64 and the
 *                       exact mismatch sum (fixed point, units of 2^-80) as
 *                       three u64 limb sums
 */
#ifndef FP_DEVICE_H
#define FP_DEVICE_H

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fpd {

// alignment_data.h:42-68
enum : uint32_t {
    F_PAIRED = 1, F_PROPER_PAIR = 2, F_UNMAP = 4, F_MATE_UNMAP = 8, F_REVERSE = 16,
    F_MATE_REVERSE = 32, F_READ1 = 64, F_READ2 = 128, F_SECONDARY = 256, F_QC_FAIL = 512,
    F_DUPLICATE = 1024, F_SUPPLEMENTARY = 2048
};
// CIGAR op codes (types.h:61-72)
enum : uint32_t { OP_M = 0, OP_I = 1, OP_D = 2, OP_N = 3, OP_S = 4, OP_H = 5, OP_P = 6, OP_EQ = 7, OP_X = 8 };
// event classes (cigar.h:41-47)
enum : uint32_t { EV_M = 0, EV_I = 1, EV_D = 2, EV_S = 3 };
// IUPAC16 codes
enum : uint32_t { B_A = 1, B_C = 2, B_G = 4, B_T = 8, B_N = 15 };

enum : uint8_t { RS_FILTERED = 0, RS_ACTIVE = 1, RS_BAQ_DROPPED = 2 };

enum : uint8_t { BF_SNP = 1, BF_INS = 2, BF_DEL = 4, BF_MASKED = 8 };

static constexpr uint16_t NONE16 = 0xffff;

// covariate table geometry: key bits 18 / 28 / 29, RG in the top 8 bits
static constexpr int TABLE_BITS[3] = {18, 28, 29};

// Per-batch counter words.  Each batch state owns a block of u32 words that the front stage of its
// batch zeroes first; words 0..3 are the BAQ band-class sizes and 8..10 the band kernels' chunk
// counters.  The words below are the batch's own read counts and error flags, so an error of one
// staged batch never outlives that batch.
enum : uint32_t {
    BW_OVERFLOW = 5,    // sort aggregation: the batch's pairs overflowed their buffer
    BW_ACTIVE = 11,     // reads left active by the filters
    BW_FILTERED = 12,   // reads dropped by the filters
    BW_BAQ = 13,        // active reads that need BAQ (at least one error)
    BW_LONGEST = 14,    // longest read of the batch (read_len)
    BW_BAD_RG = 15,     // active reads whose read group lies beyond the tables
    BW_WORDS = 16       // words read back after the front stage
};

struct BatchDev {
    uint32_t num_reads;            // reads, or the capacity when `live` is set
    const uint32_t *live;          // device {reads, qualities lo, qualities hi} of a staged batch; null: num_reads
    const uint16_t *chromosome;
    const uint32_t *alignment_start;
    const uint32_t *alignment_stop;
    const uint32_t *mate_chromosome;
    const uint32_t *mate_alignment_start;
    const int32_t *isize;
    const uint32_t *cigars;
    const uint32_t *cigar_start;
    const uint32_t *cigar_len;
    const uint32_t *reads;
    const uint32_t *read_start;
    const uint32_t *read_len;
    const uint8_t *quals;
    const uint32_t *qual_start;
    const uint32_t *qual_len;
    const uint16_t *flags;
    const uint8_t *mapq;
    const uint32_t *read_group;
};

__device__ inline uint32_t live_reads(const BatchDev &b) { return b.live ? b.live[0] : b.num_reads; }

// Wave votes over the lanes still running, as ballots, which the compiler keeps where they are
// written.  HIP's __all and __any call library functions declared const, which the optimiser may
// move to a point where other lanes run, so their answer need not describe the lanes that branch
// on it.
__device__ inline bool wave_all(bool p) { return __ballot(p) == __ballot(1); }
__device__ inline bool wave_any(bool p) { return __ballot(p) != 0; }

struct GenomeDev {
    uint32_t num_contigs;
    const uint64_t *word_offset;
    const uint32_t *length;
    const uint32_t *words;
    uint32_t num_variant_contigs;
    const uint64_t *feature_offset;
    const uint32_t *feature_start;
    const uint32_t *feature_stop;
    const uint32_t *max_end_left;
    // per-contig bucket index of the known sites (buckets of 2^SITE_BUCKET_BITS reference
    // positions): bucket j of contig c holds, relative to the contig's first site, the first site
    // whose max_end_left is >= j * 2^bits (site_lb) and the first whose start is >= j * 2^bits
    // (site_sb); a contig's last entry stands for +infinity (its site count)
    const uint64_t *bucket_offset;   // num_variant_contigs + 1
    const uint32_t *site_lb;
    const uint32_t *site_sb;
};
static constexpr uint32_t SITE_BUCKET_BITS = 9;

struct ReadStateDev {
    uint8_t *status;
    uint32_t *rwc;        // read_window_clipped   (x | y << 16)
    uint32_t *noins;      // read_window_clipped_no_insertions
    uint32_t *refw;       // reference_window_clipped
    uint16_t *num_errors;
    uint32_t *hmmw;       // hmm reference window, int16 x | int16 y << 16
    uint16_t *bandwidth;
    uint32_t *hqw;        // high quality window
    uint8_t *base_flags;  // per padded base
    uint8_t *baq;         // per quality: BQ tag (0xff = unset)
    uint32_t *baq_state;  // per quality: MAP state
    double *frac;         // per quality x3 (snp, ins, del), interleaved
};

// Dense covariate tables indexed by key.  The exact mismatch sum of an entry
// (fixed point, units of 2^-80) is kept as three u64 sums of 27-bit limbs,
// value = part[0] + part[1] * 2^27 + part[2] * 2^54, so every update is a
// plain atomic add that needs no carry and returns nothing.  A per-base value
// (<= 1.0, i.e. <= 2^80) has limbs < 2^27, which leaves room for 2^37 updates
// per entry.
struct TablesDev {
    uint64_t *count[3];
    uint64_t *part[3][3];     // [table][limb]
    uint32_t num_rg;
};

__host__ __device__ inline uint32_t pack2(uint32_t x, uint32_t y) { return (x & 0xffff) | (y << 16); }
__host__ __device__ inline uint16_t lo16(uint32_t v) { return (uint16_t)(v & 0xffff); }
__host__ __device__ inline uint16_t hi16(uint32_t v) { return (uint16_t)(v >> 16); }

__device__ inline uint32_t op_code(uint32_t w) { return (w >> 24) & 0xf; }
__device__ inline uint32_t op_len(uint32_t w) { return w & 0xffffff; }

// event class of an op (cigar.cu:68-93 semantics)
__device__ inline uint32_t op_event(uint32_t op)
{
    switch (op) {
    case OP_M: case OP_EQ: case OP_X: return EV_M;
    case OP_I: case OP_N: return EV_I;
    case OP_D: case OP_H: case OP_P: return EV_D;
    default: return EV_S;
    }
}
__device__ inline bool op_reads(uint32_t op) { uint32_t e = op_event(op); return e == EV_M || e == EV_I || e == EV_S; }
__device__ inline bool op_refs(uint32_t op) { uint32_t e = op_event(op); return e == EV_M || e == EV_D; }

__device__ inline uint8_t read_base(const uint32_t *words, uint64_t i)
{
    return (uint8_t)((words[i >> 3] >> ((i & 7) * 4)) & 0xf);
}

// the packed word holding base i (8 bases per word)
__device__ inline uint32_t read_word(const uint32_t *words, uint64_t i) { return words[i >> 3]; }

__device__ inline uint8_t contig_base(const GenomeDev &g, uint32_t c, uint32_t i)
{
    if (c >= g.num_contigs || i >= g.length[c]) return B_N;   // past the contig: N
    return (uint8_t)((g.words[g.word_offset[c] + (i >> 3)] >> ((i & 7) * 4)) & 0xf);
}

// ---------------------------------------------------------------------------
// 128-bit unsigned fixed point helpers (mismatch sums, units of 2^-80)

struct u128 {
    uint64_t lo, hi;
};

// exact v * 2^80 for v = k/n, n <= 65535 (every per-base fractional error)
__host__ __device__ inline u128 fixed_from_double(double v)
{
    u128 r = {0, 0};
    if (v == 0.0) return r;
    const uint64_t bits = __builtin_bit_cast(uint64_t, v);
    const int e = (int)((bits >> 52) & 0x7ff);
    uint64_t mant = (bits & 0xfffffffffffffull);
    int ex;
    if (e == 0) { ex = -1074; }
    else { mant |= 1ull << 52; ex = e - 1075; }
    // v = mant * 2^ex ; result = mant * 2^(ex + 80)
    const int sh = ex + 80;
    if (sh >= 0) {
        if (sh >= 64) { r.lo = 0; r.hi = mant << (sh - 64); }
        else if (sh == 0) { r.lo = mant; r.hi = 0; }
        else { r.lo = mant << sh; r.hi = mant >> (64 - sh); }
    } else {
        r.lo = (-sh >= 64) ? 0 : (mant >> (-sh));
        r.hi = 0;
    }
    return r;
}

static constexpr uint64_t LIMB_MASK = (1ull << 27) - 1;

// limbs of a fixed-point value (value < 2^118)
__host__ __device__ inline void fixed_split(const u128 &v, uint64_t &p0, uint64_t &p1, uint64_t &p2)
{
    p0 = v.lo & LIMB_MASK;
    p1 = (v.lo >> 27) & LIMB_MASK;
    p2 = (v.lo >> 54) | (v.hi << 10);
}

// p0 + p1 * 2^27 + p2 * 2^54 as a u128
__host__ __device__ inline u128 fixed_combine(uint64_t p0, uint64_t p1, uint64_t p2)
{
    u128 r;
    r.lo = p0;
    r.hi = 0;
    const uint64_t a = p1 << 27;
    r.lo += a;
    r.hi += (r.lo < a ? 1 : 0) + (p1 >> 37);
    const uint64_t b = p2 << 54;
    r.lo += b;
    r.hi += (r.lo < b ? 1 : 0) + (p2 >> 10);
    return r;
}

// round-to-nearest-even conversion of a u128 fixed-point value to double
__host__ __device__ inline double fixed_to_double(uint64_t lo, uint64_t hi)
{
    if (hi == 0 && lo == 0) return 0.0;
    // find the top bit
    int top;
    if (hi) top = 127 - __builtin_clzll(hi);
    else top = 63 - __builtin_clzll(lo);
    uint64_t mant;       // 53 significant bits
    bool round_up = false;
    if (top <= 52) {
        mant = lo;   // exact (hi == 0)
    } else {
        const int drop = top - 52;   // bits below the 53-bit mantissa
        // mant = value >> drop
        if (drop >= 64) mant = hi >> (drop - 64);
        else mant = (lo >> drop) | (drop ? (hi << (64 - drop)) : 0);
        // remainder bits
        uint64_t half_hi, half_lo, rem_hi, rem_lo;
        // rem = value & ((1 << drop) - 1), half = 1 << (drop - 1)
        if (drop >= 64) {
            rem_lo = lo;
            rem_hi = (drop == 64) ? 0 : (hi & ((1ull << (drop - 64)) - 1));
        } else {
            rem_lo = lo & ((1ull << drop) - 1);
            rem_hi = 0;
        }
        if (drop - 1 >= 64) { half_hi = 1ull << (drop - 1 - 64); half_lo = 0; }
        else { half_hi = 0; half_lo = 1ull << (drop - 1); }
        const bool gt = rem_hi > half_hi || (rem_hi == half_hi && rem_lo > half_lo);
        const bool eq = rem_hi == half_hi && rem_lo == half_lo;
        round_up = gt || (eq && (mant & 1));
        if (round_up) mant += 1;   // may carry into bit 53; the product below stays exact
        // value ~= mant * 2^drop
        double d = (double)mant;   // mant <= 2^53 exactly representable
        return ldexp(d, drop - 80);
    }
    return ldexp((double)mant, -80);
}

} // namespace fpd

#endif
