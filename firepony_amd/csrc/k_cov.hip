/*
 * K_cov: fractional errors (device/fractional_errors.cu:39-141), the high
 * quality window (device/covariates.cu:215-246) and covariate key packing
 * (covariates.cu:50-103, covariates/packer_*.h, covariates/bit_packers/*.h),
 * accumulated straight into the device-resident covariate tables.
 *
 * The tables are direct-indexed by covariate key: the read group is the most
 * significant field of every key layout, so the dense array index of a key is
 * the key itself and the rows of a table come out in key order -- the order
 * the reference obtains by radix-sorting (covariate_table.cu:43-58).  Counts
 * are u64; mismatch sums are exact u128 fixed point (units of 2^-80), so the
 * accumulated table does not depend on batch size, launch order or the order
 * atomics land in.
 */
#include "fp_device.h"

namespace fpd {

struct CovArgs {
    BatchDev b;
    ReadStateDev s;
    TablesDev t;
    uint32_t frac_f64;
    uint32_t *counters;   // [4] reads with an out-of-range read group
};

// calculateFractionalErrorArray for one error class over the clipped window
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

__device__ inline void table_add(const TablesDev &t, int tab, uint32_t key, double v, uint32_t frac_f64)
{
    atomicAdd((unsigned long long *)&t.count[tab][key], 1ull);
    if (v != 0.0) {
        const double vv = frac_f64 ? v : (double)(float)v;
        const u128 f = fixed_from_double(vv);
        const unsigned long long old = atomicAdd((unsigned long long *)&t.sum_lo[tab][key], (unsigned long long)f.lo);
        const uint64_t carry = (old + f.lo < old) ? 1 : 0;
        if (f.hi + carry) atomicAdd((unsigned long long *)&t.sum_hi[tab][key], (unsigned long long)(f.hi + carry));
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

// covariate_Context with 2 mismatch / 3 indel bases (bit_packers/context.h:104-174)
__device__ inline void context_keys(const uint32_t *reads, uint64_t rs, uint16_t hx, uint16_t hy, bool neg,
                                    uint16_t bp, uint32_t &km, uint32_t &kid)
{
    km = 2047u; kid = 2047u;
    const int dir = neg ? 1 : -1;
    int avail = neg ? (int)hy - (int)bp + 1 : (int)bp - (int)hx + 1;
    const int cb = avail < 3 ? avail : 3;
    if (cb < 2) return;
    uint32_t ctx = 0;
    int n = 0;
    for (int i = bp, stop = (int)bp + cb * dir; i != stop; i += dir) {
        const uint8_t base = read_base(reads, rs + (uint32_t)i);
        if (base != B_A && base != B_C && base != B_T && base != B_G) break;
        ctx <<= 2;
        ctx |= 31 - __builtin_clz((uint32_t)base);
        n++;
    }
    if (neg) ctx = ~ctx & ((1u << (n * 2)) - 1);
    if (n >= 2) km = ((ctx >> ((n - 2) * 2)) << 4) | 2;
    if (n >= 3) kid = ((ctx >> ((n - 3) * 2)) << 4) | 3;
}

__global__ void __launch_bounds__(256) k_cov(CovArgs a)
{
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    const BatchDev &b = a.b;
    if (r >= b.num_reads) return;
    if (a.s.status[r] != RS_ACTIVE) return;
    const uint32_t rg = b.read_group[r];
    if (rg >= a.t.num_rg) { atomicAdd(&a.counters[4], 1u); return; }

    const uint32_t rw = a.s.rwc[r];
    const uint16_t rwx = lo16(rw), rwy = hi16(rw);
    const uint32_t rs = b.read_start[r];
    const uint32_t qs = b.qual_start[r];
    const int len = (int)rwy - (int)rwx + 1;
    const uint8_t *baq = a.s.baq + qs + rwx;
    const uint8_t *bfl = a.s.base_flags + rs + rwx;
    double *frac = a.s.frac + 3ull * (qs + rwx);
    fractional_errors(baq, bfl, len, BF_SNP, frac + 0);
    fractional_errors(baq, bfl, len, BF_INS, frac + 1);
    fractional_errors(baq, bfl, len, BF_DEL, frac + 2);

    // high quality window
    uint16_t hx = rwx, hy = rwy;
    while (b.quals[qs + hx] <= 2 && hx < hy) hx++;
    while (b.quals[qs + hy] <= 2 && hy > hx) hy--;
    a.s.hqw[r] = pack2(hx, hy);

    const uint16_t flags = b.flags[r];
    const bool neg = (flags & F_REVERSE) != 0;
    const uint32_t *ops = b.cigars + b.cigar_start[r];
    const uint32_t nops = b.cigar_len[r];
    uint16_t ro = 0;
    for (uint32_t oi = 0; oi < nops; oi++) {
        const uint32_t op = op_code(ops[oi]), l = op_len(ops[oi]);
        if (!op_reads(op)) continue;
        const uint32_t ev = op_event(op);
        if (ev != EV_S) {
            for (uint32_t k = 0; k < l; k++) {
                const uint16_t bp = (uint16_t)(ro + k);
                if (bp < rwx || bp > rwy) continue;
                if (a.s.base_flags[rs + bp] & BF_MASKED) continue;
                const uint32_t q = b.quals[qs + bp];
                const double *fe = a.s.frac + 3ull * (qs + bp);
                // RecalTable1: EventTracker(2) | QualityScore(8) | ReadGroup(8)
                table_add(a.t, 0, 0u | (q << 2) | (rg << 10), fe[0], a.frac_f64);
                table_add(a.t, 0, 1u | (45u << 2) | (rg << 10), fe[1], a.frac_f64);
                table_add(a.t, 0, 2u | (45u << 2) | (rg << 10), fe[2], a.frac_f64);
                // cycle: EventTracker(2) | Cycle(10) | QualityScore(8) | ReadGroup(8)
                uint32_t cm, cid;
                cycle_keys(flags, rwx, rwy, bp, cm, cid);
                table_add(a.t, 1, 0u | (cm << 2) | (q << 12) | (rg << 20), fe[0], a.frac_f64);
                if (cid != 1023u) {
                    table_add(a.t, 1, 1u | (cid << 2) | (45u << 12) | (rg << 20), fe[1], a.frac_f64);
                    table_add(a.t, 1, 2u | (cid << 2) | (45u << 12) | (rg << 20), fe[2], a.frac_f64);
                }
                // context: EventTracker(2) | Context(11) | QualityScore(8) | ReadGroup(8)
                uint32_t xm, xid;
                context_keys(b.reads, rs, hx, hy, neg, bp, xm, xid);
                if (xm != 2047u) table_add(a.t, 2, 0u | (xm << 2) | (q << 13) | (rg << 21), fe[0], a.frac_f64);
                if (xid != 2047u) {
                    table_add(a.t, 2, 1u | (xid << 2) | (45u << 13) | (rg << 21), fe[1], a.frac_f64);
                    table_add(a.t, 2, 2u | (xid << 2) | (45u << 13) | (rg << 21), fe[2], a.frac_f64);
                }
            }
        }
        ro = (uint16_t)(ro + l);
    }
}

} // namespace fpd

extern "C" hipError_t fpd_launch_cov(const fpd::BatchDev *b, const fpd::ReadStateDev *s, const fpd::TablesDev *t,
                                     uint32_t frac_f64, uint32_t *counters, hipStream_t st)
{
    if (b->num_reads == 0) return hipSuccess;
    fpd::CovArgs a;
    a.b = *b; a.s = *s; a.t = *t; a.frac_f64 = frac_f64; a.counters = counters;
    hipLaunchKernelGGL(fpd::k_cov, dim3((b->num_reads + 255) / 256), dim3(256), 0, st, a);
    return hipGetLastError();
}
