/*
 * K_sort: the sort aggregation of covariate observations (aggregation = FP_AGG_SORT), the method of
 * the reference's covariate tables (covariate_table.cu:43-105: sort_by_key over the table's key bits,
 * then reduce_by_key summing observations and mismatches):
 *
 *   k_cov in pair mode emits one (key u32, value f64) pair per observation and table ->
 *   per table an LSD radix sort of the pairs with D-bit digits (D = 4..8) ->
 *   a segmented reduce_by_key of the sorted runs into the resident direct-indexed tables.
 *
 * Each workgroup of a fixed grid owns one contiguous range of the elements (a whole number of
 * 4096-element tiles).  One radix pass is k_sort_hist (the range's digit counts, digit-major, one
 * row of the grid per digit) -> k_sort_totals + k_sort_offsets (exclusive scan = each range's
 * global output offset per digit) -> k_sort_scatter.  The scatter walks its range tile by tile and
 * ranks a tile's elements stably in LDS: equal digits inside a wavefront are matched with D ballots
 * and counted by per-wave running digit counters, the tile is reordered by digit in LDS, each
 * digit's run is written at the range's running offset for that digit (consecutive lanes write
 * consecutive addresses), and the offsets move on by the tile's counts.  The first pass drops the
 * sentinel keys of unused pair slots.  Element counts stay on the device and the kernels derive
 * their ranges from them, so the whole chain can be captured in a graph.
 *
 * The reduce gives each lane 32 consecutive sorted pairs: the lane sums its runs sequentially
 * (count and exact fixed-point value as three 27-bit limbs, as the direct path keeps them) and adds
 * the runs that start and end inside its slice to the table at once; the runs crossing lane
 * boundaries are joined by one segmented scan across the wavefront per 2048 pairs, and each joined
 * run is added where it ends (non-returning atomics; a run that spans chunks adds once per chunk).
 * Sums are exact, so the tables equal the direct path's bit for bit whatever the digit width.
 */
#include "fp_device.h"

#include <algorithm>

namespace fpd {

static constexpr uint32_t SORT_THREADS = 512;
static constexpr uint32_t SORT_WAVES = SORT_THREADS / 64;
static constexpr uint32_t SORT_ITEMS = 8;
static constexpr uint32_t SORT_TILE = SORT_THREADS * SORT_ITEMS;   // 4096 elements
static constexpr uint32_t SORT_SENTINEL = 0xffffffffu;
static constexpr uint32_t SORT_GRID = 2048;                          // workgroups of hist / scatter
static constexpr uint32_t RED_LANE = 32;                             // consecutive pairs per reduce lane

struct SortPass {
    const uint32_t *kin;
    const double *vin;
    uint32_t *kout;
    double *vout;
    const uint32_t *n_in;     // input elements (device)
    uint32_t *n_out;          // kept elements, written by the offsets kernel (first pass), or null
    uint32_t *hist;           // [radix][grid]: range counts, then range offsets
    uint32_t *totals;         // [radix]
    uint64_t n_cap;           // capacity of the pair arrays (elements)
    uint32_t shift, bits;
    uint32_t drop;            // drop sentinel keys (first pass)
};

__device__ inline uint32_t pass_n(const SortPass &p)
{
    const uint64_t n = *p.n_in;
    return (uint32_t)(n < p.n_cap ? n : p.n_cap);
}

// the contiguous element range [lo, hi) of workgroup b (whole tiles)
__device__ inline void block_range(uint32_t n, uint32_t b, uint32_t grid, uint32_t &lo, uint32_t &hi)
{
    const uint32_t tiles = (n + SORT_TILE - 1) / SORT_TILE;
    const uint32_t per = (tiles + grid - 1) / grid;
    const uint64_t l = (uint64_t)b * per * SORT_TILE, h = l + (uint64_t)per * SORT_TILE;
    lo = (uint32_t)(l < n ? l : n);
    hi = (uint32_t)(h < n ? h : n);
}

// lanes of the wavefront holding the same BITS-bit digit (among valid lanes)
template <uint32_t BITS>
__device__ inline uint64_t match_digit(uint32_t d, bool valid)
{
    uint64_t m = __ballot(valid);
#pragma unroll
    for (uint32_t b = 0; b < BITS; b++) {
        const bool one = (d >> b) & 1;
        const uint64_t x = __ballot(one);
        m &= one ? x : ~x;
    }
    return m;
}

// the range's digit counts: LDS atomics into 16 KB of sub-histograms (per wave and lane group:
// 4 groups at 8-bit digits, up to 16 at narrower ones), which keeps same-address collisions of
// skewed digits (the read-group and quality fields) low
template <uint32_t BITS>
__global__ void __launch_bounds__(SORT_THREADS) k_sort_hist(SortPass p)
{
    constexpr uint32_t R = 1u << BITS, mask = R - 1;
    constexpr uint32_t G = (1024 / R) < 4 ? 4 : ((1024 / R) > 16 ? 16 : (1024 / R));   // groups per wave
    constexpr uint32_t C = SORT_WAVES * G;                                             // sub-histograms
    __shared__ uint32_t h[C][R + 1];   // +1: the copies of one digit fall in different banks
    const uint32_t n = pass_n(p);
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    uint32_t *mine = h[wave * G + (lane & (G - 1))];
    uint32_t lo, hi;
    block_range(n, blockIdx.x, gridDim.x, lo, hi);
    for (uint32_t i = threadIdx.x; i < C * (R + 1); i += SORT_THREADS) (&h[0][0])[i] = 0;
    __syncthreads();
    // four tiles per step: 32 loads in flight per lane (the counting order does not matter)
    constexpr uint32_t HI = 4 * SORT_ITEMS;
    for (uint32_t base = lo; base < hi; base += 4 * SORT_TILE) {
        uint32_t key[HI];
#pragma unroll
        for (uint32_t k = 0; k < HI; k++) {
            const uint32_t i = base + k * SORT_THREADS + threadIdx.x;
            key[k] = i < hi ? p.kin[i] : SORT_SENTINEL;
        }
#pragma unroll
        for (uint32_t k = 0; k < HI; k++) {
            const uint32_t i = base + k * SORT_THREADS + threadIdx.x;
            if (i < hi && !(p.drop && key[k] == SORT_SENTINEL)) atomicAdd(&mine[(key[k] >> p.shift) & mask], 1u);
        }
    }
    __syncthreads();
    for (uint32_t d = threadIdx.x; d < R; d += SORT_THREADS) {
        uint32_t t = 0;
#pragma unroll
        for (uint32_t c = 0; c < C; c++) t += h[c][d];
        p.hist[(uint64_t)d * gridDim.x + blockIdx.x] = t;
    }
}

__device__ inline uint32_t wave_incl_scan(uint32_t v)
{
    const uint32_t lane = threadIdx.x & 63;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t u = (uint32_t)__shfl_up((int)v, off);
        if (lane >= (uint32_t)off) v += u;
    }
    return v;
}

// exclusive scan of one value per thread over a 256-thread workgroup; *total gets the sum
__device__ inline uint32_t block_excl_scan(uint32_t v, uint32_t *wtot, uint32_t *total)
{
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t inc = wave_incl_scan(v);
    if (lane == 63) wtot[wave] = inc;
    __syncthreads();
    uint32_t before = 0, all = 0;
    for (uint32_t w = 0; w < SORT_THREADS / 64; w++) {
        if (w < wave) before += wtot[w];
        all += wtot[w];
    }
    __syncthreads();
    *total = all;
    return before + inc - v;
}

// totals[d] = sum over the ranges of hist[d][range]; one workgroup per digit
__global__ void __launch_bounds__(SORT_THREADS) k_sort_totals(SortPass p, uint32_t grid)
{
    __shared__ uint32_t wtot[SORT_WAVES];
    const uint32_t d = blockIdx.x;
    const uint32_t *h = p.hist + (uint64_t)d * grid;
    uint32_t s = 0;
    for (uint32_t t = threadIdx.x; t < grid; t += SORT_THREADS) s += h[t];
    uint32_t total;
    (void)block_excl_scan(s, wtot, &total);
    if (threadIdx.x == 0) p.totals[d] = total;
}

// hist[d][r] <- sum of totals below d + sum of hist[d][r' < r]; one workgroup per digit, eight
// consecutive ranges per thread and step
__global__ void __launch_bounds__(SORT_THREADS) k_sort_offsets(SortPass p, uint32_t grid)
{
    __shared__ uint32_t wtot[SORT_WAVES];
    const uint32_t d = blockIdx.x;
    const uint32_t R = gridDim.x;       // one workgroup per digit
    uint32_t below = 0, all = 0;
    for (uint32_t x = 0; x < R; x++) {
        const uint32_t t = p.totals[x];
        if (x < d) below += t;
        all += t;
    }
    if (d == 0 && threadIdx.x == 0 && p.n_out) *p.n_out = all;
    uint32_t *h = p.hist + (uint64_t)d * grid;
    uint32_t running = below;
    for (uint32_t t0 = 0; t0 < grid; t0 += SORT_THREADS * 8) {
        const uint32_t tb = t0 + threadIdx.x * 8;
        uint32_t v[8], s = 0;
#pragma unroll
        for (int j = 0; j < 8; j++) {
            v[j] = tb + j < grid ? h[tb + j] : 0u;
            s += v[j];
        }
        uint32_t step;
        uint32_t ex = block_excl_scan(s, wtot, &step) + running;
#pragma unroll
        for (int j = 0; j < 8; j++) {
            if (tb + j < grid) h[tb + j] = ex;
            ex += v[j];
        }
        running += step;
    }
}

template <uint32_t BITS>
__global__ void __launch_bounds__(SORT_THREADS) k_sort_scatter(SortPass p)
{
    __shared__ uint32_t cnt[SORT_WAVES][256];   // per-wave running digit counters, then per-wave bases
    __shared__ uint32_t gofs[256];      // the range's running global offset per digit
    __shared__ uint32_t lofs[256];      // start of each digit in the tile's digit order
    __shared__ uint32_t wtot[SORT_WAVES];
    __shared__ uint32_t kept_s;
    __shared__ uint32_t lk[SORT_TILE];
    __shared__ double lv[SORT_TILE];
    const uint32_t n = pass_n(p);
    constexpr uint32_t R = 1u << BITS, mask = R - 1;
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint64_t lt = (1ull << lane) - 1;
    uint32_t lo, hi;
    block_range(n, blockIdx.x, gridDim.x, lo, hi);
    if (lo >= hi) return;
    for (uint32_t d = threadIdx.x; d < R; d += SORT_THREADS) gofs[d] = p.hist[(uint64_t)d * gridDim.x + blockIdx.x];
    // element order inside a tile is (wave, item, lane); the next tile's elements are loaded while
    // this one is ranked, reordered and written
    uint32_t key[SORT_ITEMS], nkey[SORT_ITEMS] = {};
    double val[SORT_ITEMS], nval[SORT_ITEMS] = {};
    auto load_tile = [&](uint32_t base, uint32_t (&k_)[SORT_ITEMS], double (&v_)[SORT_ITEMS]) {
        const uint32_t t0 = base + wave * (SORT_ITEMS * 64);
#pragma unroll
        for (uint32_t k = 0; k < SORT_ITEMS; k++) {
            const uint32_t i = t0 + k * 64 + lane;
            k_[k] = i < hi ? p.kin[i] : SORT_SENTINEL;
        }
#pragma unroll
        for (uint32_t k = 0; k < SORT_ITEMS; k++) {
            const uint32_t i = t0 + k * 64 + lane;
            v_[k] = i < hi ? p.vin[i] : 0.0;
        }
    };
    load_tile(lo, key, val);
    for (uint32_t base = lo; base < hi; base += SORT_TILE) {
        for (uint32_t i = threadIdx.x; i < SORT_WAVES * 256; i += SORT_THREADS) (&cnt[0][0])[i] = 0;
        if (base + SORT_TILE < hi) load_tile(base + SORT_TILE, nkey, nval);
        __syncthreads();
        // stable ranks
        const uint32_t t0 = base + wave * (SORT_ITEMS * 64);
        uint32_t rank[SORT_ITEMS];
        uint32_t vmask = 0;
#pragma unroll
        for (uint32_t k = 0; k < SORT_ITEMS; k++) {
            const uint32_t i = t0 + k * 64 + lane;
            const bool valid = i < hi && !(p.drop && key[k] == SORT_SENTINEL);
            vmask |= (uint32_t)valid << k;
            const uint32_t d = (key[k] >> p.shift) & mask;
            const uint64_t peers = match_digit<BITS>(d, valid);
            const uint32_t c = valid ? cnt[wave][d] : 0u;
            // every lane has read the counter before the leader of each digit moves it on
            __builtin_amdgcn_wave_barrier();
            if (valid && (peers & lt) == 0) cnt[wave][d] = c + (uint32_t)__popcll(peers);
            __builtin_amdgcn_wave_barrier();
            rank[k] = c + (uint32_t)__popcll(peers & lt);
        }
        __syncthreads();
        // digit bases per wave, and the start of each digit in the tile
        uint32_t tot = 0;
        {
            const uint32_t d = threadIdx.x;
            uint32_t t[SORT_WAVES] = {};
            if (d < R) {
                for (uint32_t w = 0; w < SORT_WAVES; w++) { t[w] = cnt[w][d]; tot += t[w]; }
            }
            __syncthreads();
            if (d < R) {
                uint32_t run = 0;
                for (uint32_t w = 0; w < SORT_WAVES; w++) { cnt[w][d] = run; run += t[w]; }
            }
            uint32_t kept;
            const uint32_t start = block_excl_scan(d < R ? tot : 0u, wtot, &kept);
            if (d < R) lofs[d] = start;
            if (threadIdx.x == 0) kept_s = kept;
        }
        __syncthreads();
#pragma unroll
        for (uint32_t k = 0; k < SORT_ITEMS; k++) {
            if (!((vmask >> k) & 1)) continue;
            const uint32_t d = (key[k] >> p.shift) & mask;
            const uint32_t pos = lofs[d] + cnt[wave][d] + rank[k];
            lk[pos] = key[k];
            lv[pos] = val[k];
        }
        __syncthreads();
        const uint32_t kept = kept_s;
        for (uint32_t j = threadIdx.x; j < kept; j += SORT_THREADS) {
            const uint32_t kk = lk[j];
            const uint32_t d = (kk >> p.shift) & mask;
            const uint32_t g = gofs[d] + (j - lofs[d]);
            p.kout[g] = kk;
            p.vout[g] = lv[j];
        }
        __syncthreads();
        if (threadIdx.x < R) gofs[threadIdx.x] += tot;
#pragma unroll
        for (uint32_t k = 0; k < SORT_ITEMS; k++) { key[k] = nkey[k]; val[k] = nval[k]; }
    }
}

struct ReduceArgs {
    const uint32_t *keys;
    const double *vals;
    const uint32_t *n;
    uint64_t n_cap;
    uint32_t key_limit;       // entries of the table
    unsigned long long *count, *p0, *p1, *p2;
};

template <typename T> __device__ inline T shfl_up_t(T v, int off) { return __shfl_up(v, off); }

__device__ inline void table_add(const ReduceArgs &a, uint32_t key, uint64_t c, uint64_t v0, uint64_t v1, uint64_t v2)
{
    if (key >= a.key_limit || c == 0) return;
    atomicAdd(&a.count[key], (unsigned long long)c);
    if (v0) atomicAdd(&a.p0[key], (unsigned long long)v0);
    if (v1) atomicAdd(&a.p1[key], (unsigned long long)v1);
    if (v2) atomicAdd(&a.p2[key], (unsigned long long)v2);
}

struct RunSum {
    uint64_t c, v0, v1, v2;
};

// One wave per workgroup and per chunk of 64 x RED_LANE sorted pairs.  The chunk is loaded into LDS
// with coalesced 16-byte loads (a lane's own RED_LANE consecutive pairs would be 64 scattered
// partial lines per load instruction), padded by one entry per RED_LANE so that the lanes' reads of
// their consecutive slices fall in different banks; then each lane sums its slice run by run.
__global__ void __launch_bounds__(64) k_seg_reduce(ReduceArgs a)
{
    constexpr uint32_t CE = 64 * RED_LANE;
    constexpr uint32_t PADDED = CE + CE / RED_LANE;
    __shared__ uint32_t sk[PADDED];
    __shared__ double sv[PADDED];
    const uint64_t nn = *a.n;
    const uint32_t n = (uint32_t)(nn < a.n_cap ? nn : a.n_cap);
    const uint32_t chunks = (n + CE - 1) / CE;
    const uint32_t lane = threadIdx.x & 63;
    auto slot = [](uint32_t e) { return e + e / RED_LANE; };
    // a full chunk's pairs in registers: the next chunk is loaded while this one is summed
    uint4 rk[CE / 256];
    double2 rv[CE / 128];
    auto load_chunk = [&](uint32_t base) {
        const uint4 *k4 = (const uint4 *)(a.keys + base);
        const double2 *v2 = (const double2 *)(a.vals + base);
#pragma unroll
        for (uint32_t i = 0; i < CE / 256; i++) rk[i] = k4[lane + 64 * i];
#pragma unroll
        for (uint32_t i = 0; i < CE / 128; i++) rv[i] = v2[lane + 64 * i];
    };
    if (blockIdx.x < chunks && (uint64_t)(blockIdx.x + 1) * CE <= n) load_chunk(blockIdx.x * CE);
    for (uint32_t chunk = blockIdx.x; chunk < chunks; chunk += gridDim.x) {
        const uint32_t base = chunk * CE;
        if (base + CE <= n) {
#pragma unroll
            for (uint32_t i = 0; i < CE / 256; i++) {
                const uint32_t e = 4 * (lane + 64 * i);
                const uint4 k = rk[i];
                sk[slot(e)] = k.x; sk[slot(e + 1)] = k.y; sk[slot(e + 2)] = k.z; sk[slot(e + 3)] = k.w;
            }
#pragma unroll
            for (uint32_t i = 0; i < CE / 128; i++) {
                const uint32_t e = 2 * (lane + 64 * i);
                const double2 v = rv[i];
                sv[slot(e)] = v.x; sv[slot(e + 1)] = v.y;
            }
        } else {
            for (uint32_t e = lane; e < CE; e += 64) {
                const bool in = base + e < n;
                sk[slot(e)] = in ? a.keys[base + e] : SORT_SENTINEL;
                sv[slot(e)] = in ? a.vals[base + e] : 0.0;
            }
        }
        __syncthreads();
        {
            const uint32_t next = chunk + gridDim.x;
            if (next < chunks && (uint64_t)(next + 1) * CE <= n) load_chunk(next * CE);
        }
        // this lane's slice, summed run by run
        uint32_t hk = SORT_SENTINEL, cur = SORT_SENTINEL;
        RunSum hs = {0, 0, 0, 0}, cs = {0, 0, 0, 0};
        bool single = true;     // the slice holds one run so far
        const uint32_t s0 = lane * (RED_LANE + 1);
#pragma unroll 8
        for (uint32_t q = 0; q < RED_LANE; q++) {
            const uint32_t k = sk[s0 + q];
            if (k == SORT_SENTINEL) continue;     // past the end
            // most per-base errors are exactly 0 or 1 (1.0 = 2^80: limbs 0, 0, 2^26)
            uint64_t l0 = 0, l1 = 0, l2 = 0;
            const double v = sv[s0 + q];
            if (v == 1.0) l2 = 1ull << 26;
            else if (v != 0.0) fixed_split(fixed_from_double(v), l0, l1, l2);
            if (k != cur) {
                if (cur != SORT_SENTINEL) {
                    if (single) { hs = cs; single = false; }            // the slice's first run ended
                    else table_add(a, cur, cs.c, cs.v0, cs.v1, cs.v2);  // a run inside the slice
                } else {
                    hk = k;
                }
                cur = k;
                cs = {0, 0, 0, 0};
            }
            cs.c += 1; cs.v0 += l0; cs.v1 += l1; cs.v2 += l2;
        }
        __syncthreads();   // the next chunk overwrites the LDS copy
        // lane summary: head run (hk, hs) unless single; tail run (cur, cs) in every lane
        if (single) hk = cur;
        const uint32_t tk = cur;
        // join runs across lanes: the carry out of lane l is its tail run plus, when the slice is one
        // run continuing the left neighbour's tail, that neighbour's carry
        const uint32_t prev_tk = (uint32_t)__shfl_up((int)tk, 1);
        const bool joins_left = lane > 0 && prev_tk == hk && hk != SORT_SENTINEL;
        bool f = !(single && joins_left);
        RunSum x = cs;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint64_t uc = shfl_up_t(x.c, off), u0 = shfl_up_t(x.v0, off), u1 = shfl_up_t(x.v1, off),
                           u2 = shfl_up_t(x.v2, off);
            const bool uf = __shfl_up((int)f, off) != 0;
            if (lane >= (uint32_t)off && !f) {
                x.c += uc; x.v0 += u0; x.v1 += u1; x.v2 += u2;
                f = uf;
            }
        }
        // the left neighbour's carry, for a head run that continues it
        RunSum in = {shfl_up_t(x.c, 1), shfl_up_t(x.v0, 1), shfl_up_t(x.v1, 1), shfl_up_t(x.v2, 1)};
        if (!single) {
            if (joins_left) { hs.c += in.c; hs.v0 += in.v0; hs.v1 += in.v1; hs.v2 += in.v2; }
            table_add(a, hk, hs.c, hs.v0, hs.v1, hs.v2);
        }
        const uint32_t next_hk = (uint32_t)__shfl_down((int)hk, 1);
        if (tk != SORT_SENTINEL && (lane == 63 || next_hk != tk)) table_add(a, tk, x.c, x.v0, x.v1, x.v2);
    }
}

} // namespace fpd

// one table: radix passes over its key bits, then the reduce into the resident table.  keys/vals
// hold the emitted pairs (n_emit of them, sentinels included); tkeys/tvals are the ping-pong
// buffers.  n_kept receives the number of real pairs.  hist holds 256 x SORT_GRID counters.
extern "C" hipError_t fpd_sort_reduce(uint32_t *keys, double *vals, uint32_t *tkeys, double *tvals, uint64_t cap,
                                      const uint32_t *n_emit, uint32_t *n_kept, uint32_t *hist, uint32_t *totals,
                                      uint32_t key_bits, uint32_t digit_bits, uint32_t table_entries,
                                      unsigned long long *count, unsigned long long *const *parts, hipStream_t st)
{
    if (cap == 0) return hipSuccess;
    if (digit_bits < 4 || digit_bits > 8 || key_bits == 0 || key_bits > 32) return hipErrorInvalidValue;
    const uint32_t tiles_cap = (uint32_t)((cap + fpd::SORT_TILE - 1) / fpd::SORT_TILE);
    const uint32_t grid = std::min<uint32_t>(tiles_cap, fpd::SORT_GRID);
    uint32_t *ka = keys, *kb = tkeys;
    double *va = vals, *vb = tvals;
    // ceil(key_bits / digit_bits) passes with the key bits spread evenly over them (a narrower digit
    // ranks and writes faster, so 12 bits sort as 6 + 6 rather than 8 + 4)
    const uint32_t npass = (key_bits + digit_bits - 1) / digit_bits;
    for (uint32_t shift = 0, pass = 0; pass < npass; pass++) {
        const uint32_t bits = key_bits / npass + (pass < key_bits % npass ? 1u : 0u);
        fpd::SortPass p{};
        p.kin = ka; p.vin = va; p.kout = kb; p.vout = vb;
        p.n_in = pass == 0 ? n_emit : n_kept;
        p.n_out = pass == 0 ? n_kept : nullptr;
        p.hist = hist; p.totals = totals; p.n_cap = cap;
        p.shift = shift;
        p.bits = bits;
        p.drop = pass == 0 ? 1u : 0u;
        const uint32_t R = 1u << p.bits;
        switch (p.bits) {
#define FP_SORT_BITS(B)                                                                                      \
    case B:                                                                                                  \
        hipLaunchKernelGGL(fpd::k_sort_hist<B>, dim3(grid), dim3(fpd::SORT_THREADS), 0, st, p);              \
        hipLaunchKernelGGL(fpd::k_sort_totals, dim3(R), dim3(fpd::SORT_THREADS), 0, st, p, grid);            \
        hipLaunchKernelGGL(fpd::k_sort_offsets, dim3(R), dim3(fpd::SORT_THREADS), 0, st, p, grid);           \
        hipLaunchKernelGGL(fpd::k_sort_scatter<B>, dim3(grid), dim3(fpd::SORT_THREADS), 0, st, p);           \
        break;
            FP_SORT_BITS(1) FP_SORT_BITS(2) FP_SORT_BITS(3) FP_SORT_BITS(4)
            FP_SORT_BITS(5) FP_SORT_BITS(6) FP_SORT_BITS(7) FP_SORT_BITS(8)
#undef FP_SORT_BITS
        default: return hipErrorInvalidValue;
        }
        std::swap(ka, kb);
        std::swap(va, vb);
        shift += bits;
    }
    fpd::ReduceArgs r{};
    r.keys = ka; r.vals = va; r.n = n_kept; r.n_cap = cap; r.key_limit = table_entries;
    r.count = count; r.p0 = parts[0]; r.p1 = parts[1]; r.p2 = parts[2];
    const uint32_t chunks = (uint32_t)((cap + 64 * fpd::RED_LANE - 1) / (64 * fpd::RED_LANE));
    const uint32_t rgrid = std::max<uint32_t>(1, std::min<uint32_t>(chunks, 8192));
    hipLaunchKernelGGL(fpd::k_seg_reduce, dim3(rgrid), dim3(64), 0, st, r);
    return hipGetLastError();
}
