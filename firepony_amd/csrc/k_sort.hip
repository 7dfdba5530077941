/*
 * K_sort: the sort aggregation of covariate observations (aggregation = FP_AGG_SORT), the method of
 * the reference's covariate tables (covariate_table.cu:43-105: sort_by_key over the table's key bits,
 * then reduce_by_key summing observations and mismatches):
 *
 *   k_cov in pair mode emits one (key u32, value f64) pair per observation and table ->
 *   per table an LSD radix sort of the pairs with D-bit digits (D = 4..8) ->
 *   a segmented reduce_by_key of the sorted runs into the resident direct-indexed tables.
 *
 * One radix pass is k_sort_hist (per-tile digit counts, digit-major) -> k_sort_totals +
 * k_sort_offsets (exclusive scan = each tile's global output offset per digit) -> k_sort_scatter.
 * The scatter ranks a tile's elements stably in LDS: equal digits inside a wavefront are matched with
 * D ballots and counted by per-wave running digit counters, the tile is reordered by digit in LDS,
 * and each digit's run is written to its global offset, so consecutive lanes write consecutive
 * addresses.  The first pass drops the sentinel keys of unused pair slots.  Element counts stay on
 * the device and every kernel strides a fixed grid over the tiles that the count implies, so the
 * whole chain can be captured in a graph.
 *
 * The reduce walks the sorted pairs in wavefront chunks of 2048: a segmented scan across the lanes
 * (head = key change) sums each run's count and exact fixed-point value (three 27-bit limbs, as the
 * direct path keeps them), and the last element of each run adds the run total to the table with
 * non-returning atomics; a run that spans chunks adds once per chunk.  Sums are exact, so the tables
 * equal the direct path's bit for bit whatever the digit width.
 */
#include "fp_device.h"

#include <algorithm>

namespace fpd {

static constexpr uint32_t SORT_THREADS = 256;
static constexpr uint32_t SORT_ITEMS = 16;
static constexpr uint32_t SORT_TILE = SORT_THREADS * SORT_ITEMS;   // 4096 elements
static constexpr uint32_t SORT_SENTINEL = 0xffffffffu;
static constexpr uint32_t SORT_GRID = 2048;                          // workgroups of hist / scatter
static constexpr uint32_t RED_STEPS = 32;                            // 64 x 32 elements per reduce chunk

struct SortPass {
    const uint32_t *kin;
    const double *vin;
    uint32_t *kout;
    double *vout;
    const uint32_t *n_in;     // input elements (device)
    uint32_t *n_out;          // kept elements, written by the offsets kernel (first pass), or null
    uint32_t *hist;           // [radix][tiles_cap]: tile counts, then tile offsets
    uint32_t *totals;         // [radix]
    uint64_t n_cap;           // capacity of the pair arrays (elements)
    uint32_t tiles_cap;
    uint32_t shift, bits;
    uint32_t drop;            // drop sentinel keys (first pass)
};

__device__ inline uint32_t pass_n(const SortPass &p)
{
    const uint64_t n = *p.n_in;
    return (uint32_t)(n < p.n_cap ? n : p.n_cap);
}

// lanes of the wavefront holding the same digit (among valid lanes)
__device__ inline uint64_t match_digit(uint32_t d, uint32_t bits, bool valid)
{
    uint64_t m = __ballot(valid);
    for (uint32_t b = 0; b < bits; b++) {
        const bool one = (d >> b) & 1;
        const uint64_t x = __ballot(one);
        m &= one ? x : ~x;
    }
    return m;
}

__global__ void __launch_bounds__(SORT_THREADS) k_sort_hist(SortPass p)
{
    __shared__ uint32_t h[4][256];
    const uint32_t n = pass_n(p);
    const uint32_t tiles = (n + SORT_TILE - 1) / SORT_TILE;
    const uint32_t R = 1u << p.bits, mask = R - 1;
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint64_t lt = (1ull << lane) - 1;
    for (uint32_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        for (uint32_t i = threadIdx.x; i < 4 * 256; i += SORT_THREADS) (&h[0][0])[i] = 0;
        __syncthreads();
        const uint32_t t0 = tile * SORT_TILE + wave * (SORT_ITEMS * 64);
        uint32_t key[SORT_ITEMS];
#pragma unroll
        for (uint32_t k = 0; k < SORT_ITEMS; k++) {
            const uint32_t i = t0 + k * 64 + lane;
            key[k] = i < n ? p.kin[i] : SORT_SENTINEL;
        }
#pragma unroll
        for (uint32_t k = 0; k < SORT_ITEMS; k++) {
            const uint32_t i = t0 + k * 64 + lane;
            const bool valid = i < n && !(p.drop && key[k] == SORT_SENTINEL);
            const uint32_t d = (key[k] >> p.shift) & mask;
            const uint64_t peers = match_digit(d, p.bits, valid);
            if (valid && (peers & lt) == 0) h[wave][d] += (uint32_t)__popcll(peers);
        }
        __syncthreads();
        for (uint32_t d = threadIdx.x; d < R; d += SORT_THREADS)
            p.hist[(uint64_t)d * p.tiles_cap + tile] = h[0][d] + h[1][d] + h[2][d] + h[3][d];
        __syncthreads();
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

// totals[d] = sum over tiles of hist[d][tile]; one workgroup per digit
__global__ void __launch_bounds__(SORT_THREADS) k_sort_totals(SortPass p)
{
    __shared__ uint32_t wtot[4];
    const uint32_t d = blockIdx.x;
    const uint32_t tiles = (pass_n(p) + SORT_TILE - 1) / SORT_TILE;
    const uint32_t *h = p.hist + (uint64_t)d * p.tiles_cap;
    uint32_t s = 0;
    for (uint32_t t = threadIdx.x; t < tiles; t += SORT_THREADS) s += h[t];
    uint32_t total;
    (void)block_excl_scan(s, wtot, &total);
    if (threadIdx.x == 0) p.totals[d] = total;
}

// hist[d][t] <- sum of totals below d + sum of hist[d][t' < t]; one workgroup per digit, eight
// consecutive tiles per thread and step
__global__ void __launch_bounds__(SORT_THREADS) k_sort_offsets(SortPass p)
{
    __shared__ uint32_t wtot[4];
    const uint32_t d = blockIdx.x;
    const uint32_t R = 1u << p.bits;
    const uint32_t tiles = (pass_n(p) + SORT_TILE - 1) / SORT_TILE;
    uint32_t below = 0, all = 0;
    for (uint32_t x = 0; x < R; x++) {
        const uint32_t t = p.totals[x];
        if (x < d) below += t;
        all += t;
    }
    if (d == 0 && threadIdx.x == 0 && p.n_out) *p.n_out = all;
    uint32_t *h = p.hist + (uint64_t)d * p.tiles_cap;
    uint32_t running = below;
    for (uint32_t t0 = 0; t0 < tiles; t0 += SORT_THREADS * 8) {
        const uint32_t tb = t0 + threadIdx.x * 8;
        uint32_t v[8], s = 0;
#pragma unroll
        for (int j = 0; j < 8; j++) {
            v[j] = tb + j < tiles ? h[tb + j] : 0u;
            s += v[j];
        }
        uint32_t step;
        uint32_t ex = block_excl_scan(s, wtot, &step) + running;
#pragma unroll
        for (int j = 0; j < 8; j++) {
            if (tb + j < tiles) h[tb + j] = ex;
            ex += v[j];
        }
        running += step;
    }
}

__global__ void __launch_bounds__(SORT_THREADS) k_sort_scatter(SortPass p)
{
    __shared__ uint32_t cnt[4][256];    // per-wave running digit counters, then per-wave digit bases
    __shared__ uint32_t gofs[256];      // the tile's global offset per digit
    __shared__ uint32_t lofs[256];      // start of each digit in the tile's digit order
    __shared__ uint32_t wtot[4];
    __shared__ uint32_t kept_s;
    __shared__ uint32_t lk[SORT_TILE];
    __shared__ double lv[SORT_TILE];
    const uint32_t n = pass_n(p);
    const uint32_t tiles = (n + SORT_TILE - 1) / SORT_TILE;
    const uint32_t R = 1u << p.bits, mask = R - 1;
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint64_t lt = (1ull << lane) - 1;
    for (uint32_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        for (uint32_t i = threadIdx.x; i < 4 * 256; i += SORT_THREADS) (&cnt[0][0])[i] = 0;
        for (uint32_t d = threadIdx.x; d < R; d += SORT_THREADS) gofs[d] = p.hist[(uint64_t)d * p.tiles_cap + tile];
        __syncthreads();
        // stable ranks: element order inside the tile is (wave, item, lane)
        const uint32_t t0 = tile * SORT_TILE + wave * (SORT_ITEMS * 64);
        uint32_t key[SORT_ITEMS], rank[SORT_ITEMS];
        double val[SORT_ITEMS];
        uint32_t vmask = 0;
#pragma unroll
        for (uint32_t k = 0; k < SORT_ITEMS; k++) {
            const uint32_t i = t0 + k * 64 + lane;
            key[k] = i < n ? p.kin[i] : SORT_SENTINEL;
        }
#pragma unroll
        for (uint32_t k = 0; k < SORT_ITEMS; k++) {
            const uint32_t i = t0 + k * 64 + lane;
            val[k] = i < n ? p.vin[i] : 0.0;
        }
#pragma unroll
        for (uint32_t k = 0; k < SORT_ITEMS; k++) {
            const uint32_t i = t0 + k * 64 + lane;
            const bool valid = i < n && !(p.drop && key[k] == SORT_SENTINEL);
            vmask |= (uint32_t)valid << k;
            const uint32_t d = (key[k] >> p.shift) & mask;
            const uint64_t peers = match_digit(d, p.bits, valid);
            const uint32_t c = valid ? cnt[wave][d] : 0u;
            // every lane has read the counter before the leader of each digit moves it on
            __builtin_amdgcn_wave_barrier();
            if (valid && (peers & lt) == 0) cnt[wave][d] = c + (uint32_t)__popcll(peers);
            __builtin_amdgcn_wave_barrier();
            rank[k] = c + (uint32_t)__popcll(peers & lt);
        }
        __syncthreads();
        // digit bases per wave, and the start of each digit in the tile
        {
            const uint32_t d = threadIdx.x;
            uint32_t t[4] = {0, 0, 0, 0}, tot = 0;
            if (d < R) {
                for (int w = 0; w < 4; w++) { t[w] = cnt[w][d]; tot += t[w]; }
            }
            __syncthreads();
            if (d < R) {
                uint32_t run = 0;
                for (int w = 0; w < 4; w++) { cnt[w][d] = run; run += t[w]; }
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
    if (key >= a.key_limit) return;
    atomicAdd(&a.count[key], (unsigned long long)c);
    if (v0) atomicAdd(&a.p0[key], (unsigned long long)v0);
    if (v1) atomicAdd(&a.p1[key], (unsigned long long)v1);
    if (v2) atomicAdd(&a.p2[key], (unsigned long long)v2);
}

__global__ void __launch_bounds__(256) k_seg_reduce(ReduceArgs a)
{
    const uint64_t nn = *a.n;
    const uint32_t n = (uint32_t)(nn < a.n_cap ? nn : a.n_cap);
    const uint32_t chunk_elems = 64 * RED_STEPS;
    const uint32_t chunks = (n + chunk_elems - 1) / chunk_elems;
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t gw = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t nw = (gridDim.x * blockDim.x) >> 6;
    for (uint32_t chunk = gw; chunk < chunks; chunk += nw) {
        uint32_t ck = SORT_SENTINEL;                     // the run carried from the previous step
        uint64_t cc = 0, c0 = 0, c1 = 0, c2 = 0;
        for (uint32_t s = 0; s < RED_STEPS; s++) {
            const uint32_t i = chunk * chunk_elems + s * 64 + lane;
            if (chunk * chunk_elems + s * 64 >= n) break;    // wave-uniform
            const bool valid = i < n;
            const uint32_t k = valid ? a.keys[i] : SORT_SENTINEL;
            uint64_t l0 = 0, l1 = 0, l2 = 0;
            if (valid) fixed_split(fixed_from_double(a.vals[i]), l0, l1, l2);
            uint64_t vc = valid ? 1u : 0u;
            // the carried run ends where this step starts another key
            if ((uint32_t)__builtin_amdgcn_readlane((int)k, 0) != ck) {
                if (lane == 0 && ck != SORT_SENTINEL) table_add(a, ck, cc, c0, c1, c2);
                ck = SORT_SENTINEL;
                cc = c0 = c1 = c2 = 0;
            }
            uint32_t prevk = (uint32_t)__shfl_up((int)k, 1);
            if (lane == 0) prevk = ck;
            bool f = k != prevk;        // a run starts at this lane
            // segmented inclusive scan: sums from the start of the lane's run (inside this step)
#pragma unroll
            for (int off = 1; off < 64; off <<= 1) {
                const uint64_t uc = shfl_up_t(vc, off), u0 = shfl_up_t(l0, off), u1 = shfl_up_t(l1, off),
                               u2 = shfl_up_t(l2, off);
                const bool uf = __shfl_up((int)f, off) != 0;
                if (lane >= (uint32_t)off && !f) {
                    vc += uc; l0 += u0; l1 += u1; l2 += u2;
                    f = uf;
                }
            }
            // a run with no start in this step continues the carried run
            if (!f) { vc += cc; l0 += c0; l1 += c1; l2 += c2; }
            const uint32_t nextk = (uint32_t)__shfl_down((int)k, 1);
            if (lane < 63 && k != nextk && k != SORT_SENTINEL) table_add(a, k, vc, l0, l1, l2);
            ck = (uint32_t)__builtin_amdgcn_readlane((int)k, 63);
            cc = __shfl(vc, 63); c0 = __shfl(l0, 63); c1 = __shfl(l1, 63); c2 = __shfl(l2, 63);
        }
        if (lane == 0 && ck != SORT_SENTINEL) table_add(a, ck, cc, c0, c1, c2);
    }
}

} // namespace fpd

// one table: radix passes over its key bits, then the reduce into the resident table.  keys/vals
// hold the emitted pairs (n_emit of them, sentinels included); tkeys/tvals are the ping-pong
// buffers.  n_kept receives the number of real pairs.
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
    for (uint32_t shift = 0, pass = 0; shift < key_bits; shift += digit_bits, pass++) {
        fpd::SortPass p{};
        p.kin = ka; p.vin = va; p.kout = kb; p.vout = vb;
        p.n_in = pass == 0 ? n_emit : n_kept;
        p.n_out = pass == 0 ? n_kept : nullptr;
        p.hist = hist; p.totals = totals; p.n_cap = cap; p.tiles_cap = tiles_cap;
        p.shift = shift;
        p.bits = std::min<uint32_t>(digit_bits, key_bits - shift);
        p.drop = pass == 0 ? 1u : 0u;
        const uint32_t R = 1u << p.bits;
        hipLaunchKernelGGL(fpd::k_sort_hist, dim3(grid), dim3(fpd::SORT_THREADS), 0, st, p);
        hipLaunchKernelGGL(fpd::k_sort_totals, dim3(R), dim3(fpd::SORT_THREADS), 0, st, p);
        hipLaunchKernelGGL(fpd::k_sort_offsets, dim3(R), dim3(fpd::SORT_THREADS), 0, st, p);
        hipLaunchKernelGGL(fpd::k_sort_scatter, dim3(grid), dim3(fpd::SORT_THREADS), 0, st, p);
        std::swap(ka, kb);
        std::swap(va, vb);
    }
    fpd::ReduceArgs r{};
    r.keys = ka; r.vals = va; r.n = n_kept; r.n_cap = cap; r.key_limit = table_entries;
    r.count = count; r.p0 = parts[0]; r.p1 = parts[1]; r.p2 = parts[2];
    const uint32_t chunks = (uint32_t)((cap + 64 * fpd::RED_STEPS - 1) / (64 * fpd::RED_STEPS));
    const uint32_t rgrid = std::max<uint32_t>(1, std::min<uint32_t>((chunks + 3) / 4, 4096));
    hipLaunchKernelGGL(fpd::k_seg_reduce, dim3(rgrid), dim3(256), 0, st, r);
    return hipGetLastError();
}
