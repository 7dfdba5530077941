/*
 * Covariate table maintenance on the device-resident direct-indexed tables:
 *   - compaction of the non-empty entries into key-ordered rows
 *     (what covariate_table::sort + pack produce, covariate_table.cu:43-105);
 *   - element-wise merge of another engine's tables (gather_intermediates,
 *     device/pipeline_interface.cu:52-58 + covariates.cu:265-282).
 */
#include "fp_device.h"

namespace fpd {

__global__ void __launch_bounds__(256) k_compact_count(const uint64_t *count, uint64_t n, uint32_t *block_counts)
{
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    const bool f = i < n && count[i] != 0;
    const uint64_t m = __ballot(f);
    __shared__ uint32_t wsum[4];
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = __popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) block_counts[blockIdx.x] = wsum[0] + wsum[1] + wsum[2] + wsum[3];
}

__global__ void __launch_bounds__(256) k_compact_scatter(const uint64_t *count, const uint64_t *p0, const uint64_t *p1,
                                                         const uint64_t *p2, uint64_t n, const uint64_t *block_offsets,
                                                         uint32_t *okey, uint64_t *ocount, uint64_t *olo,
                                                         uint64_t *ohi)
{
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    const bool f = i < n && count[i] != 0;
    const uint64_t m = __ballot(f);
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __shared__ uint32_t wsum[4];
    if (lane == 0) wsum[w] = __popcll(m);
    __syncthreads();
    uint32_t base = 0;
    for (uint32_t k = 0; k < w; k++) base += wsum[k];
    const uint32_t rank = base + __popcll(m & ((1ull << lane) - 1));
    if (f) {
        const uint64_t o = block_offsets[blockIdx.x] + rank;
        okey[o] = (uint32_t)i;
        ocount[o] = count[i];
        const u128 v = fixed_combine(p0[i], p1[i], p2[i]);
        olo[o] = v.lo;
        ohi[o] = v.hi;
    }
}

__global__ void __launch_bounds__(256) k_table_add(uint64_t *dc, uint64_t *d0, uint64_t *d1, uint64_t *d2,
                                                   const uint64_t *sc, const uint64_t *s0, const uint64_t *s1,
                                                   const uint64_t *s2, uint64_t n)
{
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    if (sc[i] == 0) return;
    dc[i] += sc[i];
    d0[i] += s0[i];
    d1[i] += s1[i];
    d2[i] += s2[i];
}

} // namespace fpd

extern "C" hipError_t fpd_compact_count(const uint64_t *count, uint64_t n, uint32_t *block_counts, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(fpd::k_compact_count, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, count, n, block_counts);
    return hipGetLastError();
}

extern "C" hipError_t fpd_compact_scatter(const uint64_t *count, const uint64_t *const *parts, uint64_t n,
                                          const uint64_t *block_offsets, uint32_t *okey, uint64_t *ocount,
                                          uint64_t *olo, uint64_t *ohi, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(fpd::k_compact_scatter, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, count, parts[0],
                       parts[1], parts[2], n, block_offsets, okey, ocount, olo, ohi);
    return hipGetLastError();
}

extern "C" hipError_t fpd_table_add(uint64_t *dc, uint64_t *const *dparts, const uint64_t *sc,
                                    const uint64_t *const *sparts, uint64_t n, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(fpd::k_table_add, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, dc, dparts[0], dparts[1],
                       dparts[2], sc, sparts[0], sparts[1], sparts[2], n);
    return hipGetLastError();
}
