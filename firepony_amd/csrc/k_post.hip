/*
 * Postprocess on the device: observation rows -> empirical rows
 * (covariate_table.cu:109-136, expected_error.h:37-76,
 * empirical_quality.cu:35-197) and the read-group table
 * (read_group_table.cu:47-93).
 *
 * Two per-row terms of the empirical quality go through libm calls whose last
 * bit differs between implementations: the estimated quality
 * -10*log10(expected/obs) and the binomial coefficient built from three
 * lgammaf calls.  The host evaluates both with its libm, the one the
 * reference's host build links (empirical_quality.cu:72-75,151-156,179-189),
 * and passes them in per row; terms that depend only on small integers
 * (qualities, posterior bins, prior distances) come from host-computed tables.
 * The device folds the tables, converts the exact sums and evaluates the
 * 61-bin posterior.
 */
#include "fp_device.h"

namespace fpd {

__constant__ double c_qual_err[256];    // pow(10.0, q / -10.0)
__constant__ double c_prior[41];        // log10(gaussian(d)), d = 0..40
__constant__ double c_log1mp[61];       // log10(1 - pow(10.0, bin / -10.0))
__constant__ double c_log10e;           // log10(M_E)

// the maximum-posterior bin of the empirical quality (empirical_quality.cu:109-136,179-189) for a
// row's estimated quality and binomial coefficient
__device__ double empirical_quality(uint64_t obs, double mism, double estimated, double coeff, bool need_rounding)
{
    uint64_t nErr = uint64_t(mism + 0.5) + 1;
    uint64_t nObs = obs + 1 + 1;
    const uint32_t MAXOBS = 2147483647u - 1;
    if (nObs > MAXOBS) {
        const double fraction = double(MAXOBS) / double(nObs);
        nErr = (uint64_t)round(double(nErr) * fraction);
        nObs = MAXOBS;
    }
    int mle = 0;
    double best = 0.0;
    for (int bin = 0; bin < 61; bin++) {
        double delta = (double)bin - estimated;
        if (need_rounding) delta = round(delta);
        int d = abs(int(delta));
        d = d < 40 ? d : 40;
        const double log10p = (bin / 1.0) / -10.0;
        const double lik = coeff + log10p * nErr + c_log1mp[bin] * (nObs - nErr);
        const double post = c_prior[d] + lik;
        if (bin == 0 || post > best) { best = post; mle = bin; }
    }
    const double q = (double)mle;
    return q < 93.0 ? q : 93.0;
}

// observation rows: exact mismatch sums as doubles and expected errors (covariate_table.cu:109-136)
struct ObsArgs {
    uint64_t n;
    const uint32_t *key;
    const uint64_t *count;
    const uint64_t *lo;
    const uint64_t *hi;
    int qual_shift;            // bit offset of the quality field
    double *mism, *expected;
};

__global__ void __launch_bounds__(256) k_obs_values(ObsArgs a)
{
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= a.n) return;
    const uint32_t q = (a.key[i] >> a.qual_shift) & 255;
    a.mism[i] = fixed_to_double(a.lo[i], a.hi[i]);
    a.expected[i] = a.count[i] * c_qual_err[q];
}

struct EmpArgs {
    uint64_t n;
    const uint64_t *count;
    const double *mism;
    const double *estimated;   // host libm
    const double *coeff;       // host libm
    uint32_t need_rounding;
    double *empirical;
};

__global__ void __launch_bounds__(256) k_empirical(EmpArgs a)
{
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= a.n) return;
    a.empirical[i] = empirical_quality(a.count[i], a.mism[i], a.estimated[i], a.coeff[i], a.need_rounding != 0);
}

// one lane per (read group, event): fold the quality rows in key order
struct RgArgs {
    uint32_t num_rg;
    const uint64_t *count;   // dense quality table
    const uint64_t *p0, *p1, *p2;   // mismatch sum limbs
    uint32_t *valid;
    uint64_t *obs;
    double *mism, *expected;
};

__global__ void __launch_bounds__(256) k_rg_table(RgArgs a)
{
    const uint32_t slot = blockIdx.x * 256 + threadIdx.x;
    if (slot >= a.num_rg * 4) return;
    const uint32_t rg = slot >> 2, ev = slot & 3;
    bool any = false;
    uint64_t obs = 0;
    double m = 0.0, ex = 0.0;
    for (uint32_t q = 0; q < 256; q++) {
        const uint32_t key = ev | (q << 2) | (rg << 10);
        const uint64_t c = a.count[key];
        if (c == 0) continue;
        const u128 v = fixed_combine(a.p0[key], a.p1[key], a.p2[key]);
        const double mm = fixed_to_double(v.lo, v.hi);
        const double ee = c * c_qual_err[q];
        if (!any) { obs = c; m = mm; ex = ee; any = true; }
        else { obs += c; m = m + mm; ex = ex + ee; }
    }
    a.valid[slot] = any ? 1 : 0;
    a.obs[slot] = obs;
    a.mism[slot] = m;
    a.expected[slot] = ex;
}

} // namespace fpd

extern "C" hipError_t fpd_post_set_constants(const double *qual_err, const double *prior, const double *log1mp,
                                             double log10e)
{
    hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(fpd::c_qual_err), qual_err, 256 * sizeof(double));
    if (e != hipSuccess) return e;
    e = hipMemcpyToSymbol(HIP_SYMBOL(fpd::c_prior), prior, 41 * sizeof(double));
    if (e != hipSuccess) return e;
    e = hipMemcpyToSymbol(HIP_SYMBOL(fpd::c_log1mp), log1mp, 61 * sizeof(double));
    if (e != hipSuccess) return e;
    return hipMemcpyToSymbol(HIP_SYMBOL(fpd::c_log10e), &log10e, sizeof(double));
}

extern "C" hipError_t fpd_launch_obs_values(uint64_t n, const uint32_t *key, const uint64_t *count, const uint64_t *lo,
                                            const uint64_t *hi, int qual_shift, double *mism, double *expected,
                                            hipStream_t st)
{
    if (n == 0) return hipSuccess;
    fpd::ObsArgs a{};
    a.n = n; a.key = key; a.count = count; a.lo = lo; a.hi = hi; a.qual_shift = qual_shift;
    a.mism = mism; a.expected = expected;
    hipLaunchKernelGGL(fpd::k_obs_values, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a);
    return hipGetLastError();
}

extern "C" hipError_t fpd_launch_empirical(uint64_t n, const uint64_t *count, const double *mism,
                                           const double *estimated, const double *coeff, uint32_t need_rounding,
                                           double *empirical, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    fpd::EmpArgs a{};
    a.n = n; a.count = count; a.mism = mism; a.estimated = estimated; a.coeff = coeff;
    a.need_rounding = need_rounding; a.empirical = empirical;
    hipLaunchKernelGGL(fpd::k_empirical, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a);
    return hipGetLastError();
}

extern "C" hipError_t fpd_launch_rg_table(uint32_t num_rg, const uint64_t *count, const uint64_t *const *parts,
                                          uint32_t *valid, uint64_t *obs, double *mism, double *expected,
                                          hipStream_t st)
{
    if (num_rg == 0) return hipSuccess;
    fpd::RgArgs a{};
    a.num_rg = num_rg; a.count = count; a.p0 = parts[0]; a.p1 = parts[1]; a.p2 = parts[2];
    a.valid = valid; a.obs = obs; a.mism = mism; a.expected = expected;
    hipLaunchKernelGGL(fpd::k_rg_table, dim3((num_rg * 4 + 255) / 256), dim3(256), 0, st, a);
    return hipGetLastError();
}
