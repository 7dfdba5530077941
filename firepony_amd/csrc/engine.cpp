/*
 * firepony_amd engine: C ABI (include/firepony_amd.h) and stream orchestration.
 *
 * Per batch (firepony_process_batch, device/pipeline.cu:56-148):
 *   k_prep -> BAQ classification -> k_baq_hmm (per band class) -> k_baq_post
 *   -> k_cov (fractional errors + covariate tables)
 * all enqueued on the engine's stream.  The covariate tables stay resident
 * in device memory across batches; postprocess compacts them and computes
 * the empirical tables on the device, then copies the (small) results to the
 * host for the report.
 */
#include "../../include/firepony_amd.h"
#include "fp_device.h"
#include "report.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

// kernels (k_*.hip)
extern "C" {
hipError_t fpd_launch_prep(const fpd::BatchDev *b, const fpd::GenomeDev *g, const fpd::ReadStateDev *s,
                           uint32_t *counters, hipStream_t st);
hipError_t fpd_baq_set_constants(const double *qual2prob, const double *thr, double gap_open);
hipError_t fpd_launch_baq_classify(uint32_t num_reads, const fpd::ReadStateDev *s, uint32_t *lists,
                                   uint32_t *counts, uint32_t capacity, hipStream_t st);
hipError_t fpd_launch_baq_hmm(int cls, const fpd::BatchDev *b, const fpd::GenomeDev *g, const fpd::ReadStateDev *s,
                              const uint32_t *list, const uint32_t *count, double *scratch, uint64_t slot_doubles,
                              uint32_t rows_cap, uint32_t wmax_runtime, uint32_t waves, unsigned long long *cells,
                              int wpe, hipStream_t st);
hipError_t fpd_launch_baq_post(const fpd::BatchDev *b, const fpd::ReadStateDev *s, uint32_t band_done,
                               hipStream_t st);
hipError_t fpd_fill_u8(uint8_t *p, uint64_t n, uint8_t v, hipStream_t st);
hipError_t fpd_launch_cov(const fpd::BatchDev *b, const fpd::ReadStateDev *s, const fpd::TablesDev *t,
                          uint32_t frac_f64, uint32_t keep_frac, uint32_t *counters, uint32_t *rg_buf,
                          uint32_t max_read_len, hipStream_t st);
hipError_t fpd_compact_count(const uint64_t *count, uint64_t n, uint32_t *block_counts, hipStream_t st);
hipError_t fpd_compact_scatter(const uint64_t *count, const uint64_t *const *parts, uint64_t n,
                               const uint64_t *block_offsets, uint32_t *okey, uint64_t *ocount, uint64_t *olo,
                               uint64_t *ohi, hipStream_t st);
hipError_t fpd_table_add(uint64_t *dc, uint64_t *const *dparts, const uint64_t *sc, const uint64_t *const *sparts,
                         uint64_t n, hipStream_t st);
hipError_t fpd_post_set_constants(const double *qual_err, const double *prior, const double *log1mp, double log10e);
hipError_t fpd_launch_empirical(uint64_t n, const uint32_t *key, const uint64_t *count, const uint64_t *lo,
                                const uint64_t *hi, int qual_shift, double *mism, double *expected, double *estimated,
                                double *empirical, hipStream_t st);
hipError_t fpd_launch_rg_table(uint32_t num_rg, const uint64_t *count, const uint64_t *const *parts,
                               uint32_t *valid, uint64_t *obs, double *mism, double *expected, double *estimated,
                               double *empirical, hipStream_t st);
}

namespace {

thread_local std::string g_last_error;

fp_status fail(fp_status s, const std::string &msg)
{
    g_last_error = msg;
    return s;
}

#define HIP_TRY(expr)                                                                                   \
    do {                                                                                                \
        hipError_t _e = (expr);                                                                         \
        if (_e != hipSuccess)                                                                           \
            return fail(_e == hipErrorOutOfMemory ? FP_ERR_OUT_OF_MEMORY : FP_ERR_HIP,                 \
                        std::string(#expr) + ": " + hipGetErrorString(_e));                            \
    } while (0)

// a growable device buffer
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    hipError_t ensure(size_t need)
    {
        if (need <= bytes) return hipSuccess;
        if (p) { hipError_t e = hipFree(p); if (e != hipSuccess) return e; p = nullptr; bytes = 0; }
        size_t cap = std::max<size_t>(need, 256);
        cap = cap + cap / 4;
        hipError_t e = hipMalloc(&p, cap);
        if (e == hipSuccess) bytes = cap;
        return e;
    }
    void release() { if (p) (void)hipFree(p); p = nullptr; bytes = 0; }
    template <typename T> T *as() const { return (T *)p; }
};

// ---------------------------------------------------------------------------
// host-computed constant tables (host libm, identical to the reference's host build)

struct HostConstants {
    double qual2prob[256];    // baq.cu:356-359
    double map_thr[128];      // MAP quality thresholds
    double gap_open;          // baq.cu:81
    double qual_err[256];     // expected_error.h:47-50
    double prior[41];         // empirical_quality.cu:44-65
    double log1mp[61];        // empirical_quality.cu:84-88
    double log10e;            // empirical_quality.cu:67-70
};

static int map_quality_host(double y)
{
    // (int)(-4.343 * log(y) + .499) with x86 conversion semantics, y in (0, 1]
    volatile double vy = y;
    const double real = double(-4.343) * log(vy) + double(.499);
    if (!(real < 2147483648.0) || !(real > -2147483649.0)) return INT32_MIN;
    return (int)real;
}

static void compute_constants(HostConstants &c)
{
    for (int q = 0; q < 256; q++) {
        volatile double qd = -q / 10.0;
        c.qual2prob[q] = pow(10.0, qd);
        volatile double qe = q / -10.0;
        c.qual_err[q] = pow(10.0, qe);
    }
    {
        volatile double g = (-40.0) / 10.;
        c.gap_open = pow(10.0, g);
    }
    // T[v] = largest double y in (0, 1] with k(y) >= v; k is non-increasing in y
    for (int v = 0; v < 128; v++) c.map_thr[v] = 0.0;
    const uint64_t one_bits = 0x3ff0000000000000ull;
    for (int v = 5; v <= 101; v++) {
        uint64_t lo = 1, hi = one_bits;   // positive doubles ordered by bit pattern
        // invariant: k(lo) >= v (smallest subnormal gives a huge k), find the last such
        double y_lo;
        memcpy(&y_lo, &lo, 8);
        if (map_quality_host(y_lo) < v) { c.map_thr[v] = 0.0; continue; }
        while (lo < hi) {
            const uint64_t mid = lo + (hi - lo + 1) / 2;
            double y;
            memcpy(&y, &mid, 8);
            if (map_quality_host(y) >= v) lo = mid; else hi = mid - 1;
        }
        memcpy(&c.map_thr[v], &lo, 8);
    }
    for (int d = 0; d <= 40; d++) {
        volatile double v = double(d);
        const double g = 0.0 + 0.9 * pow(double(M_E), -(pow(v - 0.0, 2.0) / (2 * 0.5 * 0.5)));
        c.prior[d] = log10(g);
    }
    for (int bin = 0; bin < 61; bin++) {
        volatile double l = (bin / 1.0) / -10.0;
        c.log1mp[bin] = log10(1 - pow(10.0, l));
    }
    {
        volatile double e = M_E;
        c.log10e = log10(e);
    }
}

struct TableRows {
    std::vector<uint32_t> key;
    std::vector<uint64_t> obs, lo, hi;
    std::vector<double> mism, expected, estimated, empirical;
    size_t size() const { return key.size(); }
};

} // namespace

struct fp_engine {
    int device = 0;
    fp_options opts;
    hipStream_t stream = nullptr;
    std::string name;
    HostConstants consts;
    int wpe = 1;   // waves per SIMD of the 15-cell BAQ band kernel (1-4)

    // setup state
    bool is_setup = false;
    std::vector<std::string> rg_names;
    uint32_t num_rg = 0;
    DevBuf g_word_offset, g_length, g_words, g_foff, g_fstart, g_fstop, g_mel;
    fpd::GenomeDev genome{};

    // batch buffers (host batches are copied here)
    DevBuf in[18];
    // per-batch state
    DevBuf st_status, st_rwc, st_noins, st_refw, st_nerr, st_hmmw, st_bw, st_hqw, st_bflags, st_baq, st_state, st_frac;
    DevBuf counters, baq_lists, baq_counts, scratch, rg_buf;
    uint32_t last_num_reads = 0;
    uint64_t last_num_bases = 0, last_num_quals = 0;
    fpd::ReadStateDev rs{};

    // resident tables
    DevBuf tab[3][4];     // per table: count, then the three limb sums
    uint64_t tab_size[3] = {0, 0, 0};
    fpd::TablesDev tables{};

    // results
    bool postprocessed = false;
    TableRows rows[4];

    fp_stats stats;
    hipEvent_t ev[16];
};

namespace {

uint64_t table_entries(uint32_t num_rg, int t) { return (uint64_t)num_rg << (fpd::TABLE_BITS[t] - 8); }

fp_status alloc_tables(fp_engine *e)
{
    for (int t = 0; t < 3; t++) {
        e->tab_size[t] = table_entries(e->num_rg, t);
        for (int k = 0; k < 4; k++) {
            HIP_TRY(e->tab[t][k].ensure(e->tab_size[t] * 8));
            HIP_TRY(hipMemsetAsync(e->tab[t][k].p, 0, e->tab_size[t] * 8, e->stream));
        }
        e->tables.count[t] = e->tab[t][0].as<uint64_t>();
        for (int k = 0; k < 3; k++) e->tables.part[t][k] = e->tab[t][k + 1].as<uint64_t>();
    }
    e->tables.num_rg = e->num_rg;
    HIP_TRY(hipStreamSynchronize(e->stream));
    return FP_OK;
}

fp_status ensure_state(fp_engine *e, uint32_t n, uint64_t nbases, uint64_t nquals)
{
    HIP_TRY(e->st_status.ensure(n));
    HIP_TRY(e->st_rwc.ensure(4ull * n));
    HIP_TRY(e->st_noins.ensure(4ull * n));
    HIP_TRY(e->st_refw.ensure(4ull * n));
    HIP_TRY(e->st_nerr.ensure(2ull * n));
    HIP_TRY(e->st_hmmw.ensure(4ull * n));
    HIP_TRY(e->st_bw.ensure(2ull * n));
    HIP_TRY(e->st_hqw.ensure(4ull * n));
    HIP_TRY(e->st_bflags.ensure(nbases + 8));
    HIP_TRY(e->st_baq.ensure(nquals + 8));
    HIP_TRY(e->st_state.ensure(4 * (nquals + 8)));
    HIP_TRY(e->st_frac.ensure(24 * (nquals + 8)));
    HIP_TRY(e->counters.ensure(64));
    HIP_TRY(e->baq_lists.ensure(4ull * 4 * std::max<uint32_t>(n, 1)));
    HIP_TRY(e->rg_buf.ensure(4ull * (4ull * std::max<uint32_t>(e->num_rg, 1) + 2 + n)));
    HIP_TRY(e->baq_counts.ensure(64));
    fpd::ReadStateDev &s = e->rs;
    s.status = e->st_status.as<uint8_t>();
    s.rwc = e->st_rwc.as<uint32_t>();
    s.noins = e->st_noins.as<uint32_t>();
    s.refw = e->st_refw.as<uint32_t>();
    s.num_errors = e->st_nerr.as<uint16_t>();
    s.hmmw = e->st_hmmw.as<uint32_t>();
    s.bandwidth = e->st_bw.as<uint16_t>();
    s.hqw = e->st_hqw.as<uint32_t>();
    s.base_flags = e->st_bflags.as<uint8_t>();
    s.baq = e->st_baq.as<uint8_t>();
    s.baq_state = e->st_state.as<uint32_t>();
    s.frac = e->st_frac.as<double>();
    return FP_OK;
}

fpd::BatchDev to_dev(const fp_batch_view *v)
{
    fpd::BatchDev b;
    b.num_reads = v->num_reads;
    b.chromosome = v->chromosome;
    b.alignment_start = v->alignment_start;
    b.alignment_stop = v->alignment_stop;
    b.mate_chromosome = v->mate_chromosome;
    b.mate_alignment_start = v->mate_alignment_start;
    b.isize = v->inferred_insert_size;
    b.cigars = v->cigars;
    b.cigar_start = v->cigar_start;
    b.cigar_len = v->cigar_len;
    b.reads = v->reads;
    b.read_start = v->read_start;
    b.read_len = v->read_len;
    b.quals = v->qualities;
    b.qual_start = v->qual_start;
    b.qual_len = v->qual_len;
    b.flags = v->flags;
    b.mapq = v->mapq;
    b.read_group = v->read_group;
    return b;
}

float elapsed(hipEvent_t a, hipEvent_t b)
{
    float ms = 0;
    (void)hipEventElapsedTime(&ms, a, b);
    return ms;
}

static const uint32_t BAQ_CLASS_W[4] = {15, 23, 33, 0};

fp_status run_batch(fp_engine *e, const fp_batch_view *v, uint32_t max_read_len_hint)
{
    const uint32_t n = v->num_reads;
    fp_status rc = ensure_state(e, n, v->num_read_bases, v->num_qualities);
    if (rc != FP_OK) return rc;
    e->last_num_reads = n;
    e->last_num_bases = v->num_read_bases;
    e->last_num_quals = v->num_qualities;
    hipStream_t st = e->stream;
    fpd::BatchDev b = to_dev(v);

    HIP_TRY(hipEventRecord(e->ev[0], st));
    HIP_TRY(hipMemsetAsync(e->counters.p, 0, 64, st));
    HIP_TRY(hipMemsetAsync(e->baq_counts.p, 0, 64, st));
    HIP_TRY(hipEventRecord(e->ev[6], st));
    HIP_TRY(fpd_launch_prep(&b, &e->genome, &e->rs, e->counters.as<uint32_t>(), st));
    HIP_TRY(hipEventRecord(e->ev[7], st));
    HIP_TRY(hipEventRecord(e->ev[1], st));
    HIP_TRY(fpd_fill_u8(e->rs.baq, v->num_qualities, 0xff, st));
    HIP_TRY(hipMemsetAsync(e->rs.frac, 0, 24 * v->num_qualities, st));
    HIP_TRY(fpd_launch_baq_classify(n, &e->rs, e->baq_lists.as<uint32_t>(), e->baq_counts.as<uint32_t>(),
                                    std::max<uint32_t>(n, 1), st));
    uint32_t cnt[16];
    HIP_TRY(hipMemcpyAsync(cnt, e->baq_counts.p, 16, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(cnt + 4, e->counters.p, 16, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    const uint32_t rows_cap = std::max<uint32_t>(std::max(max_read_len_hint, cnt[4 + 3]), 1);

    // band-class kernels; generic class sized by the widest band
    uint32_t wmax_generic = 0;
    if (cnt[3]) {
        std::vector<uint32_t> ids(cnt[3]);
        std::vector<uint16_t> bwv(n);
        HIP_TRY(hipMemcpy(ids.data(), e->baq_lists.as<uint32_t>() + 3ull * std::max<uint32_t>(n, 1), 4ull * cnt[3],
                          hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(bwv.data(), e->rs.bandwidth, 2ull * n, hipMemcpyDeviceToHost));
        for (uint32_t id : ids) wmax_generic = std::max<uint32_t>(wmax_generic, 2u * bwv[id] + 1);
    }
    uint64_t need = 0;
    uint32_t waves[4];
    uint64_t slot[4];
    for (int c = 0; c < 4; c++) {
        const uint32_t W = c < 3 ? BAQ_CLASS_W[c] : wmax_generic;
        if (c < 3) {
            // band kernels: one resident set of 4-wave workgroups (256 CUs x 4 SIMDs x waves per SIMD)
            const uint32_t wpe = c == 0 ? (uint32_t)e->wpe : 1u;
            waves[c] = std::min<uint32_t>((cnt[c] + 63) / 64, 1024u * wpe);
            waves[c] = (waves[c] + 3) / 4 * 4;
        } else {
            waves[c] = std::min<uint32_t>((cnt[c] + 63) / 64, 2048u);
        }
        // forward rows (M, I) | scale factors | staged MAP results (band kernels) or band state (generic)
        slot[c] = ((uint64_t)(rows_cap + 1) * W * 2 + (rows_cap + 2) + (c == 3 ? 3ull * W : rows_cap + 1ull)) * 64;
        if (cnt[c]) need = std::max(need, slot[c] * waves[c]);
    }
    HIP_TRY(e->scratch.ensure(need * 8 + 8));
    uint32_t hmm_launches = 0;
    HIP_TRY(hipEventRecord(e->ev[8], st));
    for (int c = 0; c < 4; c++) {
        if (!cnt[c]) continue;
        HIP_TRY(fpd_launch_baq_hmm(c, &b, &e->genome, &e->rs,
                                   e->baq_lists.as<uint32_t>() + (uint64_t)c * std::max<uint32_t>(n, 1),
                                   e->baq_counts.as<uint32_t>() + c, e->scratch.as<double>(), slot[c], rows_cap,
                                   c == 3 ? wmax_generic : BAQ_CLASS_W[c], waves[c],
                                   (unsigned long long *)(e->counters.as<uint32_t>() + 8), e->wpe, st));
        hmm_launches++;
    }
    HIP_TRY(hipEventRecord(e->ev[9], st));
    HIP_TRY(fpd_launch_baq_post(&b, &e->rs, 1u, st));
    HIP_TRY(hipEventRecord(e->ev[2], st));
    HIP_TRY(fpd_launch_cov(&b, &e->rs, &e->tables, e->opts.frac_error_f64, e->opts.debug_arrays ? 1u : 0u,
                           e->counters.as<uint32_t>(), e->rg_buf.as<uint32_t>(), std::min<uint32_t>(rows_cap, 499),
                           st));
    HIP_TRY(hipEventRecord(e->ev[3], st));
    uint32_t post[16];
    HIP_TRY(hipMemcpyAsync(post, e->counters.p, 64, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    if (post[4]) return fail(FP_ERR_INVALID_ARG, "read group id beyond the header's read groups");

    e->stats.total_reads += n;
    e->stats.num_batches += 1;
    e->stats.filtered_reads += post[1];
    e->stats.baq_reads += post[2];
    uint64_t cells;
    memcpy(&cells, post + 8, 8);
    e->stats.baq_cells += cells;
    e->stats.prep_kernel_ms += elapsed(e->ev[6], e->ev[7]);
    e->stats.prep_launches += 1;
    if (hmm_launches) e->stats.baq_hmm_kernel_ms += elapsed(e->ev[8], e->ev[9]);
    e->stats.baq_hmm_launches += hmm_launches;
    e->stats.cov_kernel_ms += elapsed(e->ev[2], e->ev[3]);
    e->stats.cov_launches += 1;
    e->stats.read_filter_ms += elapsed(e->ev[0], e->ev[1]);
    e->stats.baq_ms += elapsed(e->ev[1], e->ev[2]);
    e->stats.covariates_ms += elapsed(e->ev[2], e->ev[3]);
    e->postprocessed = false;
    return FP_OK;
}

// compaction of one dense table into key-ordered rows on the device
fp_status compact(fp_engine *e, int t, DevBuf &okey, DevBuf &ocount, DevBuf &olo, DevBuf &ohi, uint64_t *nrows)
{
    const uint64_t n = e->tab_size[t];
    const uint64_t nb = (n + 255) / 256;
    DevBuf bc, bo;
    HIP_TRY(bc.ensure(4 * nb + 4));
    HIP_TRY(bo.ensure(8 * nb + 8));
    HIP_TRY(fpd_compact_count(e->tables.count[t], n, bc.as<uint32_t>(), e->stream));
    std::vector<uint32_t> h(nb);
    std::vector<uint64_t> off(nb);
    HIP_TRY(hipMemcpyAsync(h.data(), bc.p, 4 * nb, hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    uint64_t total = 0;
    for (uint64_t i = 0; i < nb; i++) { off[i] = total; total += h[i]; }
    HIP_TRY(hipMemcpyAsync(bo.p, off.data(), 8 * nb, hipMemcpyHostToDevice, e->stream));
    HIP_TRY(okey.ensure(4 * total + 4));
    HIP_TRY(ocount.ensure(8 * total + 8));
    HIP_TRY(olo.ensure(8 * total + 8));
    HIP_TRY(ohi.ensure(8 * total + 8));
    HIP_TRY(fpd_compact_scatter(e->tables.count[t], e->tables.part[t], n, bo.as<uint64_t>(),
                                okey.as<uint32_t>(), ocount.as<uint64_t>(), olo.as<uint64_t>(), ohi.as<uint64_t>(),
                                e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    bc.release();
    bo.release();
    *nrows = total;
    return FP_OK;
}

fp_status postprocess(fp_engine *e)
{
    hipStream_t st = e->stream;
    static const int qshift[3] = {2, 12, 13};
    for (int t = 0; t < 3; t++) {
        DevBuf k, c, lo, hi, m, ex, es, em;
        uint64_t nr = 0;
        fp_status rc = compact(e, t, k, c, lo, hi, &nr);
        if (rc != FP_OK) return rc;
        HIP_TRY(m.ensure(8 * nr + 8));
        HIP_TRY(ex.ensure(8 * nr + 8));
        HIP_TRY(es.ensure(8 * nr + 8));
        HIP_TRY(em.ensure(8 * nr + 8));
        HIP_TRY(fpd_launch_empirical(nr, k.as<uint32_t>(), c.as<uint64_t>(), lo.as<uint64_t>(), hi.as<uint64_t>(),
                                     qshift[t], m.as<double>(), ex.as<double>(), es.as<double>(), em.as<double>(), st));
        TableRows &R = e->rows[t + 1];
        R.key.resize(nr); R.obs.resize(nr); R.lo.resize(nr); R.hi.resize(nr);
        R.mism.resize(nr); R.expected.resize(nr); R.estimated.resize(nr); R.empirical.resize(nr);
        if (nr) {
            HIP_TRY(hipMemcpyAsync(R.key.data(), k.p, 4 * nr, hipMemcpyDeviceToHost, st));
            HIP_TRY(hipMemcpyAsync(R.obs.data(), c.p, 8 * nr, hipMemcpyDeviceToHost, st));
            HIP_TRY(hipMemcpyAsync(R.lo.data(), lo.p, 8 * nr, hipMemcpyDeviceToHost, st));
            HIP_TRY(hipMemcpyAsync(R.hi.data(), hi.p, 8 * nr, hipMemcpyDeviceToHost, st));
            HIP_TRY(hipMemcpyAsync(R.mism.data(), m.p, 8 * nr, hipMemcpyDeviceToHost, st));
            HIP_TRY(hipMemcpyAsync(R.expected.data(), ex.p, 8 * nr, hipMemcpyDeviceToHost, st));
            HIP_TRY(hipMemcpyAsync(R.estimated.data(), es.p, 8 * nr, hipMemcpyDeviceToHost, st));
            HIP_TRY(hipMemcpyAsync(R.empirical.data(), em.p, 8 * nr, hipMemcpyDeviceToHost, st));
        }
        HIP_TRY(hipStreamSynchronize(st));
        k.release(); c.release(); lo.release(); hi.release(); m.release(); ex.release(); es.release(); em.release();
    }
    // read group table (RecalTable0)
    {
        const uint32_t ns = e->num_rg * 4;
        DevBuf valid, obs, m, ex, es, em;
        HIP_TRY(valid.ensure(4 * ns + 4));
        HIP_TRY(obs.ensure(8 * ns + 8));
        HIP_TRY(m.ensure(8 * ns + 8));
        HIP_TRY(ex.ensure(8 * ns + 8));
        HIP_TRY(es.ensure(8 * ns + 8));
        HIP_TRY(em.ensure(8 * ns + 8));
        HIP_TRY(fpd_launch_rg_table(e->num_rg, e->tables.count[0], e->tables.part[0],
                                    valid.as<uint32_t>(), obs.as<uint64_t>(), m.as<double>(), ex.as<double>(),
                                    es.as<double>(), em.as<double>(), st));
        std::vector<uint32_t> hv(ns);
        std::vector<uint64_t> ho(ns);
        std::vector<double> hm(ns), hx(ns), hs(ns), he(ns);
        if (ns) {
            HIP_TRY(hipMemcpyAsync(hv.data(), valid.p, 4 * ns, hipMemcpyDeviceToHost, st));
            HIP_TRY(hipMemcpyAsync(ho.data(), obs.p, 8 * ns, hipMemcpyDeviceToHost, st));
            HIP_TRY(hipMemcpyAsync(hm.data(), m.p, 8 * ns, hipMemcpyDeviceToHost, st));
            HIP_TRY(hipMemcpyAsync(hx.data(), ex.p, 8 * ns, hipMemcpyDeviceToHost, st));
            HIP_TRY(hipMemcpyAsync(hs.data(), es.p, 8 * ns, hipMemcpyDeviceToHost, st));
            HIP_TRY(hipMemcpyAsync(he.data(), em.p, 8 * ns, hipMemcpyDeviceToHost, st));
        }
        HIP_TRY(hipStreamSynchronize(st));
        TableRows &R = e->rows[0];
        R = TableRows();
        for (uint32_t s = 0; s < ns; s++) {
            if (!hv[s]) continue;
            R.key.push_back((s & 3) | ((s >> 2) << 10));
            R.obs.push_back(ho[s]);
            R.lo.push_back(0);
            R.hi.push_back(0);
            R.mism.push_back(hm[s]);
            R.expected.push_back(hx[s]);
            R.estimated.push_back(hs[s]);
            R.empirical.push_back(he[s]);
        }
    }
    e->postprocessed = true;
    return FP_OK;
}

fp_status check_engine(fp_engine *e, bool need_setup)
{
    if (!e) return fail(FP_ERR_INVALID_ARG, "null engine");
    if (need_setup && !e->is_setup) return fail(FP_ERR_STATE, "engine is not set up");
    HIP_TRY(hipSetDevice(e->device));
    return FP_OK;
}

} // namespace

// ===========================================================================
extern "C" {

void fp_options_init(fp_options *o)
{
    if (!o) return;
    memset(o, 0, sizeof(*o));
    o->abi_version = FP_ABI_VERSION;
    o->engine_kind = FP_ENGINE_PLAIN;
    o->max_batch_reads = 20000;
    o->debug_arrays = 1;
}

const char *fp_last_error(void) { return g_last_error.c_str(); }

const char *fp_status_string(fp_status s)
{
    switch (s) {
    case FP_OK: return "ok";
    case FP_ERR_INVALID_ARG: return "invalid argument";
    case FP_ERR_HIP: return "HIP runtime error";
    case FP_ERR_OUT_OF_MEMORY: return "out of device memory";
    case FP_ERR_STATE: return "call out of order";
    case FP_ERR_BUDGET: return "refused by the HBM budget";
    case FP_ERR_UNSUPPORTED: return "unsupported";
    case FP_ERR_IO: return "I/O error";
    default: return "internal error";
    }
}

fp_status fp_engine_create(int hip_device, const fp_options *opts, fp_engine **out)
{
    if (!out) return fail(FP_ERR_INVALID_ARG, "null output pointer");
    *out = nullptr;
    fp_options o;
    fp_options_init(&o);
    if (opts) {
        if (opts->abi_version != FP_ABI_VERSION) return fail(FP_ERR_UNSUPPORTED, "ABI version mismatch");
        o = *opts;
    }
    int ndev = 0;
    HIP_TRY(hipGetDeviceCount(&ndev));
    if (hip_device < 0 || hip_device >= ndev) return fail(FP_ERR_INVALID_ARG, "no such HIP device");
    HIP_TRY(hipSetDevice(hip_device));
    fp_engine *e = new fp_engine();
    e->device = hip_device;
    e->opts = o;
    e->wpe = o.baq_waves_per_simd >= 1 && o.baq_waves_per_simd <= 4 ? (int)o.baq_waves_per_simd : 1;
    memset(&e->stats, 0, sizeof(e->stats));
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, hip_device) == hipSuccess) e->name = prop.name;
    hipError_t he = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking);
    if (he != hipSuccess) { delete e; return fail(FP_ERR_HIP, hipGetErrorString(he)); }
    for (int i = 0; i < 16; i++) (void)hipEventCreate(&e->ev[i]);
    compute_constants(e->consts);
    he = fpd_baq_set_constants(e->consts.qual2prob, e->consts.map_thr, e->consts.gap_open);
    if (he == hipSuccess)
        he = fpd_post_set_constants(e->consts.qual_err, e->consts.prior, e->consts.log1mp, e->consts.log10e);
    if (he != hipSuccess) { fp_engine_destroy(e); return fail(FP_ERR_HIP, hipGetErrorString(he)); }
    *out = e;
    return FP_OK;
}

fp_status fp_engine_setup(fp_engine *e, const fp_header_view *h, const fp_reference_view *r,
                          const fp_variants_view *v)
{
    fp_status rc = check_engine(e, false);
    if (rc != FP_OK) return rc;
    if (!h || !r || !v) return fail(FP_ERR_INVALID_ARG, "null view");
    if (h->num_read_groups == 0 || h->num_read_groups > 256)
        return fail(FP_ERR_UNSUPPORTED, "read group count must be 1..256");
    e->rg_names.clear();
    for (uint32_t i = 0; i < h->num_read_groups; i++)
        e->rg_names.push_back(h->read_group_names && h->read_group_names[i] ? h->read_group_names[i] : "");
    e->num_rg = h->num_read_groups;

    // reference
    HIP_TRY(e->g_word_offset.ensure(8ull * r->num_sequences + 8));
    HIP_TRY(e->g_length.ensure(4ull * r->num_sequences + 4));
    HIP_TRY(e->g_words.ensure(4ull * r->total_words + 4));
    HIP_TRY(hipMemcpy(e->g_word_offset.p, r->word_offset, 8ull * r->num_sequences, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(e->g_length.p, r->length, 4ull * r->num_sequences, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(e->g_words.p, r->words, 4ull * r->total_words, hipMemcpyHostToDevice));
    // known sites (+ max_end_point_left, loader/variants.cu:125-134)
    const uint64_t nf = v->num_features;
    std::vector<uint32_t> mel(nf);
    for (uint32_t s = 0; s < v->num_sequences; s++) {
        const uint64_t a = v->feature_offset[s], b = v->feature_offset[s + 1];
        if (b < a || b > nf) return fail(FP_ERR_INVALID_ARG, "bad feature offsets");
        uint32_t m = 0;
        for (uint64_t i = a; i < b; i++) {
            m = (i == a) ? v->feature_stop[i] : std::max(m, v->feature_stop[i]);
            mel[i] = v->max_end_point_left ? v->max_end_point_left[i] : m;
        }
    }
    HIP_TRY(e->g_foff.ensure(8ull * (v->num_sequences + 1)));
    HIP_TRY(e->g_fstart.ensure(4 * nf + 4));
    HIP_TRY(e->g_fstop.ensure(4 * nf + 4));
    HIP_TRY(e->g_mel.ensure(4 * nf + 4));
    HIP_TRY(hipMemcpy(e->g_foff.p, v->feature_offset, 8ull * (v->num_sequences + 1), hipMemcpyHostToDevice));
    if (nf) {
        HIP_TRY(hipMemcpy(e->g_fstart.p, v->feature_start, 4 * nf, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(e->g_fstop.p, v->feature_stop, 4 * nf, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(e->g_mel.p, mel.data(), 4 * nf, hipMemcpyHostToDevice));
    }
    fpd::GenomeDev &g = e->genome;
    g.num_contigs = r->num_sequences;
    g.word_offset = e->g_word_offset.as<uint64_t>();
    g.length = e->g_length.as<uint32_t>();
    g.words = e->g_words.as<uint32_t>();
    g.num_variant_contigs = v->num_sequences;
    g.feature_offset = e->g_foff.as<uint64_t>();
    g.feature_start = e->g_fstart.as<uint32_t>();
    g.feature_stop = e->g_fstop.as<uint32_t>();
    g.max_end_left = e->g_mel.as<uint32_t>();

    rc = alloc_tables(e);
    if (rc != FP_OK) return rc;
    e->is_setup = true;
    e->postprocessed = false;
    return FP_OK;
}

static fp_status validate_view(const fp_batch_view *b)
{
    if (!b) return fail(FP_ERR_INVALID_ARG, "null batch view");
    if (b->num_reads == 0) return FP_OK;
    if (!b->chromosome || !b->alignment_start || !b->alignment_stop || !b->mate_chromosome ||
        !b->mate_alignment_start || !b->inferred_insert_size || !b->cigar_start || !b->cigar_len ||
        !b->read_start || !b->read_len || !b->qual_start || !b->qual_len || !b->flags || !b->mapq ||
        !b->read_group)
        return fail(FP_ERR_INVALID_ARG, "batch view has a null array");
    if (b->num_read_bases % 8) return fail(FP_ERR_INVALID_ARG, "num_read_bases must be a multiple of 8");
    return FP_OK;
}

fp_status fp_engine_process_batch(fp_engine *e, const fp_batch_view *hb)
{
    fp_status rc = check_engine(e, true);
    if (rc != FP_OK) return rc;
    rc = validate_view(hb);
    if (rc != FP_OK) return rc;
    const uint32_t n = hb->num_reads;
    if (n == 0) { e->stats.num_batches++; return FP_OK; }
    // bounds of the host arrays (host-side validation: reads must lie inside the arrays)
    uint32_t maxlen = 0;
    for (uint32_t i = 0; i < n; i++) {
        if ((uint64_t)hb->cigar_start[i] + hb->cigar_len[i] > hb->num_cigar_ops ||
            (uint64_t)hb->read_start[i] + ((hb->read_len[i] + 7) & ~7u) > hb->num_read_bases ||
            (uint64_t)hb->qual_start[i] + hb->qual_len[i] > hb->num_qualities || (hb->read_start[i] & 7))
            return fail(FP_ERR_INVALID_ARG, "read " + std::to_string(i) + " lies outside the batch arrays");
        maxlen = std::max(maxlen, std::max(hb->read_len[i], hb->qual_len[i]));
    }
    auto t0 = std::chrono::steady_clock::now();
    hipStream_t st = e->stream;
    const size_t sz[18] = {2ull * n, 4ull * n, 4ull * n, 4ull * n, 4ull * n, 4ull * n,
                           4 * hb->num_cigar_ops + 4, 4ull * n, 4ull * n,
                           hb->num_read_bases / 2 + 4, 4ull * n, 4ull * n,
                           hb->num_qualities + 4, 4ull * n, 4ull * n, 2ull * n, 1ull * n, 4ull * n};
    const void *src[18] = {hb->chromosome, hb->alignment_start, hb->alignment_stop, hb->mate_chromosome,
                           hb->mate_alignment_start, hb->inferred_insert_size, hb->cigars, hb->cigar_start,
                           hb->cigar_len, hb->reads, hb->read_start, hb->read_len, hb->qualities, hb->qual_start,
                           hb->qual_len, hb->flags, hb->mapq, hb->read_group};
    const size_t copy_sz[18] = {2ull * n, 4ull * n, 4ull * n, 4ull * n, 4ull * n, 4ull * n,
                                4 * hb->num_cigar_ops, 4ull * n, 4ull * n,
                                hb->num_read_bases / 2, 4ull * n, 4ull * n,
                                hb->num_qualities, 4ull * n, 4ull * n, 2ull * n, 1ull * n, 4ull * n};
    for (int i = 0; i < 18; i++) {
        HIP_TRY(e->in[i].ensure(sz[i]));
        if (copy_sz[i]) HIP_TRY(hipMemcpyAsync(e->in[i].p, src[i], copy_sz[i], hipMemcpyHostToDevice, st));
    }
    fp_batch_view dv = *hb;
    dv.chromosome = e->in[0].as<uint16_t>();
    dv.alignment_start = e->in[1].as<uint32_t>();
    dv.alignment_stop = e->in[2].as<uint32_t>();
    dv.mate_chromosome = e->in[3].as<uint32_t>();
    dv.mate_alignment_start = e->in[4].as<uint32_t>();
    dv.inferred_insert_size = e->in[5].as<int32_t>();
    dv.cigars = e->in[6].as<uint32_t>();
    dv.cigar_start = e->in[7].as<uint32_t>();
    dv.cigar_len = e->in[8].as<uint32_t>();
    dv.reads = e->in[9].as<uint32_t>();
    dv.read_start = e->in[10].as<uint32_t>();
    dv.read_len = e->in[11].as<uint32_t>();
    dv.qualities = e->in[12].as<uint8_t>();
    dv.qual_start = e->in[13].as<uint32_t>();
    dv.qual_len = e->in[14].as<uint32_t>();
    dv.flags = e->in[15].as<uint16_t>();
    dv.mapq = e->in[16].as<uint8_t>();
    dv.read_group = e->in[17].as<uint32_t>();
    rc = run_batch(e, &dv, maxlen);
    auto t1 = std::chrono::steady_clock::now();
    e->stats.io_ms += 0;   // the copies are part of the batch time here
    (void)t0; (void)t1;
    return rc;
}

fp_status fp_engine_process_batch_device(fp_engine *e, const fp_batch_view *db)
{
    fp_status rc = check_engine(e, true);
    if (rc != FP_OK) return rc;
    rc = validate_view(db);
    if (rc != FP_OK) return rc;
    if (db->num_reads == 0) { e->stats.num_batches++; return FP_OK; }
    return run_batch(e, db, 0);
}

fp_status fp_engine_synchronize(fp_engine *e)
{
    fp_status rc = check_engine(e, false);
    if (rc != FP_OK) return rc;
    HIP_TRY(hipStreamSynchronize(e->stream));
    return FP_OK;
}

fp_status fp_engine_gather(fp_engine *dst, fp_engine *src)
{
    fp_status rc = check_engine(dst, true);
    if (rc != FP_OK) return rc;
    if (!src || !src->is_setup) return fail(FP_ERR_STATE, "source engine is not set up");
    if (src->num_rg != dst->num_rg) return fail(FP_ERR_INVALID_ARG, "engines disagree on read groups");
    HIP_TRY(hipStreamSynchronize(src->stream));
    for (int t = 0; t < 3; t++) {
        const uint64_t n = dst->tab_size[t];
        DevBuf tmp[4];
        const uint64_t *sp[4];
        for (int k = 0; k < 4; k++) {
            if (src->device == dst->device) {
                sp[k] = src->tab[t][k].as<uint64_t>();
            } else {
                HIP_TRY(tmp[k].ensure(8 * n));
                HIP_TRY(hipMemcpyPeerAsync(tmp[k].p, dst->device, src->tab[t][k].p, src->device, 8 * n, dst->stream));
                sp[k] = tmp[k].as<uint64_t>();
            }
        }
        HIP_TRY(fpd_table_add(dst->tables.count[t], dst->tables.part[t], sp[0], sp + 1, n, dst->stream));
        HIP_TRY(hipStreamSynchronize(dst->stream));
        for (int k = 0; k < 4; k++) tmp[k].release();
    }
    dst->stats.total_reads += src->stats.total_reads;
    dst->stats.filtered_reads += src->stats.filtered_reads;
    dst->stats.baq_reads += src->stats.baq_reads;
    dst->stats.num_batches += src->stats.num_batches;
    dst->postprocessed = false;
    return FP_OK;
}

fp_status fp_engine_postprocess(fp_engine *e)
{
    fp_status rc = check_engine(e, true);
    if (rc != FP_OK) return rc;
    HIP_TRY(hipEventRecord(e->ev[4], e->stream));
    rc = postprocess(e);
    if (rc != FP_OK) return rc;
    HIP_TRY(hipEventRecord(e->ev[5], e->stream));
    HIP_TRY(hipEventSynchronize(e->ev[5]));
    e->stats.postprocessing_ms += elapsed(e->ev[4], e->ev[5]);
    return FP_OK;
}

fp_status fp_engine_write_report(fp_engine *e, const char *path)
{
    fp_status rc = check_engine(e, true);
    if (rc != FP_OK) return rc;
    if (!path) return fail(FP_ERR_INVALID_ARG, "null path");
    if (!e->postprocessed) {
        rc = fp_engine_postprocess(e);
        if (rc != FP_OK) return rc;
    }
    auto t0 = std::chrono::steady_clock::now();
    fpr::ReportTables rt;
    for (int t = 0; t < 4; t++) {
        const TableRows &R = e->rows[t];
        fpr::ReportRows &o = rt.table[t];
        o.key = R.key; o.obs = R.obs; o.mism = R.mism; o.estimated = R.estimated; o.empirical = R.empirical;
    }
    std::string text = fpr::format_report(rt, e->rg_names, e->opts.disable_output_rounding != 0);
    FILE *fp = fopen(path, "wb");
    if (!fp) return fail(FP_ERR_IO, std::string("cannot open ") + path);
    const size_t w = fwrite(text.data(), 1, text.size(), fp);
    fclose(fp);
    if (w != text.size()) return fail(FP_ERR_IO, "short write");
    e->stats.output_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return FP_OK;
}

fp_status fp_engine_get_table(fp_engine *e, int table_id, fp_table_out *out)
{
    fp_status rc = check_engine(e, true);
    if (rc != FP_OK) return rc;
    if (!out || table_id < 0 || table_id > 3) return fail(FP_ERR_INVALID_ARG, "bad table request");
    TableRows tmp;
    const TableRows *R = nullptr;
    if (e->postprocessed) {
        R = &e->rows[table_id];
    } else {
        if (table_id == 0) { out->count = 0; return FP_OK; }
        DevBuf k, c, lo, hi;
        uint64_t nr = 0;
        rc = compact(e, table_id - 1, k, c, lo, hi, &nr);
        if (rc != FP_OK) return rc;
        tmp.key.resize(nr); tmp.obs.resize(nr); tmp.lo.resize(nr); tmp.hi.resize(nr);
        if (nr) {
            HIP_TRY(hipMemcpy(tmp.key.data(), k.p, 4 * nr, hipMemcpyDeviceToHost));
            HIP_TRY(hipMemcpy(tmp.obs.data(), c.p, 8 * nr, hipMemcpyDeviceToHost));
            HIP_TRY(hipMemcpy(tmp.lo.data(), lo.p, 8 * nr, hipMemcpyDeviceToHost));
            HIP_TRY(hipMemcpy(tmp.hi.data(), hi.p, 8 * nr, hipMemcpyDeviceToHost));
        }
        tmp.mism.resize(nr);
        for (uint64_t i = 0; i < nr; i++) tmp.mism[i] = fpd::fixed_to_double(tmp.lo[i], tmp.hi[i]);
        tmp.expected.assign(nr, 0.0); tmp.estimated.assign(nr, 0.0); tmp.empirical.assign(nr, 0.0);
        R = &tmp;
    }
    const size_t n = R->size();
    out->count = n;
    for (size_t i = 0; i < n; i++) {
        if (out->keys) out->keys[i] = R->key[i];
        if (out->observations) out->observations[i] = R->obs[i];
        if (out->mismatches) out->mismatches[i] = R->mism[i];
        if (out->mismatches_fixed_lo) out->mismatches_fixed_lo[i] = R->lo[i];
        if (out->mismatches_fixed_hi) out->mismatches_fixed_hi[i] = R->hi[i];
        if (out->expected_errors) out->expected_errors[i] = R->expected[i];
        if (out->estimated_quality) out->estimated_quality[i] = R->estimated[i];
        if (out->empirical_quality) out->empirical_quality[i] = R->empirical[i];
    }
    return FP_OK;
}

fp_status fp_engine_stats(fp_engine *e, fp_stats *out)
{
    if (!e || !out) return fail(FP_ERR_INVALID_ARG, "null argument");
    *out = e->stats;
    return FP_OK;
}

fp_status fp_engine_get_name(fp_engine *e, char *buf, size_t len)
{
    if (!e || !buf || len == 0) return fail(FP_ERR_INVALID_ARG, "null argument");
    std::string s = "firepony_amd " + e->name + " (hip device " + std::to_string(e->device) + ")";
    snprintf(buf, len, "%s", s.c_str());
    return FP_OK;
}

fp_status fp_engine_get_total_memory(fp_engine *e, uint64_t *bytes)
{
    fp_status rc = check_engine(e, false);
    if (rc != FP_OK) return rc;
    if (!bytes) return fail(FP_ERR_INVALID_ARG, "null argument");
    size_t fr = 0, tot = 0;
    HIP_TRY(hipMemGetInfo(&fr, &tot));
    *bytes = tot;
    return FP_OK;
}

static fp_status debug_bytes(fp_engine *e, int item, uint64_t *bytes)
{
    const uint64_t n = e->last_num_reads;
    switch (item) {
    case FP_DBG_ACTIVE_READS: {
        std::vector<uint8_t> st(n);
        if (n) HIP_TRY(hipMemcpy(st.data(), e->rs.status, n, hipMemcpyDeviceToHost));
        uint64_t c = 0;
        for (uint8_t s : st) c += s == fpd::RS_ACTIVE;
        *bytes = 4 * c;
        return FP_OK;
    }
    case FP_DBG_READ_WINDOW: case FP_DBG_NOINS_WINDOW: case FP_DBG_REF_WINDOW: case FP_DBG_HMM_WINDOW:
    case FP_DBG_HQ_WINDOW: *bytes = 4 * n; return FP_OK;
    case FP_DBG_NUM_ERRORS: case FP_DBG_BANDWIDTH: *bytes = 2 * n; return FP_OK;
    case FP_DBG_ERROR_FLAGS: case FP_DBG_ACTIVE_MASK: *bytes = e->last_num_bases; return FP_OK;
    case FP_DBG_BAQ: *bytes = e->last_num_quals; return FP_OK;
    case FP_DBG_FRAC_ERRORS: *bytes = 24 * e->last_num_quals; return FP_OK;
    default: return fail(FP_ERR_INVALID_ARG, "unknown debug item");
    }
}

fp_status fp_engine_debug_size(fp_engine *e, int item, uint64_t *bytes)
{
    fp_status rc = check_engine(e, true);
    if (rc != FP_OK) return rc;
    if (!bytes) return fail(FP_ERR_INVALID_ARG, "null argument");
    return debug_bytes(e, item, bytes);
}

fp_status fp_engine_debug_fetch(fp_engine *e, int item, void *dst, uint64_t bytes)
{
    fp_status rc = check_engine(e, true);
    if (rc != FP_OK) return rc;
    uint64_t need = 0;
    rc = debug_bytes(e, item, &need);
    if (rc != FP_OK) return rc;
    if (!dst || bytes < need) return fail(FP_ERR_INVALID_ARG, "debug buffer too small");
    HIP_TRY(hipStreamSynchronize(e->stream));
    const uint64_t n = e->last_num_reads;
    const void *src = nullptr;
    switch (item) {
    case FP_DBG_ACTIVE_READS: {
        std::vector<uint8_t> st(n);
        if (n) HIP_TRY(hipMemcpy(st.data(), e->rs.status, n, hipMemcpyDeviceToHost));
        uint32_t *o = (uint32_t *)dst;
        uint64_t c = 0;
        for (uint64_t i = 0; i < n; i++) if (st[i] == fpd::RS_ACTIVE) o[c++] = (uint32_t)i;
        return FP_OK;
    }
    case FP_DBG_ERROR_FLAGS: case FP_DBG_ACTIVE_MASK: {
        std::vector<uint8_t> f(need);
        if (need) HIP_TRY(hipMemcpy(f.data(), e->rs.base_flags, need, hipMemcpyDeviceToHost));
        uint8_t *o = (uint8_t *)dst;
        for (uint64_t i = 0; i < need; i++)
            o[i] = item == FP_DBG_ERROR_FLAGS ? (f[i] & 7) : ((f[i] & fpd::BF_MASKED) ? 0 : 1);
        return FP_OK;
    }
    case FP_DBG_READ_WINDOW: src = e->rs.rwc; break;
    case FP_DBG_NOINS_WINDOW: src = e->rs.noins; break;
    case FP_DBG_REF_WINDOW: src = e->rs.refw; break;
    case FP_DBG_HMM_WINDOW: src = e->rs.hmmw; break;
    case FP_DBG_HQ_WINDOW: src = e->rs.hqw; break;
    case FP_DBG_NUM_ERRORS: src = e->rs.num_errors; break;
    case FP_DBG_BANDWIDTH: src = e->rs.bandwidth; break;
    case FP_DBG_BAQ: src = e->rs.baq; break;
    case FP_DBG_FRAC_ERRORS:
        if (!e->opts.debug_arrays) return fail(FP_ERR_STATE, "fractional errors are kept only with debug_arrays = 1");
        src = e->rs.frac;
        break;
    default: return fail(FP_ERR_INVALID_ARG, "unknown debug item");
    }
    if (need) HIP_TRY(hipMemcpy(dst, src, need, hipMemcpyDeviceToHost));
    return FP_OK;
}

fp_status fp_engine_reset(fp_engine *e)
{
    fp_status rc = check_engine(e, true);
    if (rc != FP_OK) return rc;
    for (int t = 0; t < 3; t++)
        for (int k = 0; k < 4; k++) HIP_TRY(hipMemsetAsync(e->tab[t][k].p, 0, e->tab_size[t] * 8, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    memset(&e->stats, 0, sizeof(e->stats));
    e->postprocessed = false;
    for (auto &r : e->rows) r = TableRows();
    return FP_OK;
}

void fp_engine_destroy(fp_engine *e)
{
    if (!e) return;
    (void)hipSetDevice(e->device);
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    DevBuf *bufs[] = {&e->g_word_offset, &e->g_length, &e->g_words, &e->g_foff, &e->g_fstart, &e->g_fstop,
                      &e->g_mel, &e->st_status, &e->st_rwc, &e->st_noins, &e->st_refw, &e->st_nerr, &e->st_hmmw,
                      &e->st_bw, &e->st_hqw, &e->st_bflags, &e->st_baq, &e->st_state, &e->st_frac, &e->counters,
                      &e->baq_lists, &e->baq_counts, &e->scratch, &e->rg_buf};
    for (DevBuf *b : bufs) b->release();
    for (auto &b : e->in) b.release();
    for (auto &t : e->tab) for (auto &b : t) b.release();
    for (int i = 0; i < 16; i++) if (e->ev[i]) (void)hipEventDestroy(e->ev[i]);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
}

// host-side self checks usable without a GPU ------------------------------

/* MAP-quality thresholds as computed at engine creation (tests compare them
 * with the host libm formula). */
void fp_selftest_map_thresholds(double *out128)
{
    HostConstants c;
    compute_constants(c);
    memcpy(out128, c.map_thr, sizeof(c.map_thr));
}

/* fixed-point conversions used by the kernels, run on the host */
void fp_selftest_fixed(const double *v, uint64_t n, uint64_t *lo, uint64_t *hi, double *back)
{
    for (uint64_t i = 0; i < n; i++) {
        fpd::u128 f = fpd::fixed_from_double(v[i]);
        lo[i] = f.lo;
        hi[i] = f.hi;
        back[i] = fpd::fixed_to_double(f.lo, f.hi);
    }
}

double fp_selftest_fixed_to_double(uint64_t lo, uint64_t hi) { return fpd::fixed_to_double(lo, hi); }

} // extern "C"
