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
#include <deque>
#include <mutex>
#include <string>
#include <vector>

// kernels (k_*.hip)
extern "C" {
hipError_t fpd_launch_prep(const fpd::BatchDev *b, const fpd::GenomeDev *g, const fpd::ReadStateDev *s,
                           uint32_t *batch, uint32_t num_rg, uint32_t max_read_len, hipStream_t st);
hipError_t fpd_baq_set_constants(const double *qual2prob, const double *thr, double gap_open);
hipError_t fpd_launch_baq_classify(uint32_t num_reads, const uint32_t *live, const fpd::ReadStateDev *s,
                                   uint32_t *lists, uint32_t *counts, uint32_t capacity, hipStream_t st);
hipError_t fpd_launch_baq_hmm(int cls, const fpd::BatchDev *b, const fpd::GenomeDev *g, const fpd::ReadStateDev *s,
                              const uint32_t *list, const uint32_t *count, double *scratch, uint64_t slot_doubles,
                              uint32_t rows_cap, uint32_t wmax_runtime, uint32_t waves, unsigned long long *cells,
                              int wpe, int allow_stage, int recompute, uint32_t *next_chunk, hipStream_t st);
hipError_t fpd_baq_prepare();
hipError_t fpd_launch_baq_post(const fpd::BatchDev *b, const fpd::ReadStateDev *s, uint32_t band_done,
                               hipStream_t st);
hipError_t fpd_fill_u8(uint8_t *p, uint64_t n, const uint32_t *live, uint8_t v, hipStream_t st);
hipError_t fpd_fill_u32(uint32_t *p, uint64_t n, uint32_t v, hipStream_t st);
hipError_t fpd_rebase(uint32_t n, uint32_t *cigar_start, uint32_t *read_start, uint32_t *qual_start,
                      uint32_t *read_group, uint32_t cig_off, uint32_t base_off, uint32_t qual_off, uint32_t rg_off,
                      uint32_t rg_limit, hipStream_t st);
hipError_t fpd_append(const void *const *src, void *const *dst, const uint64_t *bytes, uint32_t cig_off,
                      uint32_t base_off, uint32_t qual_off, uint32_t rg_off, uint32_t rg_limit, uint64_t block_bytes,
                      hipStream_t st);
hipError_t fpd_cov_prepare();
hipError_t fpd_launch_cov(const fpd::BatchDev *b, const fpd::ReadStateDev *s, const fpd::TablesDev *t,
                          uint32_t frac_f64, uint32_t keep_frac, uint32_t *batch, uint32_t *rg_buf,
                          uint32_t max_read_len, uint32_t *const *pkey, double *const *pval, uint32_t *pcursor,
                          uint64_t pcap, hipStream_t st);
hipError_t fpd_sort_reduce(uint32_t *keys, double *vals, uint32_t *tkeys, double *tvals, uint64_t cap,
                           const uint32_t *n_emit, uint32_t *n_kept, uint32_t *hist, uint32_t *totals,
                           uint32_t key_bits, uint32_t digit_bits, uint32_t table_entries,
                           unsigned long long *count, unsigned long long *const *parts, hipStream_t st);
hipError_t fpd_compact_count(const uint64_t *count, uint64_t n, uint32_t *block_counts, hipStream_t st);
hipError_t fpd_compact_scatter(const uint64_t *count, const uint64_t *const *parts, uint64_t n,
                               const uint64_t *block_offsets, uint32_t *okey, uint64_t *ocount, uint64_t *olo,
                               uint64_t *ohi, hipStream_t st);
hipError_t fpd_table_add(uint64_t *dc, uint64_t *const *dparts, const uint64_t *sc, const uint64_t *const *sparts,
                         uint64_t n, hipStream_t st);
hipError_t fpd_post_set_constants(const double *qual_err, const double *prior, const double *log1mp, double log10e);
hipError_t fpd_launch_obs_values(uint64_t n, const uint32_t *key, const uint64_t *count, const uint64_t *lo,
                                 const uint64_t *hi, int qual_shift, double *mism, double *expected, hipStream_t st);
hipError_t fpd_launch_empirical(uint64_t n, const uint64_t *count, const double *mism, const double *estimated,
                                const double *coeff, uint32_t need_rounding, double *empirical, hipStream_t st);
hipError_t fpd_launch_rg_table(uint32_t num_rg, const uint64_t *count, const uint64_t *const *parts,
                               uint32_t *valid, uint64_t *obs, double *mism, double *expected, hipStream_t st);
hipError_t fpd_check_batch(const fpd::BatchDev *b, uint32_t rg_limit, uint32_t *out, hipStream_t st);
}

// ---------------------------------------------------------------------------
// Scratch of the allocation-free batch entry point (fp_engine_process_batch_device_noalloc): per-batch
// read state, counters and the BAQ band rows of one batch.  A file-scope device array of a fixed size;
// each GPU the library runs on holds its own copy until the process ends.  Host code never passes the
// array's name to a kernel or copy: it takes the array's device address on each device once
// (hipGetSymbolAddress) and uses that.
static constexpr uint64_t NOALLOC_SCRATCH_BYTES = 256ull << 20;
static constexpr uint32_t NOALLOC_MAX_READS = 65536;
static constexpr uint32_t NOALLOC_MAX_READ_LENGTH = 256;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

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

// per-batch device state: one set for the plain path (and batches run eagerly), one per staging
// buffer of the batched engine, so two staged batches can be in flight on two streams
struct BatchState {
    DevBuf st_status, st_rwc, st_noins, st_refw, st_nerr, st_hmmw, st_bw, st_hqw, st_bflags, st_baq, st_state, st_frac;
    DevBuf baq_lists, baq_counts, rg_buf;
    fpd::ReadStateDev rs{};
    // fork and join events around the 23- and 33-cell band classes, which run on the engine's second
    // BAQ stream beside the 15-cell class
    hipEvent_t fork = nullptr, join = nullptr;
    void release()
    {
        DevBuf *all[] = {&st_status, &st_rwc, &st_noins, &st_refw, &st_nerr, &st_hmmw, &st_bw, &st_hqw,
                         &st_bflags, &st_baq, &st_state, &st_frac, &baq_lists, &baq_counts, &rg_buf};
        for (DevBuf *b : all) b->release();
        if (fork) { (void)hipEventDestroy(fork); fork = nullptr; }
        if (join) { (void)hipEventDestroy(join); join = nullptr; }
    }
};

struct fp_engine {
    int device = 0;
    fp_options opts;
    hipStream_t stream = nullptr;
    // the BAQ stage's second stream, shared by every batch state (the BAQ stages of the staging buffers
    // take turns): a process has few hardware queues, and streams beyond them share one
    hipStream_t side = nullptr;
    // host batches of the batched engine are copied by the DMA engines on a stream of their own: the
    // copies then run while the BAQ kernels hold every CU's registers (a copy kernel would wait for
    // them), and the host waits for its own batch's copies only
    hipStream_t copy = nullptr;
    hipEvent_t copied = nullptr;
    std::string name;
    HostConstants consts;
    int wpe = 1;   // waves per SIMD of the 15-cell BAQ band kernel (1-4)
    int ckpt = 1;  // rows per forward checkpoint of the 15-cell band kernel (1: every row spilled; 3-5: recompute)

    // setup state
    bool is_setup = false;
    std::vector<std::string> rg_names;
    uint32_t num_rg = 0;
    DevBuf g_word_offset, g_length, g_words, g_foff, g_fstart, g_fstop, g_mel, g_boff, g_slb, g_ssb;
    fpd::GenomeDev genome{};

    // batch buffers (host batches are copied here)
    DevBuf in[18];
    // per-batch state
    BatchState ps;                      // the plain path's (and that of batches run eagerly)
    BatchState na;                      // allocation-free batches: views into the scratch array
    BatchState *last = &ps;             // the state of the last processed batch (inspection)
    DevBuf counters, scratch, scratch_gen;
    uint32_t *host_counts = nullptr;   // pinned readback area (64 u32)
    uint32_t last_num_reads = 0;
    uint64_t last_num_bases = 0, last_num_quals = 0;

    // sort aggregation: emitted pairs per table, one ping-pong pair buffer, radix histograms,
    // counters {append cursor x3, overflow, kept x3}
    DevBuf pk[3], pv[3], tk, tv, shist, stot, sctr;
    uint64_t pair_cap = 0;
    int sort_bits = 8;                  // radix digit width in use (4-8)
    bool sort_tuned = false;
    double sort_tune_ms[5] = {0, 0, 0, 0, 0};

    // resident tables
    DevBuf tab[3][4];     // per table: count, then the three limb sums
    uint64_t tab_size[3] = {0, 0, 0};
    fpd::TablesDev tables{};

    // results
    bool postprocessed = false;
    TableRows rows[4];

    fp_stats stats;
    hipEvent_t ev[16];

    // ---- batched engine: input batches are appended to one of two staging buffers; a full buffer
    // runs as two graph replays (stage 1: prep + classification; stage 2: BAQ bands, cap/recode,
    // covariates) with one small host readback between them (band classes, longest read)
    struct Stage {
        DevBuf arr[18];                 // the fp_batch_view arrays, capacity-sized
        DevBuf live;                    // {reads, qualities lo, qualities hi, 0}
        uint32_t *live_host = nullptr;  // pinned
        uint32_t n = 0;
        uint64_t ncig = 0, nbases = 0, nquals = 0;
        // front stage; 15-cell BAQ class; 23- and 33-cell classes; cap/recode + covariates
        hipGraphExec_t gfront = nullptr, gbaq = nullptr, gside = nullptr, gcov = nullptr;
        int g_wpe = 0, g_ckpt = 0, g_bits = 0;
        uint64_t g_epoch = 0;
        BatchState bs;                  // this buffer's per-batch state
        hipStream_t stream = nullptr;   // its appends and front stage (the engine stream)
        hipEvent_t baq_done = nullptr;  // after its 15-cell and generic BAQ kernels
        hipEvent_t cov_done = nullptr;  // after its covariate stage (the buffer may refill)
        hipEvent_t t0 = nullptr, t1 = nullptr;   // around its stage 2 (timing)
        bool back_pending = false;      // stage-2 time not yet collected
        uint64_t fill = 1;              // which fill of this buffer is staging now (a job records the
                                        // fills that hold its reads)
        bool copies_pending = false;    // copy-stream work the buffer's stream has not waited for yet
        // A flushed buffer's front stage (prep, classification) runs first; its small readback (band
        // class counts, longest read, error words) lands in hc, and the rest of the batch is launched
        // once the host sees it there: the host goes on staging the next batches meanwhile.
        uint32_t *hc = nullptr;         // pinned: [0, BW_WORDS) the front stage's counts; [64] the pair
                                        // overflow word of the buffer's previous batch
        hipEvent_t f0 = nullptr, f1 = nullptr;   // around its front stage (timing)
        hipEvent_t front_done = nullptr;         // after the front stage and its readback
        bool pending = false;           // flushed; the stages after the front not launched yet
        uint32_t pend_n = 0;            // reads of the pending batch
    };
    // staging buffers, used in turn: one fills while the others run.  A buffer refills only once its
    // previous batch has run, so with two a host-fed buffer's copies wait for its own last batch;
    // with three (the default) they overlap the other two buffers' kernels.
    static constexpr int MAX_STAGE = 4;
    Stage stg[MAX_STAGE];
    int nstage = 3;
    bool batched = false;
    // The batched engine's streams follow the stages, not the buffers: the engine stream carries
    // every buffer's device appends and front stage, bq the 15-cell and generic BAQ kernels, side the
    // 23- and 33-cell classes, cap/recode and the covariates, copy the host batches.  Each stream
    // takes the buffers in flush order, and a process has four hardware queues (queues run their
    // work in order): with one stream per stage, one buffer's front stage never waits behind
    // another buffer's BAQ or covariate kernels in a shared queue.
    hipStream_t bq = nullptr;
    int cur = 0;
    std::deque<int> pending;            // flushed buffers waiting for their readback, in flush order
    int last_back = -1;                 // the buffer whose stages were launched last (-1: none yet)
    uint32_t last_flush_n = 0;          // reads of the buffer flushed last (0 after a sync)
    double idle_since = -1;             // host time (ms) the device was first seen idle, or -1
    uint32_t cap_reads = 0, rows_cap = 0;
    uint64_t cap_cig = 0, cap_bases = 0, cap_quals = 0;
    DevBuf bcounters;                   // prep / BAQ counters accumulated over staged batches
    uint64_t epoch = 1;                 // changes when a buffer the graphs refer to is reallocated
    bool staged_ready = false;
    // autotuner: waves per SIMD of the 15-cell BAQ band kernel, timed on the first staged batches
    bool tuned = false;
    int tune_pending = -1;              // candidate whose stage-2 timing is in flight
    int tune_kind = 0;                  // 0: BAQ waves per SIMD, 1: sort digit width
    uint32_t tune_pending_reads = 0;
    double tune_ms_per_read[4] = {0, 0, 0, 0};   // one per BAQ candidate
    hipEvent_t bev[4] = {nullptr, nullptr, nullptr, nullptr};
    bool back_pending = false;          // stage-2 time of the last staged batch not yet collected
    uint64_t eager_cells = 0, eager_cells15 = 0;   // BAQ cells of oversized batches run eagerly
    bool band15_timed = false;          // profile_stages: events 11/12 bracket a 15-cell band launch

    // serving: owner of each read-group table slot (-1 = free), the HBM byte budget of the open jobs
    std::vector<int64_t> slot_owner;
    int64_t next_job_id = 1;
    uint64_t budget_bytes = 0, bytes_in_use = 0;
    // job postprocess, release and device-batch checks run on a stream of their own, ordered after
    // the staging buffers that hold the job's reads (events), so they overlap other jobs' batches
    hipStream_t post = nullptr;
    // postprocess working buffers, grown and kept: freeing device memory waits for the whole device
    DevBuf pw[12];
    DevBuf check;
};

// a serving job: its read groups occupy table slots [base, base + nrg)
struct fp_job {
    fp_engine *e = nullptr;
    int64_t id = 0;
    uint32_t base = 0, nrg = 0;
    std::vector<std::string> names;
    bool closed = false;
    uint64_t reads = 0, batches = 0;
    double post_ms = 0;
    TableRows rows[4];
    uint64_t fill[fp_engine::MAX_STAGE] = {};   // per staging buffer: the last fill that holds the job's reads (0: none)
    uint64_t bytes = 0;          // HBM footprint counted against the engine's budget
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

// per-read and per-base state for n reads; the per-quality fractional-error array is needed for
// inspection (debug_arrays) and as working memory of windows longer than 256 bases
fp_status ensure_state(fp_engine *e, BatchState &S, uint32_t n, uint64_t nbases, uint64_t nquals,
                       bool need_frac = true)
{
    if (!e->side) HIP_TRY(hipStreamCreateWithFlags(&e->side, hipStreamNonBlocking));
    if (!S.fork) HIP_TRY(hipEventCreateWithFlags(&S.fork, hipEventDisableTiming));
    if (!S.join) HIP_TRY(hipEventCreateWithFlags(&S.join, hipEventDisableTiming));
    HIP_TRY(S.st_status.ensure(n));
    HIP_TRY(S.st_rwc.ensure(4ull * n));
    HIP_TRY(S.st_noins.ensure(4ull * n));
    HIP_TRY(S.st_refw.ensure(4ull * n));
    HIP_TRY(S.st_nerr.ensure(2ull * n));
    HIP_TRY(S.st_hmmw.ensure(4ull * n));
    HIP_TRY(S.st_bw.ensure(2ull * n));
    HIP_TRY(S.st_hqw.ensure(4ull * n));
    HIP_TRY(S.st_bflags.ensure(nbases + 8));
    HIP_TRY(S.st_baq.ensure(nquals + 8));
    HIP_TRY(S.st_state.ensure(4 * (nquals + 8)));
    if (need_frac) HIP_TRY(S.st_frac.ensure(24 * (nquals + 8)));
    HIP_TRY(e->counters.ensure(64));
    HIP_TRY(S.baq_lists.ensure(4ull * 4 * std::max<uint32_t>(n, 1)));
    HIP_TRY(S.rg_buf.ensure(4ull * (4ull * std::max<uint32_t>(e->num_rg, 1) + 2 + n)));
    HIP_TRY(S.baq_counts.ensure(4 * 400));   // class sizes, (class, length bucket) sizes, offsets, cursors
    fpd::ReadStateDev &r = S.rs;
    r.status = S.st_status.as<uint8_t>();
    r.rwc = S.st_rwc.as<uint32_t>();
    r.noins = S.st_noins.as<uint32_t>();
    r.refw = S.st_refw.as<uint32_t>();
    r.num_errors = S.st_nerr.as<uint16_t>();
    r.hmmw = S.st_hmmw.as<uint32_t>();
    r.bandwidth = S.st_bw.as<uint16_t>();
    r.hqw = S.st_hqw.as<uint32_t>();
    r.base_flags = S.st_bflags.as<uint8_t>();
    r.baq = S.st_baq.as<uint8_t>();
    r.baq_state = S.st_state.as<uint32_t>();
    r.frac = S.st_frac.as<double>();
    return FP_OK;
}

fpd::BatchDev to_dev(const fp_batch_view *v)
{
    fpd::BatchDev b{};
    b.num_reads = v->num_reads;
    b.live = nullptr;          // sizes are the view's (a staged batch sets its device descriptor)
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

// the 15-cell band kernel's time in the last profiled BAQ stage (its stream work has completed)
void collect_band15_time(fp_engine *e)
{
    if (!e->band15_timed) return;
    e->stats.baq15_kernel_ms += elapsed(e->ev[11], e->ev[12]);
    e->stats.baq15_launches += 1;
    e->band15_timed = false;
}

static const uint32_t BAQ_CLASS_W[4] = {15, 23, 33, 0};
// words 8..10 of a batch's BAQ class counts (zeroed with them by the front stage): the chunk counters of
// the three band kernels
static constexpr uint32_t BAQ_CHUNK_COUNTERS = 8;

// what the launch sequence needs to know about a batch
struct Plan {
    uint32_t n;              // reads (the capacity for a staged batch)
    uint64_t nquals;         // qualities (the capacity for a staged batch)
    const uint32_t *live;    // device live sizes of a staged batch, or null
    uint32_t rows_cap;       // >= the longest read
    // allocation-free batches: the BAQ band classes take turns in this region, each with as many
    // resident waves as fit, and no per-quality fractional errors are kept
    double *band = nullptr;
    uint64_t band_doubles = 0;
    bool no_frac = false;
};

// resident waves of a band class; its per-wave scratch slot in doubles
static uint32_t band_waves(uint32_t wpe_class0, int c, uint32_t n)
{
    const uint32_t wpe = c == 0 ? wpe_class0 : 1u;
    const uint32_t w = std::min<uint32_t>((n + 63) / 64, 1024u * wpe);
    return (w + 3) / 4 * 4;
}
static uint64_t band_slot(int c, uint32_t rows_cap, uint32_t wgen)
{
    const uint64_t W = c < 3 ? BAQ_CLASS_W[c] : wgen;
    // forward rows (M, I) | scale factors | staged cap codes and tags (band kernels) or band state (generic)
    return ((uint64_t)(rows_cap + 1) * W * 2 + (rows_cap + 2) + (c == 3 ? 3ull * W : rows_cap + 1ull)) * 64;
}
// the band classes run at the same time (two streams), so each has its own region of the scratch
static uint64_t band_offset(int wpe, int c, uint32_t n, uint32_t rows_cap)
{
    uint64_t off = 0;
    for (int k = 0; k < c; k++) off += band_slot(k, rows_cap, 0) * band_waves(wpe, k, n);
    return off;
}
static uint64_t band_scratch_doubles(int wpe, uint32_t n, uint32_t rows_cap)
{
    return band_offset(wpe, 3, n, rows_cap);
}

// stage 1 (graph-safe): per-read prep, BAQ tag reset, band classification
fp_status launch_front(fp_engine *e, BatchState &S, const fpd::BatchDev &b, const Plan &p, uint32_t *counters, hipStream_t st)
{
    (void)counters;
    HIP_TRY(fpd_fill_u32(S.baq_counts.as<uint32_t>(), 144, 0u, st));   // class sizes and the batch's counter words
    HIP_TRY(fpd_launch_prep(&b, &e->genome, &S.rs, S.baq_counts.as<uint32_t>(), e->num_rg, p.rows_cap, st));
    HIP_TRY(fpd_fill_u8(S.rs.baq, p.nquals, p.live, 0xff, st));
    HIP_TRY(fpd_launch_baq_classify(p.n, p.live, &S.rs, S.baq_lists.as<uint32_t>(), S.baq_counts.as<uint32_t>(),
                                    std::max<uint32_t>(p.n, 1), st));
    return FP_OK;
}

// stage 2 (graph-safe): BAQ band kernels, cap/recode, covariate tables.  The class of bands wider
// than 33 cells runs separately (launch_generic) before it, when present.
// key bits a table's sort has to order: the fields below the read group plus the bits of the
// largest read-group id
static uint32_t sort_key_bits(uint32_t num_rg, int t)
{
    const uint32_t rg_bits = num_rg <= 1 ? 0u : 32u - (uint32_t)__builtin_clz(num_rg - 1);
    return (uint32_t)fpd::TABLE_BITS[t] - 8u + rg_bits;
}

// pair buffers of the sort aggregation for up to `cap` observations per table
fp_status ensure_pairs(fp_engine *e, uint64_t cap)
{
    if (cap > 0xf0000000ull)
        return fail(FP_ERR_INVALID_ARG, "sort aggregation: over 4e9 observations per table in one batch");
    cap = std::max<uint64_t>(cap, 4096);
    if (cap <= e->pair_cap) return FP_OK;
    for (int t = 0; t < 3; t++) {
        HIP_TRY(e->pk[t].ensure(4 * cap + 64));
        HIP_TRY(e->pv[t].ensure(8 * cap + 64));
    }
    HIP_TRY(e->tk.ensure(4 * cap + 64));
    HIP_TRY(e->tv.ensure(8 * cap + 64));
    HIP_TRY(e->shist.ensure(4ull * 256 * 2048 + 64));   // radix x the sort kernels' grid (at most 2048)
    HIP_TRY(e->stot.ensure(4 * 256));
    HIP_TRY(e->sctr.ensure(64));
    e->pair_cap = cap;
    return FP_OK;
}

// the covariate stage: direct LDS-privatised table updates, or pairs -> radix sort -> reduce_by_key
fp_status launch_cov_stage(fp_engine *e, BatchState &S, const fpd::BatchDev &b, const Plan &p, int sort_bits, uint32_t *counters,
                           hipStream_t st, hipEvent_t sort_mark)
{
    const uint32_t keep = e->opts.debug_arrays && !p.no_frac ? 1u : 0u;
    const uint32_t maxlen = std::min<uint32_t>(p.rows_cap, 499);
    if (e->opts.aggregation != FP_AGG_SORT) {
        HIP_TRY(fpd_launch_cov(&b, &S.rs, &e->tables, e->opts.frac_error_f64, keep, S.baq_counts.as<uint32_t>(),
                               S.rg_buf.as<uint32_t>(), maxlen, nullptr, nullptr, nullptr, 0, st));
        return FP_OK;
    }
    uint32_t *sc = e->sctr.as<uint32_t>();
    HIP_TRY(fpd_fill_u32(sc, 8, 0u, st));
    uint32_t *pkeys[3];
    double *pvals[3];
    for (int t = 0; t < 3; t++) { pkeys[t] = e->pk[t].as<uint32_t>(); pvals[t] = e->pv[t].as<double>(); }
    HIP_TRY(fpd_launch_cov(&b, &S.rs, &e->tables, e->opts.frac_error_f64, keep, S.baq_counts.as<uint32_t>(),
                           S.rg_buf.as<uint32_t>(), maxlen, pkeys, pvals, sc, e->pair_cap, st));
    if (sort_mark) HIP_TRY(hipEventRecord(sort_mark, st));
    for (int t = 0; t < 3; t++) {
        unsigned long long *parts[3];
        for (int k = 0; k < 3; k++) parts[k] = (unsigned long long *)e->tables.part[t][k];
        HIP_TRY(fpd_sort_reduce(pkeys[t], pvals[t], e->tk.as<uint32_t>(), e->tv.as<double>(), e->pair_cap, sc + t,
                                sc + 4 + t, e->shist.as<uint32_t>(), e->stot.as<uint32_t>(),
                                sort_key_bits(e->num_rg, t), (uint32_t)sort_bits, (uint32_t)e->tab_size[t],
                                (unsigned long long *)e->tables.count[t], parts, st));
    }
    return FP_OK;
}

// pair counts of the last sort-aggregated batch (its stream work has completed)
fp_status collect_sort_counts(fp_engine *e, int sort_bits)
{
    uint32_t c[8];
    HIP_TRY(hipMemcpy(c, e->sctr.p, sizeof c, hipMemcpyDeviceToHost));
    for (int t = 0; t < 3; t++) {
        e->stats.sort_slots += c[t];
        e->stats.sort_pairs += c[4 + t];
        e->stats.sort_passes += (sort_key_bits(e->num_rg, t) + sort_bits - 1) / sort_bits;
    }
    return FP_OK;
}

// stage 2a (graph-safe): the BAQ band classes and cap/recode.  The 15-cell class counts its band
// cells in counter words 10..11, the other classes in 8..9.
// one band class's kernel over its read list, in the class's own region of the band scratch
static hipError_t launch_band_class(fp_engine *e, BatchState &S, const fpd::BatchDev &b, const Plan &p, int wpe,
                                    int ckpt, uint32_t *counters, int c, hipStream_t s)
{
    const uint32_t cap = std::max<uint32_t>(p.n, 1);
    // the class's chunk counter: a word of the class counts that the front stage zeroes
    return fpd_launch_baq_hmm(c, &b, &e->genome, &S.rs, S.baq_lists.as<uint32_t>() + (uint64_t)c * cap,
                              S.baq_counts.as<uint32_t>() + c,
                              e->scratch.as<double>() + band_offset(wpe, c, p.n, p.rows_cap),
                              band_slot(c, p.rows_cap, 0), p.rows_cap, BAQ_CLASS_W[c], band_waves(wpe, c, p.n),
                              (unsigned long long *)(counters + (c == 0 ? 10 : 8)), wpe,
                              e->opts.disable_baq_staging ? 0 : 1, ckpt > 1 ? ckpt : 0,
                              S.baq_counts.as<uint32_t>() + BAQ_CHUNK_COUNTERS + c, s);
}

fp_status launch_baq_stage(fp_engine *e, BatchState &S, const fpd::BatchDev &b, const Plan &p, int wpe, int ckpt,
                           uint32_t *counters, hipStream_t st, hipEvent_t mark, const uint32_t *known_counts)
{
    auto launch_class = [&](int c, hipStream_t s) -> hipError_t {
        return launch_band_class(e, S, b, p, wpe, ckpt, counters, c, s);
    };
    if (p.band) {
        // allocation-free batch: the classes in turn on one stream, each in the whole band region
        const uint32_t cap = std::max<uint32_t>(p.n, 1);
        for (int c : {1, 2, 0}) {
            const uint32_t cnt = known_counts ? known_counts[c] : 0u;
            if (cnt == 0) continue;
            const uint64_t slot = band_slot(c, p.rows_cap, 0);
            const uint32_t waves = (uint32_t)std::min<uint64_t>((cnt + 63) / 64 + 3, p.band_doubles / slot) / 4 * 4;
            if (waves == 0) return fail(FP_ERR_INVALID_ARG, "allocation-free batch: a BAQ band does not fit the scratch");
            HIP_TRY(fpd_launch_baq_hmm(c, &b, &e->genome, &S.rs, S.baq_lists.as<uint32_t>() + (uint64_t)c * cap,
                                       S.baq_counts.as<uint32_t>() + c, p.band, slot, p.rows_cap, BAQ_CLASS_W[c],
                                       waves, (unsigned long long *)(counters + (c == 0 ? 10 : 8)), 1,
                                       e->opts.disable_baq_staging ? 0 : 1, 0,
                                       S.baq_counts.as<uint32_t>() + BAQ_CHUNK_COUNTERS + c, st));
        }
        if (mark) HIP_TRY(hipEventRecord(mark, st));
        HIP_TRY(fpd_launch_baq_post(&b, &S.rs, 1u, st));
        return FP_OK;
    }
    if (e->opts.profile_stages) {
        // per-kernel timing: the classes in turn on one stream, events around the 15-cell class
        for (int c = 1; c < 3; c++)
            if (!(known_counts && known_counts[c] == 0)) HIP_TRY(launch_class(c, st));
        if (!(known_counts && known_counts[0] == 0)) {
            HIP_TRY(hipEventRecord(e->ev[11], st));
            HIP_TRY(launch_class(0, st));
            HIP_TRY(hipEventRecord(e->ev[12], st));
            e->band15_timed = true;
        }
        if (mark) HIP_TRY(hipEventRecord(mark, st));
        HIP_TRY(fpd_launch_baq_post(&b, &S.rs, 1u, st));
        return FP_OK;
    }
    // The 15-cell class keeps every SIMD busy until its last round of chunks; the wider classes
    // run on the second stream beside it and fill that round (all three take their chunks from
    // counters, so waves that start late take fewer).  Eager launches skip empty classes.
    bool forked = false;
    for (int c = 1; c < 3; c++) {
        if (known_counts && known_counts[c] == 0) continue;
        if (!forked) {
            HIP_TRY(hipEventRecord(S.fork, st));
            HIP_TRY(hipStreamWaitEvent(e->side, S.fork, 0));
            forked = true;
        }
        HIP_TRY(launch_class(c, e->side));
    }
    if (!(known_counts && known_counts[0] == 0)) HIP_TRY(launch_class(0, st));
    if (forked) {
        HIP_TRY(hipEventRecord(S.join, e->side));
        HIP_TRY(hipStreamWaitEvent(st, S.join, 0));
    }
    if (mark) HIP_TRY(hipEventRecord(mark, st));
    HIP_TRY(fpd_launch_baq_post(&b, &S.rs, 1u, st));
    return FP_OK;
}

// stage 2 = 2a + 2b (covariates); marks: after the BAQ kernels, after cap/recode, after pair emission
fp_status launch_back(fp_engine *e, BatchState &S, const fpd::BatchDev &b, const Plan &p, int wpe, int ckpt,
                      int sort_bits, uint32_t *counters, hipStream_t st, hipEvent_t *marks,
                      const uint32_t *known_counts)
{
    fp_status rc = launch_baq_stage(e, S, b, p, wpe, ckpt, counters, st, marks ? marks[0] : nullptr, known_counts);
    if (rc != FP_OK) return rc;
    if (marks) HIP_TRY(hipEventRecord(marks[1], st));
    return launch_cov_stage(e, S, b, p, sort_bits, counters, st, marks ? marks[2] : nullptr);
}

// bands wider than 33 cells (rare: large indels), sized from the widest such band
fp_status launch_generic(fp_engine *e, BatchState &S, const fpd::BatchDev &b, const Plan &p, uint32_t count, uint32_t wgen,
                         uint32_t *counters, hipStream_t st)
{
    if (count == 0) return FP_OK;
    uint32_t waves = std::min<uint32_t>((count + 63) / 64, 256u);
    const uint64_t slot = band_slot(3, p.rows_cap, wgen);
    double *region;
    if (p.band) {
        waves = (uint32_t)std::min<uint64_t>(waves, p.band_doubles / slot);
        if (waves == 0) return fail(FP_ERR_INVALID_ARG, "allocation-free batch: a BAQ band does not fit the scratch");
        region = p.band;
    } else {
        HIP_TRY(e->scratch_gen.ensure(slot * waves * 8 + 8));
        region = e->scratch_gen.as<double>();
    }
    HIP_TRY(fpd_launch_baq_hmm(3, &b, &e->genome, &S.rs,
                               S.baq_lists.as<uint32_t>() + 3ull * std::max<uint32_t>(p.n, 1),
                               S.baq_counts.as<uint32_t>() + 3, region, slot, p.rows_cap, wgen,
                               waves, (unsigned long long *)(counters + 8), e->wpe, 0, 0, nullptr, st));
    return FP_OK;
}

// one batch, eagerly, with the host reading the class counts between the stages (the plain engine)
// na: an allocation-free batch, whose per-batch state, counters and BAQ band region are carved from
// the caller-independent scratch array (fp_engine_process_batch_device_noalloc); null: the engine's own
struct NoAllocPlan {
    uint32_t *counters;
    double *band;
    uint64_t band_doubles;
};

fp_status run_batch(fp_engine *e, const fp_batch_view *v, uint32_t max_read_len_hint,
                    const NoAllocPlan *na = nullptr)
{
    const uint32_t n = v->num_reads;
    BatchState &S = na ? e->na : e->ps;
    if (!na) {
        fp_status rc = ensure_state(e, S, n, v->num_read_bases, v->num_qualities);
        if (rc != FP_OK) return rc;
    }
    fp_status rc = FP_OK;
    e->last_num_reads = n;
    e->last = &S;
    e->last_num_bases = v->num_read_bases;
    e->last_num_quals = v->num_qualities;
    hipStream_t st = e->stream;
    fpd::BatchDev b = to_dev(v);
    uint32_t *counters = na ? na->counters : e->counters.as<uint32_t>();

    HIP_TRY(hipEventRecord(e->ev[0], st));
    HIP_TRY(fpd_fill_u32(counters, 16, 0u, st));
    HIP_TRY(hipEventRecord(e->ev[6], st));
    Plan p{n, v->num_qualities, nullptr, max_read_len_hint};   // rows known after the front stage
    rc = launch_front(e, S, b, p, counters, st);
    if (rc != FP_OK) return rc;
    HIP_TRY(hipEventRecord(e->ev[7], st));
    uint32_t *hc = e->host_counts;   // the batch's class sizes and counter words
    HIP_TRY(hipMemcpyAsync(hc, S.baq_counts.p, 4 * fpd::BW_WORDS, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    if (hc[fpd::BW_BAD_RG]) return fail(FP_ERR_INVALID_ARG, "read group id beyond the header's read groups");
    p.rows_cap = std::max<uint32_t>(std::max(max_read_len_hint, hc[fpd::BW_LONGEST]), 1);
    if (na) {
        if (hc[fpd::BW_LONGEST] > NOALLOC_MAX_READ_LENGTH)
            return fail(FP_ERR_INVALID_ARG, "allocation-free batch: a read is longer than " +
                                                std::to_string(NOALLOC_MAX_READ_LENGTH) + " bases");
        p.band = na->band;
        p.band_doubles = na->band_doubles;
        p.no_frac = true;
    } else {
        HIP_TRY(e->scratch.ensure(band_scratch_doubles(e->wpe, n, p.rows_cap) * 8 + 8));
    }
    const bool sorting = e->opts.aggregation == FP_AGG_SORT;
    if (sorting) {
        rc = ensure_pairs(e, 3ull * n * p.rows_cap);
        if (rc != FP_OK) return rc;
    }

    HIP_TRY(hipEventRecord(e->ev[8], st));
    rc = launch_generic(e, S, b, p, hc[3], hc[4], counters, st);
    if (rc != FP_OK) return rc;
    hipEvent_t marks[3] = {e->ev[9], e->ev[2], e->ev[10]};
    rc = launch_back(e, S, b, p, e->wpe, e->ckpt, e->sort_bits, counters, st, marks, hc);
    if (rc != FP_OK) return rc;
    HIP_TRY(hipEventRecord(e->ev[3], st));
    uint32_t *post = e->host_counts + 16;
    HIP_TRY(hipMemcpyAsync(post, counters, 64, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(post + 16, S.baq_counts.as<uint32_t>() + fpd::BW_OVERFLOW, 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    if (post[16]) return fail(FP_ERR_INTERNAL, "sort aggregation: observation pairs overflowed their buffer");

    e->stats.total_reads += n;
    e->stats.num_batches += 1;
    e->stats.filtered_reads += hc[fpd::BW_FILTERED];
    e->stats.baq_reads += hc[fpd::BW_BAQ];
    uint64_t cells, cells15;
    memcpy(&cells, post + 8, 8);
    memcpy(&cells15, post + 10, 8);
    e->stats.baq_cells += cells + cells15;
    e->stats.baq15_cells += cells15;
    collect_band15_time(e);
    const uint32_t hmm_launches = (hc[0] ? 1 : 0) + (hc[1] ? 1 : 0) + (hc[2] ? 1 : 0) + (hc[3] ? 1 : 0);
    e->stats.prep_kernel_ms += elapsed(e->ev[6], e->ev[7]);
    e->stats.prep_launches += 1;
    if (hmm_launches) e->stats.baq_hmm_kernel_ms += elapsed(e->ev[8], e->ev[9]);
    e->stats.baq_hmm_launches += hmm_launches;
    if (sorting) {
        e->stats.cov_kernel_ms += elapsed(e->ev[2], e->ev[10]);
        e->stats.covariates_sort_ms += elapsed(e->ev[10], e->ev[3]);
        rc = collect_sort_counts(e, e->sort_bits);
        if (rc != FP_OK) return rc;
    } else {
        e->stats.cov_kernel_ms += elapsed(e->ev[2], e->ev[3]);
    }
    e->stats.cov_launches += 1;
    e->stats.read_filter_ms += elapsed(e->ev[0], e->ev[7]);
    e->stats.baq_ms += elapsed(e->ev[7], e->ev[2]);
    e->stats.covariates_ms += elapsed(e->ev[2], e->ev[3]);
    e->postprocessed = false;
    return FP_OK;
}

// ---------------------------------------------------------------------------
// batched engine

// BAQ candidates of the autotuner: (waves per SIMD, rows per forward checkpoint).  One or two waves
// per SIMD of the every-row kernel, and the checkpointed kernel (one wave per SIMD) with 3 or 4 rows
// per checkpoint.
static const int TUNE_CANDIDATES[4][2] = {{1, 1}, {2, 1}, {1, 3}, {1, 4}};
static const int NUM_TUNE = 4;
static const int SORT_CANDIDATES[5] = {4, 5, 6, 7, 8};

fpd::BatchDev stage_dev(fp_engine *e, fp_engine::Stage &s)
{
    fp_batch_view v{};
    v.num_reads = e->cap_reads;
    v.chromosome = s.arr[0].as<uint16_t>();
    v.alignment_start = s.arr[1].as<uint32_t>();
    v.alignment_stop = s.arr[2].as<uint32_t>();
    v.mate_chromosome = s.arr[3].as<uint32_t>();
    v.mate_alignment_start = s.arr[4].as<uint32_t>();
    v.inferred_insert_size = s.arr[5].as<int32_t>();
    v.cigars = s.arr[6].as<uint32_t>();
    v.cigar_start = s.arr[7].as<uint32_t>();
    v.cigar_len = s.arr[8].as<uint32_t>();
    v.reads = s.arr[9].as<uint32_t>();
    v.read_start = s.arr[10].as<uint32_t>();
    v.read_len = s.arr[11].as<uint32_t>();
    v.qualities = s.arr[12].as<uint8_t>();
    v.qual_start = s.arr[13].as<uint32_t>();
    v.qual_len = s.arr[14].as<uint32_t>();
    v.flags = s.arr[15].as<uint16_t>();
    v.mapq = s.arr[16].as<uint8_t>();
    v.read_group = s.arr[17].as<uint32_t>();
    fpd::BatchDev b = to_dev(&v);
    b.live = s.live.as<uint32_t>();
    return b;
}

// capacity-sized buffers of the batched engine, allocated once (the graphs refer to them)
fp_status batched_alloc(fp_engine *e)
{
    if (e->staged_ready) return FP_OK;
    const uint64_t C = e->cap_reads;
    const uint64_t sz[18] = {2 * C, 4 * C, 4 * C, 4 * C, 4 * C, 4 * C, 4 * e->cap_cig, 4 * C, 4 * C,
                             e->cap_bases / 2, 4 * C, 4 * C, e->cap_quals, 4 * C, 4 * C, 2 * C, C, 4 * C};
    for (int si = 0; si < e->nstage; si++) {
        fp_engine::Stage &s = e->stg[si];
        for (int i = 0; i < 18; i++) HIP_TRY(s.arr[i].ensure(sz[i] + 64));
        HIP_TRY(s.live.ensure(16));
        if (!s.live_host) HIP_TRY(hipHostMalloc((void **)&s.live_host, 16, hipHostMallocDefault));
        if (!s.baq_done) HIP_TRY(hipEventCreateWithFlags(&s.baq_done, hipEventDisableTiming));
        if (!s.cov_done) HIP_TRY(hipEventCreateWithFlags(&s.cov_done, hipEventDisableTiming));
        if (!s.t0) HIP_TRY(hipEventCreate(&s.t0));
        if (!s.t1) HIP_TRY(hipEventCreate(&s.t1));
        if (!s.f0) HIP_TRY(hipEventCreate(&s.f0));
        if (!s.f1) HIP_TRY(hipEventCreate(&s.f1));
        if (!s.front_done) HIP_TRY(hipEventCreateWithFlags(&s.front_done, hipEventDisableTiming));
        if (!s.hc) {
            HIP_TRY(hipHostMalloc((void **)&s.hc, 80 * sizeof(uint32_t), hipHostMallocDefault));
            memset(s.hc, 0, 80 * sizeof(uint32_t));
        }
        // every buffer's appends and front stage run on the engine stream (profile_stages keeps all
        // of its kernels there, for clean per-kernel times)
        s.stream = e->stream;
        fp_status rc = ensure_state(e, s.bs, e->cap_reads, e->cap_bases, e->cap_quals,
                                    e->opts.debug_arrays != 0 || e->rows_cap > 256);
        if (rc != FP_OK) return rc;
        HIP_TRY(fpd_fill_u32(s.bs.baq_counts.as<uint32_t>(), 144, 0u, e->stream));   // read before the first batch
    }
    if (!e->bq) HIP_TRY(hipStreamCreateWithFlags(&e->bq, hipStreamNonBlocking));
    fp_status rc = FP_OK;
    HIP_TRY(e->scratch.ensure(band_scratch_doubles(std::max(3, e->wpe), e->cap_reads, e->rows_cap) * 8 + 8));
    if (e->opts.aggregation == FP_AGG_SORT) {
        rc = ensure_pairs(e, 3 * e->cap_quals);
        if (rc != FP_OK) return rc;
    }
    HIP_TRY(e->bcounters.ensure(64));
    HIP_TRY(fpd_fill_u32(e->bcounters.as<uint32_t>(), 16, 0u, e->stream));
    for (auto &ev : e->bev) if (!ev) HIP_TRY(hipEventCreate(&ev));
    HIP_TRY(hipStreamSynchronize(e->stream));
    e->staged_ready = true;
    e->epoch++;
    return FP_OK;
}

template <class F>
fp_status capture(hipStream_t st, hipGraphExec_t *out, F body)
{
    if (*out) { (void)hipGraphExecDestroy(*out); *out = nullptr; }
    HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    const fp_status rc = body();
    hipGraph_t g = nullptr;
    const hipError_t he = hipStreamEndCapture(st, &g);
    if (rc != FP_OK) { if (g) (void)hipGraphDestroy(g); return rc; }
    HIP_TRY(he);
    hipGraphExec_t ex = nullptr;
    const hipError_t hi = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HIP_TRY(hi);
    *out = ex;
    return FP_OK;
}

// the stage-2 time of a staging buffer's last batch, once its stream has passed it
void collect_back_time(fp_engine *e, fp_engine::Stage &s)
{
    if (!s.back_pending) return;
    (void)hipEventSynchronize(s.t1);
    e->stats.baq_ms += elapsed(s.t0, s.t1);
    s.back_pending = false;
}

// record one autotuner candidate's stage-2 time per read
void tune_record(fp_engine *e, int kind, int cand, double per_read)
{
    if (kind == 0) {
        e->tune_ms_per_read[cand] = per_read;
        bool all = true;
        for (int i = 0; i < NUM_TUNE; i++) all = all && e->tune_ms_per_read[i] > 0;
        if (all) {
            int best = 0;
            for (int i = 1; i < NUM_TUNE; i++) if (e->tune_ms_per_read[i] < e->tune_ms_per_read[best]) best = i;
            e->wpe = TUNE_CANDIDATES[best][0];
            e->ckpt = TUNE_CANDIDATES[best][1];
            e->tuned = true;
        }
    } else {
        e->sort_tune_ms[cand] = per_read;
        bool all = true;
        for (int i = 0; i < 5; i++) all = all && e->sort_tune_ms[i] > 0;
        if (all) {
            int best = 0;
            for (int i = 1; i < 5; i++) if (e->sort_tune_ms[i] < e->sort_tune_ms[best]) best = i;
            e->sort_bits = SORT_CANDIDATES[best];
            e->sort_tuned = true;
        }
    }
}

// finish a staging buffer's turn: the next buffer fills
static void stage_done(fp_engine *e, fp_engine::Stage &s)
{
    s.fill += 1;
    e->last_num_reads = s.n;
    e->last_num_bases = s.nbases;
    e->last_num_quals = s.nquals;
    e->last = &s.bs;
    s.n = 0;
    s.ncig = s.nbases = s.nquals = 0;
    e->cur = (e->cur + 1) % e->nstage;
    e->postprocessed = false;
}

// Run the current staging buffer (if it holds reads) and switch to the next one.  Each buffer has
// its own per-batch state, and each stage its own stream (engine stream: appends and front stage;
// bq: 15-cell and generic BAQ; side: wider BAQ classes, cap/recode and covariates), so one buffer's
// front stage and another's covariate stage run at the same time as a third's BAQ.  On each stream
// the buffers take turns in flush order, which is what the shared band scratch regions and the sort
// aggregation's shared pair buffers need.  The direct tables take every buffer's atomic adds in any
// order.
//
// The flush launches the front stage and its readback and returns; batched_back launches the rest
// once the readback has landed (batched_poll from later calls, or batched_drain when a call must
// wait).  Buffers go through batched_back in flush order, so the BAQ turns keep that order.
fp_status batched_back(fp_engine *e);

static double host_ms()
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// How many reads the current buffer needs before it runs without being full (0xffffffff: only when
// full).  Nothing queued behind the BAQ kernels now running (every flushed buffer's stages launched,
// and the BAQ stage of the last one begun: its start event t0 follows the previous buffer's BAQ
// kernels and its own front stage) means the BAQ stage is about to run dry; a buffer flushed then
// has its front stage done before those kernels end.  After a sync (the start of a stream) an eighth
// of a buffer is enough; after that, 1.5 times the last buffer's reads (about the ratio of the host
// link's batch rate to the device's), so that the buffers grow to full within a few turns and a full
// buffer waits while the device works (buffers that stayed small would each arrive as late as the
// last one, and a buffer refills only once its last batch has run).  A device left idle for 0.5 ms
// (a stream that arrives slower than the device runs it) again takes an eighth of a buffer.
static uint32_t batched_early_need(fp_engine *e)
{
    const uint32_t never = 0xffffffffu, eighth = e->cap_reads / 8;
    if (!e->pending.empty()) return never;
    if (e->last_flush_n == 0) return eighth;
    const hipEvent_t t0 = e->last_back >= 0 ? e->stg[e->last_back].t0 : nullptr;
    if (t0 && hipEventQuery(t0) == hipErrorNotReady) return never;
    bool idle = true;
    for (int si = 0; si < e->nstage && idle; si++) {
        const hipEvent_t c = e->stg[si].cov_done;
        idle = !(c && hipEventQuery(c) == hipErrorNotReady);
    }
    (void)hipGetLastError();
    if (!idle) {
        e->idle_since = -1;
    } else {
        const double now = host_ms();
        if (e->idle_since < 0) e->idle_since = now;
        if (now - e->idle_since > 0.5) return eighth;
    }
    return (uint32_t)std::min<uint64_t>(e->cap_reads, std::max<uint64_t>(eighth, 3ull * e->last_flush_n / 2));
}

fp_status batched_poll(fp_engine *e)
{
    while (!e->pending.empty()) {
        const hipError_t q = hipEventQuery(e->stg[e->pending.front()].front_done);
        if (q == hipErrorNotReady) return FP_OK;
        if (q != hipSuccess) { HIP_TRY(q); }
        fp_status rc = batched_back(e);
        if (rc != FP_OK) return rc;
    }
    return FP_OK;
}

// every flushed buffer's stages launched (waits for their readbacks); the first error is returned after
// all of them have been handled
fp_status batched_drain(fp_engine *e)
{
    fp_status first = FP_OK;
    std::string msg;
    while (!e->pending.empty()) {
        fp_status rc = batched_back(e);
        if (rc != FP_OK && first == FP_OK) { first = rc; msg = g_last_error; }
    }
    if (first != FP_OK) return fail(first, msg);
    return FP_OK;
}

fp_status batched_flush(fp_engine *e)
{
    fp_engine::Stage &s = e->stg[e->cur];
    if (s.n == 0) return FP_OK;
    fp_status rc = batched_alloc(e);
    if (rc != FP_OK) return rc;
    const bool prof = e->opts.profile_stages != 0;
    hipStream_t st = prof ? e->stream : s.stream;
    BatchState &S = s.bs;
    s.live_host[0] = s.n;
    s.live_host[1] = (uint32_t)s.nquals;
    s.live_host[2] = (uint32_t)(s.nquals >> 32);
    s.live_host[3] = 0;
    if (s.copies_pending) {
        // the copy stream is in order: its latest event covers every copy into this buffer
        HIP_TRY(hipStreamWaitEvent(st, e->copied, 0));
        s.copies_pending = false;
    }
    HIP_TRY(hipMemcpyAsync(s.live.p, s.live_host, 16, hipMemcpyHostToDevice, st));
    // the pair-overflow word of this buffer's previous batch, read before the front stage zeroes it
    HIP_TRY(hipMemcpyAsync(s.hc + 64, S.baq_counts.as<uint32_t>() + fpd::BW_OVERFLOW, 4, hipMemcpyDeviceToHost, st));

    const fpd::BatchDev b = stage_dev(e, s);
    Plan p{e->cap_reads, e->cap_quals, s.live.as<uint32_t>(), e->rows_cap};
    uint32_t *counters = e->bcounters.as<uint32_t>();
    if (s.g_epoch != e->epoch) {
        rc = capture(st, &s.gfront, [&] { return launch_front(e, S, b, p, counters, st); });
        if (rc != FP_OK) return rc;
        if (s.gbaq) { (void)hipGraphExecDestroy(s.gbaq); s.gbaq = nullptr; }
        if (s.gside) { (void)hipGraphExecDestroy(s.gside); s.gside = nullptr; }
        if (s.gcov) { (void)hipGraphExecDestroy(s.gcov); s.gcov = nullptr; }
        s.g_epoch = e->epoch;
    }
    HIP_TRY(hipEventRecord(s.f0, st));
    if (prof) {
        rc = launch_front(e, S, b, p, counters, st);
        if (rc != FP_OK) return rc;
    } else {
        HIP_TRY(hipGraphLaunch(s.gfront, st));
    }
    HIP_TRY(hipEventRecord(s.f1, st));
    // the band classes, the batch's read counts, longest read and error words: one small readback
    HIP_TRY(hipMemcpyAsync(s.hc, S.baq_counts.p, 4 * fpd::BW_WORDS, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipEventRecord(s.front_done, st));
    s.pending = true;
    s.pend_n = s.n;
    e->last_flush_n = s.n;
    e->idle_since = -1;
    e->pending.push_back(e->cur);
    stage_done(e, s);
    // the per-kernel profile runs every stage eagerly, in turn
    if (prof) return batched_drain(e);
    return FP_OK;
}

// the stages after the front of the oldest flushed buffer: waits for its readback, checks it, then
// launches the BAQ classes, cap/recode and the covariates
fp_status batched_back(fp_engine *e)
{
    const int si = e->pending.front();
    e->pending.pop_front();
    fp_engine::Stage &s = e->stg[si];
    s.pending = false;
    const bool prof = e->opts.profile_stages != 0;
    hipStream_t st = e->stream;
    BatchState &S = s.bs;
    HIP_TRY(hipEventSynchronize(s.front_done));
    collect_back_time(e, s);
    const uint32_t *hc = s.hc;
    e->stats.read_filter_ms += elapsed(s.f0, s.f1);
    e->stats.prep_kernel_ms += elapsed(s.f0, s.f1);
    e->stats.prep_launches += 1;
    // A failed check drops this buffer's reads before they touch the tables; the error is reported
    // once and the engine goes on with the next batch.
    auto drop = [&]() { e->stats.total_reads -= s.pend_n; };
    if (hc[64]) {
        drop();
        return fail(FP_ERR_INTERNAL, "sort aggregation: observation pairs overflowed their buffer");
    }
    if (hc[fpd::BW_BAD_RG]) {
        drop();
        return fail(FP_ERR_INVALID_ARG, "read group id beyond the header's read groups (staged batch dropped)");
    }
    if (hc[fpd::BW_LONGEST] > e->rows_cap) {
        drop();
        return fail(FP_ERR_INVALID_ARG, "a read is longer than max_read_length (" + std::to_string(e->rows_cap) +
                                            "; staged batch dropped)");
    }
    e->stats.filtered_reads += hc[fpd::BW_FILTERED];
    e->stats.baq_reads += hc[fpd::BW_BAQ];

    const fpd::BatchDev b = stage_dev(e, s);
    const Plan p{e->cap_reads, e->cap_quals, s.live.as<uint32_t>(), e->rows_cap};
    uint32_t *counters = e->bcounters.as<uint32_t>();
    fp_status rc = FP_OK;

    // autotuner: BAQ waves per SIMD first, then the radix digit width of the sort aggregation; each
    // candidate runs one staged batch eagerly, alone on the device, and is timed per read
    int wpe = e->wpe, ckpt = e->ckpt, bits = e->sort_bits;
    int cand = -1, kind = 0;
    if (!e->tuned) {
        for (int i = 0; i < NUM_TUNE; i++) if (e->tune_ms_per_read[i] == 0) { cand = i; break; }
        if (cand >= 0) { wpe = TUNE_CANDIDATES[cand][0]; ckpt = TUNE_CANDIDATES[cand][1]; }
    } else if (!e->sort_tuned) {
        for (int i = 0; i < 5; i++) if (e->sort_tune_ms[i] == 0) { cand = i; break; }
        if (cand >= 0) { bits = SORT_CANDIDATES[cand]; kind = 1; }
    }
    const bool sorting = e->opts.aggregation == FP_AGG_SORT;
    if (prof) {
        // eager stage 2 with an event after the BAQ kernels and after cap/recode
        HIP_TRY(hipEventRecord(e->ev[8], st));
        rc = launch_generic(e, S, b, p, hc[3], hc[4], counters, st);
        if (rc != FP_OK) return rc;
        hipEvent_t marks[3] = {e->ev[9], e->ev[2], e->ev[10]};
        rc = launch_back(e, S, b, p, wpe, ckpt, bits, counters, st, marks, hc);
        if (rc != FP_OK) return rc;
        HIP_TRY(hipEventRecord(e->ev[3], st));
        HIP_TRY(hipEventRecord(s.baq_done, st));
        HIP_TRY(hipEventRecord(s.cov_done, st));
        HIP_TRY(hipStreamSynchronize(st));
        collect_band15_time(e);
        const uint32_t launches = (hc[0] ? 1 : 0) + (hc[1] ? 1 : 0) + (hc[2] ? 1 : 0) + (hc[3] ? 1 : 0);
        e->stats.baq_hmm_kernel_ms += elapsed(e->ev[8], e->ev[9]);
        e->stats.baq_hmm_launches += launches;
        if (sorting) {
            e->stats.cov_kernel_ms += elapsed(e->ev[2], e->ev[10]);
            e->stats.covariates_sort_ms += elapsed(e->ev[10], e->ev[3]);
            rc = collect_sort_counts(e, bits);
            if (rc != FP_OK) return rc;
        } else {
            e->stats.cov_kernel_ms += elapsed(e->ev[2], e->ev[3]);
        }
        e->stats.cov_launches += 1;
        e->stats.baq_ms += elapsed(e->ev[8], e->ev[2]);
        e->stats.covariates_ms += elapsed(e->ev[2], e->ev[3]);
        return FP_OK;
    }
    // BAQ queue: this buffer's 15-cell and generic BAQ kernels follow its front stage and, in queue
    // order, the previous buffer's (they share the band scratch regions of their classes)
    hipStream_t bq = e->bq, sd = e->side;
    if (cand >= 0) {   // the candidate runs alone
        HIP_TRY(hipStreamSynchronize(e->stream));
        HIP_TRY(hipStreamSynchronize(bq));
        HIP_TRY(hipStreamSynchronize(sd));
    }
    HIP_TRY(hipStreamWaitEvent(bq, s.front_done, 0));
    HIP_TRY(hipEventRecord(s.t0, bq));
    e->last_back = si;
    rc = launch_generic(e, S, b, p, hc[3], hc[4], counters, bq);
    if (rc != FP_OK) return rc;
    if (cand < 0) {
        if (!s.gbaq || s.g_wpe != wpe || s.g_ckpt != ckpt || !s.gside || !s.gcov || s.g_bits != bits) {
            rc = capture(bq, &s.gbaq, [&]() -> fp_status {
                HIP_TRY(launch_band_class(e, S, b, p, wpe, ckpt, counters, 0, bq));
                return FP_OK;
            });
            if (rc != FP_OK) return rc;
            rc = capture(sd, &s.gside, [&]() -> fp_status {
                for (int c = 1; c < 3; c++) HIP_TRY(launch_band_class(e, S, b, p, wpe, ckpt, counters, c, sd));
                return FP_OK;
            });
            if (rc != FP_OK) return rc;
            rc = capture(sd, &s.gcov, [&]() -> fp_status {
                HIP_TRY(fpd_launch_baq_post(&b, &S.rs, 1u, sd));
                return launch_cov_stage(e, S, b, p, bits, counters, sd, nullptr);
            });
            if (rc != FP_OK) return rc;
            s.g_wpe = wpe;
            s.g_ckpt = ckpt;
            s.g_bits = bits;
        }
        HIP_TRY(hipGraphLaunch(s.gbaq, bq));
        HIP_TRY(hipEventRecord(s.baq_done, bq));
        // side queue: the wider classes beside the 15-cell class (their own scratch regions, in queue
        // order), then cap/recode and the covariates once every class is done; the covariate stages
        // of the sort aggregation share its pair buffers and keep queue order as well
        HIP_TRY(hipStreamWaitEvent(sd, s.front_done, 0));
        HIP_TRY(hipGraphLaunch(s.gside, sd));
        HIP_TRY(hipStreamWaitEvent(sd, s.baq_done, 0));
        HIP_TRY(hipGraphLaunch(s.gcov, sd));
        HIP_TRY(hipEventRecord(s.cov_done, sd));
        HIP_TRY(hipEventRecord(s.t1, sd));
        s.back_pending = true;
    } else {
        rc = launch_baq_stage(e, S, b, p, wpe, ckpt, counters, bq, nullptr, nullptr);   // timed candidate, eager
        if (rc != FP_OK) return rc;
        HIP_TRY(hipEventRecord(s.baq_done, bq));
        rc = launch_cov_stage(e, S, b, p, bits, counters, bq, nullptr);
        if (rc != FP_OK) return rc;
        HIP_TRY(hipEventRecord(s.cov_done, bq));
        HIP_TRY(hipEventRecord(s.t1, bq));
        HIP_TRY(hipEventSynchronize(s.t1));
        const double ms = elapsed(s.t0, s.t1);
        e->stats.baq_ms += ms;
        tune_record(e, kind, cand, ms / std::max<uint32_t>(s.pend_n, 1));
    }
    return FP_OK;
}

// wait for all staged work; collect the stage-2 times and the BAQ cell count
fp_status batched_sync(fp_engine *e)
{
    fp_status rc = batched_flush(e);
    if (rc != FP_OK) return rc;
    rc = batched_drain(e);
    if (rc != FP_OK) return rc;
    for (auto &s : e->stg)
        if (s.stream) HIP_TRY(hipStreamSynchronize(s.stream));
    if (e->bq) HIP_TRY(hipStreamSynchronize(e->bq));
    if (e->side) HIP_TRY(hipStreamSynchronize(e->side));
    HIP_TRY(hipStreamSynchronize(e->stream));
    e->last_flush_n = 0;   // the next batch starts a stream
    if (!e->staged_ready) return FP_OK;
    for (auto &s : e->stg) collect_back_time(e, s);
    uint32_t *hc = e->host_counts + 32;
    HIP_TRY(hipMemcpy(hc, e->bcounters.p, 64, hipMemcpyDeviceToHost));
    uint64_t cells, cells15;
    memcpy(&cells, hc + 8, 8);
    memcpy(&cells15, hc + 10, 8);
    e->stats.baq_cells = e->eager_cells + cells + cells15;
    e->stats.baq15_cells = e->eager_cells15 + cells15;
    if (e->opts.aggregation == FP_AGG_SORT) {
        // a staged batch's pair overflow: reported once, then cleared
        for (int si = 0; si < e->nstage; si++) {
            fp_engine::Stage &s = e->stg[si];
            uint32_t *w = s.bs.baq_counts.as<uint32_t>() + fpd::BW_OVERFLOW;
            uint32_t ov = 0;
            HIP_TRY(hipMemcpy(&ov, w, 4, hipMemcpyDeviceToHost));
            if (ov) {
                HIP_TRY(fpd_fill_u32(w, 1, 0u, e->stream));
                HIP_TRY(hipStreamSynchronize(e->stream));
                return fail(FP_ERR_INTERNAL, "sort aggregation: observation pairs overflowed their buffer");
            }
        }
    }
    return FP_OK;
}

fp_status plain_host_batch(fp_engine *e, const fp_batch_view *hb);

// Bytes of a batch array per workgroup of the append kernel.  Over the host link a few hundred waves
// keep the link busy (the kernel runs beside the other buffers' kernels, so it takes no more); from
// HBM more workgroups finish the copy sooner.
static constexpr uint64_t APPEND_HOST_BLOCK_BYTES = 256 << 10, APPEND_DEVICE_BLOCK_BYTES = 32 << 10;

// Host-side checks of a host batch: every read inside the batch arrays (and starting on a word of
// bases), no read longer than max_len, no read group at or beyond rg_limit (0xffffffff: read groups
// unchecked; the prep kernel filters reads without one, as the reference does).  One pass ORs the
// conditions over all reads (the loop vectorises); the failing read is searched for only when one
// fails, with the same messages in the same order as a per-read walk.
static fp_status check_host_reads(const fp_batch_view *b, uint32_t max_len, uint32_t rg_limit)
{
    const uint32_t n = b->num_reads;
    const uint32_t rgl = rg_limit == 0xffffffffu ? 0xffffffffu : rg_limit;
    const uint32_t rg_on = rg_limit != 0xffffffffu;
    uint32_t bad = 0;
    for (uint32_t i = 0; i < n; i++) {
        const uint64_t ce = (uint64_t)b->cigar_start[i] + b->cigar_len[i];
        const uint64_t re = (uint64_t)b->read_start[i] + ((b->read_len[i] + 7) & ~7u);
        const uint64_t qe = (uint64_t)b->qual_start[i] + b->qual_len[i];
        bad |= (uint32_t)(ce > b->num_cigar_ops) | (uint32_t)(re > b->num_read_bases) |
               (uint32_t)(qe > b->num_qualities) | (b->read_start[i] & 7) | (uint32_t)(b->read_len[i] > max_len) |
               (rg_on & (uint32_t)(b->read_group[i] >= rgl));
    }
    if (!bad) return FP_OK;
    for (uint32_t i = 0; i < n; i++) {
        if ((uint64_t)b->cigar_start[i] + b->cigar_len[i] > b->num_cigar_ops ||
            (uint64_t)b->read_start[i] + ((b->read_len[i] + 7) & ~7u) > b->num_read_bases ||
            (uint64_t)b->qual_start[i] + b->qual_len[i] > b->num_qualities || (b->read_start[i] & 7))
            return fail(FP_ERR_INVALID_ARG, "read " + std::to_string(i) + " lies outside the batch arrays");
        if (rg_on && b->read_group[i] >= rgl) return fail(FP_ERR_INVALID_ARG, "read group id beyond the job's read groups");
        if (b->read_len[i] > max_len)
            return fail(FP_ERR_INVALID_ARG, "a read is longer than max_read_length (" + std::to_string(max_len) + ")");
    }
    return FP_OK;
}

// Device-side addresses of a host batch's arrays when every one of them is pinned host memory (the
// device then reads them over the host link itself, in one launch); false when any array is
// pageable, which only the runtime's copies can move.
static bool pinned_sources(const void *const *src, const uint64_t *len, const void **dsrc)
{
    for (int i = 0; i < 18; i++) {
        dsrc[i] = nullptr;
        if (!len[i]) continue;
        hipPointerAttribute_t at;
        if (hipPointerGetAttributes(&at, src[i]) != hipSuccess) { (void)hipGetLastError(); return false; }
        if (at.type != hipMemoryTypeHost) return false;
        void *dp = nullptr;
        if (hipHostGetDevicePointer(&dp, const_cast<void *>(src[i]), 0) != hipSuccess || !dp) {
            (void)hipGetLastError();
            return false;
        }
        dsrc[i] = dp;
    }
    return true;
}

// append one batch (host or device arrays) to the current staging buffer; a serving job's reads move
// to its read-group slots [rg_off, rg_off + rg_limit)
// A host batch's checks (check_host_reads: read bounds, read length, read groups below
// host_check_rg) run while its copy is in flight; a failing batch is not staged.
fp_status batched_append(fp_engine *e, const fp_batch_view *v, bool on_device, uint32_t rg_off = 0,
                         uint32_t rg_limit = 0xffffffffu, uint32_t host_check_rg = 0xffffffffu)
{
    const uint32_t k = v->num_reads;
    if (k == 0) { e->stats.num_batches += 1; return FP_OK; }
    if (k > e->cap_reads || v->num_cigar_ops > e->cap_cig || v->num_read_bases > e->cap_bases ||
        v->num_qualities > e->cap_quals) {
        if (rg_off || rg_limit != 0xffffffffu)
            return fail(FP_ERR_INVALID_ARG, "a job's batch must fit one staging buffer (max_batch_reads)");
        // larger than a staging buffer: runs on its own, eagerly (it may grow shared buffers, so the
        // graphs are captured again afterwards)
        fp_status rc = batched_sync(e);
        if (rc != FP_OK) return rc;
        e->epoch++;
        const fp_stats before = e->stats;
        if (!on_device) {
            rc = check_host_reads(v, e->rows_cap, host_check_rg);
            if (rc != FP_OK) return rc;
        }
        rc = on_device ? run_batch(e, v, 0) : plain_host_batch(e, v);
        e->eager_cells += e->stats.baq_cells - before.baq_cells;
        e->eager_cells15 += e->stats.baq15_cells - before.baq15_cells;
        return rc;
    }
    e->stats.num_batches += 1;
    fp_engine::Stage *s = &e->stg[e->cur];
    if (s->n + k > e->cap_reads || s->ncig + v->num_cigar_ops > e->cap_cig ||
        s->nbases + v->num_read_bases > e->cap_bases || s->nquals + v->num_qualities > e->cap_quals) {
        fp_status rc = batched_flush(e);
        if (rc != FP_OK) return rc;
        s = &e->stg[e->cur];
    }
    // a buffer refills once its previous batch's stages are all launched (the copies wait for them)
    fp_status first_err = FP_OK;
    std::string first_msg;
    while (s->pending) {
        fp_status rc = batched_back(e);
        if (rc != FP_OK && first_err == FP_OK) { first_err = rc; first_msg = g_last_error; }
    }
    fp_status rc = batched_alloc(e);
    if (rc != FP_OK) return rc;
    // device appends run on the engine stream; a fill's first one waits until the buffer's last batch
    // has been consumed (its covariate stage)
    hipStream_t cs = e->stream;
    if (on_device && s->n == 0 && s->cov_done) HIP_TRY(hipStreamWaitEvent(cs, s->cov_done, 0));
    const hipMemcpyKind kind = on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    const void *src[18] = {v->chromosome, v->alignment_start, v->alignment_stop, v->mate_chromosome,
                           v->mate_alignment_start, v->inferred_insert_size, v->cigars, v->cigar_start,
                           v->cigar_len, v->reads, v->read_start, v->read_len, v->qualities, v->qual_start,
                           v->qual_len, v->flags, v->mapq, v->read_group};
    const uint64_t elem[18] = {2, 4, 4, 4, 4, 4, 0, 4, 4, 0, 4, 4, 0, 4, 4, 2, 1, 4};
    void *dst[18];
    uint64_t len[18];
    for (int i = 0; i < 18; i++) {
        uint64_t off;
        if (i == 6) { off = 4 * s->ncig; len[i] = 4 * v->num_cigar_ops; }
        else if (i == 9) { off = s->nbases / 2; len[i] = v->num_read_bases / 2; }
        else if (i == 12) { off = s->nquals; len[i] = v->num_qualities; }
        else { off = elem[i] * s->n; len[i] = elem[i] * k; }
        dst[i] = (char *)s->arr[i].p + off;
    }
    if (on_device) {
        // one launch copies and rebases the whole batch (a copy per array costs a launch each)
        HIP_TRY(fpd_append(src, dst, len, (uint32_t)s->ncig, (uint32_t)s->nbases, (uint32_t)s->nquals, rg_off,
                           rg_limit, APPEND_DEVICE_BLOCK_BYTES, cs));
    } else {
        // Host arrays, on the copy stream, which first waits for the buffer's previous batch (its
        // kernels read these arrays) when this is the fill's first batch; the host then waits for the
        // copy alone, never for other kernels on the buffer's stream.  Pinned arrays are read over the
        // host link by one copy-and-rebase kernel (18 DMA copies cost ~25 us each in the runtime:
        // 4.3e7 reads/s against 7.1e7 with the kernel); pageable ones take the runtime's copies and a
        // rebase kernel on the buffer's stream.
        if (!e->copy) {
            HIP_TRY(hipStreamCreateWithFlags(&e->copy, hipStreamNonBlocking));
            HIP_TRY(hipEventCreateWithFlags(&e->copied, hipEventDisableTiming));
        }
        if (s->n == 0 && s->cov_done) HIP_TRY(hipStreamWaitEvent(e->copy, s->cov_done, 0));
        const void *hsrc[18];
        if (pinned_sources(src, len, hsrc)) {
            HIP_TRY(fpd_append(hsrc, dst, len, (uint32_t)s->ncig, (uint32_t)s->nbases, (uint32_t)s->nquals, rg_off,
                               rg_limit, APPEND_HOST_BLOCK_BYTES, e->copy));
            HIP_TRY(hipEventRecord(e->copied, e->copy));
            s->copies_pending = true;   // the buffer's stream waits for them once, at the flush
            // the batch's checks while the copy runs: a failing batch is not staged (its bytes sit past
            // the buffer's reads and are overwritten by the next batch)
            rc = check_host_reads(v, e->rows_cap, host_check_rg);
            if (rc != FP_OK) {
                const std::string msg = g_last_error;
                HIP_TRY(hipEventSynchronize(e->copied));   // the host arrays are still being read
                return fail(rc, msg);
            }
        } else {
            rc = check_host_reads(v, e->rows_cap, host_check_rg);
            if (rc != FP_OK) return rc;
            for (int i = 0; i < 18; i++)
                if (len[i]) HIP_TRY(hipMemcpyAsync(dst[i], src[i], len[i], kind, e->copy));
            HIP_TRY(hipEventRecord(e->copied, e->copy));
            HIP_TRY(hipStreamWaitEvent(cs, e->copied, 0));
            HIP_TRY(fpd_rebase(k, s->arr[7].as<uint32_t>() + s->n, s->arr[10].as<uint32_t>() + s->n,
                               s->arr[13].as<uint32_t>() + s->n, s->arr[17].as<uint32_t>() + s->n, (uint32_t)s->ncig,
                               (uint32_t)s->nbases, (uint32_t)s->nquals, rg_off, rg_limit, cs));
        }
    }
    s->n += k;
    s->ncig += v->num_cigar_ops;
    s->nbases += v->num_read_bases;
    s->nquals += v->num_qualities;
    e->stats.total_reads += k;
    // launch the stages of flushed buffers whose readback has landed; an earlier buffer's failed check
    // is reported here (this call's batch is staged all the same)
    rc = batched_poll(e);
    if (first_err == FP_OK && rc != FP_OK) { first_err = rc; first_msg = g_last_error; }
    // When the BAQ stage is about to run dry, a buffer runs before it is full (batched_early_need):
    // its front stage then runs beside the BAQ kernels in flight and its BAQ kernels follow theirs.
    if (first_err == FP_OK && !e->opts.profile_stages && s->n >= e->cap_reads / 8 && s->n >= batched_early_need(e)) {
        rc = batched_flush(e);
        if (rc != FP_OK) { first_err = rc; first_msg = g_last_error; }
    }
    if (!on_device) HIP_TRY(hipEventSynchronize(e->copied));   // host arrays are borrowed for the call only
    if (first_err != FP_OK) return fail(first_err, first_msg);
    return FP_OK;
}

// host batch: copy to the device and run it eagerly (the plain engine)
fp_status plain_host_batch(fp_engine *e, const fp_batch_view *hb)
{
    const uint32_t n = hb->num_reads;
    uint32_t maxlen = 0;
    for (uint32_t i = 0; i < n; i++) maxlen = std::max(maxlen, std::max(hb->read_len[i], hb->qual_len[i]));
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
    for (int i = 0; i < 18; i++) HIP_TRY(e->in[i].ensure(sz[i]));
    const void *hsrc[18];
    uint64_t len[18];
    void *dst[18];
    for (int i = 0; i < 18; i++) { len[i] = copy_sz[i]; dst[i] = e->in[i].p; }
    if (pinned_sources(src, len, hsrc)) {
        // pinned host arrays: one copy launch reads all of them over the host link
        HIP_TRY(fpd_append(hsrc, dst, len, 0, 0, 0, 0, 0xffffffffu, APPEND_HOST_BLOCK_BYTES, st));
    } else {
        for (int i = 0; i < 18; i++)
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
    return run_batch(e, &dv, maxlen);
}

// compaction of one dense table into key-ordered rows on the device
// read groups [rg_base, rg_base + nrg) of table t: rows keyed relative to rg_base
fp_status compact(fp_engine *e, int t, uint32_t rg_base, uint32_t nrg, DevBuf &okey, DevBuf &ocount, DevBuf &olo,
                  DevBuf &ohi, uint64_t *nrows, hipStream_t st)
{
    const uint64_t off0 = (uint64_t)rg_base << (fpd::TABLE_BITS[t] - 8);
    const uint64_t n = (uint64_t)nrg << (fpd::TABLE_BITS[t] - 8);
    const uint64_t *count = e->tables.count[t] + off0;
    const uint64_t *parts[3] = {e->tables.part[t][0] + off0, e->tables.part[t][1] + off0,
                                e->tables.part[t][2] + off0};
    const uint64_t nb = (n + 255) / 256;
    DevBuf &bc = e->pw[0], &bo = e->pw[1];
    HIP_TRY(bc.ensure(4 * nb + 4));
    HIP_TRY(bo.ensure(8 * nb + 8));
    HIP_TRY(fpd_compact_count(count, n, bc.as<uint32_t>(), st));
    std::vector<uint32_t> h(nb);
    std::vector<uint64_t> off(nb);
    HIP_TRY(hipMemcpyAsync(h.data(), bc.p, 4 * nb, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    uint64_t total = 0;
    for (uint64_t i = 0; i < nb; i++) { off[i] = total; total += h[i]; }
    HIP_TRY(hipMemcpyAsync(bo.p, off.data(), 8 * nb, hipMemcpyHostToDevice, st));
    HIP_TRY(okey.ensure(4 * total + 4));
    HIP_TRY(ocount.ensure(8 * total + 8));
    HIP_TRY(olo.ensure(8 * total + 8));
    HIP_TRY(ohi.ensure(8 * total + 8));
    HIP_TRY(fpd_compact_scatter(count, parts, n, bo.as<uint64_t>(),
                                okey.as<uint32_t>(), ocount.as<uint64_t>(), olo.as<uint64_t>(), ohi.as<uint64_t>(),
                                st));
    HIP_TRY(hipStreamSynchronize(st));
    *nrows = total;
    return FP_OK;
}

// The estimated quality -10*log10(expected/obs) and the binomial coefficient of the 61-bin
// posterior (empirical_quality.cu:72-75,151-156,179-189), evaluated with the host libm that the
// reference's host build uses: the last bit of log10 and lgammaf differs between libm
// implementations, and either can move a posterior near-tie or a printed rounding boundary.
void host_terms(uint64_t obs, double mism, double expected, double log10e, double *est, double *coeff)
{
    *est = double(-10.0 * log10(expected / double(obs)));
    uint64_t nErr = uint64_t(mism + 0.5) + 1;
    uint64_t nObs = obs + 1 + 1;
    const uint32_t MAXOBS = 2147483647u - 1;
    if (nObs > MAXOBS) {
        const double fraction = double(MAXOBS) / double(nObs);
        nErr = (uint64_t)round(double(nErr) * fraction);
        nObs = MAXOBS;
    }
    auto lf = [&](uint64_t x) { return (double)lgammaf((float)(double)(x + 1)) * log10e; };
    *coeff = lf(nObs) - lf(nErr) - lf(nObs - nErr);
}

// postprocess of read groups [rg_base, rg_base + nrg) into rows (keys relative to rg_base), on stream st
fp_status postprocess_range(fp_engine *e, uint32_t rg_base, uint32_t nrg, TableRows *rows, hipStream_t st)
{
    static const int qshift[3] = {2, 12, 13};
    const double log10e = e->consts.log10e;
    for (int t = 0; t < 3; t++) {
        DevBuf &k = e->pw[2], &c = e->pw[3], &lo = e->pw[4], &hi = e->pw[5], &m = e->pw[6], &ex = e->pw[7],
               &es = e->pw[8], &cf = e->pw[9], &em = e->pw[10];
        uint64_t nr = 0;
        fp_status rc = compact(e, t, rg_base, nrg, k, c, lo, hi, &nr, st);
        if (rc != FP_OK) return rc;
        HIP_TRY(m.ensure(8 * nr + 8));
        HIP_TRY(ex.ensure(8 * nr + 8));
        HIP_TRY(es.ensure(8 * nr + 8));
        HIP_TRY(cf.ensure(8 * nr + 8));
        HIP_TRY(em.ensure(8 * nr + 8));
        HIP_TRY(fpd_launch_obs_values(nr, k.as<uint32_t>(), c.as<uint64_t>(), lo.as<uint64_t>(), hi.as<uint64_t>(),
                                      qshift[t], m.as<double>(), ex.as<double>(), st));
        TableRows &R = rows[t + 1];
        R.key.resize(nr); R.obs.resize(nr); R.lo.resize(nr); R.hi.resize(nr);
        R.mism.resize(nr); R.expected.resize(nr); R.estimated.resize(nr); R.empirical.resize(nr);
        std::vector<double> coeff(nr);
        if (nr) {
            HIP_TRY(hipMemcpyAsync(R.key.data(), k.p, 4 * nr, hipMemcpyDeviceToHost, st));
            HIP_TRY(hipMemcpyAsync(R.obs.data(), c.p, 8 * nr, hipMemcpyDeviceToHost, st));
            HIP_TRY(hipMemcpyAsync(R.lo.data(), lo.p, 8 * nr, hipMemcpyDeviceToHost, st));
            HIP_TRY(hipMemcpyAsync(R.hi.data(), hi.p, 8 * nr, hipMemcpyDeviceToHost, st));
            HIP_TRY(hipMemcpyAsync(R.mism.data(), m.p, 8 * nr, hipMemcpyDeviceToHost, st));
            HIP_TRY(hipMemcpyAsync(R.expected.data(), ex.p, 8 * nr, hipMemcpyDeviceToHost, st));
            HIP_TRY(hipStreamSynchronize(st));
            for (uint64_t i = 0; i < nr; i++) host_terms(R.obs[i], R.mism[i], R.expected[i], log10e, &R.estimated[i], &coeff[i]);
            HIP_TRY(hipMemcpyAsync(es.p, R.estimated.data(), 8 * nr, hipMemcpyHostToDevice, st));
            HIP_TRY(hipMemcpyAsync(cf.p, coeff.data(), 8 * nr, hipMemcpyHostToDevice, st));
            HIP_TRY(fpd_launch_empirical(nr, c.as<uint64_t>(), m.as<double>(), es.as<double>(), cf.as<double>(), 1u,
                                         em.as<double>(), st));
            HIP_TRY(hipMemcpyAsync(R.empirical.data(), em.p, 8 * nr, hipMemcpyDeviceToHost, st));
        }
        HIP_TRY(hipStreamSynchronize(st));
    }
    // read group table (RecalTable0)
    {
        const uint32_t ns = nrg * 4;
        DevBuf &valid = e->pw[11], &obs = e->pw[3], &m = e->pw[6], &ex = e->pw[7], &es = e->pw[8], &cf = e->pw[9],
               &em = e->pw[10];
        HIP_TRY(valid.ensure(4 * ns + 4));
        HIP_TRY(obs.ensure(8 * ns + 8));
        HIP_TRY(m.ensure(8 * ns + 8));
        HIP_TRY(ex.ensure(8 * ns + 8));
        HIP_TRY(es.ensure(8 * ns + 8));
        HIP_TRY(cf.ensure(8 * ns + 8));
        HIP_TRY(em.ensure(8 * ns + 8));
        const uint64_t q0 = (uint64_t)rg_base << 10;
        const uint64_t *qparts[3] = {e->tables.part[0][0] + q0, e->tables.part[0][1] + q0, e->tables.part[0][2] + q0};
        HIP_TRY(fpd_launch_rg_table(nrg, e->tables.count[0] + q0, qparts, valid.as<uint32_t>(), obs.as<uint64_t>(),
                                    m.as<double>(), ex.as<double>(), st));
        std::vector<uint32_t> hv(ns);
        std::vector<uint64_t> ho(ns);
        std::vector<double> hm(ns), hx(ns), hs(ns, 0.0), hc(ns, 0.0), he(ns);
        if (ns) {
            HIP_TRY(hipMemcpyAsync(hv.data(), valid.p, 4 * ns, hipMemcpyDeviceToHost, st));
            HIP_TRY(hipMemcpyAsync(ho.data(), obs.p, 8 * ns, hipMemcpyDeviceToHost, st));
            HIP_TRY(hipMemcpyAsync(hm.data(), m.p, 8 * ns, hipMemcpyDeviceToHost, st));
            HIP_TRY(hipMemcpyAsync(hx.data(), ex.p, 8 * ns, hipMemcpyDeviceToHost, st));
            HIP_TRY(hipStreamSynchronize(st));
            for (uint32_t s = 0; s < ns; s++)
                if (hv[s]) host_terms(ho[s], hm[s], hx[s], log10e, &hs[s], &hc[s]);
            HIP_TRY(hipMemcpyAsync(es.p, hs.data(), 8 * ns, hipMemcpyHostToDevice, st));
            HIP_TRY(hipMemcpyAsync(cf.p, hc.data(), 8 * ns, hipMemcpyHostToDevice, st));
            // slots without rows evaluate harmless zeros and are skipped below
            HIP_TRY(fpd_launch_empirical(ns, obs.as<uint64_t>(), m.as<double>(), es.as<double>(), cf.as<double>(), 0u,
                                         em.as<double>(), st));
            HIP_TRY(hipMemcpyAsync(he.data(), em.p, 8 * ns, hipMemcpyDeviceToHost, st));
        }
        HIP_TRY(hipStreamSynchronize(st));
        TableRows &R = rows[0];
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
    return FP_OK;
}

fp_status postprocess(fp_engine *e)
{
    fp_status rc = postprocess_range(e, 0, e->num_rg, e->rows, e->stream);
    if (rc != FP_OK) return rc;
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
    o->max_batch_reads = 0;
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
    case FP_ERR_FORMAT: return "malformed input file";
    default: return "internal error";
    }
}

fp_status fp_engine_set_read_group_names(fp_engine *e, uint32_t n, const char *const *names)
{
    fp_status rc = check_engine(e, true);
    if (rc != FP_OK) return rc;
    if (e->opts.serve_rg_slots) return fail(FP_ERR_STATE, "serving engines take read-group names per job");
    if (n > e->num_rg || (n && !names)) return fail(FP_ERR_INVALID_ARG, "more names than read groups");
    for (uint32_t i = 0; i < n; i++) e->rg_names[i] = names[i] ? names[i] : "";
    return FP_OK;
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
    if (o.aggregation != FP_AGG_DIRECT && o.aggregation != FP_AGG_SORT)
        return fail(FP_ERR_INVALID_ARG, "aggregation must be FP_AGG_DIRECT or FP_AGG_SORT");
    if (o.baq_checkpoint_rows > 5 || o.baq_checkpoint_rows == 2)
        return fail(FP_ERR_INVALID_ARG, "baq_checkpoint_rows must be 0 (tuned), 1, 3, 4 or 5");
    if (o.sort_digit_bits != 0 && (o.sort_digit_bits < 4 || o.sort_digit_bits > 8))
        return fail(FP_ERR_INVALID_ARG, "sort_digit_bits must be 0 (autotune) or 4..8");
    if (o.staging_buffers != 0 && (o.staging_buffers < 2 || o.staging_buffers > fp_engine::MAX_STAGE))
        return fail(FP_ERR_INVALID_ARG, "staging_buffers must be 0 (three) or 2 to 4");
    int ndev = 0;
    HIP_TRY(hipGetDeviceCount(&ndev));
    if (hip_device < 0 || hip_device >= ndev) return fail(FP_ERR_INVALID_ARG, "no such HIP device");
    HIP_TRY(hipSetDevice(hip_device));
    fp_engine *e = new fp_engine();
    e->device = hip_device;
    e->opts = o;
    e->wpe = o.baq_waves_per_simd >= 1 && o.baq_waves_per_simd <= 4 ? (int)o.baq_waves_per_simd : 1;
    e->ckpt = o.baq_checkpoint_rows >= 3 ? (int)o.baq_checkpoint_rows : 1;
    e->sort_bits = o.sort_digit_bits ? (int)o.sort_digit_bits : 8;
    e->budget_bytes = (uint64_t)o.serve_hbm_budget_mib << 20;
    e->nstage = o.staging_buffers ? (int)o.staging_buffers : 3;
    // the digit width is tuned on the batched engine's first staged batches unless it is given
    e->sort_tuned = o.sort_digit_bits != 0 || o.aggregation != FP_AGG_SORT || o.engine_kind != FP_ENGINE_BATCHED;
    memset(&e->stats, 0, sizeof(e->stats));
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, hip_device) == hipSuccess) e->name = prop.name;
    hipError_t he = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking);
    if (he != hipSuccess) { delete e; return fail(FP_ERR_HIP, hipGetErrorString(he)); }
    for (int i = 0; i < 16; i++) (void)hipEventCreate(&e->ev[i]);
    he = hipHostMalloc((void **)&e->host_counts, 128 * sizeof(uint32_t), hipHostMallocDefault);
    if (he == hipSuccess) he = hipStreamCreateWithFlags(&e->side, hipStreamNonBlocking);
    if (o.engine_kind == FP_ENGINE_BATCHED) {
        // the batched engine's four stage streams are created together, so that they take four
        // hardware queues of their own as far as the process has them free
        if (he == hipSuccess) he = hipStreamCreateWithFlags(&e->bq, hipStreamNonBlocking);
        if (he == hipSuccess) he = hipStreamCreateWithFlags(&e->copy, hipStreamNonBlocking);
        if (he == hipSuccess) he = hipEventCreateWithFlags(&e->copied, hipEventDisableTiming);
    }
    if (he == hipSuccess) he = hipEventCreateWithFlags(&e->na.fork, hipEventDisableTiming);
    if (he == hipSuccess) he = hipEventCreateWithFlags(&e->na.join, hipEventDisableTiming);
    if (he == hipSuccess) he = fpd_cov_prepare();
    if (he == hipSuccess) he = fpd_baq_prepare();
    if (he != hipSuccess) { fp_engine_destroy(e); return fail(FP_ERR_HIP, hipGetErrorString(he)); }
    if (o.engine_kind == FP_ENGINE_BATCHED) {
        e->batched = true;
        e->cap_reads = o.max_batch_reads ? std::max<uint32_t>(o.max_batch_reads, 64) : (1u << 20);
        e->rows_cap = o.max_read_length ? o.max_read_length : 500;
        e->cap_bases = (uint64_t)e->cap_reads * ((e->rows_cap + 7) & ~7u);
        e->cap_quals = (uint64_t)e->cap_reads * e->rows_cap;
        e->cap_cig = (uint64_t)e->cap_reads * 8;
        e->tuned = o.baq_waves_per_simd != 0 || o.baq_checkpoint_rows != 0;
    }
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
    e->rg_names.clear();
    if (e->opts.serve_rg_slots) {
        // serving: the tables hold the jobs' read-group slots; jobs bring their own names
        if (!e->batched || e->opts.serve_rg_slots > 256)
            return fail(FP_ERR_UNSUPPORTED, "serving needs the batched engine and 1..256 read-group slots");
        e->num_rg = e->opts.serve_rg_slots;
        for (uint32_t i = 0; i < e->num_rg; i++) e->rg_names.push_back("slot" + std::to_string(i));
        e->slot_owner.assign(e->num_rg, -1);
    } else {
        if (h->num_read_groups == 0 || h->num_read_groups > 256)
            return fail(FP_ERR_UNSUPPORTED, "read group count must be 1..256");
        for (uint32_t i = 0; i < h->num_read_groups; i++)
            e->rg_names.push_back(h->read_group_names && h->read_group_names[i] ? h->read_group_names[i] : "");
        e->num_rg = h->num_read_groups;
    }

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
    // the bucket index of the sites: the prep kernel's lower bound over max_end_left and upper bound
    // over the starts search one bucket's sites instead of the contig's
    std::vector<uint64_t> boff(v->num_sequences + 1, 0);
    std::vector<uint32_t> slb, ssb;
    for (uint32_t s = 0; s < v->num_sequences; s++) {
        const uint64_t a = v->feature_offset[s], b = v->feature_offset[s + 1];
        uint64_t span = s < r->num_sequences ? r->length[s] : 0;
        if (b > a) span = std::max<uint64_t>(span, std::max<uint64_t>(mel[b - 1], v->feature_start[b - 1]));
        const uint64_t nb = (span >> fpd::SITE_BUCKET_BITS) + 2;   // the last bucket: +infinity
        boff[s] = slb.size();
        uint64_t il = a, is = a;
        for (uint64_t j = 0; j + 1 < nb; j++) {
            const uint64_t t = j << fpd::SITE_BUCKET_BITS;
            while (il < b && mel[il] < t) il++;
            while (is < b && v->feature_start[is] < t) is++;
            slb.push_back((uint32_t)(il - a));
            ssb.push_back((uint32_t)(is - a));
        }
        slb.push_back((uint32_t)(b - a));
        ssb.push_back((uint32_t)(b - a));
    }
    boff[v->num_sequences] = slb.size();
    HIP_TRY(e->g_boff.ensure(8ull * boff.size()));
    HIP_TRY(e->g_slb.ensure(4ull * slb.size() + 4));
    HIP_TRY(e->g_ssb.ensure(4ull * ssb.size() + 4));
    HIP_TRY(hipMemcpy(e->g_boff.p, boff.data(), 8ull * boff.size(), hipMemcpyHostToDevice));
    if (!slb.empty()) {
        HIP_TRY(hipMemcpy(e->g_slb.p, slb.data(), 4ull * slb.size(), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(e->g_ssb.p, ssb.data(), 4ull * ssb.size(), hipMemcpyHostToDevice));
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
    g.bucket_offset = e->g_boff.as<uint64_t>();
    g.site_lb = e->g_slb.as<uint32_t>();
    g.site_sb = e->g_ssb.as<uint32_t>();

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
    // bounds of the host arrays (reads must lie inside the arrays) and, batched, the read lengths: on
    // the batched engine they run while the batch's copy is in flight.  Read groups are checked on the
    // device (a bad one drops its staging buffer), as before.
    if (e->batched) return batched_append(e, hb, false, 0, 0xffffffffu, 0xffffffffu);
    rc = check_host_reads(hb, 0xffffffffu, 0xffffffffu);
    if (rc != FP_OK) return rc;
    if (n == 0) { e->stats.num_batches++; return FP_OK; }
    return plain_host_batch(e, hb);
}

fp_status fp_engine_process_batch_device(fp_engine *e, const fp_batch_view *db)
{
    fp_status rc = check_engine(e, true);
    if (rc != FP_OK) return rc;
    rc = validate_view(db);
    if (rc != FP_OK) return rc;
    if (e->batched) return batched_append(e, db, true);
    if (db->num_reads == 0) { e->stats.num_batches++; return FP_OK; }
    return run_batch(e, db, 0);
}

fp_status fp_engine_synchronize(fp_engine *e)
{
    fp_status rc = check_engine(e, false);
    if (rc != FP_OK) return rc;
    if (e->batched && e->is_setup) return batched_sync(e);
    HIP_TRY(hipStreamSynchronize(e->stream));
    return FP_OK;
}

fp_status fp_engine_gather(fp_engine *dst, fp_engine *src)
{
    fp_status rc = check_engine(dst, true);
    if (rc != FP_OK) return rc;
    if (!src || !src->is_setup) return fail(FP_ERR_STATE, "source engine is not set up");
    if (src->batched) {
        rc = check_engine(src, true);
        if (rc == FP_OK) rc = batched_sync(src);
        if (rc != FP_OK) return rc;
        rc = check_engine(dst, true);
        if (rc != FP_OK) return rc;
    }
    if (dst->batched) {
        rc = batched_sync(dst);
        if (rc != FP_OK) return rc;
    }
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
    if (e->batched) {
        rc = batched_sync(e);
        if (rc != FP_OK) return rc;
    }
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
        if (e->batched) {
            rc = batched_sync(e);
            if (rc != FP_OK) return rc;
        }
        DevBuf k, c, lo, hi;
        uint64_t nr = 0;
        rc = compact(e, table_id - 1, 0, e->num_rg, k, c, lo, hi, &nr, e->stream);
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
    if (e->batched && e->is_setup) {
        fp_status rc = check_engine(e, true);
        if (rc == FP_OK) rc = batched_sync(e);
        if (rc != FP_OK) return rc;
    }
    *out = e->stats;
    out->baq_waves_per_simd = (uint64_t)e->wpe;
    out->baq_checkpoint_rows = (uint64_t)e->ckpt;
    out->sort_digit_bits = e->opts.aggregation == FP_AGG_SORT ? (uint64_t)e->sort_bits : 0;
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
        if (n) HIP_TRY(hipMemcpy(st.data(), e->last->rs.status, n, hipMemcpyDeviceToHost));
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
        if (n) HIP_TRY(hipMemcpy(st.data(), e->last->rs.status, n, hipMemcpyDeviceToHost));
        uint32_t *o = (uint32_t *)dst;
        uint64_t c = 0;
        for (uint64_t i = 0; i < n; i++) if (st[i] == fpd::RS_ACTIVE) o[c++] = (uint32_t)i;
        return FP_OK;
    }
    case FP_DBG_ERROR_FLAGS: case FP_DBG_ACTIVE_MASK: {
        std::vector<uint8_t> f(need);
        if (need) HIP_TRY(hipMemcpy(f.data(), e->last->rs.base_flags, need, hipMemcpyDeviceToHost));
        uint8_t *o = (uint8_t *)dst;
        for (uint64_t i = 0; i < need; i++)
            o[i] = item == FP_DBG_ERROR_FLAGS ? (f[i] & 7) : ((f[i] & fpd::BF_MASKED) ? 0 : 1);
        return FP_OK;
    }
    case FP_DBG_READ_WINDOW: src = e->last->rs.rwc; break;
    case FP_DBG_NOINS_WINDOW: src = e->last->rs.noins; break;
    case FP_DBG_REF_WINDOW: src = e->last->rs.refw; break;
    case FP_DBG_HMM_WINDOW: src = e->last->rs.hmmw; break;
    case FP_DBG_HQ_WINDOW: src = e->last->rs.hqw; break;
    case FP_DBG_NUM_ERRORS: src = e->last->rs.num_errors; break;
    case FP_DBG_BANDWIDTH: src = e->last->rs.bandwidth; break;
    case FP_DBG_BAQ: src = e->last->rs.baq; break;
    case FP_DBG_FRAC_ERRORS:
        if (!e->opts.debug_arrays || !e->last->rs.frac)
            return fail(FP_ERR_STATE, "fractional errors are kept only with debug_arrays = 1 (not for "
                                      "allocation-free batches)");
        src = e->last->rs.frac;
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
    // Reads still staged are dropped, not run: a reset also recovers an engine whose staged batch
    // failed a check.  Flushed buffers run (a failed check drops one, as always); work in flight is
    // waited for.
    if (e->batched) (void)batched_drain(e);
    for (auto &s : e->stg)
        if (s.stream) HIP_TRY(hipStreamSynchronize(s.stream));
    if (e->bq) HIP_TRY(hipStreamSynchronize(e->bq));
    if (e->side) HIP_TRY(hipStreamSynchronize(e->side));
    if (e->copy) HIP_TRY(hipStreamSynchronize(e->copy));
    HIP_TRY(hipStreamSynchronize(e->stream));
    if (e->batched) {
        for (int si = 0; si < e->nstage; si++) {
            fp_engine::Stage &s = e->stg[si];
            collect_back_time(e, s);
            s.n = 0;
            s.ncig = s.nbases = s.nquals = 0;
            s.fill += 1;
            if (e->staged_ready) HIP_TRY(fpd_fill_u32(s.bs.baq_counts.as<uint32_t>(), 144, 0u, e->stream));
        }
        if (e->staged_ready) HIP_TRY(fpd_fill_u32(e->bcounters.as<uint32_t>(), 16, 0u, e->stream));
    }
    for (int t = 0; t < 3; t++)
        for (int k = 0; k < 4; k++) HIP_TRY(hipMemsetAsync(e->tab[t][k].p, 0, e->tab_size[t] * 8, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    memset(&e->stats, 0, sizeof(e->stats));
    e->eager_cells = e->eager_cells15 = 0;
    e->postprocessed = false;
    for (auto &r : e->rows) r = TableRows();
    return FP_OK;
}

void fp_engine_destroy(fp_engine *e)
{
    if (!e) return;
    (void)hipSetDevice(e->device);
    for (auto &s : e->stg)
        if (s.stream) (void)hipStreamSynchronize(s.stream);
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    for (auto &s : e->stg)
        if (s.stream) (void)hipStreamSynchronize(s.stream);
    if (e->bq) (void)hipStreamSynchronize(e->bq);
    if (e->side) (void)hipStreamSynchronize(e->side);
    if (e->copy) (void)hipStreamSynchronize(e->copy);
    DevBuf *bufs[] = {&e->g_word_offset, &e->g_length, &e->g_words, &e->g_foff, &e->g_fstart, &e->g_fstop,
                      &e->g_mel, &e->g_boff, &e->g_slb, &e->g_ssb, &e->counters, &e->scratch, &e->tk, &e->tv, &e->shist, &e->stot, &e->sctr};
    for (DevBuf *b : bufs) b->release();
    for (int t = 0; t < 3; t++) { e->pk[t].release(); e->pv[t].release(); }
    e->ps.release();
    // the allocation-free state only views the scratch array: nothing of it is freed
    e->na.baq_counts.p = e->na.baq_lists.p = e->na.rg_buf.p = nullptr;
    e->na.release();
    e->scratch_gen.release();
    e->bcounters.release();
    for (auto &s : e->stg) {
        for (auto &b : s.arr) b.release();
        s.live.release();
        if (s.live_host) (void)hipHostFree(s.live_host);
        if (s.hc) (void)hipHostFree(s.hc);
        if (s.gfront) (void)hipGraphExecDestroy(s.gfront);
        if (s.gbaq) (void)hipGraphExecDestroy(s.gbaq);
        if (s.gside) (void)hipGraphExecDestroy(s.gside);
        if (s.gcov) (void)hipGraphExecDestroy(s.gcov);
        hipEvent_t evs[] = {s.baq_done, s.cov_done, s.t0, s.t1, s.f0, s.f1, s.front_done};
        for (hipEvent_t ev : evs) if (ev) (void)hipEventDestroy(ev);
        if (s.stream && s.stream != e->stream) (void)hipStreamDestroy(s.stream);
        s.bs.release();
    }
    for (auto &ev : e->bev) if (ev) (void)hipEventDestroy(ev);
    if (e->post) (void)hipStreamSynchronize(e->post);
    for (auto &b : e->pw) b.release();
    e->check.release();
    if (e->post) (void)hipStreamDestroy(e->post);
    if (e->host_counts) (void)hipHostFree(e->host_counts);
    for (auto &b : e->in) b.release();
    for (auto &t : e->tab) for (auto &b : t) b.release();
    for (int i = 0; i < 16; i++) if (e->ev[i]) (void)hipEventDestroy(e->ev[i]);
    if (e->copied) (void)hipEventDestroy(e->copied);
    if (e->copy) (void)hipStreamDestroy(e->copy);
    if (e->bq) (void)hipStreamDestroy(e->bq);
    if (e->side) (void)hipStreamDestroy(e->side);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
}

// --- allocation-free batches ------------------------------------------------

static std::mutex g_noalloc_mu;
static void *g_noalloc_ptr[64];

// the scratch array's device address on the current device, taken once per device
static fp_status noalloc_scratch(int dev, unsigned char **out)
{
    if (dev < 0 || dev >= 64) return fail(FP_ERR_UNSUPPORTED, "allocation-free batches: device index beyond 63");
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (!g_noalloc_ptr[dev]) {
        void *p = nullptr;
        const hipError_t he = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (he != hipSuccess || !p)
            return fail(FP_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(he));
        g_noalloc_ptr[dev] = p;
    }
    *out = (unsigned char *)g_noalloc_ptr[dev];
    return FP_OK;
}

fp_status fp_engine_process_batch_device_noalloc(fp_engine *e, const fp_batch_view *db)
{
    fp_status rc = check_engine(e, true);
    if (rc != FP_OK) return rc;
    rc = validate_view(db);
    if (rc != FP_OK) return rc;
    if (e->batched) return fail(FP_ERR_UNSUPPORTED, "allocation-free batches run on the plain engine");
    if (e->opts.aggregation == FP_AGG_SORT)
        return fail(FP_ERR_UNSUPPORTED, "allocation-free batches use the direct table aggregation");
    const uint32_t n = db->num_reads;
    if (n == 0) { e->stats.num_batches++; return FP_OK; }
    if (n > NOALLOC_MAX_READS)
        return fail(FP_ERR_INVALID_ARG, "allocation-free batch: more than " + std::to_string(NOALLOC_MAX_READS) +
                                            " reads");
    if (db->num_qualities > (uint64_t)n * NOALLOC_MAX_READ_LENGTH ||
        db->num_read_bases > (uint64_t)n * ((NOALLOC_MAX_READ_LENGTH + 7) & ~7u))
        return fail(FP_ERR_INVALID_ARG, "allocation-free batch: more bases than " +
                                            std::to_string(NOALLOC_MAX_READ_LENGTH) + " per read");
    unsigned char *base = nullptr;
    rc = noalloc_scratch(e->device, &base);
    if (rc != FP_OK) return rc;
    // carve the batch's state from the array, 256-byte aligned pieces
    uint64_t off = 0;
    auto take = [&](uint64_t bytes) -> void * {
        void *q = base + off;
        off += (bytes + 255) & ~255ull;
        return q;
    };
    const uint64_t nb = db->num_read_bases, nq = db->num_qualities;
    NoAllocPlan na{};
    na.counters = (uint32_t *)take(64);
    BatchState &S = e->na;
    S.baq_counts.p = take(4 * 400);
    S.baq_lists.p = take(16ull * n);
    S.rg_buf.p = take(4ull * (4ull * std::max<uint32_t>(e->num_rg, 1) + 2 + n));
    fpd::ReadStateDev &r = S.rs;
    r.status = (uint8_t *)take(n);
    r.rwc = (uint32_t *)take(4ull * n);
    r.noins = (uint32_t *)take(4ull * n);
    r.refw = (uint32_t *)take(4ull * n);
    r.num_errors = (uint16_t *)take(2ull * n);
    r.hmmw = (uint32_t *)take(4ull * n);
    r.bandwidth = (uint16_t *)take(2ull * n);
    r.hqw = (uint32_t *)take(4ull * n);
    r.base_flags = (uint8_t *)take(nb + 8);
    r.baq = (uint8_t *)take(nq + 8);
    r.baq_state = (uint32_t *)take(4 * (nq + 8));
    r.frac = nullptr;
    if (off >= NOALLOC_SCRATCH_BYTES) return fail(FP_ERR_INVALID_ARG, "allocation-free batch: larger than the scratch");
    na.band = (double *)(base + off);
    na.band_doubles = (NOALLOC_SCRATCH_BYTES - off) / 8;
    return run_batch(e, db, 0, &na);
}

// --- serving jobs -----------------------------------------------------------

// a job's HBM footprint: the resident tables of its read groups (count and three limb sums per
// entry of the three direct-indexed tables) plus a reserve for its postprocess rows
static constexpr uint64_t TABLE_BYTES_PER_RG = 32ull * ((1ull << 10) + (1ull << 20) + (1ull << 21));
static constexpr uint64_t POST_RESERVE_PER_RG = 8ull << 20;

static fp_status post_stream(fp_engine *e, hipStream_t *out)
{
    if (!e->post) HIP_TRY(hipStreamCreateWithFlags(&e->post, hipStreamNonBlocking));
    *out = e->post;
    return FP_OK;
}

fp_status fp_job_open(fp_engine *e, uint32_t num_read_groups, const char *const *read_group_names, fp_job **out)
{
    if (!out) return fail(FP_ERR_INVALID_ARG, "null output pointer");
    *out = nullptr;
    fp_status rc = check_engine(e, true);
    if (rc != FP_OK) return rc;
    if (e->slot_owner.empty()) return fail(FP_ERR_STATE, "not a serving engine (serve_rg_slots = 0)");
    const uint32_t n = num_read_groups;
    if (n == 0 || n > e->num_rg) return fail(FP_ERR_INVALID_ARG, "a job needs 1..serve_rg_slots read groups");
    const uint64_t bytes = n * (TABLE_BYTES_PER_RG + POST_RESERVE_PER_RG);
    if (e->budget_bytes && bytes > e->budget_bytes)
        return fail(FP_ERR_INVALID_ARG, "the job's footprint exceeds the whole HBM budget");
    if (e->budget_bytes && e->bytes_in_use + bytes > e->budget_bytes)
        return fail(FP_ERR_BUDGET, "no room for the job now (HBM budget: " + std::to_string(e->bytes_in_use) + " of " +
                                       std::to_string(e->budget_bytes) + " bytes in use)");
    // first fit of n contiguous free slots: the job's keys form one range per table
    uint32_t run = 0, base = 0;
    bool found = false;
    for (uint32_t i = 0; i < e->num_rg; i++) {
        if (e->slot_owner[i] >= 0) { run = 0; continue; }
        if (run == 0) base = i;
        if (++run == n) { found = true; break; }
    }
    if (!found) return fail(FP_ERR_BUDGET, "no room for the job's tables now (read-group slots in use)");
    fp_job *j = new fp_job();
    j->e = e;
    j->id = e->next_job_id++;
    j->base = base;
    j->nrg = n;
    j->bytes = bytes;
    e->bytes_in_use += bytes;
    for (uint32_t i = 0; i < n; i++)
        j->names.push_back(read_group_names && read_group_names[i] ? read_group_names[i] : "");
    for (uint32_t i = 0; i < n; i++) e->slot_owner[base + i] = j->id;
    *out = j;
    return FP_OK;
}

static fp_status job_submit(fp_job *j, const fp_batch_view *b, bool on_device)
{
    if (!j || !j->e) return fail(FP_ERR_INVALID_ARG, "null job");
    if (j->closed) return fail(FP_ERR_STATE, "the job is closed");
    fp_engine *e = j->e;
    fp_status rc = check_engine(e, true);
    if (rc != FP_OK) return rc;
    rc = validate_view(b);
    if (rc != FP_OK) return rc;
    // a job's batch is checked before it joins a staging buffer shared with other jobs, so a bad batch
    // fails its own submit and never another job's
    if (!on_device) {
        // host batch: checked by batched_append while its copy is in flight (read group included)
    } else if (b->num_reads) {
        hipStream_t ps;
        rc = post_stream(e, &ps);
        if (rc != FP_OK) return rc;
        HIP_TRY(e->check.ensure(16));
        HIP_TRY(fpd_fill_u32(e->check.as<uint32_t>(), 2, 0u, ps));
        const fpd::BatchDev db = to_dev(b);
        HIP_TRY(fpd_check_batch(&db, j->nrg, e->check.as<uint32_t>(), ps));
        uint32_t *h = e->host_counts + 96;
        HIP_TRY(hipMemcpyAsync(h, e->check.p, 8, hipMemcpyDeviceToHost, ps));
        HIP_TRY(hipStreamSynchronize(ps));
        if (h[1]) return fail(FP_ERR_INVALID_ARG, "read group id beyond the job's read groups");
        if (h[0] > e->rows_cap)
            return fail(FP_ERR_INVALID_ARG, "a read is longer than max_read_length (" + std::to_string(e->rows_cap) + ")");
    }
    rc = batched_append(e, b, on_device, j->base, j->nrg, j->nrg);
    if (rc != FP_OK) return rc;
    j->fill[e->cur] = e->stg[e->cur].fill;
    j->reads += b->num_reads;
    j->batches += 1;
    return FP_OK;
}

fp_status fp_job_submit(fp_job *j, const fp_batch_view *b) { return job_submit(j, b, false); }
fp_status fp_job_submit_device(fp_job *j, const fp_batch_view *b) { return job_submit(j, b, true); }

fp_status fp_job_close(fp_job *j)
{
    if (!j || !j->e) return fail(FP_ERR_INVALID_ARG, "null job");
    if (j->closed) return FP_OK;
    fp_engine *e = j->e;
    fp_status rc = check_engine(e, true);
    if (rc != FP_OK) return rc;
    auto t0 = std::chrono::steady_clock::now();
    // the job's last reads may still be staging: run that buffer now (with whatever else it holds)
    if (j->fill[e->cur] == e->stg[e->cur].fill && e->stg[e->cur].n > 0) {
        rc = batched_flush(e);
        if (rc != FP_OK) return rc;
    }
    // every flushed buffer's stages launched, so the events below cover the job's reads
    rc = batched_drain(e);
    if (rc != FP_OK) return rc;
    // The job's postprocess waits for the covariate stage of the buffers that held its reads (each
    // buffer's stream is in order, so its latest event covers every earlier fill), not for the whole
    // engine: other jobs' batches keep running meanwhile.
    hipStream_t ps;
    rc = post_stream(e, &ps);
    if (rc != FP_OK) return rc;
    for (int s = 0; s < e->nstage; s++)
        if (j->fill[s] && e->stg[s].cov_done) HIP_TRY(hipStreamWaitEvent(ps, e->stg[s].cov_done, 0));
    rc = postprocess_range(e, j->base, j->nrg, j->rows, ps);
    if (rc != FP_OK) return rc;
    j->post_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    j->closed = true;
    return FP_OK;
}

fp_status fp_job_write_report(fp_job *j, const char *path)
{
    if (!j || !j->e || !path) return fail(FP_ERR_INVALID_ARG, "null argument");
    if (!j->closed) return fail(FP_ERR_STATE, "close the job first");
    fpr::ReportTables rt;
    for (int t = 0; t < 4; t++) {
        const TableRows &R = j->rows[t];
        fpr::ReportRows &o = rt.table[t];
        o.key = R.key; o.obs = R.obs; o.mism = R.mism; o.estimated = R.estimated; o.empirical = R.empirical;
    }
    std::string text = fpr::format_report(rt, j->names, j->e->opts.disable_output_rounding != 0);
    FILE *fp = fopen(path, "wb");
    if (!fp) return fail(FP_ERR_IO, std::string("cannot open ") + path);
    const size_t w = fwrite(text.data(), 1, text.size(), fp);
    fclose(fp);
    if (w != text.size()) return fail(FP_ERR_IO, "short write");
    return FP_OK;
}

fp_status fp_job_get_table(fp_job *j, int table_id, fp_table_out *out)
{
    if (!j || !j->e || !out || table_id < 0 || table_id > 3) return fail(FP_ERR_INVALID_ARG, "bad table request");
    if (!j->closed) return fail(FP_ERR_STATE, "close the job first");
    const TableRows &R = j->rows[table_id];
    const size_t n = R.size();
    out->count = n;
    for (size_t i = 0; i < n; i++) {
        if (out->keys) out->keys[i] = R.key[i];
        if (out->observations) out->observations[i] = R.obs[i];
        if (out->mismatches) out->mismatches[i] = R.mism[i];
        if (out->mismatches_fixed_lo) out->mismatches_fixed_lo[i] = R.lo[i];
        if (out->mismatches_fixed_hi) out->mismatches_fixed_hi[i] = R.hi[i];
        if (out->expected_errors) out->expected_errors[i] = R.expected[i];
        if (out->estimated_quality) out->estimated_quality[i] = R.estimated[i];
        if (out->empirical_quality) out->empirical_quality[i] = R.empirical[i];
    }
    return FP_OK;
}

fp_status fp_job_stats(fp_job *j, fp_stats *out)
{
    if (!j || !out) return fail(FP_ERR_INVALID_ARG, "null argument");
    memset(out, 0, sizeof(*out));
    out->total_reads = j->reads;
    out->num_batches = j->batches;
    out->postprocessing_ms = j->post_ms;
    return FP_OK;
}

void fp_job_release(fp_job *j)
{
    if (!j) return;
    fp_engine *e = j->e;
    if (e && check_engine(e, true) == FP_OK) {
        if (!j->closed) (void)batched_sync(e);   // nothing of the job may still be in flight
        hipStream_t ps = nullptr;
        if (post_stream(e, &ps) == FP_OK) {
            // a closed job's reads are through their covariate stages (its close waited for them)
            for (int t = 0; t < 3; t++) {
                const uint64_t off0 = (uint64_t)j->base << (fpd::TABLE_BITS[t] - 8);
                const uint64_t n = (uint64_t)j->nrg << (fpd::TABLE_BITS[t] - 8);
                for (int k = 0; k < 4; k++)
                    (void)hipMemsetAsync(e->tab[t][k].as<uint64_t>() + off0, 0, 8 * n, ps);
            }
            (void)hipStreamSynchronize(ps);
        }
        for (uint32_t i = 0; i < j->nrg; i++)
            if (e->slot_owner[j->base + i] == j->id) e->slot_owner[j->base + i] = -1;
        e->bytes_in_use -= std::min(e->bytes_in_use, j->bytes);
    }
    delete j;
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
