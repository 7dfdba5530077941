/*
 * firepony_amd — MI355X-native BQSR+BAQ engine, C ABI.
 *
 * This header is the drop-in boundary for firepony's per-device pipeline
 * (device/pipeline.h:44-66, struct firepony_pipeline) and the per-batch entry
 * points beneath it (device/pipeline.cu:56 firepony_process_batch,
 * device/pipeline.cu:206 firepony_postprocess).  Only plain C types cross it:
 * opaque handles, integer status codes, pointers and sizes.
 *
 * Data layouts mirror the reference's host batch exactly
 * (alignment_data.h:125-173 alignment_batch_storage), so a host
 * alignment_batch_host can be viewed without copying:
 *   - bases: 4-bit IUPAC16 codes (A=1 C=2 G=4 T=8 N=15), 8 per uint32 word,
 *     least significant nibble first; every read starts on a word boundary
 *     (read length padded to a multiple of 8, loader/alignments.cu:367-372);
 *   - cigar ops: uint32 with len in bits 0..23 and op in bits 24..27
 *     (types.h:57-86 cigar_op{len:24, op:4}); op codes M=0 I=1 D=2 N=3 S=4
 *     H=5 P=6 '='=7 X=8;
 *   - flags: the reference's AlignmentFlags bit values (alignment_data.h:42-68).
 *
 * Threading: one engine per host thread; an engine is not re-entrant.
 * fp_last_error() is thread-local.
 */
#ifndef FIREPONY_AMD_H
#define FIREPONY_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FP_ABI_VERSION 2   /* bumped whenever fp_options or a view changes layout */

typedef enum fp_status {
    FP_OK = 0,
    FP_ERR_INVALID_ARG = 1,   /* null pointer, bad size, malformed view           */
    FP_ERR_HIP = 2,           /* a HIP runtime call failed (message in fp_last_error) */
    FP_ERR_OUT_OF_MEMORY = 3, /* device allocation failed                          */
    FP_ERR_STATE = 4,         /* call out of order (e.g. process before setup)     */
    FP_ERR_BUDGET = 5,        /* serving: job refused by the HBM admission budget   */
    FP_ERR_UNSUPPORTED = 6,   /* option or input outside what the engine supports  */
    FP_ERR_IO = 7,            /* report file could not be written                  */
    FP_ERR_INTERNAL = 8,
    FP_ERR_FORMAT = 9         /* input file malformed (loaders, firepony_amd_io.h)  */
} fp_status;

/* engine flavours (both produce bit-identical tables) */
typedef enum fp_engine_kind {
    FP_ENGINE_PLAIN = 0,    /* each batch runs as it arrives: one launch per stage,
                               with host readbacks between stages                   */
    FP_ENGINE_BATCHED = 1   /* batches are appended to large device staging buffers
                               (copy stream, double-buffered); a full buffer runs as
                               two captured hipGraph replays with persistent BAQ
                               kernels; the BAQ variant is autotuned               */
} fp_engine_kind;

typedef struct fp_options {
    uint32_t abi_version;            /* must be FP_ABI_VERSION                        */
    uint32_t engine_kind;            /* fp_engine_kind                                */
    uint32_t max_batch_reads;        /* batched engine: reads per staging buffer
                                        (0 = 1,048,576); the plain engine sizes from
                                        each batch (the reference uses 20,000,
                                        firepony.cu:148)                              */
    uint32_t max_batch_bases;        /* capacity hint for padded bases; 0 = derived   */
    uint32_t frac_error_f64;         /* 0: per-base fractional errors narrowed to float
                                        before accumulation, as the reference does
                                        (device/covariates.cu:91-101); 1: kept in fp64 */
    uint32_t disable_output_rounding;/* report floats as %.64f (table_formatter.cu:102) */
    uint32_t baq_waves_per_simd;     /* BAQ HMM register budget: 1 to 4 waves per SIMD;
                                        0 = chosen by the engine's autotuner            */
    uint32_t sort_digit_bits;        /* FP_AGG_SORT radix digit width 4..8; 0 = tuned by
                                        the batched engine (8 on the plain engine)      */
    uint32_t max_read_length;        /* upper bound of read lengths; lets batches run with
                                        no host round trip; 0 = measured per batch      */
    uint32_t aggregation;            /* fp_aggregation                                 */
    uint32_t debug_arrays;           /* 1: also keep the per-base fractional errors for
                                        fp_engine_debug_fetch (the tables never need
                                        them); 0: leave them out                       */
    uint32_t profile_stages;         /* batched engine: run the stages eagerly with an
                                        event between kernels, so fp_stats carries
                                        per-kernel times (no graph replay)             */
    uint32_t serve_rg_slots;         /* serving: read-group table slots shared by the
                                        open jobs (HBM budget = slots x ~100 MB of
                                        resident tables); 0 = not a serving engine    */
    uint32_t disable_baq_staging;    /* 1: BAQ band kernels read their per-read inputs
                                        from HBM on every row instead of staging them
                                        in LDS once per read (0 = stage when the batch's
                                        longest read fits the LDS budget)              */
    uint32_t baq_checkpoint_rows;    /* 15-cell BAQ band kernel: 1 = every forward row's
                                        state spilled to HBM for the backward pass; 3, 4
                                        or 5 = the full state of every K-th row only, the
                                        rows between recomputed into registers segment by
                                        segment (same results, 48/K instead of 32 bytes of
                                        forward-row traffic per band cell, a second
                                        forward pass); 0 = chosen by the autotuner on the
                                        batched engine (1 otherwise).  Used with LDS staging
                                        at one wave per SIMD.                            */
    uint32_t serve_hbm_budget_mib;   /* serving: HBM admission budget in MiB for the open jobs'
                                        resident tables and postprocess reserve; 0 = the
                                        read-group slots alone bound admission          */
    uint32_t staging_buffers;        /* batched engine: 2 to 4 staging buffers (0 = 3).
                                        A buffer refills once its previous batch has
                                        run; with three or more, host batches copy in
                                        while the other buffers' kernels run          */
} fp_options;

/* how covariate observations reach the resident tables (identical results) */
typedef enum fp_aggregation {
    FP_AGG_DIRECT = 0,   /* LDS-privatised direct-indexed tables (default)            */
    FP_AGG_SORT = 1      /* key/value pairs -> LDS-tiled radix sort -> segmented reduce */
} fp_aggregation;

/* Fill *opts with defaults. */
void fp_options_init(fp_options *opts);

/* --- boundary views (host or device pointers, borrowed for the call) ------ */

/* alignment_header_host (alignment_data.h:33-44): contig lengths and read-group
 * names; read-group id i prints as names[i] (RG name = PU if present else ID,
 * loader/alignments.cu:180-187). */
typedef struct fp_header_view {
    uint32_t num_chromosomes;
    const uint32_t *chromosome_lengths;
    uint32_t num_read_groups;
    const char *const *read_group_names;
} fp_header_view;

/* sequence_database_storage (sequence_database.h:57-71): contig c holds
 * length[c] bases, 4-bit packed, starting at words[word_offset[c]]. */
typedef struct fp_reference_view {
    uint32_t num_sequences;
    const uint64_t *word_offset;
    const uint32_t *length;
    const uint32_t *words;
    uint64_t total_words;
} fp_reference_view;

/* variant_database_storage (variant_database.h:41-67): features of contig c
 * are entries [feature_offset[c], feature_offset[c+1]) of the three arrays,
 * sorted by start; feature_stop is inclusive (loader/variants.cu:120-122).
 * max_end_point_left may be NULL, in which case it is computed as the
 * inclusive max-scan of feature_stop per contig (loader/variants.cu:125-134). */
typedef struct fp_variants_view {
    uint32_t num_sequences;
    const uint64_t *feature_offset;   /* num_sequences + 1 entries */
    const uint32_t *feature_start;
    const uint32_t *feature_stop;
    const uint32_t *max_end_point_left;
    uint64_t num_features;
} fp_variants_view;

/* alignment_batch_storage (alignment_data.h:125-173), field for field. */
typedef struct fp_batch_view {
    uint32_t num_reads;
    const uint16_t *chromosome;
    const uint32_t *alignment_start;       /* 0-based                          */
    const uint32_t *alignment_stop;        /* bam_endpos - 1                    */
    const uint32_t *mate_chromosome;
    const uint32_t *mate_alignment_start;
    const int32_t  *inferred_insert_size;
    const uint32_t *cigars;                /* cigar_op words                   */
    uint64_t num_cigar_ops;
    const uint32_t *cigar_start;
    const uint32_t *cigar_len;
    const uint32_t *reads;                 /* packed 4-bit bases               */
    uint64_t num_read_bases;               /* padded base count (multiple of 8) */
    const uint32_t *read_start;
    const uint32_t *read_len;
    const uint8_t  *qualities;
    uint64_t num_qualities;
    const uint32_t *qual_start;
    const uint32_t *qual_len;
    const uint16_t *flags;
    const uint8_t  *mapq;
    const uint32_t *read_group;
} fp_batch_view;

/* --- results ------------------------------------------------------------- */

typedef enum fp_table_id {
    FP_TABLE_READ_GROUP = 0,   /* RecalTable0 (device/read_group_table.cu:62-93)  */
    FP_TABLE_QUALITY = 1,      /* RecalTable1 (covariates/packer_quality_score.h) */
    FP_TABLE_CYCLE = 2,        /* RecalTable2 cycle rows                          */
    FP_TABLE_CONTEXT = 3       /* RecalTable2 context rows                        */
} fp_table_id;

/* Two-call pattern: call with all array pointers NULL to learn *count, then
 * again with arrays of at least that many entries.  Any array may be NULL.
 * Keys use the reference's packed layouts (LSB first):
 *   quality: Event(2) | Qual(8) | RG(8)
 *   cycle:   Event(2) | Cycle(10, sign in bit 0) | Qual(8) | RG(8)
 *   context: Event(2) | Context(11) | Qual(8) | RG(8)
 *   read group: quality layout with the Qual field cleared.
 * mismatches_fixed_{lo,hi} hold the exact mismatch sum as an unsigned 128-bit
 * integer in units of 2^-80 (observation tables only). Before
 * fp_engine_postprocess only keys, observations and mismatch fields are set. */
typedef struct fp_table_out {
    uint64_t count;
    uint32_t *keys;
    uint64_t *observations;
    double   *mismatches;
    uint64_t *mismatches_fixed_lo;
    uint64_t *mismatches_fixed_hi;
    double   *expected_errors;
    double   *estimated_quality;
    double   *empirical_quality;
} fp_table_out;

/* pipeline_statistics (device/firepony_context.h:53-126), times in ms */
typedef struct fp_stats {
    uint64_t total_reads;
    uint64_t filtered_reads;
    uint64_t baq_reads;
    uint64_t num_batches;
    double io_ms;
    double read_filter_ms;
    double cigar_expansion_ms;
    double bp_filter_ms;
    double snp_filter_ms;
    double baq_ms;
    double fractional_error_ms;
    double covariates_ms;
    double covariates_sort_ms;
    double postprocessing_ms;
    double output_ms;
    /* per-kernel device time (HIP events on the engine stream) and launches */
    double prep_kernel_ms;
    double baq_hmm_kernel_ms;
    double cov_kernel_ms;
    uint64_t prep_launches;
    uint64_t baq_hmm_launches;
    uint64_t cov_launches;
    uint64_t baq_cells;        /* band cells evaluated by the BAQ HMM (algorithmic work) */
    uint64_t baq_waves_per_simd; /* BAQ band-kernel variant in use (autotuned or set)    */
    uint64_t sort_digit_bits;    /* FP_AGG_SORT radix digit width in use; 0 = direct    */
    uint64_t sort_slots;         /* FP_AGG_SORT, profile_stages or plain engine: pair    */
    uint64_t sort_pairs;         /*   slots emitted, real pairs sorted, and radix       */
    uint64_t sort_passes;        /*   passes run (summed over tables and batches)       */
    uint64_t baq_checkpoint_rows; /* rows per forward checkpoint of the 15-cell BAQ kernel */
    double baq15_kernel_ms;      /* profile_stages: device time of the 15-cell band kernel alone
                                    (HIP events around each of its launches)                   */
    uint64_t baq15_launches;     /* profile_stages: its launches                              */
    uint64_t baq15_cells;        /* band cells of the 15-cell class (its algorithmic work)     */
} fp_stats;

/* Per-batch intermediates of the last processed batch, for stage-by-stage
 * parity checks.  Sizes: per read = num_reads entries, per base = padded
 * base count, per quality = num_qualities. */
typedef enum fp_debug_item {
    FP_DBG_ACTIVE_READS = 0,     /* uint32 list, count = fp_engine_debug_size     */
    FP_DBG_READ_WINDOW = 1,      /* uint16 x2 per read (read_window_clipped)       */
    FP_DBG_NOINS_WINDOW = 2,     /* uint16 x2 per read                             */
    FP_DBG_REF_WINDOW = 3,       /* uint16 x2 per read (reference_window_clipped)  */
    FP_DBG_NUM_ERRORS = 4,       /* uint16 per read                                */
    FP_DBG_ERROR_FLAGS = 5,      /* uint8 per base: bit0 snp, bit1 ins, bit2 del   */
    FP_DBG_ACTIVE_MASK = 6,      /* uint8 per base: 1 active, 0 masked             */
    FP_DBG_HMM_WINDOW = 7,       /* int16 x2 per read                              */
    FP_DBG_BANDWIDTH = 8,        /* uint16 per read                                */
    FP_DBG_BAQ = 9,              /* uint8 per quality (BQ tag values, 0xFF unset)  */
    FP_DBG_FRAC_ERRORS = 10,     /* double x3 per quality (snp, ins, del)          */
    FP_DBG_HQ_WINDOW = 11        /* uint16 x2 per read (high quality window)       */
} fp_debug_item;

/* --- engine (firepony_pipeline) ---------------------------------------- */

typedef struct fp_engine fp_engine;

/* firepony_pipeline::create (device/pipeline_interface.cu:246-265) */
fp_status fp_engine_create(int hip_device, const fp_options *opts, fp_engine **out);
/* firepony_pipeline::setup (device/pipeline_interface.cu:105-125): uploads
 * the contig lengths, reference and known sites (all host pointers). */
fp_status fp_engine_setup(fp_engine *e, const fp_header_view *header,
                          const fp_reference_view *reference,
                          const fp_variants_view *variants);
/* worker loop body (device/pipeline_interface.cu:195-224): H2D copy of a host
 * batch followed by firepony_process_batch (device/pipeline.cu:56-148).  The
 * host arrays are borrowed for the call: it returns once the batch is in device
 * memory.  On a batched engine the batch's checks (read length, read group) run
 * with its staging buffer; a failing buffer's reads are dropped and the error is
 * returned by the call that sees it (this or a later process, synchronize,
 * postprocess or job close), which still stages its own batch. */
fp_status fp_engine_process_batch(fp_engine *e, const fp_batch_view *host_batch);
/* firepony_process_batch on a batch already resident in device memory (all
 * view pointers are device pointers on this engine's device).  Stream-ordered:
 * returns after enqueueing; fp_engine_synchronize waits. */
fp_status fp_engine_process_batch_device(fp_engine *e, const fp_batch_view *device_batch);
/* No counterpart in the reference.  fp_engine_process_batch_device for callers that cannot allocate
 * device memory (code under stream capture, tight serving loops, harnesses that forbid hipMalloc):
 * it calls no hipMalloc, hipFree or allocating library routine.  Plain engines with the direct
 * aggregation; at most 65,536 reads of at most 256 bases per batch (larger batches are refused
 * with FP_ERR_INVALID_ARG).  The batch's read state, counters and BAQ band rows live in one
 * 256 MiB file-scope device array per GPU, held until the process ends and shared by every call
 * on that GPU: calls that overlap on one GPU race; each call clears what must start at zero, and
 * the last call's intermediates stay there for fp_engine_debug_fetch (fractional errors are not
 * kept).  Same tables as fp_engine_process_batch_device; returns after the batch has run. */
fp_status fp_engine_process_batch_device_noalloc(fp_engine *e, const fp_batch_view *device_batch);
fp_status fp_engine_synchronize(fp_engine *e);
/* firepony_pipeline::gather_intermediates (device/pipeline_interface.cu:52-58,
 * 137-160): merge src's observation tables into dst's. */
fp_status fp_engine_gather(fp_engine *dst, fp_engine *src);
/* firepony_postprocess without the text output (device/pipeline.cu:206-228) */
fp_status fp_engine_postprocess(fp_engine *e);
/* output half of firepony_postprocess: GATKReport v1.1 text
 * (device/pipeline.cu:150-204, read_group_table.cu:122-143, covariates.cu:285-311) */
fp_status fp_engine_write_report(fp_engine *e, const char *path);
fp_status fp_engine_get_table(fp_engine *e, int table_id, fp_table_out *out);
/* alignment_header_host::read_groups_db (alignment_data.h:51-54): the
 * reference assigns read-group ids while it loads reads and prints their names
 * at the end; this replaces the names given to fp_engine_setup (n <= the count
 * given there; ids >= n keep theirs).  Not for serving engines. */
fp_status fp_engine_set_read_group_names(fp_engine *e, uint32_t n, const char *const *names);
/* firepony_pipeline::statistics */
fp_status fp_engine_stats(fp_engine *e, fp_stats *out);
/* firepony_pipeline::get_name / get_total_memory */
fp_status fp_engine_get_name(fp_engine *e, char *buf, size_t buf_len);
fp_status fp_engine_get_total_memory(fp_engine *e, uint64_t *bytes);
fp_status fp_engine_debug_size(fp_engine *e, int item, uint64_t *bytes);
fp_status fp_engine_debug_fetch(fp_engine *e, int item, void *dst, uint64_t bytes);
/* clears the accumulated tables and statistics; keeps setup state */
fp_status fp_engine_reset(fp_engine *e);
void fp_engine_destroy(fp_engine *e);

/* --- serving: recalibration jobs sharing one batched engine --------------
 * No counterpart in the reference, which runs one job per process
 * (firepony.cu:225-423).  A serving engine (engine_kind FP_ENGINE_BATCHED,
 * serve_rg_slots > 0) keeps each open job's covariate tables resident in its
 * own read-group slots.  Reads of all open jobs share the staging buffers and
 * graph replays; a job's tables are exact and its own, so its result is the
 * same alone or in any mix.  fp_engine_setup's header gives the contigs only. */
typedef struct fp_job fp_job;

/* admits a job with its read groups (names as in fp_header_view) if its tables
 * fit the slot budget now; FP_ERR_BUDGET otherwise (retry after a release) */
fp_status fp_job_open(fp_engine *e, uint32_t num_read_groups, const char *const *read_group_names,
                      fp_job **out);
/* one batch of the job (read_group ids local to the job); host or device arrays */
fp_status fp_job_submit(fp_job *j, const fp_batch_view *host_batch);
fp_status fp_job_submit_device(fp_job *j, const fp_batch_view *device_batch);
/* all of the job's reads are in: runs what is staged, then postprocesses the
 * job's tables (firepony_postprocess for this job only) */
fp_status fp_job_close(fp_job *j);
fp_status fp_job_write_report(fp_job *j, const char *path);
fp_status fp_job_get_table(fp_job *j, int table_id, fp_table_out *out);
/* reads submitted and postprocess time of the job */
fp_status fp_job_stats(fp_job *j, fp_stats *out);
/* clears the job's tables and returns its slots to the budget */
void fp_job_release(fp_job *j);

const char *fp_last_error(void);
const char *fp_status_string(fp_status s);

#ifdef __cplusplus
}
#endif

#endif /* FIREPONY_AMD_H */
