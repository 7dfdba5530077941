/*
 * fp_oracle — CPU restatement of firepony's `system == host` pipeline.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * or the timed CPU baseline.  The product (firepony_amd) never links it.
 *
 * It follows the reference's device/ stage functors one for one, in the
 * reference's stage order, with the reference's integer types and operation
 * order (compiled with -ffp-contract=off and glibc libm, like the reference's
 * host build, gcc.cmake:7-10).  Where the reference reads uninitialised or
 * out-of-range memory the oracle defines the result (see fp_oracle.cpp,
 * "defined behaviour" notes); the engine implements the same definitions.
 */
#ifndef FP_ORACLE_H
#define FP_ORACLE_H

#include "../include/firepony_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fpo_ctx fpo_ctx;

enum {
    FPO_ACCUM_EXACT = 0,      /* fixed-point Q48.80 mismatch sums (order independent) */
    FPO_ACCUM_REF_FLOAT = 1   /* float fold in stable key order (reference-compatible) */
};

fpo_ctx *fpo_create(int frac_error_f64, int accumulate_mode, int disable_output_rounding);
int fpo_setup(fpo_ctx *c, const fp_header_view *h, const fp_reference_view *r,
              const fp_variants_view *v);
int fpo_process_batch(fpo_ctx *c, const fp_batch_view *b);
int fpo_postprocess(fpo_ctx *c);
int fpo_get_table(fpo_ctx *c, int table_id, fp_table_out *out);
int fpo_write_report(fpo_ctx *c, const char *path);
int fpo_stats(fpo_ctx *c, fp_stats *out);
/* adds src's observation tables and read counters into dst (exact accumulation only): the
 * multi-device concat + reduce of gather_intermediates (pipeline_interface.cu:52-58,137-160) */
int fpo_gather(fpo_ctx *dst, fpo_ctx *src);
/* read-group names by id as the reference's read_groups_db holds them when the report is
 * written (alignment_data.h:51-54): replaces the first n names given to fpo_setup */
int fpo_set_read_group_names(fpo_ctx *c, uint32_t n, const char *const *names);
int fpo_debug_size(fpo_ctx *c, int item, uint64_t *bytes);
int fpo_debug_fetch(fpo_ctx *c, int item, void *dst, uint64_t bytes);
/* BAQ MAP quality of the last batch as the real value before truncation,
 * -4.343*log(1-max)+.499, per quality position (NaN where unset). */
int fpo_debug_baq_real(fpo_ctx *c, double *dst, uint64_t n);
void fpo_destroy(fpo_ctx *c);

/* standalone pieces, exposed for pinning tests */
double fpo_round_n(double val, int n);
int fpo_baq_map_quality(double max_over_sum);
double fpo_empirical_quality(uint64_t observations, double mismatches, double expected_errors,
                             int need_rounding, double *estimated_out);
void fpo_hmm_glocal(const uint8_t *ref, int ref_len, const uint8_t *query, const uint8_t *qual,
                    int query_len, int bandwidth, uint8_t *out_qual, uint32_t *out_state);
void fpo_fractional_errors(const uint8_t *baq_tags, const uint8_t *error_bits, int len,
                           double *out);

#ifdef __cplusplus
}
#endif

#endif
