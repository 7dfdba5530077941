/*
 * Seeded synthetic BQSR workloads: reference genome, known sites and read
 * batches in the boundary layout of include/firepony_amd.h.
 *
 * The generator replaces the reference's htslib loaders (loader/alignments.cu,
 * loader/reference.cu, loader/variants.cu), keeping their field semantics:
 * 0-based alignment start, alignment stop = last aligned base, 8-bp padded
 * 4-bit reads, FASTA non-ACGT -> N, VCF record -> [pos, pos + rlen - 1].
 *
 * Every read is a pure function of (spec.seed, read index), so batches can be
 * produced in any slicing, in parallel, and still be identical.
 */
#ifndef FP_SYNTH_H
#define FP_SYNTH_H

#include "../../include/firepony_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fpsyn_genome_spec {
    uint64_t seed;
    uint32_t num_contigs;
    uint64_t contig_length[32];
    uint32_t num_sites;          /* known sites, uniform over the genome          */
    double n_run_fraction;       /* fraction of the genome covered by N runs      */
    double multi_base_site_frac; /* fraction of sites spanning 2..6 bases         */
} fpsyn_genome_spec;

typedef struct fpsyn_read_spec {
    uint64_t seed;
    uint32_t num_lengths;
    uint32_t lengths[8];          /* read lengths, drawn uniformly               */
    uint32_t num_read_groups;
    double rg_mismatch[32];       /* substitution rate per read group            */
    double indel_read_frac;       /* reads carrying one indel                    */
    uint32_t indel_min, indel_max;
    double clip_read_frac;        /* reads carrying a soft clip on one end        */
    uint32_t clip_min, clip_max;
    double q2_tail_frac;          /* reads whose last 1..q2_tail_max bases are Q2 */
    uint32_t q2_tail_max;
    double q2_base_frac;          /* extra isolated Q2 bases                      */
    uint32_t paired;              /* emit mate fields / pair flags               */
    double insert_mean, insert_sd;
    double edge_case_frac;        /* reads drawn from the edge-case menu          */
} fpsyn_read_spec;

typedef struct fpsyn_genome fpsyn_genome;
typedef struct fpsyn_batch fpsyn_batch;

fpsyn_genome *fpsyn_genome_create(const fpsyn_genome_spec *spec);
void fpsyn_genome_destroy(fpsyn_genome *g);
/* views stay valid for the genome's lifetime */
void fpsyn_genome_reference(const fpsyn_genome *g, fp_reference_view *out);
void fpsyn_genome_variants(const fpsyn_genome *g, fp_variants_view *out);
/* header: contig lengths + read group names "rg0".."rgN-1" */
void fpsyn_genome_header(fpsyn_genome *g, uint32_t num_read_groups, fp_header_view *out);

/* reads [first_read, first_read + num_reads) of the workload, threads >= 1 */
fpsyn_batch *fpsyn_batch_create(const fpsyn_genome *g, const fpsyn_read_spec *spec,
                                uint64_t first_read, uint32_t num_reads, uint32_t threads);
void fpsyn_batch_view(const fpsyn_batch *b, fp_batch_view *out);
void fpsyn_batch_destroy(fpsyn_batch *b);

/* the spec of one of the named workloads (0..4, see synth.cpp) */
int fpsyn_named_spec(int config, fpsyn_genome_spec *gs, fpsyn_read_spec *rs, uint64_t *num_reads);

#ifdef __cplusplus
}
#endif

#endif
