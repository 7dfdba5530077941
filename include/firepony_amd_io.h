/*
 * firepony_amd input loaders — host C ABI (no device code).
 *
 * The reference's loader/ half: a FASTA reference, a VCF of known sites and a
 * SAM or BAM alignment file, turned into the views that fp_engine_setup and
 * fp_engine_process_batch take (firepony_amd.h).  Library:
 * libfirepony_amd_io.so (g++, zlib); it does not load the HIP runtime.
 *
 * Files may be plain or gzip/BGZF-compressed (BAM always is).  Every handle
 * is single-threaded; different handles may be used from different threads.
 * Views returned by the getters point into loader-owned storage and stay
 * valid until the handle is freed.  A batch view stays valid until the
 * second fp_alignments_next_batch call after it (the loader alternates two
 * buffers, so one batch can be processed while the next loads) or
 * fp_alignments_close.
 *
 * Behaviour follows the reference where it is defined; where the reference
 * reads out of bounds or silently stops, the loaders return FP_ERR_FORMAT with
 * the file and line (or record) in fp_io_last_error().
 */
#ifndef FIREPONY_AMD_IO_H
#define FIREPONY_AMD_IO_H

#include "firepony_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* --- reference (loader/reference.cu) ------------------------------------ */

typedef struct fp_sequence_db fp_sequence_db;

/* reference_file_handle::open + load_next_sequence (loader/reference.cu:110-140,
 * 222-326).  Sequences are read in file order.  A sequence's name is its
 * header line after '>' up to the first space; a..t are upper-cased and every
 * other byte of a sequence line becomes N.  Without a "<path>.fai" index,
 * loading stops at the first sequence with no bases (that sequence stays
 * defined, with length 0), as the reference's whole-file load does; with an
 * index, only the sequences the index names are loaded, as on-demand loading
 * would make them available. */
fp_status fp_reference_load(const char *fasta_path, fp_sequence_db **out);
/* sequence_database_storage (sequence_database.h:57-71) of the loaded contigs */
fp_status fp_reference_get_view(const fp_sequence_db *db, fp_reference_view *out);
/* string_database::lookup (string_database.cu:33-45): contig id of a name, or
 * UINT32_MAX.  Names are keyed by string_database::hash (string_database.cu:83-94). */
uint32_t fp_reference_lookup(const fp_sequence_db *db, const char *name);
/* name of contig id (NUL-terminated, truncated to buf_len - 1) */
fp_status fp_reference_name(const fp_sequence_db *db, uint32_t id, char *buf, size_t buf_len);
void fp_reference_free(fp_sequence_db *db);

/* --- known sites (loader/variants.cu) ------------------------------------ */

typedef struct fp_variant_db fp_variant_db;

/* load_vcf (loader/variants.cu:44-136): one feature per VCF record,
 * [POS-1, POS-1+rlen-1] with rlen the REF length (INFO END=, when present,
 * sets the last position to END-1), appended per contig in file order;
 * records on contigs absent from db are skipped.  max_end_point_left is the
 * inclusive max-scan of feature_stop per contig (loader/variants.cu:125-134). */
fp_status fp_variants_load(const char *vcf_path, const fp_sequence_db *db, fp_variant_db **out);
fp_status fp_variants_get_view(const fp_variant_db *v, fp_variants_view *out);
/* records read, and records skipped because their contig is not in the reference */
fp_status fp_variants_counts(const fp_variant_db *v, uint64_t *records, uint64_t *skipped);
void fp_variants_free(fp_variant_db *v);

/* --- alignments (loader/alignments.cu) ----------------------------------- */

typedef struct fp_alignment_file fp_alignment_file;

/* alignment_file::init (loader/alignments.cu:124-198): reads the header and
 * builds the read-group ID -> name map of its @RG records (name = PU if
 * present, else ID; loader/alignments.cu:160-187).  db must outlive the file. */
fp_status fp_alignments_open(const char *path, const fp_sequence_db *db, fp_alignment_file **out);
/* alignment_file::next_batch (loader/alignments.cu:238-407): up to batch_size
 * records into loader-owned storage (one of two alternating buffers).
 * out->num_reads == 0 at end of file.
 * chromosome / mate_chromosome are contig ids of db; read-group ids are
 * assigned in order of first appearance (loader/alignments.cu:396-401). */
fp_status fp_alignments_next_batch(fp_alignment_file *f, uint32_t batch_size, fp_batch_view *out);
/* alignment_header_host (alignment_data.h:45-54): contig lengths of the file's
 * header, and the read-group names by id assigned so far.  max_read_groups is
 * the number of distinct names among the header's @RG records, an upper bound
 * of the ids any read can get: fp_engine_setup sizes the tables from it, and
 * fp_engine_set_read_group_names gives the final names before the report. */
fp_status fp_alignments_header(const fp_alignment_file *f, fp_header_view *out, uint32_t *max_read_groups);
/* records read so far, and records whose RG tag named no @RG record */
fp_status fp_alignments_counts(const fp_alignment_file *f, uint64_t *records, uint64_t *unknown_read_group);
void fp_alignments_close(fp_alignment_file *f);

/* thread-local message of the last failed loader call */
const char *fp_io_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* FIREPONY_AMD_IO_H */
