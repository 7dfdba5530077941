"""ctypes mirror of include/firepony_amd.h (the C ABI of the engine).

Structures and enums here must match the header field for field; the
`test_abi_layout` test checks sizes and offsets against the compiled library.
"""
import ctypes as C

FP_ABI_VERSION = 2

FP_OK = 0
FP_ERR_INVALID_ARG = 1
FP_ERR_HIP = 2
FP_ERR_OUT_OF_MEMORY = 3
FP_ERR_STATE = 4
FP_ERR_BUDGET = 5
FP_ERR_UNSUPPORTED = 6
FP_ERR_IO = 7
FP_ERR_INTERNAL = 8
FP_ERR_FORMAT = 9

FP_ENGINE_PLAIN = 0
FP_ENGINE_BATCHED = 1

FP_AGG_DIRECT = 0
FP_AGG_SORT = 1

FP_TABLE_READ_GROUP = 0
FP_TABLE_QUALITY = 1
FP_TABLE_CYCLE = 2
FP_TABLE_CONTEXT = 3

(FP_DBG_ACTIVE_READS, FP_DBG_READ_WINDOW, FP_DBG_NOINS_WINDOW, FP_DBG_REF_WINDOW,
 FP_DBG_NUM_ERRORS, FP_DBG_ERROR_FLAGS, FP_DBG_ACTIVE_MASK, FP_DBG_HMM_WINDOW,
 FP_DBG_BANDWIDTH, FP_DBG_BAQ, FP_DBG_FRAC_ERRORS, FP_DBG_HQ_WINDOW) = range(12)

u8p = C.POINTER(C.c_uint8)
u16p = C.POINTER(C.c_uint16)
u32p = C.POINTER(C.c_uint32)
i32p = C.POINTER(C.c_int32)
u64p = C.POINTER(C.c_uint64)
f64p = C.POINTER(C.c_double)


class FpOptions(C.Structure):
    _fields_ = [
        ("abi_version", C.c_uint32),
        ("engine_kind", C.c_uint32),
        ("max_batch_reads", C.c_uint32),
        ("max_batch_bases", C.c_uint32),
        ("frac_error_f64", C.c_uint32),
        ("disable_output_rounding", C.c_uint32),
        ("baq_waves_per_simd", C.c_uint32),
        ("sort_digit_bits", C.c_uint32),
        ("max_read_length", C.c_uint32),
        ("aggregation", C.c_uint32),
        ("debug_arrays", C.c_uint32),
        ("profile_stages", C.c_uint32),
        ("serve_rg_slots", C.c_uint32),
        ("disable_baq_staging", C.c_uint32),
        ("baq_checkpoint_rows", C.c_uint32),
        ("serve_hbm_budget_mib", C.c_uint32),
        ("staging_buffers", C.c_uint32),
    ]


class FpHeaderView(C.Structure):
    _fields_ = [
        ("num_chromosomes", C.c_uint32),
        ("chromosome_lengths", u32p),
        ("num_read_groups", C.c_uint32),
        ("read_group_names", C.POINTER(C.c_char_p)),
    ]


class FpReferenceView(C.Structure):
    _fields_ = [
        ("num_sequences", C.c_uint32),
        ("word_offset", u64p),
        ("length", u32p),
        ("words", u32p),
        ("total_words", C.c_uint64),
    ]


class FpVariantsView(C.Structure):
    _fields_ = [
        ("num_sequences", C.c_uint32),
        ("feature_offset", u64p),
        ("feature_start", u32p),
        ("feature_stop", u32p),
        ("max_end_point_left", u32p),
        ("num_features", C.c_uint64),
    ]


class FpBatchView(C.Structure):
    _fields_ = [
        ("num_reads", C.c_uint32),
        ("chromosome", u16p),
        ("alignment_start", u32p),
        ("alignment_stop", u32p),
        ("mate_chromosome", u32p),
        ("mate_alignment_start", u32p),
        ("inferred_insert_size", i32p),
        ("cigars", u32p),
        ("num_cigar_ops", C.c_uint64),
        ("cigar_start", u32p),
        ("cigar_len", u32p),
        ("reads", u32p),
        ("num_read_bases", C.c_uint64),
        ("read_start", u32p),
        ("read_len", u32p),
        ("qualities", u8p),
        ("num_qualities", C.c_uint64),
        ("qual_start", u32p),
        ("qual_len", u32p),
        ("flags", u16p),
        ("mapq", u8p),
        ("read_group", u32p),
    ]


class FpTableOut(C.Structure):
    _fields_ = [
        ("count", C.c_uint64),
        ("keys", u32p),
        ("observations", u64p),
        ("mismatches", f64p),
        ("mismatches_fixed_lo", u64p),
        ("mismatches_fixed_hi", u64p),
        ("expected_errors", f64p),
        ("estimated_quality", f64p),
        ("empirical_quality", f64p),
    ]


class FpStats(C.Structure):
    _fields_ = [
        ("total_reads", C.c_uint64),
        ("filtered_reads", C.c_uint64),
        ("baq_reads", C.c_uint64),
        ("num_batches", C.c_uint64),
        ("io_ms", C.c_double),
        ("read_filter_ms", C.c_double),
        ("cigar_expansion_ms", C.c_double),
        ("bp_filter_ms", C.c_double),
        ("snp_filter_ms", C.c_double),
        ("baq_ms", C.c_double),
        ("fractional_error_ms", C.c_double),
        ("covariates_ms", C.c_double),
        ("covariates_sort_ms", C.c_double),
        ("postprocessing_ms", C.c_double),
        ("output_ms", C.c_double),
        ("prep_kernel_ms", C.c_double),
        ("baq_hmm_kernel_ms", C.c_double),
        ("cov_kernel_ms", C.c_double),
        ("prep_launches", C.c_uint64),
        ("baq_hmm_launches", C.c_uint64),
        ("cov_launches", C.c_uint64),
        ("baq_cells", C.c_uint64),
        ("baq_waves_per_simd", C.c_uint64),
        ("sort_digit_bits", C.c_uint64),
        ("sort_slots", C.c_uint64),
        ("sort_pairs", C.c_uint64),
        ("sort_passes", C.c_uint64),
        ("baq_checkpoint_rows", C.c_uint64),
        ("baq15_kernel_ms", C.c_double),
        ("baq15_launches", C.c_uint64),
        ("baq15_cells", C.c_uint64),
    ]


class FpsynGenomeSpec(C.Structure):
    _fields_ = [
        ("seed", C.c_uint64),
        ("num_contigs", C.c_uint32),
        ("contig_length", C.c_uint64 * 32),
        ("num_sites", C.c_uint32),
        ("n_run_fraction", C.c_double),
        ("multi_base_site_frac", C.c_double),
    ]


class FpsynReadSpec(C.Structure):
    _fields_ = [
        ("seed", C.c_uint64),
        ("num_lengths", C.c_uint32),
        ("lengths", C.c_uint32 * 8),
        ("num_read_groups", C.c_uint32),
        ("rg_mismatch", C.c_double * 32),
        ("indel_read_frac", C.c_double),
        ("indel_min", C.c_uint32),
        ("indel_max", C.c_uint32),
        ("clip_read_frac", C.c_double),
        ("clip_min", C.c_uint32),
        ("clip_max", C.c_uint32),
        ("q2_tail_frac", C.c_double),
        ("q2_tail_max", C.c_uint32),
        ("q2_base_frac", C.c_double),
        ("paired", C.c_uint32),
        ("insert_mean", C.c_double),
        ("insert_sd", C.c_double),
        ("edge_case_frac", C.c_double),
    ]


def view_arrays(v):
    """numpy views of every array of an fp_batch_view over host memory (no copies)."""
    import numpy as np
    n = v.num_reads

    def arr(ptr, count, dtype):
        if count == 0 or not ptr:
            return np.zeros(0, dtype=dtype)
        return np.ctypeslib.as_array(ptr, shape=(count,)).view(dtype)

    return {
        "chromosome": arr(v.chromosome, n, np.uint16),
        "alignment_start": arr(v.alignment_start, n, np.uint32),
        "alignment_stop": arr(v.alignment_stop, n, np.uint32),
        "mate_chromosome": arr(v.mate_chromosome, n, np.uint32),
        "mate_alignment_start": arr(v.mate_alignment_start, n, np.uint32),
        "inferred_insert_size": arr(v.inferred_insert_size, n, np.int32),
        "cigars": arr(v.cigars, v.num_cigar_ops, np.uint32),
        "cigar_start": arr(v.cigar_start, n, np.uint32),
        "cigar_len": arr(v.cigar_len, n, np.uint32),
        "reads": arr(v.reads, v.num_read_bases // 8, np.uint32),
        "read_start": arr(v.read_start, n, np.uint32),
        "read_len": arr(v.read_len, n, np.uint32),
        "qualities": arr(v.qualities, v.num_qualities, np.uint8),
        "qual_start": arr(v.qual_start, n, np.uint32),
        "qual_len": arr(v.qual_len, n, np.uint32),
        "flags": arr(v.flags, n, np.uint16),
        "mapq": arr(v.mapq, n, np.uint8),
        "read_group": arr(v.read_group, n, np.uint32),
    }
