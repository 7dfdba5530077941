/*
 * fp_oracle.cpp — CPU restatement of firepony's host pipeline (test
 * infrastructure only; see fp_oracle.h).
 *
 * Every stage below cites the reference functor it restates.  Integer types
 * follow the reference (uint16 read/ref coordinates, ushort2/short2 windows)
 * so wrap-around behaves identically.
 *
 * Defined behaviour where the reference is undefined (the engine implements
 * the same definitions, and valid inputs never reach them):
 *   D1  per-read arrays the reference leaves uninitialised (no-insertion and
 *       reference windows when no event matches) start at 0 for every batch;
 *   D2  reference bases past a contig end read as N (15);
 *   D3  a read whose CIGAR query length differs from read_len, or whose
 *       chromosome id is not a loaded contig, is dropped by the malformed-read
 *       filters (the reference writes/reads other reads' memory there);
 *   D4  event walks never leave the read's own event range: the adapter
 *       left-tail search (cigar.cu:191-201) treats stepping before the first
 *       event as "no coordinate";
 *   D5  the BAQ MAP quality of a position whose posterior is exactly 1
 *       (log(0) = -inf) follows x86 cvttsd2si: INT_MIN, hence clamped to 4;
 *   D6  zero-length I or D ops mark no error bit (the reference would read the
 *       next op's first event), and op codes above 8 make a read malformed
 *       (htslib never produces them; the reference leaves their events unset).
 */
#include "fp_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

namespace {

typedef uint8_t uint8;
typedef uint16_t uint16;
typedef uint32_t uint32;
typedef uint64_t uint64;
typedef int16_t int16;
typedef int32_t int32;
typedef unsigned __int128 u128;

struct ushort2 { uint16 x, y; };
struct short2 { int16 x, y; };
struct uint2 { uint32 x, y; };

static inline ushort2 mk_us2(uint16 x, uint16 y) { ushort2 r; r.x = x; r.y = y; return r; }

// alignment_data.h:42-68
enum {
    F_PAIRED = 1, F_PROPER_PAIR = 2, F_UNMAP = 4, F_MATE_UNMAP = 8, F_REVERSE = 16,
    F_MATE_REVERSE = 32, F_READ1 = 64, F_READ2 = 128, F_SECONDARY = 256, F_QC_FAIL = 512,
    F_DUPLICATE = 1024, F_SUPPLEMENTARY = 2048
};
// types.h:61-72
enum { OP_M = 0, OP_I = 1, OP_D = 2, OP_N = 3, OP_S = 4, OP_H = 5, OP_P = 6, OP_MATCH = 7, OP_X = 8 };
// cigar.h:41-47
enum { EV_M = 0, EV_I = 1, EV_D = 2, EV_S = 3 };
// from_nvbio/alphabet.h:114-135
enum { IUPAC_A = 1, IUPAC_C = 2, IUPAC_G = 4, IUPAC_T = 8, IUPAC_N = 15 };

static const uint16 U16_NONE = 0xffff;

// ---------------------------------------------------------------------------
// reference data (sequence_database.h, variant_database.h)
struct Contig {
    std::vector<uint32> words;
    uint32 length;
    uint8 base(uint32 i) const {
        if (i >= length) return IUPAC_N;  // D2
        return (words[i >> 3] >> ((i & 7) * 4)) & 0xf;
    }
};

struct VariantContig {
    std::vector<uint32> feature_start, feature_stop, max_end_point_left;
};

// ---------------------------------------------------------------------------
// covariate tables
struct ExactAcc {
    uint64 obs;
    u128 mism;
};

struct FloatRow {
    uint32 key;
    uint64 obs;
    float mism;
};

struct EmpRow {
    uint32 key;
    uint64 obs;
    double mism;
    double expected;
    double estimated;
    double empirical;
    u128 mism_fixed;
};

// fixed-point conversion: every per-base fractional error is k/n with
// n <= 65535, so v * 2^80 is an integer below 2^81.
static u128 to_fixed(double v)
{
    if (v == 0.0) return 0;
    int ex;
    double f = std::frexp(v, &ex);              // v = f * 2^ex, f in [0.5, 1)
    uint64 mant = (uint64)std::ldexp(f, 53);    // exact
    int sh = ex - 53 + 80;
    if (sh < 0) {
        // not representable exactly (never for n <= 65535)
        return (u128)(mant >> (-sh));
    }
    return ((u128)mant) << sh;
}

static double fixed_to_double(u128 v)
{
    // unsigned __int128 -> double rounds to nearest even; scaling by 2^-80 is exact
    return std::ldexp((double)v, -80);
}

// ---------------------------------------------------------------------------
// table_formatter (table_formatter.cu:35-233) restated over a std::string sink
struct Formatter {
    enum Fmt { STRING, CHAR, UINT64, FLOAT_2, FLOAT_4 };
    enum Align { DEFAULT, LEFT, RIGHT };

    std::string *out;
    bool disable_rounding;
    std::string name, description;
    uint32 num_columns = 0, num_rows = 0, col_idx = 0;
    bool preprocess = true;
    std::vector<std::string> column_names;
    std::vector<Fmt> column_formats;
    std::vector<uint32> column_widths;
    std::vector<Align> column_alignments, header_alignments;

    Formatter(std::string *out, bool dr, const std::string &n, const std::string &d = "")
        : out(out), disable_rounding(dr), name(n), description(d) {}

    void printf_(const char *fmt, ...) __attribute__((format(printf, 2, 3)))
    {
        char buf[4096];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof(buf), fmt, ap);
        va_end(ap);
        out->append(buf);
    }

    void add_column(const std::string &n, Fmt f, Align a = DEFAULT, Align ha = DEFAULT)
    {
        column_names.push_back(n);
        column_formats.push_back(f);
        column_widths.push_back((uint32)strlen(n.c_str()));
        if (a == DEFAULT) a = (f == STRING || f == CHAR) ? LEFT : RIGHT;
        column_alignments.push_back(a);
        if (ha == DEFAULT) ha = a;
        header_alignments.push_back(ha);
        num_columns++;
    }
    void start_row() { col_idx = 0; num_rows++; }
    void end_row() { if (!preprocess) printf_("\n"); }
    void end_table()
    {
        if (preprocess) {
            preprocess = false;
            printf_("#:GATKTable:%d:%d:", num_columns, num_rows);
            for (uint32 i = 0; i < num_columns; i++) {
                switch (column_formats[i]) {
                case STRING: case CHAR: printf_("%%s:"); break;
                case UINT64: printf_("%%d:"); break;
                case FLOAT_2: printf_(disable_rounding ? "%%.64f:" : "%%.2f:"); break;
                case FLOAT_4: printf_(disable_rounding ? "%%.64f:" : "%%.4f:"); break;
                }
            }
            printf_(";\n");
            printf_("#:GATKTable:%s:%s\n", name.c_str(), description.c_str());
            for (uint32 i = 0; i < num_columns; i++) {
                int w = (header_alignments[i] == RIGHT ? 1 : -1) * (int)column_widths[i];
                printf_("%*s", w, column_names[i].c_str());
                printf_(i == num_columns - 1 ? "\n" : "  ");
            }
        } else {
            printf_("\n");
        }
    }
    void data_str(const std::string &val)
    {
        if (preprocess) {
            column_widths[col_idx] = std::max((uint32)column_widths[col_idx], (uint32)strlen(val.c_str()));
        } else {
            int w = (column_alignments[col_idx] == RIGHT ? 1 : -1) * (int)column_widths[col_idx];
            printf_("%*s", w, val.c_str());
            if (col_idx < num_columns - 1) printf_("  ");
        }
        col_idx++;
    }
    void data_char(char c) { char b[8]; snprintf(b, sizeof(b), "%c", c); data_str(b); }
    void data_u64(uint64 v) { char b[64]; snprintf(b, sizeof(b), "%lu", (unsigned long)v); data_str(b); }
    void data_double(double v)
    {
        char b[256];
        if (column_formats[col_idx] == FLOAT_2) {
            if (disable_rounding) snprintf(b, sizeof(b), "%.64f", v);
            else snprintf(b, sizeof(b), "%.2f", fpo_round_n(v, 2));
        } else {
            if (disable_rounding) snprintf(b, sizeof(b), "%.64f", v);
            else snprintf(b, sizeof(b), "%.4f", fpo_round_n(v, 4));
        }
        data_str(b);
    }
    void data_int_as_string(int v) { char b[64]; snprintf(b, sizeof(b), "%d", v); data_str(b); }
};

} // namespace

// ---------------------------------------------------------------------------
// the oracle context (device/firepony_context.h:128-188, restated)
struct fpo_ctx {
    int frac_error_f64;
    int accum_mode;
    int disable_rounding;

    // setup
    std::vector<uint32> chromosome_lengths;
    std::vector<std::string> rg_names;
    std::vector<Contig> reference;
    std::vector<VariantContig> variants;
    bool is_setup = false;

    // tables
    std::map<uint32, ExactAcc> exact[3];     // quality, cycle, context
    std::vector<FloatRow> flt[3];
    std::vector<EmpRow> emp[4];              // rg, quality, cycle, context
    bool postprocessed = false;

    fp_stats stats;

    // per-batch state ------------------------------------------------------
    const fp_batch_view *b = nullptr;
    uint32 num_reads = 0;
    std::vector<uint32> active_read_list;
    std::vector<uint8> active_loc;          // per padded base: 0xf active, 0 inactive
    std::vector<uint16> read_offset_list;
    std::vector<uint2> alignment_windows;
    std::vector<uint32> cigar_offsets;
    std::vector<uint8> cigar_events;
    std::vector<uint32> ev_read_index;
    std::vector<uint16> ev_read_coord;
    std::vector<uint16> ev_ref_coord;
    std::vector<ushort2> read_window_clipped, read_window_noins, reference_window_clipped;
    std::vector<uint8> is_snp, is_ins, is_del;
    std::vector<uint16> num_errors;
    std::vector<uint2> vcf_ranges;
    std::vector<short2> hmm_windows;
    std::vector<uint16> bandwidth;
    std::vector<uint8> baq_qual;
    std::vector<uint32> baq_state;
    std::vector<double> baq_real;
    std::vector<double> frac[3];
    std::vector<ushort2> hq_window;

    // batch accessors ---------------------------------------------------
    uint8 read_base(uint64 i) const { return (b->reads[i >> 3] >> ((i & 7) * 4)) & 0xf; }
    uint32 op_len(uint32 i) const { return b->cigars[i] & 0xffffff; }
    uint32 op_code(uint32 i) const { return (b->cigars[i] >> 24) & 0xf; }
};

// qual2prob (baq.cu:356-359) and the GAP_OPEN constant (baq.cu:81), glibc pow
static double qual2prob(uint8 q) { return pow(10.0, -q / 10.0); }
#define MAX_PHRED_SCORE 93
#define EM 0.33333333333
#define EI 0.25
#define MIN_BAND_WIDTH 7
#define MIN_BASE_QUAL 4
#define GAP_OPEN_PROBABILITY (pow(10.0, (-40.0) / 10.))
#define GAP_EXTENSION_PROBABILITY 0.1
#define NO_BAQ_UNCERTAINTY 64

static uint32 uabs(int32 v) { return v < 0 ? uint32(-v) : uint32(v); }   // baq.cu:84-87

// ===========================================================================
// read filters (device/read_filters.cu)

// filter_mapq (read_filters.cu:60-75)
static bool filter_mapq(const fpo_ctx &c, uint32 r)
{
    return !(c.b->mapq[r] == 0 || c.b->mapq[r] == 255);
}

// filter_if_any_set<DUP|QCFAIL|UNMAP|SECONDARY> (read_filters.cu:43-57, 282-286)
static bool filter_flags(const fpo_ctx &c, uint32 r)
{
    const uint32 flags = F_DUPLICATE | F_QC_FAIL | F_UNMAP | F_SECONDARY;
    return (c.b->flags[r] & flags) == 0;
}

// filter_if_cigar_malformed (read_filters.cu:157-258) + D3 query-length check
static bool filter_cigar(const fpo_ctx &c, uint32 r)
{
    enum { LEADING_HARD, LEADING_SOFT, NO_CLIPS, TRAILING_SOFT, TRAILING_HARD } st = LEADING_HARD;
    const uint32 cs = c.b->cigar_start[r], cl = c.b->cigar_len[r];
    uint64 qlen = 0;
    for (uint32 i = cs; i < cs + cl; i++) {
        const uint32 op = c.op_code(i);
        if (op == OP_N) return false;
        if (op > OP_X) return false;   // D6
        if (op == OP_H) {
            switch (st) {
            case LEADING_HARD: break;
            case LEADING_SOFT: return false;
            case NO_CLIPS: case TRAILING_SOFT: st = TRAILING_HARD; break;
            case TRAILING_HARD: break;
            }
        } else if (op == OP_S) {
            switch (st) {
            case LEADING_HARD: st = LEADING_SOFT; break;
            case LEADING_SOFT: break;
            case NO_CLIPS: st = TRAILING_SOFT; break;
            case TRAILING_SOFT: break;
            case TRAILING_HARD: return false;
            }
        } else {
            switch (st) {
            case LEADING_HARD: case LEADING_SOFT: st = NO_CLIPS; break;
            case NO_CLIPS: break;
            case TRAILING_SOFT: case TRAILING_HARD: return false;
            }
        }
        if (op == OP_M || op == OP_I || op == OP_S || op == OP_MATCH || op == OP_X || op == OP_N)
            qlen += c.op_len(i);
    }
    // D3: the offset list below writes exactly qlen entries
    if (qlen != c.b->read_len[r]) return false;
    return true;
}

// filter_if_read_malformed (read_filters.cu:78-154) + D3 chromosome check
static bool filter_read_malformed(const fpo_ctx &c, uint32 r)
{
    const fp_batch_view *b = c.b;
    if (!(b->flags[r] & F_UNMAP)) {
        // uint16 chromosome never equals uint32(-1) (read_filters.cu:92)
        if (b->alignment_start[r] == uint32(-1)) return false;
        if (c.alignment_windows[r].y <= c.alignment_windows[r].x) return false;
        if (b->chromosome[r] >= c.reference.size()) return false;   // D3
        if (c.alignment_windows[r].x >= c.reference[b->chromosome[r]].length) return false;
        if (b->cigar_len[r] == 0) return false;
    }
    if (b->read_group[r] == uint32(-1)) return false;
    if (b->qual_len[r] != b->read_len[r]) return false;
    if (b->read_len[r] == 0) return false;
    return true;
}

// filter_invalid_reads (read_filters.cu:262-336)
static void filter_invalid_reads(fpo_ctx &c)
{
    const uint32 start = (uint32)c.active_read_list.size();
    std::vector<uint32> out;
    for (uint32 r : c.active_read_list) if (filter_mapq(c, r)) out.push_back(r);
    c.active_read_list.swap(out); out.clear();
    for (uint32 r : c.active_read_list) if (filter_flags(c, r)) out.push_back(r);
    c.active_read_list.swap(out); out.clear();
    for (uint32 r : c.active_read_list) if (filter_cigar(c, r)) out.push_back(r);
    c.active_read_list.swap(out);
    c.stats.filtered_reads += start - c.active_read_list.size();
}

// compute_read_offset_list (snp_filter.cu:44-101)
static void build_read_offset_list(fpo_ctx &c)
{
    for (uint32 r : c.active_read_list) {
        const uint32 cs = c.b->cigar_start[r];
        uint16 *out = &c.read_offset_list[c.b->read_start[r]];
        uint16 offset = 0;
        for (uint32 k = 0; k < c.b->cigar_len[r]; k++) {
            const uint32 len = c.op_len(cs + k);
            switch (c.op_code(cs + k)) {
            case OP_M: case OP_MATCH: case OP_X:
                for (uint32 i = 0; i < len; i++) { *out++ = offset; offset++; }
                break;
            case OP_I: case OP_N:
                for (uint32 i = 0; i < len; i++) *out++ = offset;
                break;
            case OP_D: case OP_H: case OP_P:
                offset += len;
                break;
            case OP_S:
                for (uint32 i = 0; i < len; i++) *out++ = U16_NONE;
                break;
            }
        }
    }
}

// compute_alignment_window (snp_filter.cu:118-147)
static void build_alignment_windows(fpo_ctx &c)
{
    for (uint32 r : c.active_read_list) {
        const uint16 *ol = &c.read_offset_list[c.b->read_start[r]];
        int k;
        for (k = (int)c.b->read_len[r] - 1; k >= 0; k--)
            if (ol[k] != U16_NONE) break;
        if (k < 0) {
            c.alignment_windows[r].x = uint32(-1);
            c.alignment_windows[r].y = uint32(-1);
        } else {
            c.alignment_windows[r].x = c.b->alignment_start[r];
            c.alignment_windows[r].y = c.alignment_windows[r].x + ol[k];
        }
    }
}

// filter_malformed_reads (read_filters.cu:341-372)
static void filter_malformed_reads(fpo_ctx &c)
{
    const uint32 start = (uint32)c.active_read_list.size();
    std::vector<uint32> out;
    for (uint32 r : c.active_read_list) if (filter_read_malformed(c, r)) out.push_back(r);
    c.active_read_list.swap(out);
    c.stats.filtered_reads += start - c.active_read_list.size();
}

// ===========================================================================
// CIGAR expansion (device/cigar.cu)

struct AdapterClip {
    fpo_ctx &c;
    uint32 r;
    AdapterClip(fpo_ctx &c, uint32 r) : c(c), r(r) {}

    // cigar.cu:112-142
    bool hasWellDefinedFragmentSize()
    {
        const uint32 flags = c.b->flags[r];
        const int32 isize = c.b->inferred_insert_size[r];
        if (isize == 0) return false;
        if (!(flags & F_PAIRED)) return false;
        if ((flags & F_UNMAP) || (flags & F_MATE_UNMAP)) return false;
        const bool rev = flags & F_REVERSE;
        const bool mate_rev = flags & F_MATE_REVERSE;
        if (rev == mate_rev) return false;
        if (flags & F_REVERSE)
            return c.b->alignment_stop[r] > c.b->mate_alignment_start[r];
        else
            return c.b->alignment_start[r] <= c.b->mate_alignment_start[r] + isize;
    }

    // cigar.cu:146-160
    uint32 getAdaptorBoundary()
    {
        if (!hasWellDefinedFragmentSize()) return 0xffffffffu;
        if (c.b->flags[r] & F_REVERSE) {
            return uint32(c.b->mate_alignment_start[r] - 1);
        } else {
            const int32 is = c.b->inferred_insert_size[r];
            const int insertSize = is < 0 ? -is : is;
            return uint32(c.b->alignment_start[r] + insertSize + 1);
        }
    }

    // cigar.cu:168-215
    int getReadCoordinateForReferenceCoordinate(uint32 ref_coord, bool right_tail)
    {
        const uint32 cs = c.cigar_offsets[c.b->cigar_start[r]];
        const uint32 ce = c.cigar_offsets[c.b->cigar_start[r] + c.b->cigar_len[r]];
        for (uint32 ev = cs; ev < ce; ev++) {
            if (c.b->alignment_start[r] + c.ev_ref_coord[ev] == ref_coord) {
                uint16 rc = c.ev_read_coord[ev];
                if (rc == U16_NONE) {
                    if (right_tail) {
                        while (ev < ce) {
                            rc = c.ev_read_coord[ev];
                            if (rc != U16_NONE) break;
                            ev++;
                        }
                    } else {
                        while (ev >= cs && ev > 0) {
                            ev--;
                            if (ev < cs) { rc = U16_NONE; break; }   // D4
                            rc = c.ev_read_coord[ev];
                            if (rc != U16_NONE) break;
                        }
                    }
                    if (rc == U16_NONE) return -1;
                }
                return rc;
            }
        }
        return -1;
    }

    // cigar.cu:217-233
    ushort2 hardClipByReferenceCoordinates(int refStart, int refStop)
    {
        int start, stop;
        if (refStart < 0) {
            start = 0;
            stop = getReadCoordinateForReferenceCoordinate(uint32(refStop), false);
        } else {
            start = getReadCoordinateForReferenceCoordinate(uint32(refStart), true);
            stop = c.read_window_clipped[r].y - c.read_window_clipped[r].x - 1;
        }
        return mk_us2((uint16)start, (uint16)stop);
    }

    // cigar.cu:252-301 (get_current_reference_window)
    ushort2 get_current_reference_window()
    {
        const ushort2 rwc = c.read_window_clipped[r];
        ushort2 ref = mk_us2(0, 0);   // D1
        const uint32 cs = c.cigar_offsets[c.b->cigar_start[r]];
        const uint32 ce = c.cigar_offsets[c.b->cigar_start[r] + c.b->cigar_len[r]];
        for (uint32 i = cs; i < ce; i++) {
            if (c.ev_read_coord[i] == rwc.x) {
                while (i < ce && c.ev_ref_coord[i] == U16_NONE) i++;
                if (i == ce) return mk_us2(U16_NONE, U16_NONE);
                ref.x = c.ev_ref_coord[i];
                break;
            }
        }
        if (ce > cs) {
            for (uint32 i = ce - 1;; i--) {
                if (c.ev_read_coord[i] == rwc.y) {
                    while (c.ev_ref_coord[i] == U16_NONE && i > cs) i--;
                    if (i == cs) return mk_us2(U16_NONE, U16_NONE);
                    ref.y = c.ev_ref_coord[i];
                    break;
                }
                if (i == cs) break;
            }
        }
        return ref;
    }

    // cigar.cu:303-325
    void run()
    {
        const uint32 boundary = getAdaptorBoundary();
        if (boundary == 0xffffffffu) return;
        const ushort2 rw = get_current_reference_window();
        if (boundary < c.b->alignment_start[r] + rw.x || boundary > c.b->alignment_start[r] + rw.y)
            return;
        ushort2 &rwc = c.read_window_clipped[r];
        if (c.b->flags[r] & F_REVERSE) {
            ushort2 a = hardClipByReferenceCoordinates(-1, (int)boundary);
            rwc.x = std::max<uint16>(rwc.x, (uint16)(a.y + 1));
        } else {
            ushort2 a = hardClipByReferenceCoordinates((int)boundary, -1);
            rwc.y = std::min<uint16>(rwc.y, (uint16)(a.x - 1));
        }
    }
};

static void expand_cigars(fpo_ctx &c)
{
    const fp_batch_view *b = c.b;
    const uint32 nops = (uint32)b->num_cigar_ops;

    // cigar_offsets: inclusive scan of op lengths over *all* ops (cigar.cu:823-834)
    c.cigar_offsets.assign(nops + 1, 0);
    for (uint32 i = 0; i < nops; i++) c.cigar_offsets[i + 1] = c.cigar_offsets[i] + c.op_len(i);
    const uint32 E = c.cigar_offsets[nops];

    // cigar_op_expand (cigar.cu:56-94)
    c.cigar_events.assign(E, 0);
    for (uint32 i = 0; i < nops; i++) {
        uint8 ev = 0;
        switch (c.op_code(i)) {
        case OP_M: case OP_MATCH: case OP_X: ev = EV_M; break;
        case OP_I: case OP_N: ev = EV_I; break;
        case OP_D: case OP_H: case OP_P: ev = EV_D; break;
        case OP_S: ev = EV_S; break;
        default: ev = EV_M; break;   // unknown ops: 2-bit vector keeps 0 (uninitialised in the reference)
        }
        for (uint32 k = c.cigar_offsets[i]; k < c.cigar_offsets[i + 1]; k++) c.cigar_events[k] = ev;
    }

    c.ev_read_index.assign(E, uint32(-1));          // cigar.cu:862-863
    c.ev_read_coord.assign(E, 0);                   // D1
    c.ev_ref_coord.assign(E, 0);                    // D1
    c.num_errors.assign(c.num_reads, 0);

    // cigar_coordinates_expand (cigar.cu:538-609)
    for (uint32 r : c.active_read_list) {
        const uint32 cs = b->cigar_start[r];
        uint32 base = c.cigar_offsets[cs];
        uint16 ro = 0, fo = 0;
        for (uint32 k = 0; k < b->cigar_len[r]; k++) {
            const uint32 len = c.op_len(cs + k);
            switch (c.op_code(cs + k)) {
            case OP_M: case OP_MATCH: case OP_X:
                for (uint32 i = 0; i < len; i++) {
                    c.ev_read_index[base] = r; c.ev_read_coord[base] = ro; c.ev_ref_coord[base] = fo;
                    base++; ro++; fo++;
                }
                break;
            case OP_S:
            case OP_N: case OP_I:
                for (uint32 i = 0; i < len; i++) {
                    c.ev_read_index[base] = r; c.ev_read_coord[base] = ro; c.ev_ref_coord[base] = U16_NONE;
                    base++; ro++;
                }
                break;
            case OP_D: case OP_H: case OP_P:
                for (uint32 i = 0; i < len; i++) {
                    c.ev_read_index[base] = r; c.ev_read_coord[base] = U16_NONE; c.ev_ref_coord[base] = fo;
                    base++; fo++;
                }
                break;
            }
        }
    }

    // read_window_init (cigar.cu:98-108)
    for (uint32 r : c.active_read_list) c.read_window_clipped[r] = mk_us2(0, (uint16)(b->read_len[r] - 1));

    // remove_adapters (cigar.cu:110-326)
    for (uint32 r : c.active_read_list) AdapterClip(c, r).run();

    // remove_soft_clips (cigar.cu:329-389)
    for (uint32 r : c.active_read_list) {
        ushort2 rwc = c.read_window_clipped[r];
        const uint32 cs = b->cigar_start[r], cl = b->cigar_len[r];
        uint32 ci = cs, ro = 0;
        while (ci < cs + cl && ro < b->read_len[r]) {
            const uint32 op = c.op_code(ci), len = c.op_len(ci);
            if (op == OP_H || op == OP_S) {
                if (ro + len > rwc.x) rwc.x = std::min<uint16>((uint16)(ro + len), rwc.y);
                ro += len; ci++;
            } else break;
        }
        if (cl > 0) {
            ci = cs + cl - 1;
            ro = b->read_len[r] - 1;
            while (ro > rwc.x) {
                const uint32 op = c.op_code(ci), len = c.op_len(ci);
                if (op == OP_H || op == OP_S) {
                    if (ro - len < rwc.y) rwc.y = std::max<uint16>((uint16)(ro - len), rwc.x);
                    ro -= len;
                    if (ci == cs) break;   // the reference would index before the read's ops
                    ci--;
                } else break;
            }
        }
        c.read_window_clipped[r] = rwc;
    }

    // compute_no_insertions_window (cigar.cu:392-451)
    for (uint32 r : c.active_read_list) {
        const ushort2 rwc = c.read_window_clipped[r];
        ushort2 &ni = c.read_window_noins[r];
        const uint32 cs = c.cigar_offsets[b->cigar_start[r]];
        const uint32 ce = c.cigar_offsets[b->cigar_start[r] + b->cigar_len[r]];
        uint32 ev = cs;
        while (ev < ce) {
            const uint16 ro = c.ev_read_coord[ev];
            if (ro == U16_NONE || ro < rwc.x) { ev++; continue; }
            if (c.cigar_events[ev] == EV_I) { ev++; continue; }
            ni.x = ro;
            break;
        }
        if (ce > cs) {
            ev = ce - 1;
            while (ev >= cs && ev < ce) {
                const uint16 ro = c.ev_read_coord[ev];
                if (ro == U16_NONE || ro > rwc.y || c.cigar_events[ev] == EV_I) {
                    if (ev == cs) break;
                    ev--;
                    continue;
                }
                ni.y = ro;
                break;
            }
        }
    }

    // compute_reference_window (cigar.cu:454-535)
    for (uint32 r : c.active_read_list) {
        const ushort2 ni = c.read_window_noins[r];
        ushort2 &rw = c.reference_window_clipped[r];
        const uint32 cs = c.cigar_offsets[b->cigar_start[r]];
        const uint32 ce = c.cigar_offsets[b->cigar_start[r] + b->cigar_len[r]];
        bool invalid = false;
        for (uint32 i = cs; i < ce; i++) {
            if (c.ev_read_coord[i] == ni.x) {
                while (i < ce && c.ev_ref_coord[i] == U16_NONE) i++;
                if (i == ce) { rw = mk_us2(U16_NONE, U16_NONE); invalid = true; break; }
                rw.x = c.ev_ref_coord[i];
                break;
            }
        }
        if (invalid || ce == cs) continue;
        for (uint32 i = ce - 1;; i--) {
            if (c.ev_read_coord[i] == ni.y) {
                while (c.ev_ref_coord[i] == U16_NONE && i > cs) i--;
                if (i == cs) { rw = mk_us2(U16_NONE, U16_NONE); break; }
                rw.y = c.ev_ref_coord[i];
                break;
            }
            if (i == cs) break;
        }
    }

    // compute_error_vectors (cigar.cu:612-733)
    for (uint32 r : c.active_read_list) {
        const bool neg = b->flags[r] & F_REVERSE;
        const Contig &ref = c.reference[b->chromosome[r]];
        const uint32 astart = b->alignment_start[r];
        const ushort2 rwc = c.read_window_clipped[r];
        const ushort2 fwc = c.reference_window_clipped[r];
        const uint32 rs = b->read_start[r];
        uint16 cur = 0;
        uint16 nerr = 0;
        for (uint32 op = b->cigar_start[r]; op < b->cigar_start[r] + b->cigar_len[r]; op++) {
            const uint32 es = c.cigar_offsets[op], ee = c.cigar_offsets[op + 1];
            switch (c.op_code(op)) {
            case OP_M:
                for (uint32 i = es; i < ee; i++) {
                    cur = c.ev_read_coord[i];
                    const uint8 rb = c.read_base(rs + cur);
                    const uint32 fi = c.ev_ref_coord[i];
                    const uint8 fb = ref.base(astart + fi);
                    if (fb != rb) {
                        c.is_snp[rs + cur] = 1;
                        if (cur >= rwc.x && cur <= rwc.y) nerr++;
                    }
                }
                break;
            case EV_I: {   // == OP_I
                if (es == ee) break;   // D6
                cur = c.ev_read_coord[es];
                if (cur >= rwc.x && cur <= rwc.y) {
                    int off;
                    if (!neg) off = cur - 1;
                    else off = cur + (int)c.op_len(op);
                    if (off >= 0 && off <= rwc.y) {
                        c.is_ins[rs + off] = 1;
                        nerr++;
                    }
                }
                break;
            }
            case EV_D: {   // == OP_D
                if (es == ee) break;   // D6
                const uint16 cri = c.ev_ref_coord[es];
                if (cri >= fwc.x && cri <= fwc.y) {
                    if (!neg) {
                        c.is_del[rs + cur] = 1;
                        nerr++;
                    } else {
                        uint16 off = cur + 1;
                        if (off < b->read_len[r]) {
                            c.is_del[rs + off] = 1;
                            nerr++;
                        }
                    }
                }
                break;
            }
            default:
                break;
            }
        }
        c.num_errors[r] = nerr;
    }
}

// ===========================================================================
// per-base filters (read_filters.cu:376-438)
static void filter_bases(fpo_ctx &c)
{
    for (uint32 r : c.active_read_list) {
        const uint32 rs = c.b->read_start[r];
        for (uint32 i = rs; i < rs + c.b->read_len[r]; i++) {
            const uint8 bp = c.read_base(i);
            if (bp != IUPAC_A && bp != IUPAC_C && bp != IUPAC_G && bp != IUPAC_T) c.active_loc[i] = 0;
        }
    }
    for (uint32 r : c.active_read_list) {
        const uint32 rs = c.b->read_start[r], qs = c.b->qual_start[r];
        for (uint32 i = 0; i < c.b->read_len[r]; i++)
            if (c.b->qualities[qs + i] < 6) c.active_loc[rs + i] = 0;    // MIN_USABLE_Q_SCORE
    }
}

// ===========================================================================
// known-site mask (snp_filter.cu:162-428), lower/upper_bound (primitives/algorithms.h:39-100)

static size_t lower_bound_ref(uint32 x, const uint32 *a, size_t n)
{
    if (n == 0) return 0;
    if (x < a[0]) return 0;
    if (a[n - 1] < x) return n;
    size_t base = 0, count = n;
    while (count > 0) {
        size_t c2 = count / 2;
        size_t mid = base + c2;
        if (a[mid] < x) { base = mid + 1; count -= c2 + 1; }
        else count = c2;
    }
    return base;
}

static size_t upper_bound_ref(uint32 x, const uint32 *a, size_t n)
{
    size_t base = 0, count = n;
    while (count > 0) {
        size_t step = count / 2;
        size_t it = base + step;
        if (!(x < a[it])) { base = it + 1; count -= step + 1; }
        else count = step;
    }
    return base;
}

static void filter_known_snps(fpo_ctx &c)
{
    const fp_batch_view *b = c.b;
    // compute_vcf_ranges (snp_filter.cu:162-222)
    for (uint32 r : c.active_read_list) {
        const uint32 ch = b->chromosome[r];
        if (ch >= c.variants.size()) { c.vcf_ranges[r].x = c.vcf_ranges[r].y = uint32(-1); continue; }
        const VariantContig &db = c.variants[ch];
        const ushort2 fw = c.reference_window_clipped[r];
        const uint32 off = b->alignment_start[r];
        const uint32 wx = off + uint32(fw.x), wy = off + uint32(fw.y);
        uint2 vr;
        vr.x = (uint32)lower_bound_ref(wx, db.max_end_point_left.data(), db.max_end_point_left.size());
        if (vr.x >= db.feature_start.size()) { c.vcf_ranges[r].x = c.vcf_ranges[r].y = uint32(-1); continue; }
        vr.y = (uint32)upper_bound_ref(wy, db.feature_start.data(), db.feature_start.size());
        if (vr.y <= vr.x) { c.vcf_ranges[r].x = c.vcf_ranges[r].y = uint32(-1); continue; }
        if (vr.y >= db.feature_start.size()) vr.y = (uint32)db.feature_start.size() - 1;
        c.vcf_ranges[r] = vr;
    }

    // filter_bps over reads with an active range (snp_filter.cu:239-395, 409-427)
    for (uint32 r : c.active_read_list) {
        if (c.vcf_ranges[r].x == uint32(-1)) continue;
        const VariantContig &db = c.variants[b->chromosome[r]];
        const ushort2 rwc = c.read_window_clipped[r];
        const ushort2 fwc = c.reference_window_clipped[r];
        const uint32 astart = b->alignment_start[r];
        const uint2 vr = c.vcf_ranges[r];
        const uint32 cs = c.cigar_offsets[b->cigar_start[r]];
        const uint32 ce = c.cigar_offsets[b->cigar_start[r] + b->cigar_len[r]];

        for (uint32 feature = vr.x; feature <= vr.y; feature++) {
            const int feature_start = (int)(db.feature_start[feature] - astart);
            const int feature_end = (int)(db.feature_stop[feature] - astart);
            uint32 ev;
            uint16 ref_coord = U16_NONE;
            for (ev = cs; ev < ce; ev++) {
                ref_coord = c.ev_ref_coord[ev];
                if (ref_coord != U16_NONE && ref_coord >= feature_start) break;
            }
            if (ev == ce) continue;

            int feature_bp_left = ref_coord - feature_start;
            uint32 evs = ev;
            if (feature_bp_left > 0) {
                while (feature_bp_left && evs > cs) {
                    const uint8 e = c.cigar_events[evs];
                    if (e == EV_S) break;
                    if (e == EV_M || e == EV_D) feature_bp_left--;
                    evs--;
                }
                while (evs > cs && c.cigar_events[evs - 1] == EV_I) evs--;
                while (evs < ce && c.ev_read_coord[evs] == U16_NONE) evs++;
            } else {
                if (c.ev_ref_coord[evs] <= fwc.y) {
                    while (evs > cs && c.cigar_events[evs] == EV_D) evs--;
                }
            }

            int feature_bp_right = feature_end - ref_coord;
            while (feature_bp_right > 0 && ev < ce - 1) {
                ev++;
                const uint8 e = c.cigar_events[ev];
                if (e == EV_S) break;
                if (e == EV_M || e == EV_D) feature_bp_right--;
            }
            while (ev > cs && c.ev_read_coord[ev] == U16_NONE) ev--;
            if (ev < evs) continue;

            uint32 rs = c.ev_read_coord[evs];
            uint32 re = c.ev_read_coord[ev];
            if ((rs < rwc.x && re < rwc.x) || (rs > rwc.y && re > rwc.y)) continue;
            rs = std::max<uint32>(rs, rwc.x);
            rs = std::min<uint32>(rs, rwc.y);
            re = std::max<uint32>(re, rwc.x);
            re = std::min<uint32>(re, rwc.y);
            for (uint32 d = rs; d <= re; d++) c.active_loc[b->read_start[r] + d] = 0;
        }
    }
}

// ===========================================================================
// BAQ (device/baq.cu)

// hmm_glocal (baq.cu:214-697), forward + backward + MAP, with the
// reference's own matrix layout (set_u, off) over zero-filled matrices.
struct Hmm {
    int bandWidth, bandWidth2;
    int referenceLength, queryLen;
    double sM, sI, bM, bI;
    double m[9];
    std::vector<double> F, B, S;
    const uint8 *refb;    // reference window bases (4-bit codes)
    const uint8 *qb;      // query bases
    const uint8 *qual;    // query qualities

    static int set_u(int b, int i, int k) { int x = i - b; x = x > 0 ? x : 0; return (k + 1 - x) * 3; }
    int off(int i, int j = 0) const { return i * (bandWidth2 * 3 + 6) + j; }

    static double calcEpsilon(uint8 ref, uint8 read, uint8 qualB)
    {
        if (ref == IUPAC_N || read == IUPAC_N) return 1.0;
        double q = qual2prob(qualB < MIN_BASE_QUAL ? MIN_BASE_QUAL : qualB);
        return ref == read ? 1 - q : q * EM;
    }

    // baq.cu:268-315: the transition terms use the window length before the
    // contig-start clamp decrements it (refLenLoops = refLenWindow - 1 there)
    void setup(int refLenWindow, int refLenLoops, int qLen, int bw, int matrix_rows)
    {
        referenceLength = refLenWindow; queryLen = qLen; bandWidth = bw; bandWidth2 = bw * 2 + 1;
        sM = 1.0 / (2 * queryLen + 2);
        sI = sM;
        bM = (1 - GAP_OPEN_PROBABILITY) / referenceLength;
        bI = GAP_OPEN_PROBABILITY / referenceLength;
        referenceLength = refLenLoops;
        m[0] = (1 - GAP_OPEN_PROBABILITY - GAP_OPEN_PROBABILITY) * (1 - sM);
        m[1] = GAP_OPEN_PROBABILITY * (1 - sM);
        m[2] = m[1];
        m[3] = (1 - GAP_EXTENSION_PROBABILITY) * (1 - sI);
        m[4] = GAP_EXTENSION_PROBABILITY * (1 - sI);
        m[5] = 0.0;
        m[6] = 1 - GAP_EXTENSION_PROBABILITY;
        m[7] = 0.0;
        m[8] = GAP_EXTENSION_PROBABILITY;
        const size_t msz = (size_t)(matrix_rows + 1) * (bandWidth2 * 3 + 6);
        F.assign(msz, 0.0);
        B.assign(msz, 0.0);
        S.assign(matrix_rows + 2, 0.0);
    }

    void forward()
    {
        int i, k;
        F[off(0, set_u(bandWidth, 0, 0))] = 1.0;
        S[0] = 1.0;
        {
            double *fi = &F[off(1)];
            double sum;
            int beg = 1, end = referenceLength < bandWidth + 1 ? referenceLength : bandWidth + 1;
            sum = 0.0;
            for (k = beg; k <= end; ++k) {
                double e = calcEpsilon(refb[k - 1], qb[0], qual[0]);
                int u = set_u(bandWidth, 1, k);
                fi[u + 0] = e * bM;
                fi[u + 1] = EI * bI;
                sum += fi[u] + fi[u + 1];
            }
            S[1] = sum;
            int _beg = set_u(bandWidth, 1, beg);
            int _end = set_u(bandWidth, 1, end) + 2;
            for (k = _beg; k <= _end; ++k) fi[k] /= sum;
        }
        for (i = 2; i <= queryLen; ++i) {
            double *fi = &F[off(i)];
            double *fi1 = &F[off(i - 1)];
            double sum;
            int beg = 1, end = referenceLength, x;
            uint8 qyi = qb[i - 1];
            x = i - bandWidth; beg = beg > x ? beg : x;
            x = i + bandWidth; end = end < x ? end : x;
            sum = 0.0;
            for (k = beg; k <= end; ++k) {
                double e = calcEpsilon(refb[k - 1], qyi, qual[i - 1]);
                int u = set_u(bandWidth, i, k);
                int v11 = set_u(bandWidth, i - 1, k - 1);
                int v10 = set_u(bandWidth, i - 1, k);
                int v01 = set_u(bandWidth, i, k - 1);
                fi[u + 0] = e * (m[0] * fi1[v11 + 0] + m[3] * fi1[v11 + 1] + m[6] * fi1[v11 + 2]);
                fi[u + 1] = EI * (m[1] * fi1[v10 + 0] + m[4] * fi1[v10 + 1]);
                fi[u + 2] = m[2] * fi[v01 + 0] + m[8] * fi[v01 + 2];
                sum += fi[u] + fi[u + 1] + fi[u + 2];
            }
            S[i] = sum;
            int _beg = set_u(bandWidth, i, beg);
            int _end = set_u(bandWidth, i, end) + 2;
            for (k = _beg, sum = 1. / sum; k <= _end; ++k) fi[k] *= sum;
        }
        {
            double sum = 0.0;
            for (k = 1; k <= referenceLength; ++k) {
                int u = set_u(bandWidth, queryLen, k);
                if (u < 3 || u >= bandWidth2 * 3 + 3) continue;
                sum += F[off(queryLen, u + 0)] * sM + F[off(queryLen, u + 1)] * sI;
            }
            S[queryLen + 1] = sum;
        }
    }

    void backward()
    {
        int i, k;
        for (k = 1; k <= referenceLength; ++k) {
            int u = set_u(bandWidth, queryLen, k);
            double *bi = &B[off(queryLen)];
            if (u < 3 || u >= bandWidth2 * 3 + 3) continue;
            bi[u + 0] = sM / S[queryLen] / S[queryLen + 1];
            bi[u + 1] = sI / S[queryLen] / S[queryLen + 1];
        }
        for (i = queryLen - 1; i >= 1; --i) {
            int beg = 1, end = referenceLength, x, _beg, _end;
            double *bi = &B[off(i)];
            double *bi1 = &B[off(i + 1)];
            double y = (i > 1) ? 1. : 0.;
            uint8 qyi1 = qb[i];
            x = i - bandWidth; beg = beg > x ? beg : x;
            x = i + bandWidth; end = end < x ? end : x;
            for (k = end; k >= beg; --k) {
                int u = set_u(bandWidth, i, k);
                int v11 = set_u(bandWidth, i + 1, k + 1);
                int v10 = set_u(bandWidth, i + 1, k);
                int v01 = set_u(bandWidth, i, k + 1);
                double e;
                if (k >= referenceLength) e = 0;
                else e = calcEpsilon(refb[k], qyi1, qual[i]) * bi1[v11];
                bi[u + 0] = e * m[0] + EI * m[1] * bi1[v10 + 1] + m[2] * bi[v01 + 2];
                bi[u + 1] = e * m[3] + EI * m[4] * bi1[v10 + 1];
                bi[u + 2] = (e * m[6] + m[8] * bi[v01 + 2]) * y;
            }
            _beg = set_u(bandWidth, i, beg);
            _end = set_u(bandWidth, i, end) + 2;
            y = 1.0 / S[i];
            for (k = _beg; k <= _end; ++k) bi[k] *= y;
        }
        {
            int beg = 1, end = referenceLength < bandWidth + 1 ? referenceLength : bandWidth + 1;
            double sum = 0.0;
            for (k = end; k >= beg; --k) {
                int u = set_u(bandWidth, 1, k);
                double e = calcEpsilon(refb[k - 1], qb[0], qual[0]);
                if (u < 3 || u >= bandWidth2 * 3 + 3) continue;
                sum += e * B[off(1, u + 0)] * bM + EI * B[off(1, u + 1)] * bI;
            }
            B[off(0, set_u(bandWidth, 0, 0))] = sum / S[0];
        }
    }

    void map(uint8 *outq, uint32 *outs, double *outreal)
    {
        int i, k;
        for (i = 1; i <= queryLen; ++i) {
            double sum = 0.0, max = 0.0;
            const double *fi = &F[off(i)];
            const double *bi = &B[off(i)];
            int beg = 1, end = referenceLength, x, max_k = -1;
            x = i - bandWidth; beg = beg > x ? beg : x;
            x = i + bandWidth; end = end < x ? end : x;
            for (k = beg; k <= end; ++k) {
                const int u = set_u(bandWidth, i, k);
                double z;
                z = fi[u + 0] * bi[u + 0];
                sum += z;
                if (z > max) { max = z; max_k = (k - 1) << 2 | 0; }
                z = fi[u + 1] * bi[u + 1];
                sum += z;
                if (z > max) { max = z; max_k = (k - 1) << 2 | 1; }
            }
            max /= sum;
            if (outs) outs[i - 1] = (uint32)max_k;
            const double real = double(-4.343) * log(double(1.0) - double(max)) + double(.499);
            if (outreal) outreal[i - 1] = real;
            outq[i - 1] = (uint8)fpo_baq_map_quality(max);
        }
    }
};

int fpo_baq_map_quality(double max)
{
    // baq.cu:673-674; (int) of +inf follows x86 cvttsd2si (INT_MIN), see D5
    const double real = double(-4.343) * log(double(1.0) - double(max)) + double(.499);
    int k;
    if (!(real < 2147483648.0) || !(real > -2147483649.0)) k = INT32_MIN;
    else k = (int)real;
    return (char)(k > 100 ? 99 : (k < MIN_BASE_QUAL ? MIN_BASE_QUAL : k));
}

void fpo_hmm_glocal(const uint8_t *ref, int ref_len, const uint8_t *query, const uint8_t *qual,
                    int query_len, int bandwidth, uint8_t *out_qual, uint32_t *out_state)
{
    Hmm h;
    h.refb = ref; h.qb = query; h.qual = qual;
    h.setup(ref_len, ref_len, query_len, bandwidth, query_len);
    h.forward();
    h.backward();
    h.map(out_qual, out_state, nullptr);
}

static void baq_reads(fpo_ctx &c)
{
    const fp_batch_view *b = c.b;
    c.baq_qual.assign(b->num_qualities, 0xff);                           // baq.cu:1100-1101
    c.baq_state.assign(b->num_qualities, uint32(-1));
    c.baq_real.assign(b->num_qualities, NAN);

    std::vector<uint32> baq_list;                                         // read_needs_baq (baq.cu:828-840)
    for (uint32 r : c.active_read_list) if (c.num_errors[r] != 0) baq_list.push_back(r);

    if (!baq_list.empty()) {
        // compute_hmm_windows over the full active list (baq.cu:89-187)
        for (uint32 r : c.active_read_list) {
            const bool needs = c.num_errors[r] != 0;
            const ushort2 rwc = c.read_window_clipped[r];
            const ushort2 ni = c.read_window_noins[r];
            const ushort2 fwc = c.reference_window_clipped[r];
            const int offset = MIN_BAND_WIDTH / 2;
            const bool left_clip = (rwc.x != 0);
            const bool right_clip = (rwc.y != b->read_len[r] - 1);
            const short left_ins = (short)(left_clip ? 0 : ni.x - rwc.x);
            const short right_ins = (short)(right_clip ? 0 : rwc.y - ni.y);
            short2 hw;
            hw.x = (int16)(fwc.x - left_ins - offset);
            hw.y = (int16)(fwc.y + right_ins + offset);
            const Contig &ref = c.reference[b->chromosome[r]];
            int left_shift = 0;
            if (hw.x < 0 && b->alignment_start[r] < uabs(hw.x)) left_shift = -1;
            if (needs && (uint32)(b->alignment_start[r] + hw.y + left_shift) >= ref.length) {
                const uint32 es = c.cigar_offsets[b->cigar_start[r]];
                const uint32 ee = c.cigar_offsets[b->cigar_start[r] + b->cigar_len[r]];
                for (uint32 i = es; i < ee; i++) { c.ev_read_index[i] = uint32(-1); c.ev_read_coord[i] = U16_NONE; }
                c.hmm_windows[r].x = -1; c.hmm_windows[r].y = -1;
                continue;
            }
            c.hmm_windows[r] = hw;
            int refLen = hw.y - hw.x + 1;
            int qLen = rwc.y - rwc.x + 1;
            uint16 bw = (uint16)std::max(refLen, qLen);
            if (MIN_BAND_WIDTH < uabs(refLen - qLen)) bw = (uint16)(uabs(refLen - qLen) + 3);
            if (bw > MIN_BAND_WIDTH) bw = MIN_BAND_WIDTH;
            if (bw < uabs(refLen - qLen)) bw = (uint16)uabs(refLen - qLen);
            c.bandwidth[r] = bw;
        }
        auto valid = [&](uint32 r) { return !(c.hmm_windows[r].x == -1 && c.hmm_windows[r].y == -1); };
        std::vector<uint32> t;
        for (uint32 r : baq_list) if (valid(r)) t.push_back(r);
        baq_list.swap(t);
        if (!baq_list.empty()) {
            t.clear();
            for (uint32 r : c.active_read_list) if (valid(r)) t.push_back(r);
            c.active_read_list.swap(t);
        }
    }

    // hmm_glocal per BAQ read (baq.cu:260-358 setup, 374-681)
    for (uint32 r : baq_list) {
        const short2 hw = c.hmm_windows[r];
        const ushort2 rwc = c.read_window_clipped[r];
        const int refLenWindow = hw.y - hw.x + 1;
        int refLen = refLenWindow;
        const int qStart = rwc.x;
        const int qLen = rwc.y - rwc.x + 1;
        const int bw = c.bandwidth[r];
        uint32 hoff;
        if (hw.x < 0 && b->alignment_start[r] < uabs(hw.x)) { hoff = 0; refLen--; }
        else hoff = b->alignment_start[r] + hw.x;
        const Contig &ref = c.reference[b->chromosome[r]];
        std::vector<uint8> refb(std::max(refLen, 0) + 1), qb(qLen), ql(qLen);
        for (int k = 0; k < refLen; k++) refb[k] = ref.base(hoff + k);
        for (int k = 0; k < qLen; k++) {
            qb[k] = c.read_base(b->read_start[r] + qStart + k);
            ql[k] = b->qualities[b->qual_start[r] + qStart + k];
        }
        Hmm h;
        h.refb = refb.data(); h.qb = qb.data(); h.qual = ql.data();
        h.setup(refLenWindow, refLen, qLen, bw, (int)b->read_len[r]);
        h.forward();
        h.backward();
        const uint32 qo = b->qual_start[r] + qStart;
        h.map(&c.baq_qual[qo], &c.baq_state[qo], &c.baq_real[qo]);
    }

    // read_flat_baq (baq.cu:842-868)
    for (uint32 r : c.active_read_list) {
        if (c.num_errors[r] != 0) continue;
        const ushort2 rw = c.read_window_clipped[r];
        const uint32 qs = rw.x, ql = rw.y - rw.x + 1;
        memset(&c.baq_qual[b->qual_start[r] + qs], NO_BAQ_UNCERTAINTY, ql);
    }

    // cap_baq_qualities (baq.cu:870-1056)
    for (uint32 r : baq_list) {
        const uint32 cs = c.cigar_offsets[b->cigar_start[r]];
        const uint32 ce = c.cigar_offsets[b->cigar_start[r] + b->cigar_len[r]];
        const ushort2 rwc = c.read_window_clipped[r];
        const ushort2 fwc = c.reference_window_clipped[r];
        const short2 hw = c.hmm_windows[r];
        const uint32 qs = b->qual_start[r];

        uint32 baq_start = 0;
        for (uint32 i = 0; i < ce - cs; i++) {
            const uint16 rb = c.ev_read_coord[cs + i];
            if (rb != U16_NONE && rb >= rwc.x) { baq_start = i; break; }
        }

        auto cap = [&](bool update) -> bool {
            uint32 readI = 0, refI = 0, numD = 0;
            int16 reference_shift, read_shift;
            if (hw.x < 0 && b->alignment_start[r] < uabs(hw.x)) {
                reference_shift = (int16)uabs(hw.x);
                read_shift = (int16)b->alignment_start[r];
            } else {
                reference_shift = 0;
                read_shift = 0;
            }
            const int16 refOffset = (int16)(hw.x - fwc.x + reference_shift);
            for (uint32 i = baq_start; i < ce - cs; i++) {
                const uint16 rb = c.ev_read_coord[cs + i];
                const uint32 qi = qs + rb;
                if (rb != U16_NONE) {
                    if (rb < rwc.x) continue;
                    if (rb > rwc.y) break;
                }
                switch (c.cigar_events[i + cs]) {
                case EV_S: refI++; break;
                case EV_I:
                    if (update) c.baq_qual[qi] = b->qualities[qi];
                    readI++;
                    break;
                case EV_D: refI++; numD++; break;
                case EV_M:
                    if (update) {
                        const uint32 expectedPos = refI - refOffset + (i - numD - baq_start - readI) + read_shift;
                        const uint8 oq = b->qualities[qs + rb];
                        const uint8 bq = c.baq_qual[qs + rb];
                        const uint32 state = c.baq_state[qs + rb];
                        const bool isIndel = (state & 3) != 0;
                        const uint32 pos = state >> 2;
                        uint8 v;
                        if (isIndel || pos != expectedPos) v = MIN_BASE_QUAL;
                        else v = std::min(bq, oq);
                        c.baq_qual[qi] = v;
                    }
                    readI++; refI++;
                    break;
                }
            }
            if (!update) {
                const uint32 rl = rwc.y - rwc.x + 1;
                if (readI != rl) return false;
            }
            return true;
        };

        if (cap(false)) {
            cap(true);
        } else {
            const uint32 bo = qs + rwc.x, len = rwc.y - rwc.x + 1;
            memcpy(&c.baq_qual[bo], &b->qualities[bo], len);
        }
    }

    // recode_baq_qualities (baq.cu:1058-1081)
    for (uint32 r : baq_list) {
        const uint32 qs = b->qual_start[r];
        for (uint32 i = qs; i < qs + b->qual_len[r]; i++) {
            const uint8 bi = c.baq_qual[i];
            if (bi == 0xff) continue;
            const uint8 bq = (uint8)(b->qualities[i] + 64);
            c.baq_qual[i] = (uint8)(bq - bi);
        }
    }
    c.stats.baq_reads += baq_list.size();
}

// ===========================================================================
// fractional errors (fractional_errors.cu:39-117)
void fpo_fractional_errors(const uint8_t *baqArray, const uint8_t *errorArray, int len, double *fe)
{
    bool inBlock = false;
    int blockStart = -1;
    int iii;
    auto block = [&](int end, int bs) {
        int total = 0;
        const int s = std::max(0, bs - 1);
        for (int j = s; j <= end; j++) total += errorArray[j];
        for (int j = s; j <= end; j++) fe[j] = ((double)total) / ((double)(end - s + 1));
    };
    for (iii = 0; iii < len; iii++) {
        if (baqArray[iii] == NO_BAQ_UNCERTAINTY) {
            if (!inBlock) {
                fe[iii] = (double)errorArray[iii];
            } else {
                block(iii, blockStart);
                inBlock = false;
                blockStart = -1;
            }
        } else {
            inBlock = true;
            if (blockStart == -1) blockStart = iii;
        }
    }
    if (inBlock) block(iii - 1, blockStart);
}

static void build_fractional_errors(fpo_ctx &c)
{
    const fp_batch_view *b = c.b;
    for (int t = 0; t < 3; t++) c.frac[t].assign(b->num_qualities, 0.0);
    const std::vector<uint8> *errs[3] = {&c.is_snp, &c.is_ins, &c.is_del};
    for (uint32 r : c.active_read_list) {
        const ushort2 rw = c.read_window_clipped[r];
        const int len = rw.y - rw.x + 1;
        for (int t = 0; t < 3; t++)
            fpo_fractional_errors(&c.baq_qual[b->qual_start[r] + rw.x],
                                  &(*errs[t])[b->read_start[r] + rw.x], len,
                                  &c.frac[t][b->qual_start[r] + rw.x]);
    }
}

// ===========================================================================
// covariates (covariates.cu, covariates/*)

// iupac16_to_dna (from_nvbio/dna.h:58-65)
static uint32 iupac16_to_dna(uint8 c) { return 31 - __builtin_clz((uint32)c); }

struct KeySet { uint32 M, I, D; };

// covariate_Cycle_Illumina::encode (bit_packers/cycle_illumina.h:46-95)
static KeySet cycle_field(const fpo_ctx &c, uint32 r, uint16 bp)
{
    const uint32 flags = c.b->flags[r];
    const bool paired = flags & F_PAIRED, second = flags & F_READ2, neg = flags & F_REVERSE;
    const ushort2 w = c.read_window_clipped[r];
    const int readLength = w.y - w.x + 1;
    const int readOrderFactor = (paired && second) ? -1 : 1;
    const int i = bp - w.x;
    int cycle, inc;
    if (neg) { cycle = readLength * readOrderFactor; inc = -1 * readOrderFactor; }
    else { cycle = readOrderFactor; inc = readOrderFactor; }
    cycle += i * inc;
    const int MAX_CYCLE_FOR_INDELS = readLength - 4 - 1;
    uint32 key = uabs(cycle);
    key = key << 1;
    if (cycle < 0) key++;
    const uint32 indel = (i < 4 || i > MAX_CYCLE_FOR_INDELS) ? 1023u : key;
    return KeySet{key, indel, indel};
}

// covariate_Context::encode (bit_packers/context.h:104-174), mismatch 2, indel 3
static KeySet context_field(const fpo_ctx &c, uint32 r, uint16 bp)
{
    const ushort2 w = c.hq_window[r];
    const bool neg = c.b->flags[r] & F_REVERSE;
    uint32 context = 0, cm = 2047, ci = 2047;
    const int direction = neg ? 1 : -1;
    const int context_bases = neg ? std::min<int>(w.y - bp + 1, 3) : std::min<int>(bp - w.x + 1, 3);
    const int start = bp;
    const int stop = bp + context_bases * direction;
    if (context_bases >= 2) {
        int n = 0;
        for (int i = start; i != stop; i += direction) {
            const uint8 base = c.read_base(c.b->read_start[r] + i);
            if (base != IUPAC_A && base != IUPAC_C && base != IUPAC_T && base != IUPAC_G) break;
            context <<= 2;
            context |= iupac16_to_dna(base);
            n++;
        }
        if (neg) context = ~context & ((1u << (n * 2)) - 1);
        if (n >= 2) { cm = context >> (n - 2) * 2; cm = (cm << 4) | 2; }
        if (n >= 3) { ci = context >> (n - 3) * 2; ci = (ci << 4) | 3; }
    }
    return KeySet{cm, ci, ci};
}

// key chains (packer_quality_score.h:42-57, packer_cycle_illumina.h:42-60, packer_context.h:42-74)
static KeySet encode_key(const fpo_ctx &c, int table, uint32 r, uint16 bp)
{
    const uint32 rg = c.b->read_group[r];
    const uint32 q = c.b->qualities[c.b->qual_start[r] + bp];
    KeySet k;
    if (table == 0) {
        k.M = (0u) | (q << 2) | (rg << 10);
        k.I = (1u) | (45u << 2) | (rg << 10);
        k.D = (2u) | (45u << 2) | (rg << 10);
    } else if (table == 1) {
        KeySet f = cycle_field(c, r, bp);
        k.M = 0u | (f.M << 2) | (q << 12) | (rg << 20);
        k.I = 1u | (f.I << 2) | (45u << 12) | (rg << 20);
        k.D = 2u | (f.D << 2) | (45u << 12) | (rg << 20);
    } else {
        KeySet f = context_field(c, r, bp);
        k.M = 0u | (f.M << 2) | (q << 13) | (rg << 21);
        k.I = 1u | (f.I << 2) | (45u << 13) | (rg << 21);
        k.D = 2u | (f.D << 2) | (45u << 13) | (rg << 21);
    }
    return k;
}

static bool key_valid(int table, uint32 key)
{
    if (key == 0xffffffffu) return false;
    if (table == 1 && ((key >> 2) & 1023) == 1023) return false;
    if (table == 2 && ((key >> 2) & 2047) == 2047) return false;
    return true;
}

static void gather_covariates(fpo_ctx &c)
{
    const fp_batch_view *b = c.b;
    // compute_high_quality_windows (covariates.cu:215-246)
    for (uint32 r : c.active_read_list) {
        ushort2 w = c.read_window_clipped[r];
        const uint32 qs = b->qual_start[r];
        while (b->qualities[qs + w.x] <= 2 && w.x < w.y) w.x++;
        while (b->qualities[qs + w.y] <= 2 && w.y > w.x) w.y--;
        c.hq_window[r] = w;
    }

    const uint32 E = (uint32)c.cigar_events.size();
    for (int t = 0; t < 3; t++) {
        std::vector<FloatRow> fresh;
        // covariate_gatherer (covariates.cu:50-103) in event order
        for (uint32 ev = 0; ev < E; ev++) {
            const uint32 r = c.ev_read_index[ev];
            if (r == uint32(-1)) continue;
            const uint16 bp = c.ev_read_coord[ev];
            if (bp == U16_NONE) continue;
            if (bp < c.read_window_clipped[r].x || bp > c.read_window_clipped[r].y) continue;
            if (c.active_loc[b->read_start[r] + bp] == 0) continue;
            if (c.cigar_events[ev] == EV_S) continue;
            const KeySet k = encode_key(c, t, r, bp);
            const uint32 qi = b->qual_start[r] + bp;
            const uint32 keys[3] = {k.M, k.I, k.D};
            for (int e = 0; e < 3; e++) {
                if (!key_valid(t, keys[e])) continue;
                const double v = c.frac[e][qi];
                if (c.accum_mode == FPO_ACCUM_EXACT) {
                    ExactAcc &a = c.exact[t][keys[e]];
                    a.obs += 1;
                    a.mism += c.frac_error_f64 ? to_fixed(v) : to_fixed((double)(float)v);
                } else {
                    fresh.push_back(FloatRow{keys[e], 1, (float)v});
                }
            }
        }
        if (c.accum_mode == FPO_ACCUM_REF_FLOAT && !fresh.empty()) {
            // append, stable sort, sequential fold (covariates.cu:172-206, covariate_table.cu:43-105)
            std::vector<FloatRow> &tab = c.flt[t];
            tab.insert(tab.end(), fresh.begin(), fresh.end());
            std::stable_sort(tab.begin(), tab.end(), [](const FloatRow &a, const FloatRow &b) { return a.key < b.key; });
            std::vector<FloatRow> packed;
            for (const FloatRow &row : tab) {
                if (!packed.empty() && packed.back().key == row.key) {
                    packed.back().obs += row.obs;
                    packed.back().mism = packed.back().mism + row.mism;
                } else packed.push_back(row);
            }
            tab.swap(packed);
        }
    }
}

// ===========================================================================
// postprocess (read_group_table.cu, covariates.cu:314-339, empirical_quality.cu)

static double log10QempPrior(double Qempirical, double Qreported, bool need_rounding)
{
    double delta = Qempirical - Qreported;
    if (need_rounding) delta = round(delta);
    int difference = std::min(abs(int(delta)), 40);
    // gaussian (empirical_quality.cu:44-54)
    const double v = double(difference);
    const double g = 0.0 + 0.9 * pow(double(M_E), -(pow(v - 0.0, 2.0) / (2 * 0.5 * 0.5)));
    return log10(g);
}

static double log10Gamma(double x) { return (double)lgammaf((float)x) * log10(M_E); }
static double log10Factorial(uint64 x) { return log10Gamma((double)(x + 1)); }

static double log10BinomialProbability(uint64 n, uint64 k, double log10p)
{
    double log10OneMinusP = log10(1 - pow(10.0, log10p));
    double coeff = log10Factorial(n) - log10Factorial(k) - log10Factorial(n - k);
    return coeff + log10p * k + log10OneMinusP * (n - k);
}

static double log10QempLikelihood(double Qempirical, uint64 nObs, uint64 nErr)
{
    if (nObs == 0) return 0.0;
    const uint32 MAXOBS = 2147483647u - 1;
    if (nObs > MAXOBS) {
        double fraction = double(MAXOBS) / double(nObs);
        nErr = (uint64)round(double(nErr) * fraction);
        nObs = MAXOBS;
    }
    return log10BinomialProbability(nObs, nErr, Qempirical / -10.0);
}

double fpo_empirical_quality(uint64_t obs, double mism, double expected, int need_rounding, double *est_out)
{
    // calc_empirical_quality (empirical_quality.cu:179-189)
    const double estimated = double(-10.0 * log10(expected / double(obs)));
    if (est_out) *est_out = estimated;
    const uint64 mismatches = uint64(mism + 0.5) + 1;
    const uint64 observations = obs + 1 + 1;
    double post[61];
    for (int bin = 0; bin < 61; bin++) {
        const double q = bin / 1.0;
        post[bin] = log10QempPrior(q, estimated, need_rounding) + log10QempLikelihood(q, observations, mismatches);
    }
    int mle = 0;
    for (int i = 1; i < 61; i++) if (post[i] > post[mle]) mle = i;
    return std::min((double)mle, 93.0);
}

double fpo_round_n(double val, int n)
{
    // util.cu:107-116
    val = val * pow(10.0, n);
    val = round(val);
    val = val / pow(10.0, n);
    return val;
}

static void postprocess(fpo_ctx &c)
{
    // observation tables -> empirical rows (covariate_table.cu:109-136)
    std::vector<EmpRow> obs[3];
    for (int t = 0; t < 3; t++) {
        if (c.accum_mode == FPO_ACCUM_EXACT) {
            for (auto &kv : c.exact[t])
                obs[t].push_back(EmpRow{kv.first, kv.second.obs, fixed_to_double(kv.second.mism), 0, 0, 0, kv.second.mism});
        } else {
            for (auto &row : c.flt[t]) obs[t].push_back(EmpRow{row.key, row.obs, (double)row.mism, 0, 0, 0, 0});
        }
    }

    auto qual_of = [](int t, uint32 key) -> uint32 {
        if (t == 0) return (key >> 2) & 255;
        if (t == 1) return (key >> 12) & 255;
        return (key >> 13) & 255;
    };

    // build_read_group_table (read_group_table.cu:62-93)
    c.emp[0].clear();
    if (!obs[0].empty()) {
        std::vector<EmpRow> rg = obs[0];
        for (auto &row : rg) {
            const uint8 q = (uint8)qual_of(0, row.key);
            row.expected = row.obs * pow(10.0, q / -10.0);
            row.key &= ~(255u << 2);
        }
        std::stable_sort(rg.begin(), rg.end(), [](const EmpRow &a, const EmpRow &b) { return a.key < b.key; });
        for (const EmpRow &row : rg) {
            if (!c.emp[0].empty() && c.emp[0].back().key == row.key) {
                EmpRow &a = c.emp[0].back();
                a.obs += row.obs;
                a.mism = a.mism + row.mism;
                a.expected = a.expected + row.expected;
            } else {
                EmpRow a = row;
                a.estimated = 0; a.empirical = 0; a.mism_fixed = 0;
                c.emp[0].push_back(a);
            }
        }
        for (auto &row : c.emp[0])
            row.empirical = fpo_empirical_quality(row.obs, row.mism, row.expected, 0, &row.estimated);
    }

    // compute_empirical_quality_scores (covariates.cu:314-339)
    for (int t = 0; t < 3; t++) {
        c.emp[t + 1] = obs[t];
        for (auto &row : c.emp[t + 1]) {
            const uint8 q = (uint8)qual_of(t, row.key);
            row.expected = row.obs * pow(10.0, q / -10.0);
            row.empirical = fpo_empirical_quality(row.obs, row.mism, row.expected, 1, &row.estimated);
        }
    }
    c.postprocessed = true;
}

// GATK report (pipeline.cu:150-204, read_group_table.cu:96-143, covariates.cu:285-311)
static std::string make_report(const fpo_ctx &c)
{
    std::string s;
    const bool dr = c.disable_rounding;
    s += "#:GATKReport.v1.1:5\n";
    s += "#:GATKTable:2:18:%s:%s:;\n";
    s += "#:GATKTable:Arguments:Recalibration argument collection values used in this run\n";
    s += "Argument                    Value                                                                   \n";
    s += "binary_tag_name             null                                                                    \n";
    s += "covariate                   ReadGroupCovariate,QualityScoreCovariate,ContextCovariate,CycleCovariate\n";
    s += "default_platform            null                                                                    \n";
    s += "deletions_default_quality   45                                                                      \n";
    s += "force_platform              null                                                                    \n";
    s += "indels_context_size         3                                                                       \n";
    s += "insertions_default_quality  45                                                                      \n";
    s += "low_quality_tail            2                                                                       \n";
    s += "maximum_cycle_value         500                                                                     \n";
    s += "mismatches_context_size     2                                                                       \n";
    s += "mismatches_default_quality  -1                                                                      \n";
    s += "no_standard_covs            false                                                                   \n";
    s += "quantizing_levels           16                                                                      \n";
    s += "recalibration_report        null                                                                    \n";
    s += "run_without_dbsnp           false                                                                   \n";
    s += "solid_nocall_strategy       THROW_EXCEPTION                                                         \n";
    s += "solid_recal_mode            SET_Q_ZERO                                                              \n";
    {
        char buf[256];
        snprintf(buf, sizeof(buf), "software                    Firepony_%d.%d.%d                                                       \n", 1, 1, 1);
        s += buf;
    }
    s += "\n";
    {
        Formatter f(&s, dr, "Quantized", "Quality quantization map");
        f.add_column("QualityScore", Formatter::STRING, Formatter::RIGHT);
        f.add_column("Count", Formatter::STRING, Formatter::RIGHT);
        f.add_column("QuantizedScore", Formatter::STRING, Formatter::RIGHT);
        for (int pass = 0; pass < 2; pass++) {
            for (int i = 0; i < 94; i++) {
                char buf[64];
                f.start_row();
                snprintf(buf, sizeof(buf), "%d", i);
                f.data_str(buf);
                f.data_str("0");
                f.data_str(buf);
                f.end_row();
            }
            f.end_table();
        }
    }
    auto rg_name = [&](uint32 id) -> std::string {
        return id < c.rg_names.size() ? c.rg_names[id] : std::string("<invalid>");
    };
    auto evc = [](uint32 e) -> char { return e == 0 ? 'M' : e == 1 ? 'I' : e == 2 ? 'D' : e == 3 ? 'S' : '?'; };
    {
        Formatter f(&s, dr, "RecalTable0");
        f.add_column("ReadGroup", Formatter::STRING);
        f.add_column("EventType", Formatter::CHAR);
        f.add_column("EmpiricalQuality", Formatter::FLOAT_4);
        f.add_column("EstimatedQReported", Formatter::FLOAT_4);
        f.add_column("Observations", Formatter::UINT64);
        f.add_column("Errors", Formatter::FLOAT_2, Formatter::RIGHT, Formatter::LEFT);
        for (int pass = 0; pass < 2; pass++) {
            for (const EmpRow &r : c.emp[0]) {
                f.start_row();
                f.data_str(rg_name((r.key >> 10) & 255));
                f.data_char(evc(r.key & 3));
                f.data_double(r.empirical);
                f.data_double(r.estimated);
                f.data_u64(r.obs);
                f.data_double(r.mism);
                f.end_row();
            }
            f.end_table();
        }
    }
    {
        Formatter f(&s, dr, "RecalTable1");
        f.add_column("ReadGroup", Formatter::STRING);
        f.add_column("QualityScore", Formatter::STRING);
        f.add_column("EventType", Formatter::CHAR);
        f.add_column("EmpiricalQuality", Formatter::FLOAT_4);
        f.add_column("Observations", Formatter::UINT64);
        f.add_column("Errors", Formatter::FLOAT_2, Formatter::RIGHT, Formatter::LEFT);
        for (int pass = 0; pass < 2; pass++) {
            for (const EmpRow &r : c.emp[1]) {
                if (r.obs == 0) continue;
                char qs[64];
                snprintf(qs, sizeof(qs), "%d", (uint8)((r.key >> 2) & 255));
                f.start_row();
                f.data_str(rg_name((r.key >> 10) & 255));
                f.data_str(qs);
                f.data_char(evc(r.key & 3));
                f.data_double(r.empirical);
                f.data_u64(r.obs);
                f.data_double(r.mism);
                f.end_row();
            }
            f.end_table();
        }
    }
    {
        Formatter f(&s, dr, "RecalTable2");
        f.add_column("ReadGroup", Formatter::STRING);
        f.add_column("QualityScore", Formatter::STRING);
        f.add_column("CovariateValue", Formatter::STRING);
        f.add_column("CovariateName", Formatter::STRING);
        f.add_column("EventType", Formatter::CHAR);
        f.add_column("EmpiricalQuality", Formatter::FLOAT_4);
        f.add_column("Observations", Formatter::UINT64);
        f.add_column("Errors", Formatter::FLOAT_2, Formatter::RIGHT, Formatter::LEFT);
        for (int pass = 0; pass < 2; pass++) {
            for (const EmpRow &r : c.emp[3]) {   // context rows (packer_context.h:83-128)
                if (r.obs == 0) continue;
                const uint32 raw = (r.key >> 2) & 2047;
                const int size = raw & 15;
                const uint32 ctx = raw >> 4;
                std::string seq;
                for (int j = 0; j < size; j++) seq += "ACGT"[(ctx >> (j * 2)) & 3];
                f.start_row();
                f.data_str(rg_name((r.key >> 21) & 255));
                f.data_int_as_string((uint8)((r.key >> 13) & 255));
                f.data_str(seq);
                f.data_str("Context");
                f.data_char(evc(r.key & 3));
                f.data_double(r.empirical);
                f.data_u64(r.obs);
                f.data_double(r.mism);
                f.end_row();
            }
            for (const EmpRow &r : c.emp[2]) {   // cycle rows (packer_cycle_illumina.h:70-110)
                if (r.obs == 0) continue;
                const uint32 raw = (r.key >> 2) & 1023;
                int group = raw >> 1;
                if (raw & 1) group = -group;
                f.start_row();
                f.data_str(rg_name((r.key >> 20) & 255));
                f.data_int_as_string((uint8)((r.key >> 12) & 255));
                f.data_int_as_string(group);
                f.data_str("Cycle");
                f.data_char(evc(r.key & 3));
                f.data_double(r.empirical);
                f.data_u64(r.obs);
                f.data_double(r.mism);
                f.end_row();
            }
            f.end_table();
        }
    }
    return s;
}

// ===========================================================================
// C API

extern "C" {

fpo_ctx *fpo_create(int frac_error_f64, int accumulate_mode, int disable_output_rounding)
{
    fpo_ctx *c = new fpo_ctx();
    c->frac_error_f64 = frac_error_f64;
    c->accum_mode = accumulate_mode;
    c->disable_rounding = disable_output_rounding;
    memset(&c->stats, 0, sizeof(c->stats));
    return c;
}

int fpo_set_read_group_names(fpo_ctx *c, uint32_t n, const char *const *names)
{
    if (!c || n > c->rg_names.size() || (n && !names)) return FP_ERR_INVALID_ARG;
    for (uint32 i = 0; i < n; i++) c->rg_names[i] = names[i] ? names[i] : "";
    return FP_OK;
}

int fpo_setup(fpo_ctx *c, const fp_header_view *h, const fp_reference_view *r, const fp_variants_view *v)
{
    if (!c || !h || !r || !v) return FP_ERR_INVALID_ARG;
    c->chromosome_lengths.assign(h->chromosome_lengths, h->chromosome_lengths + h->num_chromosomes);
    c->rg_names.clear();
    for (uint32 i = 0; i < h->num_read_groups; i++) c->rg_names.push_back(h->read_group_names[i]);
    c->reference.assign(r->num_sequences, Contig());
    for (uint32 s = 0; s < r->num_sequences; s++) {
        Contig &ct = c->reference[s];
        ct.length = r->length[s];
        const uint64 nw = (ct.length + 7) / 8;
        ct.words.assign(r->words + r->word_offset[s], r->words + r->word_offset[s] + nw);
    }
    c->variants.assign(v->num_sequences, VariantContig());
    for (uint32 s = 0; s < v->num_sequences; s++) {
        VariantContig &vc = c->variants[s];
        const uint64 a = v->feature_offset[s], e = v->feature_offset[s + 1];
        vc.feature_start.assign(v->feature_start + a, v->feature_start + e);
        vc.feature_stop.assign(v->feature_stop + a, v->feature_stop + e);
        vc.max_end_point_left.resize(e - a);
        uint32 m = 0;
        for (uint64 i = 0; i < e - a; i++) {           // loader/variants.cu:125-134
            m = i == 0 ? vc.feature_stop[0] : std::max(m, vc.feature_stop[i]);
            vc.max_end_point_left[i] = m;
        }
    }
    c->is_setup = true;
    return FP_OK;
}

int fpo_process_batch(fpo_ctx *c, const fp_batch_view *b)
{
    if (!c || !b || !c->is_setup) return FP_ERR_INVALID_ARG;
    c->b = b;
    const uint32 n = b->num_reads;
    c->num_reads = n;
    // start_batch (firepony_context.cu:43-60)
    c->active_read_list.resize(n);
    for (uint32 i = 0; i < n; i++) c->active_read_list[i] = i;
    c->active_loc.assign(b->num_read_bases, 0xf);
    c->read_offset_list.assign(b->num_read_bases, 0);
    c->alignment_windows.assign(n, uint2{0, 0});
    c->read_window_clipped.assign(n, mk_us2(0, 0));
    c->read_window_noins.assign(n, mk_us2(0, 0));
    c->reference_window_clipped.assign(n, mk_us2(0, 0));
    c->is_snp.assign(b->num_read_bases, 0);
    c->is_ins.assign(b->num_read_bases, 0);
    c->is_del.assign(b->num_read_bases, 0);
    c->vcf_ranges.assign(n, uint2{0, 0});
    c->hmm_windows.assign(n, short2{0, 0});
    c->bandwidth.assign(n, 0);
    c->hq_window.assign(n, mk_us2(0, 0));
    c->num_errors.assign(n, 0);
    c->cigar_events.clear();
    c->ev_read_index.clear();
    c->ev_read_coord.clear();
    c->ev_ref_coord.clear();
    c->baq_qual.assign(b->num_qualities, 0xff);
    for (int t = 0; t < 3; t++) c->frac[t].assign(b->num_qualities, 0.0);

    // firepony_process_batch (pipeline.cu:56-148)
    filter_invalid_reads(*c);
    if (!c->active_read_list.empty()) {
        build_read_offset_list(*c);
        build_alignment_windows(*c);
        filter_malformed_reads(*c);
    }
    if (!c->active_read_list.empty()) {
        expand_cigars(*c);
        filter_bases(*c);
        filter_known_snps(*c);
        baq_reads(*c);
        build_fractional_errors(*c);
        gather_covariates(*c);
    }
    c->stats.total_reads += n;
    c->stats.num_batches++;
    c->postprocessed = false;
    c->b = nullptr;
    return FP_OK;
}

// gather_intermediates (pipeline_interface.cu:52-58,137-160) followed by the covariate re-pack of
// postprocess_covariates (covariates.cu:265-282): with exact sums the merge is a per-key add
int fpo_gather(fpo_ctx *dst, fpo_ctx *src)
{
    if (!dst || !src) return FP_ERR_INVALID_ARG;
    if (dst->accum_mode != FPO_ACCUM_EXACT || src->accum_mode != FPO_ACCUM_EXACT) return FP_ERR_UNSUPPORTED;
    for (int t = 0; t < 3; t++) {
        for (const auto &kv : src->exact[t]) {
            ExactAcc &a = dst->exact[t][kv.first];
            a.obs += kv.second.obs;
            a.mism += kv.second.mism;
        }
    }
    dst->stats.total_reads += src->stats.total_reads;
    dst->stats.filtered_reads += src->stats.filtered_reads;
    dst->stats.baq_reads += src->stats.baq_reads;
    dst->stats.num_batches += src->stats.num_batches;
    dst->postprocessed = false;
    return FP_OK;
}

int fpo_postprocess(fpo_ctx *c)
{
    if (!c) return FP_ERR_INVALID_ARG;
    postprocess(*c);
    return FP_OK;
}

int fpo_get_table(fpo_ctx *c, int table_id, fp_table_out *out)
{
    if (!c || !out || table_id < 0 || table_id > 3) return FP_ERR_INVALID_ARG;
    if (c->postprocessed) {
        const std::vector<EmpRow> &t = c->emp[table_id];
        out->count = t.size();
        for (size_t i = 0; i < t.size(); i++) {
            if (out->keys) out->keys[i] = t[i].key;
            if (out->observations) out->observations[i] = t[i].obs;
            if (out->mismatches) out->mismatches[i] = t[i].mism;
            if (out->mismatches_fixed_lo) out->mismatches_fixed_lo[i] = (uint64)t[i].mism_fixed;
            if (out->mismatches_fixed_hi) out->mismatches_fixed_hi[i] = (uint64)(t[i].mism_fixed >> 64);
            if (out->expected_errors) out->expected_errors[i] = t[i].expected;
            if (out->estimated_quality) out->estimated_quality[i] = t[i].estimated;
            if (out->empirical_quality) out->empirical_quality[i] = t[i].empirical;
        }
        return FP_OK;
    }
    if (table_id == 0) { out->count = 0; return FP_OK; }
    const int t = table_id - 1;
    if (c->accum_mode == FPO_ACCUM_EXACT) {
        out->count = c->exact[t].size();
        size_t i = 0;
        for (auto &kv : c->exact[t]) {
            if (out->keys) out->keys[i] = kv.first;
            if (out->observations) out->observations[i] = kv.second.obs;
            if (out->mismatches) out->mismatches[i] = fixed_to_double(kv.second.mism);
            if (out->mismatches_fixed_lo) out->mismatches_fixed_lo[i] = (uint64)kv.second.mism;
            if (out->mismatches_fixed_hi) out->mismatches_fixed_hi[i] = (uint64)(kv.second.mism >> 64);
            i++;
        }
    } else {
        out->count = c->flt[t].size();
        for (size_t i = 0; i < c->flt[t].size(); i++) {
            if (out->keys) out->keys[i] = c->flt[t][i].key;
            if (out->observations) out->observations[i] = c->flt[t][i].obs;
            if (out->mismatches) out->mismatches[i] = c->flt[t][i].mism;
        }
    }
    return FP_OK;
}

int fpo_write_report(fpo_ctx *c, const char *path)
{
    if (!c || !path) return FP_ERR_INVALID_ARG;
    if (!c->postprocessed) postprocess(*c);
    std::string s = make_report(*c);
    FILE *fp = fopen(path, "wb");
    if (!fp) return FP_ERR_IO;
    fwrite(s.data(), 1, s.size(), fp);
    fclose(fp);
    return FP_OK;
}

int fpo_stats(fpo_ctx *c, fp_stats *out)
{
    if (!c || !out) return FP_ERR_INVALID_ARG;
    *out = c->stats;
    return FP_OK;
}

static const void *dbg_ptr(fpo_ctx *c, int item, uint64_t *bytes, std::vector<uint8> &tmp)
{
    switch (item) {
    case FP_DBG_ACTIVE_READS: *bytes = c->active_read_list.size() * 4; return c->active_read_list.data();
    case FP_DBG_READ_WINDOW: *bytes = c->read_window_clipped.size() * 4; return c->read_window_clipped.data();
    case FP_DBG_NOINS_WINDOW: *bytes = c->read_window_noins.size() * 4; return c->read_window_noins.data();
    case FP_DBG_REF_WINDOW: *bytes = c->reference_window_clipped.size() * 4; return c->reference_window_clipped.data();
    case FP_DBG_NUM_ERRORS: *bytes = c->num_errors.size() * 2; return c->num_errors.data();
    case FP_DBG_ERROR_FLAGS:
        tmp.resize(c->is_snp.size());
        for (size_t i = 0; i < tmp.size(); i++) tmp[i] = c->is_snp[i] | (c->is_ins[i] << 1) | (c->is_del[i] << 2);
        *bytes = tmp.size();
        return tmp.data();
    case FP_DBG_ACTIVE_MASK:
        tmp.resize(c->active_loc.size());
        for (size_t i = 0; i < tmp.size(); i++) tmp[i] = c->active_loc[i] ? 1 : 0;
        *bytes = tmp.size();
        return tmp.data();
    case FP_DBG_HMM_WINDOW: *bytes = c->hmm_windows.size() * 4; return c->hmm_windows.data();
    case FP_DBG_BANDWIDTH: *bytes = c->bandwidth.size() * 2; return c->bandwidth.data();
    case FP_DBG_BAQ: *bytes = c->baq_qual.size(); return c->baq_qual.data();
    case FP_DBG_FRAC_ERRORS: {
        const size_t n = c->frac[0].size();
        tmp.resize(n * 3 * 8);
        double *d = (double *)tmp.data();
        for (size_t i = 0; i < n; i++) { d[3 * i] = c->frac[0][i]; d[3 * i + 1] = c->frac[1][i]; d[3 * i + 2] = c->frac[2][i]; }
        *bytes = tmp.size();
        return tmp.data();
    }
    case FP_DBG_HQ_WINDOW: *bytes = c->hq_window.size() * 4; return c->hq_window.data();
    default: *bytes = 0; return nullptr;
    }
}

int fpo_debug_size(fpo_ctx *c, int item, uint64_t *bytes)
{
    if (!c || !bytes) return FP_ERR_INVALID_ARG;
    std::vector<uint8> tmp;
    if (!dbg_ptr(c, item, bytes, tmp) && *bytes) return FP_ERR_INVALID_ARG;
    return FP_OK;
}

int fpo_debug_fetch(fpo_ctx *c, int item, void *dst, uint64_t bytes)
{
    if (!c || !dst) return FP_ERR_INVALID_ARG;
    std::vector<uint8> tmp;
    uint64_t n = 0;
    const void *p = dbg_ptr(c, item, &n, tmp);
    if (bytes < n) return FP_ERR_INVALID_ARG;
    if (n) memcpy(dst, p, n);
    return FP_OK;
}

int fpo_debug_baq_real(fpo_ctx *c, double *dst, uint64_t n)
{
    if (!c || !dst || n < c->baq_real.size()) return FP_ERR_INVALID_ARG;
    memcpy(dst, c->baq_real.data(), c->baq_real.size() * 8);
    return FP_OK;
}

void fpo_destroy(fpo_ctx *c) { delete c; }

} // extern "C"
