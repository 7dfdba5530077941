/*
 * GATKReport v1.1 text for the recalibration tables.
 *
 * Output contract (what GATK's report parser and firepony's diffreport read,
 * device/pipeline.cu:150-204, table_formatter.cu:35-233):
 *   - report header "#:GATKReport.v1.1:5", then the Arguments table as fixed
 *     text (18 argument rows including the software row), a dummy Quantized
 *     table (0..93), RecalTable0, RecalTable1, RecalTable2;
 *   - each table: "#:GATKTable:<cols>:<rows>:<fmt>:...:;", the name line, a
 *     column-name line, one line per row, then an empty line;
 *   - a column is as wide as its widest cell or name; cells are separated by
 *     two spaces; text and event columns are left aligned, numbers right
 *     aligned (the Errors column header is left aligned);
 *   - floats are rounded to 2 or 4 decimals first and printed with %.2f/%.4f,
 *     or printed with %.64f when rounding is disabled.
 * Rows with zero observations are left out of RecalTable1 and RecalTable2;
 * RecalTable2 lists all context rows, then all cycle rows.
 */
#include "report.h"

#include <cmath>
#include <cstdio>

namespace fpr {

double round_decimals(double v, int n)
{
    const double scale = pow(10.0, n);
    v = v * scale;
    v = round(v);
    return v / scale;
}

namespace {

enum Kind { TEXT, EVENT, COUNT, REAL2, REAL4 };

struct Column {
    std::string name;
    Kind kind;
    bool right;          // cell alignment
    bool header_right;   // header alignment
};

struct Table {
    std::string name, description;
    std::vector<Column> cols;
    std::vector<std::vector<std::string>> rows;
};

Column col(const std::string &n, Kind k)
{
    const bool right = !(k == TEXT || k == EVENT);
    return Column{n, k, right, right};
}

std::string fmt_real(double v, int decimals, bool raw)
{
    char b[512];
    if (raw) snprintf(b, sizeof(b), "%.64f", v);
    else if (decimals == 2) snprintf(b, sizeof(b), "%.2f", round_decimals(v, 2));
    else snprintf(b, sizeof(b), "%.4f", round_decimals(v, 4));
    return b;
}

std::string fmt_u64(uint64_t v)
{
    char b[32];
    snprintf(b, sizeof(b), "%lu", (unsigned long)v);
    return b;
}

std::string fmt_int(int v)
{
    char b[32];
    snprintf(b, sizeof(b), "%d", v);
    return b;
}

void emit(std::string &out, const Table &t, bool raw)
{
    std::vector<size_t> width;
    for (const Column &c : t.cols) width.push_back(c.name.size());
    for (const auto &r : t.rows)
        for (size_t i = 0; i < r.size(); i++) width[i] = std::max(width[i], r[i].size());

    out += "#:GATKTable:" + std::to_string(t.cols.size()) + ":" + std::to_string(t.rows.size()) + ":";
    for (const Column &c : t.cols) {
        switch (c.kind) {
        case TEXT: case EVENT: out += "%s:"; break;
        case COUNT: out += "%d:"; break;
        case REAL2: out += raw ? "%.64f:" : "%.2f:"; break;
        case REAL4: out += raw ? "%.64f:" : "%.4f:"; break;
        }
    }
    out += ";\n";
    out += "#:GATKTable:" + t.name + ":" + t.description + "\n";

    auto cell = [&](const std::string &s, size_t w, bool right) {
        if (s.size() >= w) { out += s; return; }
        if (right) out.append(w - s.size(), ' ');
        out += s;
        if (!right) out.append(w - s.size(), ' ');
    };
    for (size_t i = 0; i < t.cols.size(); i++) {
        cell(t.cols[i].name, width[i], t.cols[i].header_right);
        out += i + 1 == t.cols.size() ? "\n" : "  ";
    }
    for (const auto &r : t.rows) {
        for (size_t i = 0; i < r.size(); i++) {
            cell(r[i], width[i], t.cols[i].right);
            if (i + 1 < r.size()) out += "  ";
        }
        out += "\n";
    }
    out += "\n";
}

const char *ARGUMENTS_BLOCK =
    "#:GATKReport.v1.1:5\n"
    "#:GATKTable:2:18:%s:%s:;\n"
    "#:GATKTable:Arguments:Recalibration argument collection values used in this run\n";

struct Arg {
    const char *name, *value;
};

const Arg ARGUMENTS[] = {
    {"Argument", "Value"},
    {"binary_tag_name", "null"},
    {"covariate", "ReadGroupCovariate,QualityScoreCovariate,ContextCovariate,CycleCovariate"},
    {"default_platform", "null"},
    {"deletions_default_quality", "45"},
    {"force_platform", "null"},
    {"indels_context_size", "3"},
    {"insertions_default_quality", "45"},
    {"low_quality_tail", "2"},
    {"maximum_cycle_value", "500"},
    {"mismatches_context_size", "2"},
    {"mismatches_default_quality", "-1"},
    {"no_standard_covs", "false"},
    {"quantizing_levels", "16"},
    {"recalibration_report", "null"},
    {"run_without_dbsnp", "false"},
    {"solid_nocall_strategy", "THROW_EXCEPTION"},
    {"solid_recal_mode", "SET_Q_ZERO"},
};

char event_char(uint32_t e) { return e == 0 ? 'M' : e == 1 ? 'I' : e == 2 ? 'D' : e == 3 ? 'S' : '?'; }

} // namespace

std::string format_report(const ReportTables &tt, const std::vector<std::string> &rg_names, bool raw)
{
    std::string out = ARGUMENTS_BLOCK;
    // the Arguments table is fixed-width text: names padded to 28 columns, values to 72
    for (const Arg &a : ARGUMENTS) {
        char line[256];
        snprintf(line, sizeof(line), "%-28s%-72s\n", a.name, a.value);
        out += line;
    }
    {
        // the software row keeps the reference's literal padding (55 spaces after the version)
        out += "software                    Firepony_1.1.1";
        out.append(55, ' ');
        out += "\n\n";
    }

    auto rg = [&](uint32_t id) -> std::string { return id < rg_names.size() ? rg_names[id] : "<invalid>"; };
    auto ev = [](uint32_t e) { return std::string(1, event_char(e)); };

    {
        Table t{"Quantized", "Quality quantization map", {}, {}};
        t.cols = {Column{"QualityScore", TEXT, true, true}, Column{"Count", TEXT, true, true},
                  Column{"QuantizedScore", TEXT, true, true}};
        for (int i = 0; i < 94; i++) t.rows.push_back({fmt_int(i), "0", fmt_int(i)});
        emit(out, t, raw);
    }
    {
        Table t{"RecalTable0", "", {}, {}};
        t.cols = {col("ReadGroup", TEXT), col("EventType", EVENT), col("EmpiricalQuality", REAL4),
                  col("EstimatedQReported", REAL4), col("Observations", COUNT), col("Errors", REAL2)};
        t.cols.back().header_right = false;
        const ReportRows &r = tt.table[0];
        for (size_t i = 0; i < r.key.size(); i++)
            t.rows.push_back({rg((r.key[i] >> 10) & 255), ev(r.key[i] & 3), fmt_real(r.empirical[i], 4, raw),
                              fmt_real(r.estimated[i], 4, raw), fmt_u64(r.obs[i]), fmt_real(r.mism[i], 2, raw)});
        emit(out, t, raw);
    }
    {
        Table t{"RecalTable1", "", {}, {}};
        t.cols = {col("ReadGroup", TEXT), col("QualityScore", TEXT), col("EventType", EVENT),
                  col("EmpiricalQuality", REAL4), col("Observations", COUNT), col("Errors", REAL2)};
        t.cols.back().header_right = false;
        const ReportRows &r = tt.table[1];
        for (size_t i = 0; i < r.key.size(); i++) {
            if (r.obs[i] == 0) continue;
            t.rows.push_back({rg((r.key[i] >> 10) & 255), fmt_int((r.key[i] >> 2) & 255), ev(r.key[i] & 3),
                              fmt_real(r.empirical[i], 4, raw), fmt_u64(r.obs[i]), fmt_real(r.mism[i], 2, raw)});
        }
        emit(out, t, raw);
    }
    {
        Table t{"RecalTable2", "", {}, {}};
        t.cols = {col("ReadGroup", TEXT), col("QualityScore", TEXT), col("CovariateValue", TEXT),
                  col("CovariateName", TEXT), col("EventType", EVENT), col("EmpiricalQuality", REAL4),
                  col("Observations", COUNT), col("Errors", REAL2)};
        t.cols.back().header_right = false;
        const ReportRows &cx = tt.table[3];
        for (size_t i = 0; i < cx.key.size(); i++) {
            if (cx.obs[i] == 0) continue;
            const uint32_t raw_ctx = (cx.key[i] >> 2) & 2047;
            const int size = raw_ctx & 15;
            const uint32_t bases = raw_ctx >> 4;
            std::string seq;
            for (int j = 0; j < size; j++) seq += "ACGT"[(bases >> (2 * j)) & 3];
            t.rows.push_back({rg((cx.key[i] >> 21) & 255), fmt_int((cx.key[i] >> 13) & 255), seq, "Context",
                              ev(cx.key[i] & 3), fmt_real(cx.empirical[i], 4, raw), fmt_u64(cx.obs[i]),
                              fmt_real(cx.mism[i], 2, raw)});
        }
        const ReportRows &cy = tt.table[2];
        for (size_t i = 0; i < cy.key.size(); i++) {
            if (cy.obs[i] == 0) continue;
            const uint32_t rawc = (cy.key[i] >> 2) & 1023;
            int cycle = (int)(rawc >> 1);
            if (rawc & 1) cycle = -cycle;
            t.rows.push_back({rg((cy.key[i] >> 20) & 255), fmt_int((cy.key[i] >> 12) & 255), fmt_int(cycle), "Cycle",
                              ev(cy.key[i] & 3), fmt_real(cy.empirical[i], 4, raw), fmt_u64(cy.obs[i]),
                              fmt_real(cy.mism[i], 2, raw)});
        }
        emit(out, t, raw);
    }
    return out;
}

} // namespace fpr
