/*
 * GATKReport v1.1 writer for the recalibration tables.
 */
#ifndef FP_REPORT_H
#define FP_REPORT_H

#include <stdint.h>

#include <string>
#include <vector>

namespace fpr {

struct ReportRows {
    std::vector<uint32_t> key;
    std::vector<uint64_t> obs;
    std::vector<double> mism, estimated, empirical;
};

struct ReportTables {
    ReportRows table[4];   // read group, quality, cycle, context (fp_table_id order)
};

std::string format_report(const ReportTables &t, const std::vector<std::string> &rg_names, bool disable_rounding);

// round to n decimals the way the reference does before printing (util.cu:107-116)
double round_decimals(double v, int n);

} // namespace fpr

#endif
