// Text form of a query result (terminal_app.cpp:263-302, formatResultCell
// 102-127, printTimingLine 24-26).
#include "result_text.h"

#include <cstdio>
#include <sstream>

namespace db {

namespace {

std::string cellText(const QueryResult& r, size_t row, size_t col) {
    const ResultColumnDesc& cd = r.columns[col];
    // a column the plan cannot make NULL prints its value (terminal_app.cpp:63-65)
    if (cd.nullable && !r.validity[col][row]) return "NULL";
    const uint64_t raw = r.data[col][row];
    if (cd.type == LogicalType::Float64) {
        double d;
        static_assert(sizeof d == sizeof raw, "double is 64 bits");
        __builtin_memcpy(&d, &raw, sizeof d);
        std::ostringstream o;
        o.precision(15);  // std::setprecision(15), default floatfield
        o << d;
        return o.str();
    }
    if (cd.type == LogicalType::Int64) return std::to_string(static_cast<long long>(raw));
    return std::to_string(static_cast<unsigned long long>(raw));
}

void timingLine(std::ostringstream& o, const char* label, double ms) {
    char buf[64];
    std::snprintf(buf, sizeof buf, "%.3f", ms);
    o << label << ": " << buf << " ms\n";
}

}  // namespace

std::string formatResultText(const QueryResult& r, const ResultTextOptions& opt) {
    std::ostringstream o;
    const size_t tuple = r.columns.size();
    const size_t total = r.row_count;
    // the last `limit` rows (the console keeps the tail of a long result)
    const size_t begin = (opt.limit_enabled && total > opt.limit) ? total - opt.limit : 0;
    o << "--- Result ---\n";
    if (tuple > 0) {
        o << "Columns: ";
        for (size_t j = 0; j < tuple; ++j) o << r.columns[j].name << (j + 1 < tuple ? " | " : "");
        o << "\n";
    }
    if (total == 0) {
        o << "0\n";
    } else {
        for (size_t row = begin; row < total; ++row) {
            if (tuple <= 1) {
                o << cellText(r, row, 0) << "\n";
                continue;
            }
            o << "Row " << row << ": ";
            for (size_t j = 0; j < tuple; ++j) o << cellText(r, row, j) << (j + 1 < tuple ? " | " : "");
            o << "\n";
        }
    }
    o << "Rows returned: " << total << "\n";
    if (opt.limit_enabled && total > opt.limit) o << "Rows shown: " << (total - begin) << "\n";
    timingLine(o, "GPU execution time", r.timing.gpu_execute_ms);
    if (opt.extended_timing) {
        timingLine(o, "Code generation time", r.timing.codegen_ms);
        // the reference's label for its code-compilation step (hipRTC here)
        timingLine(o, "ACPP compilation time", r.timing.compile_ms);
        timingLine(o, "Library load time", r.timing.library_load_ms);
        timingLine(o, "Result materialization/fetch time", r.timing.host_fetch_ms);
        timingLine(o, "Engine processing time", r.timing.engine_ms);
    }
    o << "--- End ---\n";
    return o.str();
}

}  // namespace db
