// Text form of a query result as the reference's console prints it
// (src/cli/terminal_app.cpp:263-302): the block that its olap_bench runner
// parses (Rows returned / Rows shown / GPU execution time and the extended
// timing lines, olap_bench/tools/run_olap_bench.py:12-23).
#pragma once

#include <cstddef>
#include <string>

#include "execution.h"

namespace db {

struct ResultTextOptions {
    bool limit_enabled = true;  // show only the last `limit` rows
    size_t limit = 1000;
    bool extended_timing = false;
};

std::string formatResultText(const QueryResult& result, const ResultTextOptions& options);

}  // namespace db
