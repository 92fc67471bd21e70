// Batched execution of many queries over the resident columns.
#pragma once

#include <string>
#include <vector>

#include "../runtime/database.h"

namespace db {

class BatchExecutor {
public:
    explicit BatchExecutor(DatabaseInstance& db) : db_(db) {}
    // Executes every query; returns results in input order.
    std::vector<QueryResult> execute(const std::vector<std::string>& sqls, const std::string& options);
    const std::string& lastReport() const { return report_; }

private:
    DatabaseInstance& db_;
    std::string report_;
};

}  // namespace db
