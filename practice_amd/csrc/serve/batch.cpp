// Batched execution (per-query JIT path for every query).
#include "batch.h"

#include <chrono>
#include <sstream>

namespace db {

std::vector<QueryResult> BatchExecutor::execute(const std::vector<std::string>& sqls, const std::string& options) {
    (void)options;
    std::vector<QueryResult> out;
    out.reserve(sqls.size());
    const auto t0 = std::chrono::steady_clock::now();
    for (const auto& q : sqls) out.push_back(db_.executeQuery(q));
    const double ms = elapsedMs(t0, std::chrono::steady_clock::now());
    std::ostringstream r;
    r << "{\"queries\":" << sqls.size() << ",\"shared_scan\":0,\"jit\":" << sqls.size() << ",\"wall_ms\":" << ms << "}";
    report_ = r.str();
    return out;
}

}  // namespace db
