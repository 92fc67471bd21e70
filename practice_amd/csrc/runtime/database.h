// DatabaseInstance: device residency, catalog, query entry point, result fetch.
//
// Counterpart of the reference DatabaseInstance (include/db/database_instance.h:14-39,
// src/db/database_instance.cpp): every loaded column stays resident in HBM as
// int32 plus an optional LSB-first u64 validity bitmap; the catalog carries
// min / max / exact distinct count per column (computed on the GPU instead of
// the reference's host scan, catalog_cache.h:76-123) and PK flags.  The
// synthetic SSB schema is generated directly in HBM by a counter-based kernel.
#pragma once

#include <cstdint>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "execution.h"

namespace db {

struct GenSpec {
    int sf = 1;
    int64_t fact_rows = 0;     // 0: 6,000,000 * sf
    uint64_t seed = 42;
    std::string nullable;      // "TABLE.column:bp,..."
};

// Synthetic schema definition (same rules as the oracle generator).
int64_t genTableRows(const GenSpec& spec, const std::string& table);
std::vector<std::string> genTableColumns(const std::string& table);
std::vector<std::string> genTables();

class DatabaseInstance {
public:
    // device >= 0: GPU residency; device < 0: host-only catalog (code
    // generation and hipRTC compilation work, execution does not).
    explicit DatabaseInstance(int device = 0, EngineConfig config = {});
    ~DatabaseInstance();

    bool hostOnly() const { return device_ < 0; }
    int device() const { return device_; }

    void createTable(const std::string& name, int64_t rows, bool is_fact);
    void loadColumn(const std::string& table, const std::string& column, const int32_t* data, const uint64_t* validity);
    void setPrimaryKey(const std::string& table, const std::string& column);
    // Catalog-only column (no data): planning and code generation against
    // declared statistics, as the reference's catalog cache provides them.
    void declareColumn(const std::string& table, const std::string& column, bool nullable);
    void setColumnStats(const std::string& column, int64_t min_value, int64_t max_value, uint64_t cardinality);
    // AggregateNode::calculateResultSize of the optimized plan (0 without an aggregate)
    uint64_t resultSize(const std::string& sql) const;
    // Generates every table / column of the synthetic schema in device memory.
    void generateSSB(const GenSpec& spec);
    // Statistics for every column (min, max, distinct count, nullability).
    void finalizeCatalog();
    // The reference's on-disk column store (runtime/storage.h; reference
    // DatabaseInstance::initCatalog + loadData, database_instance.cpp:57-174):
    // every SSB column (or those named in `only`) from `dir`, rows[5] per
    // table in storage::referenceTables() order, validity bitmaps where a
    // file exists (the column is then nullable), the dimension primary keys,
    // and statistics from the text cache at `stats_cache` where its line
    // matches the file (recomputed and written back otherwise; "" for none).
    // Finalizes the catalog.
    void loadReferenceFiles(const std::string& dir, const int64_t rows[5], const std::string& stats_cache,
                            const std::vector<std::string>& only);
    // Every loaded column of the SSB schema written in that format, with
    // validity files (suffix ".valid64", ".null8", ...) for nullable columns.
    void saveReferenceFiles(const std::string& dir, const std::string& validity_suffix) const;

    QueryResult executeQuery(const std::string& sql);
    // Scan only fact rows [row_begin, row_end) (serving / batching chunks).
    QueryResult executeQueryRange(const std::string& sql, int64_t row_begin, int64_t row_end);
    // executeQuery with no device or pinned allocation: scratch from the
    // caller's device block (HipModuleExecutor::executeInto); the plan's
    // module must have been compiled (compile()) beforehand.
    QueryResult executeQueryInto(const std::string& sql, void* scratch, size_t bytes);
    std::string explain(const std::string& sql) const;           // optimized plan, S-expression
    std::string explainTree(const std::string& sql) const;       // optimized plan, indented
    std::string generateQueryCode(const std::string& sql) const; // HIP source
    QueryProgram prepare(const std::string& sql) const;
    // Compile (or fetch from the cache) without executing.
    ModuleHandle compile(const std::string& sql);

    const Catalog& catalog() const { return *catalog_; }
    std::shared_ptr<const Catalog> catalogPtr() const { return catalog_; }
    const std::unordered_map<std::string, DeviceColumn>& columns() const { return columns_; }
    void fetchColumn(const std::string& column, int32_t* out, uint64_t* validity_out) const;
    size_t loadedDeviceBytes() const { return loaded_bytes_; }
    QueryEngine& engine() { return *engine_; }
    const QueryEngine& engine() const { return *engine_; }
    EngineConfig& config() { return ctx_->config; }
    const EngineConfig& config() const { return ctx_->config; }
    PlanSignatureCache& moduleCache() { return *cache_; }
    HipRtcCompiler& compiler() { return *compiler_; }
    ExecutionContext& context() { return *ctx_; }
    // Another context on its own stream over the same resident columns,
    // catalog and module cache (serving lanes).  Columns must not be loaded
    // or dropped while such contexts run queries.
    std::unique_ptr<ExecutionContext> newContext();
    hipStream_t stream() const { return ctx_->stream; }
    // Contexts on `base`'s stream over its columns, with its configuration
    // (the members of an interpreted group); `pool` grows to n and keeps
    // them.  A pool must be destroyed before its base context.
    static void groupContexts(ExecutionContext& base, std::vector<std::unique_ptr<ExecutionContext>>& pool, size_t n);
    // Statements run together on the database's stream
    // (QueryEngine::executeInterpretedGroup): with the interpreted tier on,
    // plans with a linear form share scans.  errs[k]: statement k's error
    // ("" when it ran; its result is then out[k]).
    std::vector<QueryResult> executeGroup(const std::vector<std::string>& sqls, std::vector<std::string>& errs);
    // Serializes use of the execution context (queries, batches, generation).
    std::mutex& contextMutex() const { return mu_; }

private:
    int device_;
    std::shared_ptr<Catalog> catalog_;
    std::unordered_map<std::string, DeviceColumn> columns_;
    std::vector<void*> allocations_;
    std::unordered_map<std::string, std::vector<int32_t>> host_data_;        // host-only mode
    std::unordered_map<std::string, std::vector<uint64_t>> host_validity_;  // host-only mode
    size_t loaded_bytes_ = 0;
    // min / max / distinct count known ahead of finalizeCatalog (statistics cache)
    std::unordered_map<std::string, ColumnStatistics> preset_stats_;
    std::shared_ptr<PlanSignatureCache> cache_;
    std::shared_ptr<HipRtcCompiler> compiler_;
    std::shared_ptr<HipModuleExecutor> executor_;
    std::unique_ptr<QueryEngine> engine_;
    std::unique_ptr<ExecutionContext> ctx_;
    std::vector<std::unique_ptr<ExecutionContext>> group_ctx_;  // on ctx_'s stream (destroyed before it)
    mutable std::mutex mu_;

    void ensureTable(const std::string& table) const;
    void* deviceAlloc(size_t bytes);
    ColumnStatistics computeStats(const std::string& column, bool& nullable);
};

}  // namespace db
