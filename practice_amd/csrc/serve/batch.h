// Batched execution: many queries, one shared fact scan.
//
// The reference runs every query as its own JIT kernel sequence
// (QueryEngine::executeQuery, src/core/execution.cpp:652-725), so N queries
// stream the fact table N times.  Here a batch is split into shared scans:
// each one groups its queries by plan signature ("templates"), loads the
// union of their fact columns once per tile in one persistent launch
// (cq_batch_scan) and runs every template's pipeline for each of its queries
// over those registers.  Dimension builds and result shaping stay per query;
// all device work of a batch between the parameter upload and the row-count
// readback is captured once into a HIP graph and replayed for batches of the
// same shape (only the device-resident parameter blocks change).
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../runtime/database.h"

namespace db {

struct BatchOptions {
    int block_threads = 1024;     // cq_batch_scan block size (64..1024)
    int vectors_per_thread = 1;   // 128-bit loads per column per lane per tile (1..4)
    bool graph = true;            // capture / replay the device phases as a HIP graph
    bool autotune = false;        // time block / load-width variants, keep the fastest
    int max_templates = 16;       // templates per shared scan
    int max_queries = 256;        // queries per shared scan
    bool shared_tables = true;    // one probe per template and dimension (packed attributes)
    bool lds_aggregation = true;  // per-query LDS aggregation for small group domains
    int lds_kb = 160;             // dynamic LDS of one shared-scan block (KiB, 16..160)
    bool masks = true;            // query-mask form for templates whose filters allow it
    // shared dimension tables staged into LDS: bit k for the k-th candidate
    // (generateBatchScanSource); the autotuner sweeps the first three
    unsigned table_mask = ~0u;
    // static LDS for staged dimension tables and query-mask tables (KiB,
    // 16..96); the per-query regions get what the block has left
    int table_kb = 32;
    // shared-scan grid as a percentage of the persistent grid (blocks that
    // fit at once on every CU): below 100 leaves CUs to other streams
    int grid_percent = 100;
};

// "key=value,key=value" (keys: bt, vpt, graph, autotune, max_templates,
// max_queries, shared, lds_agg, lds_kb, masks, table_mask, table_kb, grid_percent)
BatchOptions parseBatchOptions(const std::string& text);

struct BatchScanReport {
    int queries = 0, templates = 0, union_columns = 0;
    int block_threads = 0, vectors_per_thread = 0;
    unsigned lds_table_bytes = 0;  // static LDS of the staged shared tables and query-mask tables
    unsigned table_mask = ~0u;     // LDS placement of the candidate tables
    int table_kb = 32;             // LDS budget of the staged tables
    bool graph_replayed = false;
    const char* graph_mode = "eager";  // eager / captured / replayed
    double scan_ms = 0;            // cq_batch_scan alone (events)
    double device_ms = 0;          // builds + scan + shaping (events)
    uint64_t fact_bytes = 0;       // algorithmic fact bytes read by the shared scan
    // host time per phase
    double modules_ms = 0, bind_ms = 0, enqueue_ms = 0, wait_ms = 0, fetch_ms = 0;
    size_t arena_bytes = 0;        // device scratch of the batch (one block)
};

// Generated sources of a batch, host-only databases included: part < 0 is
// the shared-scan module of all the queries (one scan), part = k the
// template module (dimension builds + result flags) of query k.
std::string generateBatchCode(const DatabaseInstance& db, const std::vector<std::string>& sqls, const BatchOptions& options,
                              int part);

class BatchExecutor {
public:
    // Runs on the database's own context (serialized with its queries).
    explicit BatchExecutor(DatabaseInstance& db);
    // Runs on `ctx` (its stream and scratch), serialized by `ctx_mutex`.
    BatchExecutor(DatabaseInstance& db, ExecutionContext* ctx, std::mutex* ctx_mutex);
    ~BatchExecutor();
    BatchExecutor(const BatchExecutor&) = delete;
    BatchExecutor& operator=(const BatchExecutor&) = delete;

    // Executes every query; results come back in input order and are the
    // same as DatabaseInstance::executeQuery's for each query alone.
    std::vector<QueryResult> execute(const std::vector<std::string>& sqls, const std::string& options);
    std::vector<QueryResult> execute(const std::vector<std::string>& sqls, const BatchOptions& options);
    // Prepared programs (CodegenOptions::batch_body); `module_templates`, when
    // given, lists the templates the shared-scan module is generated for (a
    // superset of the programs' templates: ones without queries are skipped).
    std::vector<QueryResult> executePrograms(std::vector<QueryProgram>& programs, const BatchOptions& options,
                                             const std::vector<const QueryProgram*>* module_templates = nullptr);
    // Codegen options a batch prepares its programs with.
    CodegenOptions codegenOptions(const BatchOptions& options) const;
    // JSON: per shared scan counts, timings, tuning choice, graph use.
    const std::string& lastReport() const { return report_; }
    // graph captures, replays and failed captures since construction
    long long graphCaptures() const { return graph_captures_.load(); }
    long long graphReplays() const { return graph_replays_.load(); }
    long long graphFailures() const { return graph_failures_.load(); }

private:
    struct Graph {
        std::string key;
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        // code objects the captured kernel nodes run: held so that a module
        // cache eviction cannot unload them while the graph may replay
        std::vector<ModuleHandle> modules;
    };
    struct Tuned {
        int block_threads = 1024, vectors_per_thread = 1;
        unsigned table_mask = ~0u;
        int table_kb = 32;
    };
    DatabaseInstance& db_;
    ExecutionContext* ctx_;
    std::mutex* ctx_mu_;
    std::string report_;
    std::map<std::string, Graph> graphs_;   // keyed by batch shape
    std::map<std::string, int> graph_seen_; // shapes seen (captured on the second)
    std::map<std::string, Tuned> tuned_;    // keyed by template-set signature
    double plan_ms_ = 0;                    // SQL -> programs time of the current call
    std::atomic<long long> graph_failures_{0};  // captures that failed (the batch then ran eagerly)
    std::atomic<long long> graph_captures_{0}, graph_replays_{0};
    std::string last_graph_error_;
    void* pinned_ = nullptr;                // staging for parameter blocks / row counts
    size_t pinned_bytes_ = 0;
    hipEvent_t ev_[4] = {nullptr, nullptr, nullptr, nullptr};

    void* pinned(size_t bytes);
    void runScan(const std::vector<int>& members, std::vector<QueryProgram>& programs, const BatchOptions& opt, int scan_index,
                 std::vector<QueryResult>& results, BatchScanReport& rep,
                 const std::vector<const QueryProgram*>* module_templates);
    // aggregate-state initialization of a batch's queries: the eligible ones
    // (compiled, shared tables) in one launch over their descriptors
    struct BatchInit {
        const cq_init_desc* d = nullptr;
        std::vector<char> eligible;
        long long max_slots = 0;
    };
    void prepareTables(const std::vector<const QueryProgram*>& templates, const std::vector<std::vector<int>>& per_template,
                       std::vector<QueryRun>& runs, const cq_query_params* dqp, const BatchInit& init, ExecutionContext* ctx);
    Tuned autotune(const std::vector<const QueryProgram*>& templates, const std::vector<std::vector<int>>& per_template,
                   std::vector<QueryRun>& runs, const cq_query_params* dqp, const BatchInit& init, const BatchOptions& opt,
                   ExecutionContext* ctx);
    void dropGraph(Graph& g);
};

}  // namespace db
