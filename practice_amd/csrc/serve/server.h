// Query server: continuous batching over resident columns.
//
// No reference counterpart (the reference executes one query per call,
// src/core/execution.cpp:652-725).  Requests queue on submit().  Each lane
// (own stream, scratch and shared-scan executor) repeatedly takes every
// queued request it can admit, up to max_batch, and runs them together;
// while one lane's work is on the GPU the next lane admits the requests that
// arrived meanwhile, so the lanes' work overlaps on the device.
//
// Execution tiers per plan signature:
//   hot    promoted signatures: the requests of one signature in a batch
//          share one cq_batch_scan launch (one pass over its fact columns
//          for all of them; per_template) or all hot signatures share one
//          launch of a module holding the whole hot set
//   cold   a signature not yet hot runs at once on the ahead-of-time
//          interpreter on a cold lane (no compilation on the request's
//          path, no wait behind hot batches); once seen `promote_after`
//          times its modules are compiled on a background thread and it
//          turns hot.  With interpret_cold off, cold requests wait for their
//          plain module instead.
//
// Admission: a request's device footprint (hash tables, aggregate slots,
// materialization buffers; estimateQueryTemporaryBytes) is reserved against
// the HBM budget while it is in flight; requests that do not fit wait,
// requests larger than the whole budget fail.
#pragma once

#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <string>
#include <thread>
#include <vector>

#include "batch.h"

namespace db {

struct ServerOptions {
    // concurrent batches on the GPU (own stream each).  One lane with
    // large batches measured best at scale 100: a shared scan costs about
    // the same for 16 or 128 queries, and two lanes' scans only share the
    // device (p99 68 vs 224 ms at 2,700 requests/s)
    int lanes = 1;
    int max_batch = 128;           // requests per shared scan
    int max_templates = 16;        // plan signatures per shared-scan module
    int promote_after = 2;         // requests of a signature before it joins the shared scan
    int compile_threads = 2;       // background hipRTC threads
    size_t budget_bytes = 0;       // HBM for query temporaries (0: quarter of free memory)
    int window_us = 0;             // wait for more arrivals before taking a batch
    bool async_compile = true;     // compile off the lanes
    bool interpret_cold = true;    // run new signatures at once on the interpreted tier
    int cold_lanes = 1;            // lanes (own stream each) running cold requests
    // interpreted requests a cold lane takes at once: those with a linear
    // form share scans (16 per scan) and read the fact columns once.  One
    // at a time by default: at scale 100 the shared scan's per-plan
    // arithmetic costs more than the column pass it saves beyond pairs
    int cold_group = 1;
    // cold lanes' interpreted scans on this percentage of a full-device
    // grid, leaving CUs to the shared scans running meanwhile
    int cold_grid_percent = 100;
    // 1: hot signatures each get their own shared-scan module and launch (a
    // batch runs one launch per signature present); 0: one module holds the
    // whole hot set (up to max_templates) and a batch is one launch; 2: both
    // modules, and a batch is one launch while fewer than cold_backlog cold
    // requests are queued, one launch per signature otherwise (a single
    // launch holds every CU for its whole scan, so cold scans wait behind
    // it and a backlog would grow)
    int per_template = 1;
    int cold_backlog = 2;
    int max_hot = 64;              // hot signatures kept (least recently promoted dropped)
    BatchOptions batch;
};

// "key=value,..." with lanes, max_batch, max_templates, promote_after,
// compile_threads, budget_mb, window_us, async_compile, interpret,
// cold_lanes, cold_group, cold_grid_percent, per_template, cold_backlog, max_hot, bt, vpt,
// graph, lds_kb, grid_percent
ServerOptions parseServerOptions(const std::string& text);

class QueryServer {
public:
    QueryServer(DatabaseInstance& db, const std::string& options);
    QueryServer(DatabaseInstance& db, const ServerOptions& options);
    ~QueryServer();
    QueryServer(const QueryServer&) = delete;
    QueryServer& operator=(const QueryServer&) = delete;

    long long submit(const std::string& sql);
    // 0 on success, negative error code otherwise (err receives the message)
    int wait(long long ticket, int timeout_ms, QueryResult& out, std::string& err);
    std::string statsJson();
    void resetStats();

private:
    struct Request {
        long long ticket = 0;
        std::string sql;
        std::chrono::steady_clock::time_point arrival;
        std::shared_ptr<QueryProgram> program;   // prepared with the per-tile body
        size_t footprint = 0;
        bool interpreted = false;                 // cold request: interpreted tier (else its plain module)
        bool plain_only = false;                  // the shared scan cannot take the plan (OR-key joins): cold lanes, compiled
        bool cold = false;                        // ran on a cold lane (interpreted or plain), not in a shared scan
    };
    struct Done {
        int rc = 0;
        std::string err;
        QueryResult result;
        double latency_ms = 0;
    };
    struct Lane {
        std::unique_ptr<ExecutionContext> ctx;
        std::mutex mu;
        std::unique_ptr<BatchExecutor> exec;   // replaced under the server mutex
        std::vector<std::unique_ptr<ExecutionContext>> group;  // cold lanes: contexts of a grouped run, on ctx's stream
        std::thread thread;
        long long graph_captures = 0, graph_replays = 0, graph_failures = 0;  // of replaced executors
    };

    DatabaseInstance& db_;
    ServerOptions opt_;
    std::mutex mu_;
    std::condition_variable cv_, done_cv_, compile_cv_;
    std::deque<Request> queue_;         // new requests and hot-tier requests
    std::deque<Request> plain_ready_;   // cold requests whose plain module is compiled
    std::deque<Request> compile_q_;     // cold requests waiting for their plain module
    std::deque<std::shared_ptr<QueryProgram>> promote_q_;  // signatures to add to the shared scan
    std::deque<Request> cold_q_;        // admitted cold requests for the cold lanes
    std::condition_variable cold_cv_;
    std::map<std::string, int> seen_;   // requests per signature
    std::set<std::string> promoting_;
    std::map<long long, Done> done_;
    long long next_ = 0;
    bool stop_ = false;
    size_t in_flight_bytes_ = 0;

    // templates of the compiled shared-scan set (order = kernel order)
    std::vector<std::shared_ptr<QueryProgram>> scan_set_;
    // per_template = 2: the hot set whose one-launch module is compiled
    std::vector<std::shared_ptr<QueryProgram>> combined_set_;

    std::vector<std::unique_ptr<Lane>> lanes_;
    std::vector<std::unique_ptr<Lane>> cold_lanes_;
    std::vector<std::thread> compilers_;

    // statistics
    std::vector<double> latencies_;
    std::vector<double> cold_latencies_;  // the cold lanes' share of latencies_
    double batch_wall_ms_ = 0;            // shared-scan batches: host wall time, summed
    long long cold_groups_ = 0, cold_grouped_ = 0;  // interpreted requests run together: runs, requests
    long long combined_batches_ = 0;    // per_template = 2: batches run as one launch
    std::chrono::steady_clock::time_point stats_start_;
    std::string last_error_;            // message of the last failed batch or cold request
    long long batches_ = 0, batched_queries_ = 0, plain_queries_ = 0, interpreted_queries_ = 0, admission_waits_ = 0,
              compiles_ = 0, promotions_ = 0, rejected_ = 0;

    void start();
    void laneLoop(Lane* lane);
    void coldLoop(Lane* lane);
    void release(const Request& r);
    void compileLoop();
    bool compileOne();  // one queued compilation job; false when there is none
    void compileCombined();
    void finish(Request& r, int rc, const std::string& err, QueryResult* result);
    bool prepare(Request& r);
    bool inScanSet(const std::string& signature) const;
};

}  // namespace db
