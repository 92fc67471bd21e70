// Query server: lanes of shared-scan batches, footprint admission, async compile.
#include "server.h"

#include <algorithm>
#include <sstream>
#include <stdexcept>

namespace db {

namespace {
std::string jsonEscape(const std::string& t) {
    std::string o;
    for (char c : t.substr(0, 300)) {
        if (c == '"' || c == '\\') o += '\\', o += c;
        else if (static_cast<unsigned char>(c) < 0x20) o += ' ';
        else o += c;
    }
    return o;
}
}  // namespace

ServerOptions parseServerOptions(const std::string& text) {
    ServerOptions o;
    std::string batch_opts;
    std::stringstream ss(text);
    std::string item;
    while (std::getline(ss, item, ',')) {
        if (item.empty()) continue;
        const size_t eq = item.find('=');
        if (eq == std::string::npos) throw std::runtime_error("server option without value: " + item);
        const std::string k = item.substr(0, eq);
        const long long v = std::stoll(item.substr(eq + 1));
        if (k == "lanes") o.lanes = static_cast<int>(v);
        else if (k == "max_batch") o.max_batch = static_cast<int>(v);
        else if (k == "max_templates") o.max_templates = static_cast<int>(v);
        else if (k == "promote_after") o.promote_after = static_cast<int>(v);
        else if (k == "compile_threads") o.compile_threads = static_cast<int>(v);
        else if (k == "budget_mb") o.budget_bytes = static_cast<size_t>(v) << 20;
        else if (k == "window_us") o.window_us = static_cast<int>(v);
        else if (k == "async_compile") o.async_compile = v != 0;
        else if (k == "interpret") o.interpret_cold = v != 0;
        else if (k == "cold_lanes") o.cold_lanes = static_cast<int>(v);
        else if (k == "cold_group") o.cold_group = static_cast<int>(v);
        else if (k == "cold_grid_percent") o.cold_grid_percent = static_cast<int>(v);
        else if (k == "per_template") o.per_template = static_cast<int>(v);
        else if (k == "cold_backlog") o.cold_backlog = static_cast<int>(v);
        else if (k == "max_hot") o.max_hot = static_cast<int>(v);
        else if (k == "bt" || k == "vpt" || k == "graph" || k == "lds_kb" || k == "grid_percent") batch_opts += (batch_opts.empty() ? "" : ",") + item;
        else throw std::runtime_error("unknown server option: " + k);
    }
    // served batches vary in composition and buffer placement, so captured
    // graphs are rarely replayed: eager launches unless asked for
    o.batch = parseBatchOptions(batch_opts.find("graph=") == std::string::npos ? (batch_opts.empty() ? "graph=0" : batch_opts + ",graph=0")
                                                                                 : batch_opts);
    if (o.lanes < 1 || o.lanes > 8) throw std::runtime_error("server lanes must be in [1, 8]");
    if (o.max_batch < 1) throw std::runtime_error("server max_batch must be positive");
    o.max_templates = std::max(1, std::min(o.max_templates, CQ_MAX_TEMPLATES));
    o.promote_after = std::max(1, o.promote_after);
    o.compile_threads = std::max(1, std::min(o.compile_threads, 16));
    if (o.cold_lanes < 1 || o.cold_lanes > 8) throw std::runtime_error("server cold_lanes must be in [1, 8]");
    if (o.per_template < 0 || o.per_template > 2) throw std::runtime_error("server per_template must be 0, 1 or 2");
    o.max_hot = std::max(1, o.max_hot);
    return o;
}

QueryServer::QueryServer(DatabaseInstance& db, const std::string& options) : QueryServer(db, parseServerOptions(options)) {}

QueryServer::QueryServer(DatabaseInstance& db, const ServerOptions& options) : db_(db), opt_(options) {
    if (db_.hostOnly()) throw std::runtime_error("the query server needs a GPU device (host-only database)");
    if (opt_.budget_bytes == 0) {
        size_t free_b = 0, total = 0;
        hipCheck(hipSetDevice(db_.device()), "hipSetDevice");
        hipCheck(hipMemGetInfo(&free_b, &total), "hipMemGetInfo");
        opt_.budget_bytes = free_b / 4;
    }
    stats_start_ = std::chrono::steady_clock::now();
    start();
}

void QueryServer::start() {
    for (int i = 0; i < opt_.lanes; ++i) {
        auto lane = std::make_unique<Lane>();
        lane->ctx = db_.newContext();
        lane->exec = std::make_unique<BatchExecutor>(db_, lane->ctx.get(), &lane->mu);
        lanes_.push_back(std::move(lane));
    }
    for (int i = 0; i < opt_.cold_lanes; ++i) {
        auto lane = std::make_unique<Lane>();
        lane->ctx = db_.newContext();
        lane->ctx->config.vm_grid_percent = std::max(1, std::min(100, opt_.cold_grid_percent));
        cold_lanes_.push_back(std::move(lane));
    }
    for (auto& l : lanes_) {
        Lane* lp = l.get();
        l->thread = std::thread([this, lp] { laneLoop(lp); });
    }
    for (auto& l : cold_lanes_) {
        Lane* lp = l.get();
        l->thread = std::thread([this, lp] { coldLoop(lp); });
    }
    if (opt_.async_compile)
        for (int i = 0; i < opt_.compile_threads; ++i) compilers_.emplace_back([this] { compileLoop(); });
}

QueryServer::~QueryServer() {
    {
        std::lock_guard<std::mutex> lock(mu_);
        stop_ = true;
    }
    cv_.notify_all();
    compile_cv_.notify_all();
    cold_cv_.notify_all();
    for (auto& l : lanes_)
        if (l->thread.joinable()) l->thread.join();
    for (auto& l : cold_lanes_)
        if (l->thread.joinable()) l->thread.join();
    for (auto& t : compilers_)
        if (t.joinable()) t.join();
    // requests still queued when the server stops fail
    std::lock_guard<std::mutex> lock(mu_);
    for (auto* q : {&queue_, &compile_q_, &plain_ready_, &cold_q_}) {
        for (auto& r : *q) {
            Done d;
            d.rc = -3;
            d.err = "server stopped";
            done_[r.ticket] = std::move(d);
        }
        q->clear();
    }
    done_cv_.notify_all();
}

long long QueryServer::submit(const std::string& sql) {
    Request r;
    r.sql = sql;
    r.arrival = std::chrono::steady_clock::now();
    {
        std::lock_guard<std::mutex> lock(mu_);
        if (stop_) throw std::runtime_error("server stopped");
        r.ticket = next_++;
    }
    // plan and code generation on the caller's thread, off the lanes; a
    // request that fails here is already finished with its error
    const long long ticket = r.ticket;
    if (!prepare(r)) return ticket;
    std::lock_guard<std::mutex> lock(mu_);
    if (stop_) {  // stopped while this request was being planned
        Done d;
        d.rc = -3;
        d.err = "server stopped";
        done_[ticket] = std::move(d);
        done_cv_.notify_all();
        return ticket;
    }
    // a cold request (a signature outside the shared scan, or a plan the
    // shared scan cannot take) goes straight to the cold lanes when its
    // footprint fits, instead of waiting for a hot lane between batches
    if (opt_.interpret_cold && opt_.async_compile && (r.plain_only || !inScanSet(r.program->signature)) &&
        in_flight_bytes_ + r.footprint <= opt_.budget_bytes) {
        bool compile_work = false;
        if (!r.plain_only) {
            const std::string& sig = r.program->signature;
            const int n = ++seen_[sig];
            if (n >= opt_.promote_after && !promoting_.count(sig)) {
                promoting_.insert(sig);
                promote_q_.push_back(r.program);
                compile_work = true;
            }
        }
        in_flight_bytes_ += r.footprint;
        r.interpreted = !r.plain_only;
        r.cold = true;
        cold_q_.push_back(std::move(r));
        cold_cv_.notify_one();
        if (compile_work) compile_cv_.notify_all();
        return ticket;
    }
    queue_.push_back(std::move(r));
    cv_.notify_one();
    return ticket;
}

void QueryServer::finish(Request& r, int rc, const std::string& err, QueryResult* result) {
    Done d;
    d.rc = rc;
    d.err = err;
    if (result) d.result = std::move(*result);
    d.latency_ms = elapsedMs(r.arrival, std::chrono::steady_clock::now());
    std::lock_guard<std::mutex> lock(mu_);
    if (rc == 0) {
        latencies_.push_back(d.latency_ms);
        if (r.cold) cold_latencies_.push_back(d.latency_ms);
    }
    else ++rejected_;
    done_[r.ticket] = std::move(d);
    done_cv_.notify_all();
}

bool QueryServer::prepare(Request& r) {
    try {
        BatchOptions bo = opt_.batch;
        CodegenOptions co = db_.config().codegen;
        co.batch_body = true;
        co.block_threads = bo.block_threads;
        co.vectors_per_thread = bo.vectors_per_thread;
        try {
            r.program = std::make_shared<QueryProgram>(db_.engine().prepareCached(r.sql, co));
        } catch (const std::runtime_error&) {
            // not for the shared scan (an OR-key join): the plain module on a
            // cold lane; a plan the plain path rejects as well fails here
            CodegenOptions plain = db_.config().codegen;
            plain.vm = false;
            r.program = std::make_shared<QueryProgram>(db_.engine().prepareCached(r.sql, plain));
            r.plain_only = true;
        }
        r.footprint = estimateQueryTemporaryBytes(*r.program);
        if (r.footprint > opt_.budget_bytes) {
            finish(r, -5, "query needs " + std::to_string(r.footprint) + " bytes of device memory, over the serving budget of " +
                              std::to_string(opt_.budget_bytes),
                   nullptr);
            return false;
        }
        return true;
    } catch (const std::exception& e) {
        finish(r, -4, e.what(), nullptr);
        return false;
    }
}

bool QueryServer::inScanSet(const std::string& signature) const {
    for (const auto& t : scan_set_)
        if (t->signature == signature) return true;
    return false;
}

void QueryServer::laneLoop(Lane* lane) {
    for (;;) {
        std::vector<Request> taken, plain_taken;
        {
            std::unique_lock<std::mutex> lock(mu_);
            cv_.wait(lock, [&] { return stop_ || !queue_.empty() || !plain_ready_.empty(); });
            if (stop_) return;
            if (opt_.window_us > 0 && static_cast<int>(queue_.size()) < opt_.max_batch) {
                lock.unlock();
                std::this_thread::sleep_for(std::chrono::microseconds(opt_.window_us));
                lock.lock();
                if (stop_) return;
            }
            while (!queue_.empty() && static_cast<int>(taken.size()) < opt_.max_batch) {
                taken.push_back(std::move(queue_.front()));
                queue_.pop_front();
            }
            while (!plain_ready_.empty() && static_cast<int>(plain_taken.size()) < opt_.max_batch) {
                plain_taken.push_back(std::move(plain_ready_.front()));
                plain_ready_.pop_front();
            }
        }
        // plan + codegen outside the lock
        std::vector<Request> ready;
        for (auto& r : taken)
            if (r.program || prepare(r)) ready.push_back(std::move(r));

        // hot signatures join the shared scan, cold ones go to compilation;
        // admission reserves each request's footprint against the budget
        std::vector<Request> batch, plain, later, later_plain, interp;
        std::vector<std::shared_ptr<QueryProgram>> set;
        bool compile_work = false;
        {
            std::lock_guard<std::mutex> lock(mu_);
            set = scan_set_;
            for (auto& r : ready) {
                const std::string& sig = r.program->signature;
                if (r.plain_only) {
                    if (in_flight_bytes_ + r.footprint <= opt_.budget_bytes) {
                        in_flight_bytes_ += r.footprint;
                        plain.push_back(std::move(r));
                    } else {
                        later.push_back(std::move(r));
                    }
                    continue;
                }
                if (!inScanSet(sig)) {
                    const int n = ++seen_[sig];
                    if (n >= opt_.promote_after && !promoting_.count(sig)) {
                        promoting_.insert(sig);
                        promote_q_.push_back(r.program);
                        compile_work = true;
                    }
                    if (!opt_.interpret_cold) {
                        compile_q_.push_back(std::move(r));
                        compile_work = true;
                    } else if (in_flight_bytes_ + r.footprint <= opt_.budget_bytes) {
                        in_flight_bytes_ += r.footprint;
                        interp.push_back(std::move(r));
                    } else {
                        later.push_back(std::move(r));
                    }
                } else if (in_flight_bytes_ + r.footprint <= opt_.budget_bytes) {
                    in_flight_bytes_ += r.footprint;
                    batch.push_back(std::move(r));
                } else {
                    later.push_back(std::move(r));
                }
            }
            for (auto& r : plain_taken) {
                if (in_flight_bytes_ + r.footprint <= opt_.budget_bytes) {
                    in_flight_bytes_ += r.footprint;
                    plain.push_back(std::move(r));
                } else {
                    later_plain.push_back(std::move(r));
                }
            }
            if (!later.empty() || !later_plain.empty()) ++admission_waits_;
            for (auto it = later.rbegin(); it != later.rend(); ++it) queue_.push_front(std::move(*it));
            for (auto it = later_plain.rbegin(); it != later_plain.rend(); ++it) plain_ready_.push_front(std::move(*it));
        }
        if (compile_work) {
            if (opt_.async_compile) compile_cv_.notify_all();
            else
                while (compileOne()) {
                }
        }
        if (batch.empty() && plain.empty() && interp.empty()) {
            if (!later.empty() || !later_plain.empty()) {
                // nothing admissible right now: wait for memory to free up
                std::unique_lock<std::mutex> lock(mu_);
                const size_t before = in_flight_bytes_;
                cv_.wait_for(lock, std::chrono::milliseconds(5), [&] { return stop_ || in_flight_bytes_ < before; });
            }
            continue;
        }

        {
            // cold requests (footprint already reserved) run on the cold
            // lanes, concurrently with this lane's batch
            std::lock_guard<std::mutex> lock(mu_);
            for (auto& r : interp) {
                r.interpreted = true;
                r.cold = true;
                cold_q_.push_back(std::move(r));
            }
            for (auto& r : plain) {
                r.interpreted = false;
                r.cold = true;
                cold_q_.push_back(std::move(r));
            }
        }
        if (!interp.empty() || !plain.empty()) cold_cv_.notify_all();
        if (!batch.empty()) {
            size_t released = 0;
            std::vector<QueryProgram> programs;
            programs.reserve(batch.size());
            for (auto& r : batch) programs.push_back(*r.program);
            std::vector<const QueryProgram*> templates;
            for (const auto& t : set) templates.push_back(t.get());
            std::vector<QueryResult> results;
            std::string err;
            const auto tb = std::chrono::steady_clock::now();
            // per_template = 2: one launch for the batch when no cold backlog
            // builds and the compiled hot-set module covers every request
            std::vector<std::shared_ptr<QueryProgram>> combined_hold;
            std::vector<const QueryProgram*> combined;
            if (opt_.per_template == 2) {
                std::lock_guard<std::mutex> lock(mu_);
                bool covers = static_cast<int>(cold_q_.size()) < opt_.cold_backlog && !combined_set_.empty();
                for (size_t i = 0; covers && i < programs.size(); ++i) {
                    bool have = false;
                    for (const auto& t : combined_set_) have = have || t->signature == programs[i].signature;
                    covers = have;
                }
                if (covers) {
                    combined_hold = combined_set_;
                    for (const auto& t : combined_hold) combined.push_back(t.get());
                    ++combined_batches_;
                }
            }
            try {
                if (!combined.empty()) {
                    results = lane->exec->executePrograms(programs, opt_.batch, &combined);
                } else if (opt_.per_template) {
                    // one launch per signature, each on its own (precompiled) module
                    BatchOptions bo = opt_.batch;
                    bo.max_templates = 1;
                    results = lane->exec->executePrograms(programs, bo, nullptr);
                } else {
                    results = lane->exec->executePrograms(programs, opt_.batch, &templates);
                }
            } catch (const std::exception& e) {
                err = e.what();
                // start the lane over with a fresh executor (arena, graphs, events)
                try {
                    auto fresh = std::make_unique<BatchExecutor>(db_, lane->ctx.get(), &lane->mu);
                    std::lock_guard<std::mutex> lock(mu_);
                    lane->graph_captures += lane->exec->graphCaptures();
                    lane->graph_replays += lane->exec->graphReplays();
                    lane->graph_failures += lane->exec->graphFailures();
                    lane->exec = std::move(fresh);
                } catch (const std::exception&) {
                }
                std::lock_guard<std::mutex> lock(mu_);
                last_error_ = err;
            }
            // the batch's reservation is returned before its results are
            // published, so a client that has every result sees it released
            for (const auto& r : batch) released += r.footprint;
            {
                std::lock_guard<std::mutex> lock(mu_);
                in_flight_bytes_ -= released;
                batch_wall_ms_ += elapsedMs(tb, std::chrono::steady_clock::now());
                ++batches_;
                batched_queries_ += static_cast<long long>(batch.size());
            }
            for (size_t i = 0; i < batch.size(); ++i) {
                if (err.empty()) finish(batch[i], 0, "", &results[i]);
                else finish(batch[i], -3, err, nullptr);
            }
        }
        cv_.notify_all();
    }
}

// returns a cold request's reservation (before its result is published)
void QueryServer::release(const Request& r) {
    std::lock_guard<std::mutex> lock(mu_);
    in_flight_bytes_ -= r.footprint;
    if (r.interpreted) ++interpreted_queries_;
    else ++plain_queries_;
}

void QueryServer::coldLoop(Lane* lane) {
    for (;;) {
        std::vector<Request> rs;
        {
            std::unique_lock<std::mutex> lock(mu_);
            cold_cv_.wait(lock, [&] { return stop_ || !cold_q_.empty(); });
            if (stop_) return;
            rs.push_back(std::move(cold_q_.front()));
            cold_q_.pop_front();
            // interpreted requests queued behind it run with it
            if (rs.front().interpreted)
                for (auto it = cold_q_.begin(); it != cold_q_.end() && static_cast<int>(rs.size()) < opt_.cold_group;) {
                    if (!it->interpreted) {
                        ++it;
                        continue;
                    }
                    rs.push_back(std::move(*it));
                    it = cold_q_.erase(it);
                }
        }
        ExecutionContext* ctx = lane->ctx.get();
        if (rs.size() > 1) {
            // interpreted requests together: linear forms share scans
            std::vector<std::string> sqls;
            for (const auto& r : rs) sqls.push_back(r.sql);
            std::vector<std::string> errs;
            std::vector<ExecutionContext*> cs;
            try {
                DatabaseInstance::groupContexts(*ctx, lane->group, rs.size());
                for (size_t k = 0; k < rs.size(); ++k) {
                    lane->group[k]->config.codegen.vm = true;
                    cs.push_back(lane->group[k].get());
                }
                errs = db_.engine().executeInterpretedGroup(sqls, cs);
            } catch (const std::exception& e) {
                errs.assign(rs.size(), e.what());
            }
            {
                std::lock_guard<std::mutex> lock(mu_);
                ++cold_groups_;
                cold_grouped_ += static_cast<long long>(rs.size());
            }
            for (size_t k = 0; k < rs.size(); ++k) {
                QueryResult res;
                if (errs[k].empty()) res = std::move(cs[k]->result);
                release(rs[k]);
                if (errs[k].empty()) {
                    finish(rs[k], 0, "", &res);
                } else {
                    finish(rs[k], -3, errs[k], nullptr);
                    std::lock_guard<std::mutex> lock(mu_);
                    last_error_ = errs[k];
                }
            }
            cv_.notify_all();
            continue;
        }
        // a cold signature: its own scan on the interpreted tier (no
        // compilation) or on its plain module once that is compiled
        Request& r = rs.front();
        bool returned = false;
        try {
            ctx->row_begin = 0;
            ctx->row_end = -1;
            ctx->config.codegen.vm = r.interpreted;
            db_.engine().executeQuery(r.sql, ctx);
            ctx->config.codegen.vm = false;
            QueryResult res = std::move(ctx->result);
            release(r);
            returned = true;
            finish(r, 0, "", &res);
        } catch (const std::exception& e) {
            ctx->config.codegen.vm = false;
            if (!returned) release(r);
            finish(r, -3, e.what(), nullptr);
            std::lock_guard<std::mutex> lock(mu_);
            last_error_ = e.what();
        }
        cv_.notify_all();
    }
}

void QueryServer::compileLoop() {
    for (;;) {
        {
            std::unique_lock<std::mutex> lock(mu_);
            compile_cv_.wait(lock, [&] { return stop_ || !compile_q_.empty() || !promote_q_.empty(); });
            if (stop_) return;
        }
        compileOne();
    }
}

bool QueryServer::compileOne() {
    Request r;
    std::shared_ptr<QueryProgram> promote;
    {
        std::lock_guard<std::mutex> lock(mu_);
        if (!compile_q_.empty()) {  // latency first: a request is waiting for it
            r = std::move(compile_q_.front());
            compile_q_.pop_front();
        } else if (!promote_q_.empty()) {
            promote = promote_q_.front();
            promote_q_.pop_front();
        } else {
            return false;
        }
    }
    if (!promote) {
        // plain module of a cold request, then it waits for a lane
        try {
            QueryProgram p = db_.engine().prepare(r.sql, db_.config().codegen);
            QueryTimingStats t;
            db_.engine().moduleFor(p, &t);
            std::lock_guard<std::mutex> lock(mu_);
            if (!t.cache_hit) ++compiles_;
            plain_ready_.push_back(std::move(r));
        } catch (const std::exception& e) {
            finish(r, -2, e.what(), nullptr);
            return true;
        }
        cv_.notify_all();
        return true;
    }
    if (opt_.per_template) {
        // promotion: the template's module and its own shared-scan module
        try {
            db_.engine().moduleFor(*promote, nullptr);
            std::vector<const QueryProgram*> tp{promote.get()};
            std::vector<std::string> cols;
            std::vector<bool> nullable;
            const std::string src =
                generateBatchScanSource(tp, opt_.batch.block_threads, opt_.batch.vectors_per_thread, cols, nullable);
            db_.engine().moduleForSource(src, signatureOf(src), nullptr);
            std::lock_guard<std::mutex> lock(mu_);
            if (!inScanSet(promote->signature)) {
                if (static_cast<int>(scan_set_.size()) >= opt_.max_hot) scan_set_.erase(scan_set_.begin());
                scan_set_.push_back(promote);
                ++promotions_;
            }
            promoting_.erase(promote->signature);
        } catch (const std::exception&) {
            std::lock_guard<std::mutex> lock(mu_);
            promoting_.erase(promote->signature);  // stays on the cold tier
        }
        if (opt_.per_template == 2) compileCombined();
        cv_.notify_all();
        return true;
    }
    // promotion: template module, then the shared-scan module of the grown set
    std::vector<std::shared_ptr<QueryProgram>> set;
    {
        std::lock_guard<std::mutex> lock(mu_);
        set = scan_set_;
    }
    try {
        if (!set.empty() && set.front()->fact_table != promote->fact_table)
            throw std::runtime_error("a server's shared scan covers one fact table (" + set.front()->fact_table + ")");
        db_.engine().moduleFor(*promote, nullptr);
        bool present = false;
        for (const auto& t : set) present = present || t->signature == promote->signature;
        if (!present) {
            if (static_cast<int>(set.size()) >= opt_.max_templates) set.erase(set.begin());
            set.push_back(promote);
        }
        std::vector<const QueryProgram*> tp;
        for (const auto& t : set) tp.push_back(t.get());
        std::vector<std::string> cols;
        std::vector<bool> nullable;
        const std::string src = generateBatchScanSource(tp, opt_.batch.block_threads, opt_.batch.vectors_per_thread, cols, nullable);
        db_.engine().moduleForSource(src, signatureOf(src), nullptr);
        std::lock_guard<std::mutex> lock(mu_);
        // keep templates promoted meanwhile by another thread
        for (const auto& t : scan_set_) {
            bool have = false;
            for (const auto& u : set) have = have || u->signature == t->signature;
            if (!have) {
                set.clear();
                break;
            }
        }
        if (!set.empty()) {
            scan_set_ = set;
            ++promotions_;
            promoting_.erase(promote->signature);
        } else {
            promote_q_.push_back(promote);  // raced with another promotion: redo on the newer set
        }
    } catch (const std::exception&) {
        std::lock_guard<std::mutex> lock(mu_);
        promoting_.erase(promote->signature);  // stays on the plain tier
    }
    cv_.notify_all();
    return true;
}

int QueryServer::wait(long long ticket, int timeout_ms, QueryResult& out, std::string& err) {
    std::unique_lock<std::mutex> lock(mu_);
    auto ready = [&] { return done_.count(ticket) > 0; };
    if (timeout_ms < 0) done_cv_.wait(lock, ready);
    else if (!done_cv_.wait_for(lock, std::chrono::milliseconds(timeout_ms), ready)) {
        err = "timeout";
        return -3;
    }
    Done d = std::move(done_[ticket]);
    done_.erase(ticket);
    if (d.rc != 0) {
        err = d.err;
        return d.rc;
    }
    out = std::move(d.result);
    out.timing.engine_ms = d.latency_ms;
    return 0;
}

// per_template = 2: the one-launch module of the current hot set (one fact
// table, at most max_templates signatures); it takes over once compiled,
// unless the set changed meanwhile (the later promotion compiles its own)
void QueryServer::compileCombined() {
    std::vector<std::shared_ptr<QueryProgram>> set;
    {
        std::lock_guard<std::mutex> lock(mu_);
        set = scan_set_;
    }
    if (set.empty() || static_cast<int>(set.size()) > opt_.max_templates) return;
    for (const auto& t : set)
        if (t->fact_table != set.front()->fact_table) return;
    try {
        std::vector<const QueryProgram*> tp;
        for (const auto& t : set) tp.push_back(t.get());
        std::vector<std::string> cols;
        std::vector<bool> nullable;
        const std::string src = generateBatchScanSource(tp, opt_.batch.block_threads, opt_.batch.vectors_per_thread, cols, nullable);
        db_.engine().moduleForSource(src, signatureOf(src), nullptr);
    } catch (const std::exception&) {
        return;  // batches keep one launch per signature
    }
    std::lock_guard<std::mutex> lock(mu_);
    bool same = set.size() == scan_set_.size();
    for (size_t i = 0; same && i < set.size(); ++i) same = set[i]->signature == scan_set_[i]->signature;
    if (same) combined_set_ = std::move(set);
}

void QueryServer::resetStats() {
    std::lock_guard<std::mutex> lock(mu_);
    latencies_.clear();
    cold_latencies_.clear();
    batch_wall_ms_ = 0;
    cold_groups_ = cold_grouped_ = 0;
    combined_batches_ = 0;
    batches_ = batched_queries_ = plain_queries_ = interpreted_queries_ = admission_waits_ = compiles_ = promotions_ = rejected_ = 0;
    stats_start_ = std::chrono::steady_clock::now();
}

std::string QueryServer::statsJson() {
    std::lock_guard<std::mutex> lock(mu_);
    std::vector<double> l = latencies_, lc = cold_latencies_;
    std::sort(l.begin(), l.end());
    std::sort(lc.begin(), lc.end());
    auto pctOf = [](const std::vector<double>& v, double q) {
        if (v.empty()) return 0.0;
        const size_t i = static_cast<size_t>(q * static_cast<double>(v.size() - 1) + 0.5);
        return v[std::min(v.size() - 1, i)];
    };
    auto pct = [&](double q) { return pctOf(l, q); };
    const double secs = elapsedMs(stats_start_, std::chrono::steady_clock::now()) / 1000.0;
    long long g_capt = 0, g_replay = 0, g_fail = 0;
    for (const auto& l : lanes_) {
        g_capt += l->graph_captures + (l->exec ? l->exec->graphCaptures() : 0);
        g_replay += l->graph_replays + (l->exec ? l->exec->graphReplays() : 0);
        g_fail += l->graph_failures + (l->exec ? l->exec->graphFailures() : 0);
    }
    std::ostringstream o;
    o << "{\"completed\":" << l.size() << ",\"p50_ms\":" << pct(0.50) << ",\"p99_ms\":" << pct(0.99)
      << ",\"max_ms\":" << (l.empty() ? 0.0 : l.back()) << ",\"qps\":" << (secs > 0 ? static_cast<double>(l.size()) / secs : 0.0)
      << ",\"cold_completed\":" << lc.size() << ",\"cold_p50_ms\":" << pctOf(lc, 0.50) << ",\"cold_p99_ms\":" << pctOf(lc, 0.99)
      << ",\"cold_groups\":" << cold_groups_ << ",\"cold_grouped\":" << cold_grouped_
      << ",\"combined_batches\":" << combined_batches_
      << ",\"batch_wall_ms\":" << (batches_ ? batch_wall_ms_ / static_cast<double>(batches_) : 0.0)
      << ",\"batches\":" << batches_ << ",\"mean_batch\":"
      << (batches_ ? static_cast<double>(batched_queries_) / static_cast<double>(batches_) : 0.0)
      << ",\"plain_queries\":" << plain_queries_ << ",\"interpreted_queries\":" << interpreted_queries_
      << ",\"admission_waits\":" << admission_waits_
      << ",\"compiles\":" << compiles_ << ",\"promotions\":" << promotions_ << ",\"rejected\":" << rejected_
      << ",\"graph_captures\":" << g_capt << ",\"graph_replays\":" << g_replay << ",\"graph_failures\":" << g_fail
      << ",\"last_error\":\"" << jsonEscape(last_error_) << "\""
      << ",\"lanes\":" << opt_.lanes << ",\"budget_bytes\":" << opt_.budget_bytes << ",\"in_flight_bytes\":" << in_flight_bytes_
      << ",\"scan_set\":" << scan_set_.size() << "}";
    return o.str();
}

}  // namespace db
