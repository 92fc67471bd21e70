// Batched execution over one shared fact scan per group of templates.
#include "batch.h"

#include "../kernels/aot_kernels.h"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <set>
#include <sstream>
#include <stdexcept>

namespace db {

namespace {
constexpr int kHashMinBits = 7, kHashMaxBits = 10;  // per-query LDS hash tables: 128..1024 slots
constexpr int kMaskQueries = 64;                    // queries of a mask-form template per launch
constexpr int kMaxGraphs = 64;                      // captured batch shapes kept per executor
// dynamic LDS for the per-query regions: what lds_kb allows, minus the room
// kept for shared dimension tables staged into LDS
unsigned ldsLimitFor(const BatchOptions& opt, int table_kb) {
    // 1 KiB more for the kernel's other static LDS
    return std::min(static_cast<unsigned>(opt.lds_kb) * 1024u, 160u * 1024u - static_cast<unsigned>(table_kb) * 1024u - 1024u);
}
unsigned ldsLimit(const BatchOptions& opt) { return ldsLimitFor(opt, opt.table_kb); }
}  // namespace

BatchOptions parseBatchOptions(const std::string& text) {
    BatchOptions o;
    std::stringstream ss(text);
    std::string item;
    while (std::getline(ss, item, ',')) {
        if (item.empty()) continue;
        const size_t eq = item.find('=');
        if (eq == std::string::npos) throw std::runtime_error("batch option without value: " + item);
        const std::string k = item.substr(0, eq);
        const long v = std::stol(item.substr(eq + 1));
        if (k == "bt") o.block_threads = static_cast<int>(v);
        else if (k == "vpt") o.vectors_per_thread = static_cast<int>(v);
        else if (k == "graph") o.graph = v != 0;
        else if (k == "autotune") o.autotune = v != 0;
        else if (k == "max_templates") o.max_templates = static_cast<int>(v);
        else if (k == "max_queries") o.max_queries = static_cast<int>(v);
        else if (k == "shared") o.shared_tables = v != 0;
        else if (k == "lds_agg") o.lds_aggregation = v != 0;
        else if (k == "lds_kb") o.lds_kb = static_cast<int>(v);
        else if (k == "masks") o.masks = v != 0;
        else if (k == "table_mask") o.table_mask = static_cast<unsigned>(v);
        else if (k == "table_kb") o.table_kb = static_cast<int>(v);
        else if (k == "grid_percent") o.grid_percent = static_cast<int>(v);
        else throw std::runtime_error("unknown batch option: " + k);
    }
    if (o.block_threads < 64 || o.block_threads > 1024 || o.block_threads % 64 != 0)
        throw std::runtime_error("batch block size must be a multiple of 64 in [64, 1024]");
    if (o.vectors_per_thread < 1 || o.vectors_per_thread > 4) throw std::runtime_error("batch vpt must be in [1, 4]");
    if (o.lds_kb < 16 || o.lds_kb > 160) throw std::runtime_error("batch lds_kb must be in [16, 160]");
    if (o.table_kb < 16 || o.table_kb > static_cast<int>(kBatchTableLdsMax / 1024)) throw std::runtime_error("batch table_kb must be in [16, 96]");
    if (o.grid_percent < 1 || o.grid_percent > 100) throw std::runtime_error("batch grid_percent must be in [1, 100]");
    o.max_templates = std::max(1, std::min(o.max_templates, CQ_MAX_TEMPLATES));
    o.max_queries = std::max(1, o.max_queries);
    return o;
}

std::string generateBatchCode(const DatabaseInstance& db, const std::vector<std::string>& sqls, const BatchOptions& opt,
                              int part) {
    CodegenOptions co = db.config().codegen;
    co.batch_body = true;
    co.block_threads = opt.block_threads;
    co.vectors_per_thread = opt.vectors_per_thread;
    co.batch_shared_tables = opt.shared_tables;
    co.batch_masks = opt.masks;
    co.allow_lds_aggregation = co.allow_lds_aggregation && opt.lds_aggregation;
    std::vector<QueryProgram> programs;
    for (const auto& q : sqls) programs.push_back(db.engine().prepare(q, co));
    if (part >= 0) {
        if (static_cast<size_t>(part) >= programs.size()) throw std::runtime_error("batch part out of range");
        return programs[static_cast<size_t>(part)].source;
    }
    std::vector<const QueryProgram*> templates;
    for (const auto& p : programs) {
        bool seen = false;
        for (const auto* t : templates) seen = seen || t->signature == p.signature;
        if (!seen) templates.push_back(&p);
    }
    // the order a launch uses (runScan): by signature
    std::sort(templates.begin(), templates.end(),
              [](const QueryProgram* a, const QueryProgram* b) { return a->signature < b->signature; });
    std::vector<std::string> cols;
    std::vector<bool> nullable;
    return generateBatchScanSource(templates, opt.block_threads, opt.vectors_per_thread, cols, nullable, nullptr, opt.table_mask, nullptr,
                                   static_cast<unsigned>(opt.table_kb) * 1024u);
}

BatchExecutor::BatchExecutor(DatabaseInstance& db)
    : db_(db), ctx_(db.hostOnly() ? nullptr : &db.context()), ctx_mu_(&db.contextMutex()) {}

BatchExecutor::BatchExecutor(DatabaseInstance& db, ExecutionContext* ctx, std::mutex* ctx_mutex)
    : db_(db), ctx_(ctx), ctx_mu_(ctx_mutex) {}

CodegenOptions BatchExecutor::codegenOptions(const BatchOptions& opt) const {
    CodegenOptions co = db_.config().codegen;
    co.batch_body = true;
    co.block_threads = opt.block_threads;
    co.vectors_per_thread = opt.vectors_per_thread;
    co.batch_shared_tables = opt.shared_tables;
    co.batch_masks = opt.masks;
    co.allow_lds_aggregation = co.allow_lds_aggregation && opt.lds_aggregation;
    return co;
}

BatchExecutor::~BatchExecutor() {
    if (db_.hostOnly()) return;
    (void)hipSetDevice(db_.device());
    for (auto& kv : graphs_) dropGraph(kv.second);
    for (auto& e : ev_)
        if (e) (void)hipEventDestroy(e);
    if (pinned_) {
        DeviceMapLock lock;
        (void)hipHostFree(pinned_);
    }
}

void BatchExecutor::dropGraph(Graph& g) {
    if (g.exec || g.graph) {
        DeviceMapLock lock;
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        if (g.graph) (void)hipGraphDestroy(g.graph);
    }
    g.exec = nullptr;
    g.graph = nullptr;
    g.key.clear();
    g.modules.clear();
}

void* BatchExecutor::pinned(size_t bytes) {
    if (bytes <= pinned_bytes_) return pinned_;
    DeviceMapLock lock;
    if (pinned_) hipCheck(hipHostFree(pinned_), "hipHostFree");
    pinned_ = nullptr;
    pinned_bytes_ = 0;
    const size_t want = std::max<size_t>(bytes, 64 * 1024);
    hipCheck(hipHostMalloc(&pinned_, want, hipHostMallocDefault), "hipHostMalloc");
    pinned_bytes_ = want;
    return pinned_;
}

std::vector<QueryResult> BatchExecutor::execute(const std::vector<std::string>& sqls, const std::string& options) {
    return execute(sqls, parseBatchOptions(options));
}

std::vector<QueryResult> BatchExecutor::execute(const std::vector<std::string>& sqls, const BatchOptions& opt) {
    if (db_.hostOnly()) throw std::runtime_error("batch execution needs a GPU device (host-only database)");
    // plans -> programs carrying the per-tile pipeline body
    const auto t0 = std::chrono::steady_clock::now();
    const CodegenOptions co = codegenOptions(opt);
    std::vector<QueryProgram> programs(sqls.size());
    for (size_t i = 0; i < sqls.size(); ++i) {
        try {
            programs[i] = db_.engine().prepareCached(sqls[i], co);
        } catch (const std::exception& e) {
            throw std::runtime_error("batch query " + std::to_string(i) + ": " + e.what());
        }
    }
    plan_ms_ = elapsedMs(t0, std::chrono::steady_clock::now());
    return executePrograms(programs, opt, nullptr);
}

std::vector<QueryResult> BatchExecutor::executePrograms(std::vector<QueryProgram>& programs, const BatchOptions& opt,
                                                        const std::vector<const QueryProgram*>* module_templates) {
    if (db_.hostOnly() || !ctx_) throw std::runtime_error("batch execution needs a GPU device (host-only database)");
    std::lock_guard<std::mutex> lock(*ctx_mu_);
    hipCheck(hipSetDevice(db_.device()), "hipSetDevice");
    for (auto& e : ev_)
        if (!e) hipCheck(hipEventCreate(&e), "hipEventCreate");
    const auto t0 = std::chrono::steady_clock::now();
    const double plan_ms = plan_ms_;
    plan_ms_ = 0;
    for (const auto& p : programs)
        if (p.batch_body.empty()) throw std::runtime_error("batch program generated without its per-tile body");

    // queries of one template stay together; templates are packed into
    // shared scans of the same fact table within the template / column /
    // query limits of one launch
    std::vector<std::string> order;  // (fact table, signature) keys in first-seen order
    std::map<std::string, std::vector<int>> members;
    for (size_t i = 0; i < programs.size(); ++i) {
        const std::string key = programs[i].fact_table + "|" + programs[i].signature;
        if (!members.count(key)) order.push_back(key);
        members[key].push_back(static_cast<int>(i));
    }
    std::vector<std::vector<int>> scans;
    if (module_templates) {
        // one launch of the caller's module per chunk that fits the LDS regions
        std::vector<int> cur;
        uint64_t lds = 0;
        std::map<std::string, int> per_sig;  // queries of each signature in `cur`
        for (size_t i = 0; i < programs.size(); ++i) {
            const uint64_t need = programs[i].batch_lds_stride;
            const bool mask_full = programs[i].batch_mask && per_sig[programs[i].signature] >= kMaskQueries;
            if (!cur.empty() && (lds + need > ldsLimit(opt) || mask_full)) {
                scans.push_back(cur);
                cur.clear();
                per_sig.clear();
                lds = 0;
            }
            cur.push_back(static_cast<int>(i));
            ++per_sig[programs[i].signature];
            lds += need;
        }
        if (!cur.empty()) scans.push_back(cur);
        order.clear();
    }
    std::vector<int> cur;
    std::vector<std::string> cur_cols;
    std::string cur_table;
    std::set<std::string> cur_keys;  // templates in `cur` (a template's chunks go to different launches)
    int cur_templates = 0;
    uint64_t cur_lds = 0;
    for (const auto& key : order) {
        const auto& qs = members[key];
        const QueryProgram& t = programs[static_cast<size_t>(qs.front())];
        // queries of one template per launch: bounded by max_queries and by
        // their LDS aggregation regions
        size_t chunk = static_cast<size_t>(opt.max_queries);
        if (t.batch_lds_stride > 0) chunk = std::min<size_t>(chunk, std::max<uint64_t>(1, ldsLimit(opt) / t.batch_lds_stride));
        if (t.batch_mask) chunk = std::min<size_t>(chunk, kMaskQueries);  // one bit per query in a 64-bit mask
        for (size_t start = 0; start < qs.size(); start += chunk) {
            const size_t stop = std::min(qs.size(), start + chunk);
            const uint64_t lds_need = t.batch_lds_stride * (stop - start);
            std::vector<std::string> cols = cur_cols;
            for (const auto& c : t.fact_cols)
                if (std::find(cols.begin(), cols.end(), c) == cols.end()) cols.push_back(c);
            const bool fits = !cur.empty() && !cur_keys.count(key) && cur_table == t.fact_table && cur_templates + 1 <= opt.max_templates &&
                              cols.size() <= CQ_MAX_BATCH_COLS &&
                              cur.size() + (stop - start) <= static_cast<size_t>(opt.max_queries) &&
                              cur_lds + lds_need <= ldsLimit(opt);
            if (!fits && !cur.empty()) {
                scans.push_back(cur);
                cur.clear();
                cur_keys.clear();
                cur_templates = 0;
                cur_lds = 0;
                cols.assign(t.fact_cols.begin(), t.fact_cols.end());
            }
            cur_table = t.fact_table;
            cur_cols = cols;
            cur_keys.insert(key);
            ++cur_templates;
            cur_lds += lds_need;
            for (size_t k = start; k < stop; ++k) cur.push_back(qs[k]);
        }
    }
    if (!cur.empty()) scans.push_back(cur);

    std::vector<QueryResult> results(programs.size());
    std::ostringstream r;
    r << "{\"queries\":" << programs.size() << ",\"shared_scans\":" << scans.size() << ",\"plan_ms\":" << plan_ms << ",\"scans\":[";
    for (size_t s = 0; s < scans.size(); ++s) {
        BatchScanReport rep;
        runScan(scans[s], programs, opt, static_cast<int>(s), results, rep, module_templates);
        r << (s ? "," : "") << "{\"queries\":" << rep.queries << ",\"templates\":" << rep.templates
          << ",\"union_columns\":" << rep.union_columns << ",\"block_threads\":" << rep.block_threads
          << ",\"vectors_per_thread\":" << rep.vectors_per_thread << ",\"table_mask\":" << rep.table_mask
          << ",\"lds_table_bytes\":" << rep.lds_table_bytes << ",\"table_kb\":" << rep.table_kb << ",\"graph_replayed\":" << (rep.graph_replayed ? 1 : 0)
          << ",\"graph\":\"" << rep.graph_mode << "\""
          << ",\"scan_ms\":" << rep.scan_ms << ",\"device_ms\":" << rep.device_ms << ",\"fact_bytes\":" << rep.fact_bytes
          << ",\"host_ms\":{\"modules\":" << rep.modules_ms << ",\"bind\":" << rep.bind_ms << ",\"enqueue\":" << rep.enqueue_ms
          << ",\"wait\":" << rep.wait_ms << ",\"fetch\":" << rep.fetch_ms << "},\"arena_bytes\":" << rep.arena_bytes << "}";
    }
    r << "],\"graphs\":" << graphs_.size() << ",\"graph_failures\":" << graph_failures_.load() << ",\"last_graph_error\":\"";
    for (char c : last_graph_error_.substr(0, 200)) r << ((c == '"' || c == '\\' || static_cast<unsigned char>(c) < 0x20) ? ' ' : c);
    r << "\",\"wall_ms\":" << elapsedMs(t0, std::chrono::steady_clock::now()) << "}";
    report_ = r.str();
    return results;
}

namespace {

std::string batchSource(const std::vector<const QueryProgram*>& templates, int bt, int vpt, std::vector<std::string>& cols,
                        std::vector<bool>& nullable, unsigned table_mask = ~0u, unsigned* table_lds_bytes = nullptr,
                        int* lds_candidates = nullptr, int table_kb = 32) {
    return generateBatchScanSource(templates, bt, vpt, cols, nullable, table_lds_bytes, table_mask, lds_candidates,
                                   static_cast<unsigned>(table_kb) * 1024u);
}

}  // namespace

namespace {

// Per-query LDS regions of one launch: offset and stride per template, and
// for hash-aggregating templates the table size, grown one bit at a time for
// all of them while the launch stays within the limit.  Returns the bytes.
unsigned layoutLds(const std::vector<const QueryProgram*>& templates, const std::vector<std::vector<int>>& per_template,
                   unsigned limit, cq_batch_params& B) {
    const size_t nt = templates.size();
    std::vector<int> bits(nt, kHashMinBits);
    auto stride = [&](size_t t) -> uint64_t {
        const QueryProgram* p = templates[t];
        return p->batch_hash_slot_bytes > 0 ? p->batch_hash_slot_bytes << bits[t] : p->batch_lds_stride;
    };
    auto total = [&]() {
        uint64_t b = 0;
        for (size_t t = 0; t < nt; ++t) b += stride(t) * per_template[t].size();
        return b;
    };
    if (total() > limit) throw std::runtime_error("batch LDS aggregation regions exceed the limit");
    for (int b = kHashMinBits + 1; b <= kHashMaxBits; ++b) {
        for (size_t t = 0; t < nt; ++t) {
            // only tables that reached b - 1 grow again
            if (templates[t]->batch_hash_slot_bytes == 0 || per_template[t].empty() || bits[t] != b - 1) continue;
            bits[t] = b;
            if (total() > limit) bits[t] = b - 1;
        }
    }
    unsigned off = 0;
    for (size_t t = 0; t < nt; ++t) {
        B.lds_off[t] = off;
        B.lds_stride[t] = static_cast<unsigned>(stride(t));
        B.hbits[t] = static_cast<unsigned>(bits[t]);
        off += static_cast<unsigned>(stride(t) * per_template[t].size());
    }
    if (off > limit) throw std::logic_error("batch LDS layout exceeds its limit");
    return off;
}

// persistent grid for a launch with dynamic shared memory, `percent` of the
// blocks that fit on the device at once (at least one)
unsigned scanGridShared(hipFunction_t f, int block_threads, int device, long long num_tiles, unsigned lds_bytes,
                        int percent) {
    int per_cu = 0;
    if (hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, block_threads, lds_bytes) != hipSuccess || per_cu < 1)
        per_cu = 1;
    hipDeviceProp_t prop;
    hipCheck(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties");
    const long long max_blocks = static_cast<long long>(per_cu) * prop.multiProcessorCount * percent / 100;
    return static_cast<unsigned>(std::max<long long>(1, std::min<long long>(num_tiles, max_blocks)));
}

}  // namespace

// Zero and build every table of the batch: a template's shared tables once
// (from its first query's parameter block), per-query tables in setup().
void BatchExecutor::prepareTables(const std::vector<const QueryProgram*>& templates,
                                  const std::vector<std::vector<int>>& per_template, std::vector<QueryRun>& runs,
                                  const cq_query_params* dqp, const BatchInit& init, ExecutionContext* ctx) {
    HipModuleExecutor* ex = db_.engine().hipExecutor();
    int first = 0;
    for (size_t t = 0; t < templates.size(); ++t) {
        const int cnt = static_cast<int>(per_template[t].size());
        if (cnt > 0 && templates[t]->shared_tables) {
            QueryRun& r0 = runs[static_cast<size_t>(first)];
            for (size_t j = 0; j < templates[t]->joins.size(); ++j)
                hipCheck(cq_launch_fill_u32(r0.P.ht[j], 0u, static_cast<long long>(templates[t]->joins[j].ht_words), ctx->stream),
                         "fill shared table");
            ex->buildTables(r0, ctx, dqp + first);
        }
        first += cnt;
    }
    hipCheck(cq_launch_batch_init(init.d, static_cast<int>(runs.size()), init.max_slots, ctx->stream), "batch init");
    for (size_t k = 0; k < runs.size(); ++k)
        if (!init.eligible[k]) ex->setup(runs[k], ctx, dqp + k);
}

void BatchExecutor::runScan(const std::vector<int>& members, std::vector<QueryProgram>& programs, const BatchOptions& opt,
                            int scan_index, std::vector<QueryResult>& results, BatchScanReport& rep,
                            const std::vector<const QueryProgram*>* module_templates) {
    ExecutionContext* ctx = ctx_;
    QueryEngine& engine = db_.engine();
    HipModuleExecutor* ex = engine.hipExecutor();
    if (!ex) throw std::runtime_error("batch execution needs the HIP module executor");
    hipStream_t s = ctx->stream;

    // module templates (given, or the members' in first-seen order) and
    // the members grouped by template
    std::vector<const QueryProgram*> templates;
    if (module_templates) templates = *module_templates;
    std::vector<std::vector<int>> per_template(templates.size());
    for (int q : members) {
        const QueryProgram& p = programs[static_cast<size_t>(q)];
        size_t t = 0;
        while (t < templates.size() && templates[t]->signature != p.signature) ++t;
        if (t == templates.size()) {
            if (module_templates) throw std::runtime_error("batch query template missing from the shared-scan module");
            templates.push_back(&p);
            per_template.emplace_back();
        }
        per_template[t].push_back(q);
    }
    for (size_t t = 0; t < templates.size(); ++t)
        if (templates[t]->batch_mask && per_template[t].size() > static_cast<size_t>(kMaskQueries))
            throw std::logic_error("mask-form template with more than 64 queries in one launch");
    if (templates.empty()) return;
    if (templates.size() > static_cast<size_t>(CQ_MAX_TEMPLATES)) throw std::runtime_error("too many templates in one scan");
    if (!module_templates) {
        // canonical order: the module depends on the set of templates only
        std::vector<size_t> ord(templates.size());
        for (size_t t = 0; t < ord.size(); ++t) ord[t] = t;
        std::sort(ord.begin(), ord.end(), [&](size_t a, size_t b) { return templates[a]->signature < templates[b]->signature; });
        std::vector<const QueryProgram*> t2;
        std::vector<std::vector<int>> p2;
        for (size_t t : ord) {
            t2.push_back(templates[t]);
            p2.push_back(std::move(per_template[t]));
        }
        templates.swap(t2);
        per_template.swap(p2);
    }
    std::vector<int> qorder;
    for (const auto& v : per_template) qorder.insert(qorder.end(), v.begin(), v.end());
    const int n = static_cast<int>(qorder.size());
    std::string set_key;
    for (const auto* t : templates) set_key += t->signature + ";";

    // per-template modules (dimension builds + result flags)
    const auto h0 = std::chrono::steady_clock::now();
    std::vector<ModuleHandle> tmodules;
    for (size_t t = 0; t < templates.size(); ++t)
        tmodules.push_back(per_template[t].empty() ? nullptr : engine.moduleFor(*templates[t], nullptr));

    // bind every query: a measuring pass sizes one arena for all of their
    // scratch, the second pass hands out the addresses
    const auto h1 = std::chrono::steady_clock::now();
    ctx->row_begin = 0;
    ctx->row_end = -1;
    std::vector<QueryRun> runs(static_cast<size_t>(n));
    std::vector<size_t> tix(static_cast<size_t>(n));
    for (int k = 0; k < n; ++k) {
        const int q = qorder[static_cast<size_t>(k)];
        size_t t = 0;
        while (templates[t]->signature != programs[static_cast<size_t>(q)].signature) ++t;
        tix[static_cast<size_t>(k)] = t;
    }
    const std::string pre = "s" + std::to_string(scan_index) + "/";
    const size_t qp_bytes = static_cast<size_t>(n) * sizeof(cq_query_params);
    const size_t init_bytes = static_cast<size_t>(n) * sizeof(cq_init_desc);
    const size_t shape_bytes = static_cast<size_t>(n) * sizeof(cq_shape_desc);
    const size_t upload_bytes = qp_bytes + init_bytes + shape_bytes;
    ScratchArena arena;
    cq_query_params* dqp = nullptr;
    unsigned long long* dtot = nullptr;
    for (int pass = 0; pass < 2; ++pass) {
        if (pass == 0) arena.measure();
        else {
            DeviceBuffer& ab = ctx->scratch(pre + "arena");
            ab.ensure(arena.used());
            arena.assign(ab.data(), ab.capacity());
        }
        for (int k = 0; k < n; ++k)
            ex->bind(tmodules[tix[static_cast<size_t>(k)]], programs[static_cast<size_t>(qorder[static_cast<size_t>(k)])], ctx, pre,
                     runs[static_cast<size_t>(k)], &arena);
        // one table per join for each template with shared tables
        int first = 0;
        for (size_t t = 0; t < templates.size(); ++t) {
            const int cnt = static_cast<int>(per_template[t].size());
            if (cnt > 0 && templates[t]->shared_tables) {
                for (size_t j = 0; j < templates[t]->joins.size(); ++j) {
                    auto* tab = static_cast<unsigned int*>(arena.take(std::max<size_t>(templates[t]->joins[j].ht_words * 4, 16)));
                    for (int k = first; k < first + cnt; ++k) runs[static_cast<size_t>(k)].P.ht[j] = tab;
                }
            }
            first += cnt;
        }
        // parameter blocks, then the init and shaping descriptors: one upload
        dqp = static_cast<cq_query_params*>(arena.take(upload_bytes));
        dtot = static_cast<unsigned long long*>(arena.take(static_cast<size_t>(n) * 8));
    }
    auto* dinit = reinterpret_cast<cq_init_desc*>(reinterpret_cast<char*>(dqp) + qp_bytes);
    auto* dshape = reinterpret_cast<cq_shape_desc*>(reinterpret_cast<char*>(dqp) + qp_bytes + init_bytes);

    // Per-query aggregate-state init and result shaping run as one launch
    // for the whole batch where the query allows it: compiled, shared
    // tables (init); at most kBatchShapeMaxGroups groups and
    // kBatchShapeMaxCols columns, no AVG finalization, not a single cell
    // (shaping).  The others take the per-query path.
    BatchInit init;
    init.d = dinit;
    init.eligible.assign(static_cast<size_t>(n), 0);
    std::vector<char> shape_ok(static_cast<size_t>(n), 0);
    std::vector<cq_init_desc> hinit(static_cast<size_t>(n));
    std::vector<cq_shape_desc> hshape(static_cast<size_t>(n));
    for (int k = 0; k < n; ++k) {
        const QueryRun& rr = runs[static_cast<size_t>(k)];
        const QueryProgram& p = *rr.program;
        cq_init_desc& id = hinit[static_cast<size_t>(k)];
        std::memset(&id, 0, sizeof id);
        if (p.shared_tables && !p.vm) {
            id.res = rr.P.result;
            id.rv = rr.P.result_valid;
            id.slots = static_cast<long long>(rr.slots);
            id.G = rr.G;
            id.sts = p.storage_ts;
            for (int slot : p.min_slots) id.min_mask |= 1u << slot;
            if (!p.grouped && rr.P.row_end > rr.P.row_begin)
                for (int slot : p.count_slots_scalar)
                    if (slot < 64) id.count_bits |= 1ull << slot;
            init.eligible[static_cast<size_t>(k)] = 1;
            init.max_slots = std::max(init.max_slots, id.slots);
        }
        cq_shape_desc& sd = hshape[static_cast<size_t>(k)];
        std::memset(&sd, 0, sizeof sd);
        sd.G = -1;
        if (!p.vm && !p.projection && !p.has_avg && !rr.single_cell && rr.G <= kBatchShapeMaxGroups &&
            p.visible_ts <= kBatchShapeMaxCols) {
            sd.res = rr.res;
            sd.rv = rr.rv;
            sd.fl = rr.fl;
            sd.o = rr.o;
            sd.ov = rr.ov;
            sd.rows = dtot + k;
            sd.G = rr.G;
            sd.cap_rows = rr.cap_rows;
            sd.vwords = rr.vwords;
            sd.vts = p.visible_ts;
            sd.sparse = p.fast_sparse ? 1 : 0;
            sd.use_validity = p.fast_sparse ? 0 : 1;
            for (int c = 0; c < p.visible_ts; ++c)
                if (p.out[static_cast<size_t>(c)].nullable) sd.nullable_mask |= 1u << c;
            shape_ok[static_cast<size_t>(k)] = 1;
        }
    }

    // parameter blocks and descriptors -> device (pinned staging, stream ordered)
    char* stage = static_cast<char*>(pinned(upload_bytes + static_cast<size_t>(n) * 8));
    for (int k = 0; k < n; ++k) std::memcpy(stage + static_cast<size_t>(k) * sizeof(cq_query_params), &runs[static_cast<size_t>(k)].P,
                                            sizeof(cq_query_params));
    std::memcpy(stage + qp_bytes, hinit.data(), init_bytes);
    std::memcpy(stage + qp_bytes + init_bytes, hshape.data(), shape_bytes);
    hipCheck(hipMemcpyAsync(dqp, stage, upload_bytes, hipMemcpyHostToDevice, s), "upload parameter blocks");

    // launch shape: tuned per template set when asked
    Tuned cfg{opt.block_threads, opt.vectors_per_thread, opt.table_mask, opt.table_kb};
    auto tit = tuned_.find(set_key);
    if (tit != tuned_.end()) cfg = tit->second;
    else if (opt.autotune) {
        cfg = autotune(templates, per_template, runs, dqp, init, opt, ctx);
        tuned_[set_key] = cfg;
    }
    // a tuned table budget leaves less LDS to the per-query regions: a
    // launch whose regions no longer fit takes the default budget
    if (cfg.table_kb != opt.table_kb) {
        cq_batch_params probe;
        std::memset(&probe, 0, sizeof probe);
        try {
            (void)layoutLds(templates, per_template, ldsLimitFor(opt, cfg.table_kb), probe);
        } catch (const std::runtime_error&) {
            cfg.table_kb = opt.table_kb;
            cfg.table_mask = opt.table_mask;
        }
    }
    std::vector<std::string> ucols;
    std::vector<bool> unull;
    unsigned table_bytes = 0;
    const std::string src = batchSource(templates, cfg.block_threads, cfg.vectors_per_thread, ucols, unull, cfg.table_mask, &table_bytes,
                                        nullptr, cfg.table_kb);
    ModuleHandle bm = engine.moduleForSource(src, signatureOf(src), nullptr);
    hipFunction_t fn = bm->function("cq_batch_scan");

    cq_batch_params B;
    std::memset(&B, 0, sizeof B);
    const QueryProgram& first = *templates.front();
    for (size_t u = 0; u < ucols.size(); ++u) {
        auto it = ctx->columns->find(ucols[u]);
        if (it == ctx->columns->end()) throw std::runtime_error("column not loaded on the device: " + ucols[u]);
        B.col[u] = it->second.data;
        B.valid[u] = reinterpret_cast<const unsigned int*>(it->second.validity);
        if (unull[u] && !it->second.validity) throw std::runtime_error("nullable column without bitmap: " + ucols[u]);
    }
    B.qp = dqp;
    int at = 0;
    for (size_t t = 0; t < templates.size(); ++t) {
        B.tq_begin[t] = at;
        at += static_cast<int>(per_template[t].size());
        B.tq_end[t] = at;
    }
    B.row_begin = 0;
    B.row_end = first.fact_rows;
    const long long tile = static_cast<long long>(cfg.block_threads) * 4 * cfg.vectors_per_thread;
    B.num_tiles = (B.row_end - B.row_begin + tile - 1) / tile;
    const unsigned lds_bytes = layoutLds(templates, per_template, ldsLimitFor(opt, cfg.table_kb), B);
    const unsigned grid = scanGridShared(fn, cfg.block_threads, ctx->device, B.num_tiles, lds_bytes, opt.grid_percent);

    auto enqueue = [&](bool timed) {
        prepareTables(templates, per_template, runs, dqp, init, ctx);
        if (timed) hipCheck(hipEventRecord(ev_[2], s), "hipEventRecord");
        void* args[] = {&B};
        if (B.num_tiles > 0) {
            const hipError_t e =
                hipModuleLaunchKernel(fn, grid, 1, 1, static_cast<unsigned>(cfg.block_threads), 1, 1, lds_bytes, s, args, nullptr);
            if (e != hipSuccess) {
                (void)hipGetLastError();  // a failed launch must not fail the next kernel launch
                int max_threads = 0, regs = 0, max_dyn = 0;
                (void)hipFuncGetAttribute(&max_threads, HIP_FUNC_ATTRIBUTE_MAX_THREADS_PER_BLOCK, fn);
                (void)hipFuncGetAttribute(&regs, HIP_FUNC_ATTRIBUTE_NUM_REGS, fn);
                (void)hipFuncGetAttribute(&max_dyn, HIP_FUNC_ATTRIBUTE_MAX_DYNAMIC_SHARED_SIZE_BYTES, fn);
                throw std::runtime_error("launch batch scan (block " + std::to_string(cfg.block_threads) + ", grid " + std::to_string(grid) +
                                         ", LDS " + std::to_string(lds_bytes) + " B; kernel: max block " + std::to_string(max_threads) +
                                         ", registers " + std::to_string(regs) + ", max dynamic LDS " + std::to_string(max_dyn) +
                                         "): " + hipGetErrorString(e));
            }
        }
        if (timed) hipCheck(hipEventRecord(ev_[3], s), "hipEventRecord");
        // result shaping: per template, group keys and present / HAVING
        // flags of all its queries in one launch each, then the rows of every
        // eligible query of the batch in one launch
        int tfirst = 0;
        for (size_t t = 0; t < templates.size(); ++t) {
            const int cnt = static_cast<int>(per_template[t].size());
            if (cnt > 0 && shape_ok[static_cast<size_t>(tfirst)]) {
                const QueryRun& r0 = runs[static_cast<size_t>(tfirst)];
                const unsigned gx = static_cast<unsigned>((r0.G + 255) / 256);
                const cq_query_params* qp = dqp + tfirst;
                const cq_shape_desc* sdp = dshape + tfirst;
                if (r0.keys) {
                    void* kargs[] = {&qp};
                    hipCheck(hipModuleLaunchKernel(tmodules[t]->function("cq_keys_b"), gx, static_cast<unsigned>(cnt), 1, 256, 1, 1, 0, s,
                                                   kargs, nullptr),
                             "launch batched keys");
                }
                if (!r0.program->fast_sparse) {
                    void* dargs[] = {&qp, &sdp};
                    hipCheck(hipModuleLaunchKernel(tmodules[t]->function("cq_dense_count_b"), gx, static_cast<unsigned>(cnt), 1, 256, 1, 1,
                                                   0, s, dargs, nullptr),
                             "launch batched dense count");
                }
            }
            tfirst += cnt;
        }
        hipCheck(cq_launch_batch_shape(dshape, n, s), "batch shape");
        for (int k = 0; k < n; ++k)
            if (!shape_ok[static_cast<size_t>(k)]) ex->shape(runs[static_cast<size_t>(k)], ctx, dqp + k, dtot + k);
    };

    // A graph pays off for a batch shape that comes back: the first time a
    // shape is seen it runs eagerly, the second time it is captured, later
    // it is replayed (only the uploaded parameter blocks differ).
    const auto h2 = std::chrono::steady_clock::now();
    hipCheck(hipEventRecord(ev_[0], s), "hipEventRecord");
    bool replayed = false, timed = true;
    if (opt.graph) {
        // everything a replay depends on besides the uploaded parameter blocks
        std::ostringstream key;
        key << bm->signature() << "|" << grid << "|" << cfg.block_threads << "|" << lds_bytes << "|" << B.num_tiles << "|" << dqp
            << "|" << dtot;
        for (size_t u = 0; u < ucols.size(); ++u) key << "|" << B.col[u] << "," << B.valid[u];
        for (int k = 0; k < n; ++k) {
            const QueryRun& rr = runs[static_cast<size_t>(k)];
            key << "|" << rr.program->signature << ":" << rr.P.result << "," << rr.P.row_end;
        }
        const std::string ks = key.str();
        auto git = graphs_.find(ks);
        if (git != graphs_.end()) {
            hipCheck(hipGraphLaunch(git->second.exec, s), "hipGraphLaunch");
            replayed = true;
            timed = false;
            rep.graph_mode = "replayed";
            ++graph_replays_;
        } else if (++graph_seen_[ks] >= 2) {
            if (graphs_.size() >= static_cast<size_t>(kMaxGraphs)) {  // keep a bounded set of shapes
                dropGraph(graphs_.begin()->second);
                graphs_.erase(graphs_.begin());
            }
            // Capture under the shared device-map lock: no thread of this
            // process allocates, frees or synchronizes the device meanwhile.
            // A capture that fails leaves no entry behind and the batch runs
            // eagerly instead (the captured work was never launched).
            Graph g;
            bool ok = false;
            std::string why;
            {
                CaptureLock capture;
                hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
                if (e == hipSuccess) {
                    try {
                        enqueue(false);
                    } catch (const std::exception& ex) {
                        why = ex.what();
                    }
                    hipGraph_t captured = nullptr;
                    const hipError_t ee = hipStreamEndCapture(s, &captured);
                    if (why.empty() && ee != hipSuccess) why = std::string("hipStreamEndCapture: ") + hipGetErrorString(ee);
                    if (why.empty() && captured) {
                        e = hipGraphInstantiate(&g.exec, captured, nullptr, nullptr, 0);
                        if (e != hipSuccess) why = std::string("hipGraphInstantiate: ") + hipGetErrorString(e);
                    }
                    g.graph = captured;
                    ok = why.empty() && g.exec != nullptr;
                } else {
                    why = std::string("hipStreamBeginCapture: ") + hipGetErrorString(e);
                }
            }
            if (ok) {
                g.key = ks;
                g.modules = tmodules;
                g.modules.push_back(bm);
                for (const QueryRun& rr : runs) g.modules.push_back(rr.module);
                hipCheck(hipGraphLaunch(g.exec, s), "hipGraphLaunch");
                graphs_.emplace(ks, std::move(g));
                ++graph_captures_;
                timed = false;
                rep.graph_mode = "captured";
            } else {
                (void)hipGetLastError();
                dropGraph(g);
                ++graph_failures_;
                last_graph_error_ = why;
                enqueue(true);
                rep.graph_mode = "capture-failed";
            }
        } else {
            enqueue(true);
        }
        if (graph_seen_.size() > 1024) graph_seen_.clear();
    } else {
        enqueue(true);
    }
    hipCheck(hipEventRecord(ev_[1], s), "hipEventRecord");
    auto* htot = reinterpret_cast<unsigned long long*>(stage + upload_bytes);
    hipCheck(hipMemcpyAsync(htot, dtot, static_cast<size_t>(n) * 8, hipMemcpyDeviceToHost, s), "copy row counts");
    const auto h3 = std::chrono::steady_clock::now();
    hipCheck(hipStreamSynchronize(s), "batch sync");
    const auto h4 = std::chrono::steady_clock::now();
    float dms = 0, sms = 0;
    (void)hipEventElapsedTime(&dms, ev_[0], ev_[1]);
    if (timed) (void)hipEventElapsedTime(&sms, ev_[2], ev_[3]);

    // results: sort / gather and copies for every query, one synchronization
    std::vector<unsigned long long> rows(htot, htot + n);
    std::vector<size_t> off(static_cast<size_t>(n) + 1, 0);
    for (int k = 0; k < n; ++k)
        off[static_cast<size_t>(k) + 1] = off[static_cast<size_t>(k)] +
                                          ((ex->fetchBytes(runs[static_cast<size_t>(k)], rows[static_cast<size_t>(k)]) + 255) & ~size_t(255));
    // every result range gathered on the device into one buffer, then one
    // device-to-host copy (instead of a copy per column and query)
    const size_t fetch_bytes = off.back();
    std::vector<cq_copy_seg> segs;
    for (int k = 0; k < n; ++k)
        ex->fetchSegments(runs[static_cast<size_t>(k)], ctx, rows[static_cast<size_t>(k)], nullptr, off[static_cast<size_t>(k)] / 8, segs);
    const size_t seg_bytes = segs.size() * sizeof(cq_copy_seg);
    char* host = static_cast<char*>(pinned(fetch_bytes + seg_bytes + 256));
    char* hseg = host + ((fetch_bytes + 255) & ~size_t(255));
    if (!segs.empty()) std::memcpy(hseg, segs.data(), seg_bytes);
    DeviceBuffer& dfetch = ctx->scratch("batch_fetch");
    dfetch.ensure(((fetch_bytes + 255) & ~size_t(255)) + seg_bytes + 8);
    char* dbase = static_cast<char*>(dfetch.data());
    cq_copy_seg* dseg = reinterpret_cast<cq_copy_seg*>(dbase + ((fetch_bytes + 255) & ~size_t(255)));
    if (!segs.empty()) {
        hipCheck(hipMemcpyAsync(dseg, hseg, seg_bytes, hipMemcpyHostToDevice, s), "upload fetch segments");
        hipCheck(cq_launch_copy_segments(dseg, static_cast<int>(segs.size()), reinterpret_cast<unsigned long long*>(dbase), s),
                 "gather results");
        hipCheck(hipMemcpyAsync(host, dbase, fetch_bytes, hipMemcpyDeviceToHost, s), "fetch results");
    }
    for (int k = 0; k < n; ++k) runs[static_cast<size_t>(k)].fetch_host = host + off[static_cast<size_t>(k)];
    hipCheck(hipStreamSynchronize(s), "fetch sync");
    for (int k = 0; k < n; ++k) {
        QueryRun& rr = runs[static_cast<size_t>(k)];
        ex->fetchFinish(rr);
        rr.result.timing.gpu_execute_ms = dms;
        results[static_cast<size_t>(qorder[static_cast<size_t>(k)])] = std::move(rr.result);
    }
    const auto h5 = std::chrono::steady_clock::now();

    rep.queries = n;
    rep.templates = static_cast<int>(templates.size());
    rep.union_columns = static_cast<int>(ucols.size());
    rep.block_threads = cfg.block_threads;
    rep.vectors_per_thread = cfg.vectors_per_thread;
    rep.table_mask = cfg.table_mask;
    rep.table_kb = cfg.table_kb;
    rep.lds_table_bytes = table_bytes;
    rep.graph_replayed = replayed;
    rep.scan_ms = sms;
    rep.device_ms = dms;
    rep.modules_ms = elapsedMs(h0, h1);
    rep.bind_ms = elapsedMs(h1, h2);
    rep.enqueue_ms = elapsedMs(h2, h3);
    rep.wait_ms = elapsedMs(h3, h4);
    rep.fetch_ms = elapsedMs(h4, h5);
    rep.arena_bytes = arena.used();
    const uint64_t N = static_cast<uint64_t>(first.fact_rows);
    for (size_t u = 0; u < ucols.size(); ++u) rep.fact_bytes += 4 * N + (unull[u] ? 8 * ((N + 63) / 64) : 0);
}

BatchExecutor::Tuned BatchExecutor::autotune(const std::vector<const QueryProgram*>& templates,
                                             const std::vector<std::vector<int>>& per_template, std::vector<QueryRun>& runs,
                                             const cq_query_params* dqp, const BatchInit& init, const BatchOptions& opt,
                                             ExecutionContext* ctx) {
    QueryEngine& engine = db_.engine();
    hipStream_t s = ctx->stream;
    // one variant: its module, then two timed launches (the first warms the
    // module and the caches); -1 when the launch covers no rows
    auto timeVariant = [&](const Tuned& v, int* candidates) -> float {
        std::vector<std::string> ucols;
        std::vector<bool> unull;
        const std::string src =
            batchSource(templates, v.block_threads, v.vectors_per_thread, ucols, unull, v.table_mask, nullptr, candidates, v.table_kb);
        ModuleHandle bm = engine.moduleForSource(src, signatureOf(src), nullptr);
        hipFunction_t fn = bm->function("cq_batch_scan");
        cq_batch_params B;
        std::memset(&B, 0, sizeof B);
        for (size_t u = 0; u < ucols.size(); ++u) {
            const DeviceColumn& c = ctx->columns->at(ucols[u]);
            B.col[u] = c.data;
            B.valid[u] = reinterpret_cast<const unsigned int*>(c.validity);
        }
        B.qp = dqp;
        int at = 0;
        for (size_t t = 0; t < templates.size(); ++t) {
            B.tq_begin[t] = at;
            at += static_cast<int>(per_template[t].size());
            B.tq_end[t] = at;
        }
        B.row_end = templates.front()->fact_rows;
        const long long tile = static_cast<long long>(v.block_threads) * 4 * v.vectors_per_thread;
        B.num_tiles = (B.row_end + tile - 1) / tile;
        if (B.num_tiles == 0) return -1.0f;
        unsigned lds_bytes = 0;
        try {
            lds_bytes = layoutLds(templates, per_template, ldsLimitFor(opt, v.table_kb), B);
        } catch (const std::runtime_error&) {
            return -2.0f;  // the per-query regions do not fit beside the tables
        }
        const unsigned grid = scanGridShared(fn, v.block_threads, ctx->device, B.num_tiles, lds_bytes, opt.grid_percent);
        float ms = 0;
        for (int rep = 0; rep < 2; ++rep) {
            prepareTables(templates, per_template, runs, dqp, init, ctx);
            hipCheck(hipEventRecord(ev_[2], s), "hipEventRecord");
            void* args[] = {&B};
            hipCheck(hipModuleLaunchKernel(fn, grid, 1, 1, static_cast<unsigned>(v.block_threads), 1, 1, lds_bytes, s, args, nullptr),
                     "launch batch scan");
            hipCheck(hipEventRecord(ev_[3], s), "hipEventRecord");
            hipCheck(hipEventSynchronize(ev_[3]), "sync");
            (void)hipEventElapsedTime(&ms, ev_[2], ev_[3]);
        }
        return ms;
    };
    // launch shape with every candidate table in LDS
    Tuned best;
    best.table_kb = opt.table_kb;
    float best_ms = -1;
    int candidates = 0;
    const int bts[] = {64, 128, 256, 512, 1024};
    const int vpts[] = {1, 2, 4};
    for (int bt : bts) {
        for (int vpt : vpts) {
            Tuned v;
            v.block_threads = bt;
            v.vectors_per_thread = vpt;
            v.table_kb = opt.table_kb;
            const float ms = timeVariant(v, &candidates);
            if (ms == -1.0f) return best;
            if (ms < 0) continue;
            if (best_ms < 0 || ms < best_ms) {
                best_ms = ms;
                best = v;
            }
        }
    }
    // a larger LDS budget for the tables, when it admits more of them (a
    // table staged in LDS takes its probe off the L2; the per-query
    // regions get less LDS)
    for (int tk : {64, 96}) {
        if (tk <= opt.table_kb) continue;
        Tuned v = best;
        v.table_kb = tk;
        int more = 0;
        std::vector<std::string> uc;
        std::vector<bool> un;
        (void)batchSource(templates, v.block_threads, v.vectors_per_thread, uc, un, v.table_mask, nullptr, &more, tk);
        if (more <= candidates) continue;
        const float ms = timeVariant(v, nullptr);
        if (ms >= 0 && ms < best_ms) {
            best_ms = ms;
            best = v;
            candidates = more;
        }
    }
    // LDS-versus-HBM placement of the first three candidate tables at that
    // shape (a table left in HBM frees LDS for the others and costs L2
    // gathers instead of LDS bank conflicts)
    const int m = std::min(candidates, 3);
    const Tuned shape = best;
    for (unsigned bits = 0; bits + 1 < (1u << m); ++bits) {
        Tuned v = shape;
        v.table_mask = (~0u << m) | bits;
        const float ms = timeVariant(v, nullptr);
        if (ms >= 0 && ms < best_ms) {
            best_ms = ms;
            best = v;
        }
    }
    return best;
}

}  // namespace db
