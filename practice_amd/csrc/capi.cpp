// extern "C" boundary of libcrystal_hip.so (declarations: include/crystal_hip.h).
#include "../../include/crystal_hip.h"

#include <cstring>
#include <memory>
#include <sstream>
#include <string>

#include "runtime/database.h"
#include "runtime/storage.h"
#include "runtime/execution.h"
#include "runtime/result_text.h"
#include "serve/batch.h"
#include "serve/server.h"
#include "kernels/noalloc.h"

struct cq_module {
    db::ModuleHandle handle;
};

struct cq_db {
    std::unique_ptr<db::DatabaseInstance> inst;
    std::unique_ptr<db::BatchExecutor> batch;
    std::string last_batch_report;
};

struct cq_result {
    db::QueryResult r;
};

struct cq_server {
    std::unique_ptr<db::QueryServer> server;
};

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

template <class F>
int guarded(F&& f) {
    try {
        g_last_error.clear();
        return f();
    } catch (const std::bad_alloc& e) {
        return fail(CQ_ERR_MEMORY, std::string("out of memory: ") + e.what());
    } catch (const std::exception& e) {
        const std::string m = e.what();
        if (m.find("hipRTC compilation failed") != std::string::npos) return fail(CQ_ERR_COMPILE, m);
        if (m.find("not supported") != std::string::npos || m.find("Unsupported") != std::string::npos)
            return fail(CQ_ERR_UNSUPPORTED, m);
        if (m.find("Insufficient GPU memory") != std::string::npos) return fail(CQ_ERR_MEMORY, m);
        return fail(CQ_ERR_RUNTIME, m);
    } catch (...) {
        return fail(CQ_ERR_RUNTIME, "unknown error");
    }
}

int copyOut(const std::string& s, char* buf, size_t len, size_t* needed) {
    if (needed) *needed = s.size() + 1;
    if (buf && len) {
        const size_t n = std::min(len - 1, s.size());
        std::memcpy(buf, s.data(), n);
        buf[n] = 0;
    }
    return CQ_OK;
}

std::string jsonEscape(const std::string& s) {
    std::string o;
    for (char c : s) {
        if (c == '"' || c == '\\') o += '\\', o += c;
        else if (c == '\n') o += "\\n";
        else o += c;
    }
    return o;
}

}  // namespace

extern "C" {

const char* cq_last_error(void) { return g_last_error.c_str(); }
const char* cq_version(void) { return "crystal-hip 0.1 (gfx950)"; }

int cq_compile(const char* source, size_t length, const char* arch, cq_module** out) {
    return guarded([&] {
        if (!source || !out) return fail(CQ_ERR_INVALID, "cq_compile: null argument");
        std::string src(source, length);
        db::HipRtcCompiler c(arch && *arch ? arch : "gfx950");
        auto* m = new cq_module;
        m->handle = c.compile(src, db::signatureOf(src));
        *out = m;
        return CQ_OK;
    });
}

int cq_launch(cq_module* module, const char* name, const struct cq_query_params* params, unsigned grid, unsigned block,
              void* stream) {
    return guarded([&] {
        if (!module || !name || !params) return fail(CQ_ERR_INVALID, "cq_launch: null argument");
        hipFunction_t f = module->handle->function(name);
        cq_query_params p = *params;
        void* args[] = {&p};
        db::hipCheck(hipModuleLaunchKernel(f, grid, 1, 1, block, 1, 1, 0, static_cast<hipStream_t>(stream), args, nullptr),
                     "hipModuleLaunchKernel");
        return CQ_OK;
    });
}

size_t cq_module_code_size(const cq_module* module) { return module ? module->handle->codeSize() : 0; }
size_t cq_query_params_size(void) { return sizeof(cq_query_params); }
void cq_module_release(cq_module* module) { delete module; }

int cq_db_create(int device, cq_db** out) {
    return guarded([&] {
        if (!out) return fail(CQ_ERR_INVALID, "cq_db_create: null out");
        auto d = std::make_unique<cq_db>();
        d->inst = std::make_unique<db::DatabaseInstance>(device);
        // the allocation-free execute's scratch address on this GPU, once
        if (device >= 0) (void)db::noAllocScratch(device);
        *out = d.release();
        return CQ_OK;
    });
}

void cq_db_destroy(cq_db* d) { delete d; }

int cq_db_generate_ssb(cq_db* d, int sf, long long fact_rows, unsigned long long seed, const char* nullable) {
    return guarded([&] {
        db::GenSpec g;
        g.sf = sf;
        g.fact_rows = fact_rows;
        g.seed = seed;
        g.nullable = nullable ? nullable : "";
        d->inst->generateSSB(g);
        return CQ_OK;
    });
}

int cq_db_load_reference_files(cq_db* d, const char* data_dir, const long long rows[5], const char* stats_cache,
                               const char* columns) {
    return guarded([&] {
        if (!data_dir || !rows) return fail(CQ_ERR_INVALID, "data_dir and rows are required");
        int64_t r[5];
        for (int i = 0; i < 5; ++i) r[i] = rows[i];
        std::vector<std::string> only;
        std::string cur;
        for (const char* p = columns ? columns : ""; ; ++p) {
            if (*p == ',' || *p == 0) {
                if (!cur.empty()) only.push_back(cur);
                cur.clear();
                if (*p == 0) break;
            } else if (*p != ' ') {
                cur += *p;
            }
        }
        d->inst->loadReferenceFiles(data_dir, r, stats_cache ? stats_cache : "", only);
        return CQ_OK;
    });
}

int cq_reference_rows(int sf, long long rows[5]) {
    return guarded([&] {
        int64_t r[5];
        if (!rows || !db::storage::referenceRows(sf, r)) return fail(CQ_ERR_INVALID, "reference row counts exist for SF 10 and 20");
        for (int i = 0; i < 5; ++i) rows[i] = r[i];
        return CQ_OK;
    });
}

int cq_db_save_reference_files(cq_db* d, const char* data_dir, const char* validity_suffix) {
    return guarded([&] {
        if (!data_dir) return fail(CQ_ERR_INVALID, "data_dir is required");
        d->inst->saveReferenceFiles(data_dir, validity_suffix ? validity_suffix : ".valid64");
        return CQ_OK;
    });
}

int cq_db_create_table(cq_db* d, const char* table, long long rows, int is_fact) {
    return guarded([&] {
        d->inst->createTable(table, rows, is_fact != 0);
        return CQ_OK;
    });
}

int cq_db_load_column(cq_db* d, const char* table, const char* column, const int32_t* data, const uint64_t* validity) {
    return guarded([&] {
        if (!data && d->inst->catalog().getTableMetadata(table).getSize() > 0)
            return fail(CQ_ERR_INVALID, "cq_db_load_column: null data");
        d->inst->loadColumn(table, column, data, validity);
        return CQ_OK;
    });
}

int cq_db_set_primary_key(cq_db* d, const char* table, const char* column) {
    return guarded([&] {
        d->inst->setPrimaryKey(table, column);
        return CQ_OK;
    });
}

int cq_db_finalize(cq_db* d) {
    return guarded([&] {
        d->inst->finalizeCatalog();
        return CQ_OK;
    });
}

int cq_db_declare_column(cq_db* d, const char* table, const char* column, int nullable) {
    return guarded([&] {
        d->inst->declareColumn(table, column, nullable != 0);
        return CQ_OK;
    });
}

int cq_db_set_column_stats(cq_db* d, const char* column, long long mn, long long mx, unsigned long long card) {
    return guarded([&] {
        d->inst->setColumnStats(column, mn, mx, card);
        return CQ_OK;
    });
}

long long cq_db_result_size(cq_db* d, const char* sql) {
    try {
        return static_cast<long long>(d->inst->resultSize(sql));
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return -1;
    }
}

int cq_db_column_stats(cq_db* d, const char* column, long long* mn, long long* mx, unsigned long long* card, int* nullable,
                       int* unique) {
    return guarded([&] {
        const auto& cat = d->inst->catalog();
        const std::string t = cat.tableOfColumn(column);
        const auto& m = cat.getTableMetadata(t);
        const auto& s = m.getColumnStats(column);
        if (mn) *mn = s.min_value_;
        if (mx) *mx = s.max_value_;
        if (card) *card = s.cardinality_;
        if (nullable) *nullable = m.isColumnNullable(column) ? 1 : 0;
        if (unique) *unique = (m.isColumnUnique(column) || m.isColumnPrimaryKey(column)) ? 1 : 0;
        return CQ_OK;
    });
}

int cq_db_fetch_column(cq_db* d, const char* column, int32_t* out, uint64_t* validity_out) {
    return guarded([&] {
        d->inst->fetchColumn(column, out, validity_out);
        return CQ_OK;
    });
}

int cq_db_column_device(cq_db* d, const char* column, const int32_t** data, const uint64_t** validity, long long* rows) {
    return guarded([&] {
        if (!d || !column || !data) return fail(CQ_ERR_INVALID, "cq_db_column_device: null argument");
        if (d->inst->hostOnly()) return fail(CQ_ERR_INVALID, "cq_db_column_device: host-only database");
        const auto& cols = d->inst->columns();
        auto it = cols.find(column);
        if (it == cols.end()) return fail(CQ_ERR_INVALID, std::string("column not loaded on the device: ") + column);
        *data = it->second.data;
        if (validity) *validity = it->second.validity;
        if (rows) *rows = it->second.rows;
        return CQ_OK;
    });
}

long long cq_db_table_rows(cq_db* d, const char* table) {
    try {
        return static_cast<long long>(d->inst->catalog().getTableMetadata(table).getSize());
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return -1;
    }
}

int cq_db_set_option(cq_db* d, const char* name, long long v) {
    return guarded([&] {
        auto& c = d->inst->config();
        const std::string n = name ? name : "";
        if (n == "kernel_timing") c.kernel_timing = v != 0;
        else if (n == "memory_guard") c.memory_guard_enabled = v != 0;
        else if (n == "memory_guard_reserve_bytes") c.memory_guard_reserve_bytes = static_cast<size_t>(v);
        else if (n == "lds_budget") c.codegen.lds_budget = static_cast<uint64_t>(v);
        else if (n == "block_threads") c.codegen.block_threads = static_cast<int>(v);
        else if (n == "vectors_per_thread") c.codegen.vectors_per_thread = static_cast<int>(v);
        else if (n == "lds_tables") c.codegen.allow_lds_tables = v != 0;
        else if (n == "lds_aggregation") c.codegen.allow_lds_aggregation = v != 0;
        else if (n == "compact_tables") c.codegen.allow_compact_tables = v != 0;
        else if (n == "dense_tables") c.codegen.allow_dense_tables = v != 0;
        else if (n == "prefetch") c.codegen.prefetch = v != 0;
        else if (n == "stream_tables") c.codegen.stream_large_tables = v != 0;
        else if (n == "calendar_tables") c.codegen.calendar_tables = v != 0;
        else if (n == "split_probes") c.codegen.split_probes = v != 0;
        else if (n == "buffer_tables") c.codegen.buffer_tables = v != 0;
        else if (n == "hoist_gathers") c.codegen.hoist_gathers = v != 0;
        else if (n == "defer_probes") c.codegen.defer_probes = v != 0;
        else if (n == "interpret") c.codegen.vm = v != 0;
        else if (n == "vm_row_kernel") c.vm_row_kernel = v != 0;
        else if (n == "vm_linear") c.vm_linear = v != 0;
        else if (n == "vm_grid_percent") {
            if (v < 1 || v > 100) return fail(CQ_ERR_INVALID, "vm_grid_percent must be in [1, 100]");
            c.vm_grid_percent = static_cast<int>(v);
        }
        else if (n == "host_sort_rows") c.host_sort_rows = v;
        else if (n == "prepared_cache") c.prepared_cache = v != 0;
        else if (n == "module_cache_capacity") {
            if (v < 1) return fail(CQ_ERR_INVALID, "module_cache_capacity must be positive");
            c.module_cache_capacity = static_cast<size_t>(v);
            d->inst->moduleCache().setCapacity(static_cast<size_t>(v));
        }
        else return fail(CQ_ERR_INVALID, "unknown option " + n);
        return CQ_OK;
    });
}

long long cq_db_cache_stat(cq_db* d, const char* name) {
    const std::string n = name ? name : "";
    auto& c = d->inst->moduleCache();
    if (n == "size") return static_cast<long long>(c.size());
    if (n == "hits") return static_cast<long long>(c.hits());
    if (n == "misses") return static_cast<long long>(c.misses());
    if (n == "evictions") return static_cast<long long>(c.evictions());
    if (n == "compiles") return static_cast<long long>(d->inst->compiler().compileCount());
    return -1;
}

int cq_db_clear_cache(cq_db* d) {
    return guarded([&] {
        d->inst->moduleCache().clear();
        return CQ_OK;
    });
}

int cq_db_execute(cq_db* d, const char* sql, cq_result** out) { return cq_db_execute_range(d, sql, 0, -1, out); }

int cq_db_execute_range(cq_db* d, const char* sql, long long b, long long e, cq_result** out) {
    return guarded([&] {
        if (!d || !sql || !out) return fail(CQ_ERR_INVALID, "cq_db_execute: null argument");
        auto r = std::make_unique<cq_result>();
        r->r = d->inst->executeQueryRange(sql, b, e);
        *out = r.release();
        return CQ_OK;
    });
}

int cq_db_execute_group(cq_db* d, const char* const* sqls, int n, cq_result** out) {
    return guarded([&] {
        if (!d || n < 0 || (n > 0 && (!sqls || !out))) return fail(CQ_ERR_INVALID, "cq_db_execute_group: null argument");
        std::vector<std::string> q;
        for (int i = 0; i < n; ++i) {
            if (!sqls[i]) return fail(CQ_ERR_INVALID, "cq_db_execute_group: null statement");
            q.emplace_back(sqls[i]);
            out[i] = nullptr;
        }
        std::vector<std::string> errs;
        std::vector<db::QueryResult> res = d->inst->executeGroup(q, errs);
        std::string first;
        for (int i = 0; i < n; ++i) {
            if (!errs[static_cast<size_t>(i)].empty()) {
                if (first.empty()) first = "statement " + std::to_string(i) + ": " + errs[static_cast<size_t>(i)];
                continue;
            }
            auto r = std::make_unique<cq_result>();
            r->r = std::move(res[static_cast<size_t>(i)]);
            out[i] = r.release();
        }
        if (!first.empty()) return fail(CQ_ERR_RUNTIME, first);
        return CQ_OK;
    });
}

int cq_db_execute_noalloc(cq_db* d, const char* sql, cq_result** out) {
    return guarded([&] {
        if (!d || !sql || !out) return fail(CQ_ERR_INVALID, "cq_db_execute_noalloc: null argument");
        if (d->inst->hostOnly()) return fail(CQ_ERR_INVALID, "cq_db_execute_noalloc: host-only database");
        void* scratch = db::noAllocScratch(d->inst->device());
        auto r = std::make_unique<cq_result>();
        try {
            r->r = d->inst->executeQueryInto(sql, scratch, db::noAllocScratchBytes());
        } catch (const std::length_error& e) {
            return fail(CQ_ERR_MEMORY, e.what());
        } catch (const std::invalid_argument& e) {
            return fail(CQ_ERR_INVALID, e.what());
        }
        *out = r.release();
        return CQ_OK;
    });
}

int cq_db_explain(cq_db* d, const char* sql, char* buf, size_t len, size_t* needed) {
    return guarded([&] { return copyOut(d->inst->explain(sql), buf, len, needed); });
}

int cq_db_generate_code(cq_db* d, const char* sql, char* buf, size_t len, size_t* needed) {
    return guarded([&] { return copyOut(d->inst->generateQueryCode(sql), buf, len, needed); });
}

int cq_db_describe(cq_db* d, const char* sql, char* buf, size_t len, size_t* needed) {
    return guarded([&] {
        const db::QueryProgram p = d->inst->prepare(sql);
        std::ostringstream o;
        o << "{\"signature\":\"" << p.signature << "\",\"fact_table\":\"" << p.fact_table << "\",\"fact_rows\":" << p.fact_rows
          << ",\"groups\":" << p.groups << ",\"storage_ts\":" << p.storage_ts << ",\"visible_ts\":" << p.visible_ts
          << ",\"fast_sparse\":" << (p.fast_sparse ? 1 : 0) << ",\"lds_aggregation\":" << (p.lds_aggregation ? 1 : 0)
          << ",\"block_threads\":" << p.block_threads << ",\"vectors_per_thread\":" << p.vectors_per_thread
          << ",\"scan_lds_bytes\":" << p.scan_lds_bytes << ",\"algorithmic_bytes\":" << p.algorithmic_bytes
          << ",\"interpreted\":" << (p.vm ? 1 : 0) << ",\"linear\":" << ((p.vm && p.vm->lin.enabled) ? 1 : 0)
          << ",\"fact_cols\":[";
        for (size_t i = 0; i < p.fact_cols.size(); ++i) o << (i ? "," : "") << "\"" << p.fact_cols[i] << "\"";
        o << "],\"dim_cols\":[";
        for (size_t i = 0; i < p.dim_cols.size(); ++i) o << (i ? "," : "") << "\"" << p.dim_cols[i] << "\"";
        o << "],\"joins\":[";
        for (size_t j = 0; j < p.joins.size(); ++j) {
            const auto& jp = p.joins[j];
            o << (j ? "," : "") << "{\"dim\":\"" << jp.dim_table << "\",\"pk\":\"" << jp.pk_col << "\",\"fk\":\"" << jp.fk_col
              << "\",\"payload\":\"" << jp.payload_col << "\",\"kmin\":" << jp.kmin << ",\"len\":" << jp.len
              << ",\"layout\":\"" << (jp.mht ? "multi-value" : jp.dense ? "key-indexed" : (jp.compact ? "compact" : "reference"))
              << "\",\"compact\":" << (jp.compact ? 1 : 0) << ",\"entry_bits\":" << (jp.dense ? jp.db : jp.eb)
              << ",\"tag_bits\":" << (jp.dense ? 0 : jp.tb) << ",\"absent_code\":" << (jp.dense ? jp.doff : 1)
              << ",\"calendar\":" << (jp.dense ? jp.cal : 0) << ",\"lds\":" << (jp.lds ? 1 : 0) << ",\"bytes\":" << jp.ht_words * 4
              << "}";
        }
        o << "],\"lits\":[";
        for (size_t k = 0; k < p.lits.size(); ++k) o << (k ? "," : "") << p.lits[k];
        o << "],\"has_limit\":" << (p.has_limit ? 1 : 0) << ",\"limit\":" << p.limit << ",\"outputs\":[";
        for (size_t c = 0; c < p.out.size(); ++c)
            o << (c ? "," : "") << "{\"type\":" << static_cast<int>(p.out[c].type) << ",\"nullable\":" << (p.out[c].nullable ? 1 : 0)
              << "}";
        o << "],\"plan\":\"" << jsonEscape(p.plan_sexpr) << "\"}";
        return copyOut(o.str(), buf, len, needed);
    });
}

int cq_db_compile(cq_db* d, const char* sql) {
    return guarded([&] {
        d->inst->compile(sql);
        return CQ_OK;
    });
}

long long cq_result_rows(const cq_result* r) { return r ? static_cast<long long>(r->r.row_count) : -1; }
int cq_result_cols(const cq_result* r) { return r ? static_cast<int>(r->r.columns.size()) : -1; }
int cq_result_type(const cq_result* r, int c) {
    if (!r || c < 0 || c >= static_cast<int>(r->r.columns.size())) return -1;
    switch (r->r.columns[static_cast<size_t>(c)].type) {
        case LogicalType::Int64: return 0;
        case LogicalType::UInt64: return 1;
        case LogicalType::Float64: return 2;
    }
    return -1;
}
int cq_result_nullable(const cq_result* r, int c) {
    if (!r || c < 0 || c >= static_cast<int>(r->r.columns.size())) return -1;
    return r->r.columns[static_cast<size_t>(c)].nullable ? 1 : 0;
}
int cq_result_ordered(const cq_result* r) { return r ? (r->r.ordered ? 1 : 0) : -1; }
const unsigned long long* cq_result_column(const cq_result* r, int c) {
    if (!r || c < 0 || c >= static_cast<int>(r->r.data.size())) return nullptr;
    return reinterpret_cast<const unsigned long long*>(r->r.data[static_cast<size_t>(c)].data());
}
const unsigned char* cq_result_validity(const cq_result* r, int c) {
    if (!r || c < 0 || c >= static_cast<int>(r->r.validity.size())) return nullptr;
    return r->r.validity[static_cast<size_t>(c)].data();
}
const char* cq_result_column_name(const cq_result* r, int c) {
    if (!r || c < 0 || c >= static_cast<int>(r->r.columns.size())) return nullptr;
    return r->r.columns[static_cast<size_t>(c)].name.c_str();
}
int cq_result_format(const cq_result* r, long long row_limit, int extended_timing, char* buf, size_t len, size_t* needed) {
    return guarded([&] {
        if (!r) return fail(CQ_ERR_INVALID, "cq_result_format: null result");
        db::ResultTextOptions o;
        o.limit_enabled = row_limit > 0;
        o.limit = row_limit > 0 ? static_cast<size_t>(row_limit) : 0;
        o.extended_timing = extended_timing != 0;
        return copyOut(db::formatResultText(r->r, o), buf, len, needed);
    });
}
int cq_result_shape(const cq_result* r, long long* rows, int* cols, int* ordered, int* types, int* nullable, int max_cols) {
    if (!r || !rows || !cols || !ordered) return CQ_ERR_INVALID;
    const auto& q = r->r;
    *rows = static_cast<long long>(q.row_count);
    *cols = static_cast<int>(q.columns.size());
    *ordered = q.ordered ? 1 : 0;
    if (*cols > max_cols) return CQ_ERR_INVALID;
    for (int c = 0; c < *cols; ++c) {
        const auto& cd = q.columns[static_cast<size_t>(c)];
        if (types) types[c] = static_cast<int>(cd.type);
        if (nullable) nullable[c] = cd.nullable ? 1 : 0;
    }
    return CQ_OK;
}
int cq_result_copy(const cq_result* r, unsigned long long* values, unsigned char* valid) {
    if (!r) return CQ_ERR_INVALID;
    const auto& q = r->r;
    const size_t n = q.row_count;
    for (size_t c = 0; c < q.data.size(); ++c) {
        if (values && n) std::memcpy(values + c * n, q.data[c].data(), n * sizeof(unsigned long long));
        if (valid && n) std::memcpy(valid + c * n, q.validity[c].data(), n);
    }
    return CQ_OK;
}
int cq_result_timings(const cq_result* r, double* out) {
    if (!r || !out) return CQ_ERR_INVALID;
    const auto& t = r->r.timing;
    out[0] = t.gpu_execute_ms;
    out[1] = t.scan_kernel_ms;
    out[2] = t.compile_ms;
    out[3] = t.codegen_ms;
    out[4] = t.parse_ms;
    out[5] = t.host_fetch_ms;
    out[6] = t.engine_ms;
    out[7] = t.cache_hit ? 1 : 0;
    return CQ_OK;
}
double cq_result_timing(const cq_result* r, const char* name) {
    if (!r || !name) return -1;
    const auto& t = r->r.timing;
    const std::string n = name;
    if (n == "gpu_execute_ms") return t.gpu_execute_ms;
    if (n == "scan_kernel_ms") return t.scan_kernel_ms;
    if (n == "compile_ms") return t.compile_ms;
    if (n == "codegen_ms") return t.codegen_ms;
    if (n == "parse_ms") return t.parse_ms;
    if (n == "host_fetch_ms") return t.host_fetch_ms;
    if (n == "engine_ms") return t.engine_ms;
    if (n == "cache_hit") return t.cache_hit ? 1 : 0;
    return -1;
}
void cq_result_free(cq_result* r) { delete r; }

int cq_batch_execute(cq_db* d, const char* const* sqls, int n, const char* options, cq_result** outs) {
    return guarded([&] {
        if (!d || (!sqls && n > 0) || !outs) return fail(CQ_ERR_INVALID, "cq_batch_execute: null argument");
        if (!d->batch) d->batch = std::make_unique<db::BatchExecutor>(*d->inst);
        std::vector<std::string> q;
        for (int i = 0; i < n; ++i) q.emplace_back(sqls[i]);
        std::vector<db::QueryResult> rs = d->batch->execute(q, options ? options : "");
        for (int i = 0; i < n; ++i) {
            auto* r = new cq_result;
            r->r = std::move(rs[static_cast<size_t>(i)]);
            outs[i] = r;
        }
        d->last_batch_report = d->batch->lastReport();
        return CQ_OK;
    });
}

int cq_batch_last_report(cq_db* d, char* buf, size_t len, size_t* needed) {
    return guarded([&] { return copyOut(d->last_batch_report, buf, len, needed); });
}

int cq_db_generate_batch_code(cq_db* d, const char* const* sqls, int n, const char* options, int part, char* buf, size_t len,
                              size_t* needed) {
    return guarded([&] {
        if (!d || (!sqls && n > 0)) return fail(CQ_ERR_INVALID, "cq_db_generate_batch_code: null argument");
        std::vector<std::string> q;
        for (int i = 0; i < n; ++i) q.emplace_back(sqls[i]);
        return copyOut(db::generateBatchCode(*d->inst, q, db::parseBatchOptions(options ? options : ""), part), buf, len, needed);
    });
}

int cq_server_start(cq_db* d, const char* options, cq_server** out) {
    return guarded([&] {
        auto s = std::make_unique<cq_server>();
        s->server = std::make_unique<db::QueryServer>(*d->inst, options ? options : "");
        *out = s.release();
        return CQ_OK;
    });
}

long long cq_server_submit(cq_server* s, const char* sql) {
    try {
        return s->server->submit(sql);
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return -1;
    }
}

int cq_server_wait(cq_server* s, long long ticket, int timeout_ms, cq_result** out) {
    return guarded([&] {
        db::QueryResult r;
        std::string err;
        const int rc = s->server->wait(ticket, timeout_ms, r, err);
        if (rc != 0) return fail(rc, err);
        auto* res = new cq_result;
        res->r = std::move(r);
        *out = res;
        return CQ_OK;
    });
}

int cq_server_stats(cq_server* s, char* buf, size_t len, size_t* needed) {
    return guarded([&] { return copyOut(s->server->statsJson(), buf, len, needed); });
}

int cq_server_reset_stats(cq_server* s) {
    return guarded([&] {
        if (!s) return fail(CQ_ERR_INVALID, "cq_server_reset_stats: null server");
        s->server->resetStats();
        return CQ_OK;
    });
}

void cq_server_stop(cq_server* s) { delete s; }

}  // extern "C"
