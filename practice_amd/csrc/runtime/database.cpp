// DatabaseInstance implementation and the synthetic schema definition.
#include "database.h"
#include "storage.h"

#include <algorithm>
#include <cstring>
#include <sstream>
#include <stdexcept>

#include "../frontend/sql.h"
#include "../kernels/aot_kernels.h"

namespace db {

namespace {

// all values dates written yyyymmdd: month 1..12, day 1..31, 8 digits
bool calendarColumn(const int32_t* d, int64_t rows, int64_t mn, int64_t mx) {
    if (rows <= 0 || mn < 10000101 || mx > 99991231) return false;
    for (int64_t i = 0; i < rows; ++i) {
        const int32_t k = d[i];
        const int32_t m = (k / 100) % 100, day = k % 100;
        if (m < 1 || m > 12 || day < 1 || day > 31) return false;
    }
    return true;
}

}  // namespace


// ------------------------------------------------------------------ synthetic schema

namespace {

constexpr uint64_t kGamma = 0x9E3779B97F4A7C15ULL;
constexpr int64_t kDateRows = 2556;

uint64_t mix64(uint64_t x) {
    x += kGamma;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

uint64_t fnv1a(const std::string& s) {
    uint64_t h = 0xCBF29CE484222325ULL;
    for (unsigned char c : s) {
        h ^= c;
        h *= 0x100000001B3ULL;
    }
    return h;
}

uint64_t streamSeed(uint64_t seed, const std::string& table, const std::string& column, const char* suffix = "") {
    return mix64(fnv1a(table + "." + column + suffix) ^ seed);
}

int floorLog2(int v) {
    int r = 0;
    while (v > 1) v >>= 1, ++r;
    return r;
}

int nullBasisPoints(const GenSpec& spec, const std::string& table, const std::string& column) {
    const std::string want = table + "." + column + ":";
    std::stringstream ss(spec.nullable);
    std::string item;
    while (std::getline(ss, item, ',')) {
        item.erase(0, item.find_first_not_of(' '));
        if (item.compare(0, want.size(), want) == 0) return std::stoi(item.substr(want.size()));
    }
    return 0;
}

cq_gen_recipe recipeFor(const GenSpec& sp, const std::string& t, const std::string& c) {
    cq_gen_recipe r;
    std::memset(&r, 0, sizeof r);
    auto S = [&](const std::string& col) { return streamSeed(sp.seed, t, col); };
    auto uniform = [&](int64_t lo, int64_t hi) {
        r.kind = CQ_GEN_UNIFORM;
        r.s0 = S(c);
        r.lo = lo;
        r.hi = hi;
    };
    if (t == "LINEORDER") {
        if (c == "lo_custkey") uniform(1, genTableRows(sp, "CUSTOMER"));
        else if (c == "lo_partkey") uniform(1, genTableRows(sp, "PART"));
        else if (c == "lo_suppkey") uniform(1, genTableRows(sp, "SUPPLIER"));
        else if (c == "lo_orderdate" || c == "lo_commitdate") {
            r.kind = CQ_GEN_DATE_FK;
            r.s0 = S(c);
            r.lo = 0;
            r.hi = kDateRows - 1;
        } else if (c == "lo_quantity") uniform(1, 50);
        else if (c == "lo_extendedprice") uniform(90000, 10500000);
        else if (c == "lo_discount") uniform(0, 10);
        else if (c == "lo_revenue") uniform(100000, 10000000);
        else if (c == "lo_supplycost") uniform(60000, 600000);
        else if (c == "lo_tax") uniform(0, 8);
        else if (c == "lo_orderkey") r.kind = CQ_GEN_ORDERKEY;  // orders of 4 lines
        else if (c == "lo_linenumber") r.kind = CQ_GEN_LINENUMBER;
        else throw std::runtime_error("unknown generated column " + c);
    } else if (t == "PART") {
        r.s0 = S("p_category");
        if (c == "p_partkey") r.kind = CQ_GEN_KEY;
        else if (c == "p_category") r.kind = CQ_GEN_PCAT;
        else if (c == "p_mfgr") r.kind = CQ_GEN_PMFGR;
        else if (c == "p_brand1") r.kind = CQ_GEN_PBRAND, r.s1 = S(c);
        else if (c == "p_size") uniform(1, 50);
        else throw std::runtime_error("unknown generated column " + c);
    } else if (t == "SUPPLIER" || t == "CUSTOMER") {
        const bool supp = t == "SUPPLIER";
        r.s0 = S(supp ? "s_nation" : "c_nation");
        const std::string col = c.substr(2);
        if (col == (supp ? "suppkey" : "custkey")) r.kind = CQ_GEN_KEY;
        else if (col == "nation") r.kind = CQ_GEN_NATION;
        else if (col == "region") r.kind = CQ_GEN_REGION;
        else if (col == "city") r.kind = CQ_GEN_CITY, r.s1 = S(c);
        else if (!supp && col == "mktsegment") uniform(0, 4);
        else throw std::runtime_error("unknown generated column " + c);
    } else if (t == "DDATE") {
        static const std::pair<const char*, int> kinds[] = {
            {"d_datekey", CQ_GEN_DKEY}, {"d_year", CQ_GEN_DYEAR}, {"d_month", CQ_GEN_DMONTH},
            {"d_yearmonthnum", CQ_GEN_DYM}, {"d_daynuminweek", CQ_GEN_DDOW}, {"d_daynuminmonth", CQ_GEN_DDOM},
            {"d_daynuminyear", CQ_GEN_DDOY}, {"d_weeknuminyear", CQ_GEN_DWEEK}};
        bool found = false;
        for (const auto& k : kinds)
            if (c == k.first) r.kind = k.second, found = true;
        if (!found) throw std::runtime_error("unknown generated column " + c);
    } else {
        throw std::runtime_error("unknown generated table " + t);
    }
    r.null_bp = nullBasisPoints(sp, t, c);
    r.vseed = streamSeed(sp.seed, t, c, "#valid");
    return r;
}

}  // namespace

int64_t genTableRows(const GenSpec& spec, const std::string& table) {
    const int sf = std::max(spec.sf, 1);
    if (table == "LINEORDER") return spec.fact_rows > 0 ? spec.fact_rows : 6000000LL * sf;
    if (table == "PART") return 200000LL * (1 + floorLog2(sf));
    if (table == "SUPPLIER") return 2000LL * sf;
    if (table == "CUSTOMER") return 30000LL * sf;
    if (table == "DDATE") return kDateRows;
    throw std::runtime_error("unknown table " + table);
}

std::vector<std::string> genTables() { return {"LINEORDER", "PART", "SUPPLIER", "CUSTOMER", "DDATE"}; }

std::vector<std::string> genTableColumns(const std::string& table) {
    if (table == "LINEORDER")
        return {"lo_custkey", "lo_partkey", "lo_suppkey", "lo_orderdate", "lo_commitdate", "lo_quantity",
                "lo_extendedprice", "lo_discount", "lo_revenue", "lo_supplycost", "lo_tax", "lo_orderkey", "lo_linenumber"};
    if (table == "PART") return {"p_partkey", "p_mfgr", "p_category", "p_brand1", "p_size"};
    if (table == "SUPPLIER") return {"s_suppkey", "s_city", "s_nation", "s_region"};
    if (table == "CUSTOMER") return {"c_custkey", "c_city", "c_nation", "c_region", "c_mktsegment"};
    if (table == "DDATE")
        return {"d_datekey", "d_year", "d_month", "d_yearmonthnum", "d_daynuminweek", "d_daynuminmonth", "d_daynuminyear",
                "d_weeknuminyear"};
    throw std::runtime_error("unknown table " + table);
}

// ------------------------------------------------------------------ DatabaseInstance

DatabaseInstance::DatabaseInstance(int device, EngineConfig config) : device_(device) {
    catalog_ = std::make_shared<Catalog>();
    cache_ = std::make_shared<PlanSignatureCache>(config.module_cache_capacity);
    compiler_ = std::make_shared<HipRtcCompiler>();
    executor_ = std::make_shared<HipModuleExecutor>();
    engine_ = std::make_unique<QueryEngine>(cache_, compiler_, executor_, catalog_);
    ctx_ = std::make_unique<ExecutionContext>();
    ctx_->config = config;
    ctx_->device = device;
    ctx_->columns = &columns_;
    if (device_ >= 0) {
        int n = 0;
        hipCheck(hipGetDeviceCount(&n), "hipGetDeviceCount");
        if (device_ >= n) throw std::runtime_error("HIP device " + std::to_string(device_) + " not present");
        hipCheck(hipSetDevice(device_), "hipSetDevice");
        DeviceMapLock lock;
        hipCheck(hipStreamCreate(&ctx_->stream), "hipStreamCreate");
        ctx_->owns_stream = true;
    }
}

std::unique_ptr<ExecutionContext> DatabaseInstance::newContext() {
    if (device_ < 0) throw std::runtime_error("execution contexts need a GPU device (host-only database)");
    auto c = std::make_unique<ExecutionContext>();
    c->config = ctx_->config;
    c->device = device_;
    c->columns = &columns_;
    c->loaded_device_bytes = loaded_bytes_;
    hipCheck(hipSetDevice(device_), "hipSetDevice");
    DeviceMapLock lock;
    hipCheck(hipStreamCreate(&c->stream), "hipStreamCreate");
    c->owns_stream = true;
    return c;
}

DatabaseInstance::~DatabaseInstance() {
    ctx_.reset();
    engine_.reset();
    cache_.reset();
    if (device_ >= 0) {
        DeviceMapLock lock;
        (void)hipDeviceSynchronize();
        for (void* p : allocations_) (void)hipFree(p);
    }
}

void* DatabaseInstance::deviceAlloc(size_t bytes) {
    void* p = nullptr;
    DeviceMapLock lock;
    hipCheck(hipMalloc(&p, std::max<size_t>(bytes, 256)), "hipMalloc column");
    allocations_.push_back(p);
    loaded_bytes_ += bytes;
    ctx_->loaded_device_bytes = loaded_bytes_;
    return p;
}

void DatabaseInstance::ensureTable(const std::string& table) const {
    if (!catalog_->hasTable(table)) throw std::runtime_error("Table not found: " + table);
}

void DatabaseInstance::createTable(const std::string& name, int64_t rows, bool is_fact) {
    engine_->invalidate();  // cached programs were planned on the old catalog
    std::lock_guard<std::mutex> lock(mu_);
    if (catalog_->hasTable(name)) throw std::runtime_error("table already exists: " + name);
    catalog_->pushTableMetadata(TableMetadata(name, {}, static_cast<uint64_t>(rows), is_fact));
}

void DatabaseInstance::loadColumn(const std::string& table, const std::string& column, const int32_t* data,
                                  const uint64_t* validity) {
    engine_->invalidate();  // cached programs were planned on the old catalog
    std::lock_guard<std::mutex> lock(mu_);
    ensureTable(table);
    auto& meta = catalog_->getTableMetadataMutable(table);
    const int64_t rows = static_cast<int64_t>(meta.getSize());
    if (columns_.count(column)) throw std::runtime_error("column already loaded: " + column);
    const size_t words = static_cast<size_t>((rows + 63) / 64);
    DeviceColumn dc;
    dc.rows = rows;
    if (device_ >= 0) {
        void* d = deviceAlloc(static_cast<size_t>(rows) * 4);
        {
            DeviceMapLock lock;  // null-stream copy: synchronizes with every stream
            if (rows) hipCheck(hipMemcpy(d, data, static_cast<size_t>(rows) * 4, hipMemcpyHostToDevice), "upload column");
        }
        dc.data = static_cast<const int*>(d);
        if (validity) {
            void* v = deviceAlloc(words * 8);
            // mask bits past the last row (maskUnusedValidityBits, utils.h:111-115)
            std::vector<uint64_t> vv(validity, validity + words);
            if (rows % 64 && !vv.empty()) vv.back() &= (1ULL << (rows % 64)) - 1;
            {
                DeviceMapLock lock;  // null-stream copy: synchronizes with every stream
                if (words) hipCheck(hipMemcpy(v, vv.data(), words * 8, hipMemcpyHostToDevice), "upload validity");
            }
            dc.validity = static_cast<const uint64_t*>(v);
        }
    } else {
        host_data_[column].assign(data, data + rows);
        dc.data = host_data_[column].data();
        if (validity) {
            host_validity_[column].assign(validity, validity + words);
            if (rows % 64 && words) host_validity_[column].back() &= (1ULL << (rows % 64)) - 1;
            dc.validity = host_validity_[column].data();
        }
    }
    columns_[column] = dc;
    // the catalog keeps column order of the table
    std::vector<std::string> names = meta.getColumnNames();
    names.push_back(column);
    TableMetadata updated(meta.getName(), names, meta.getSize(), meta.isFactTable());
    for (const auto& n : meta.getColumnNames())
        if (meta.hasColumnStats(n)) updated.setColumnStats(n, meta.getColumnStats(n));
    if (validity) updated.setColumnNullable(column, true);
    meta = updated;
}

void DatabaseInstance::setPrimaryKey(const std::string& table, const std::string& column) {
    engine_->invalidate();  // cached programs were planned on the old catalog
    std::lock_guard<std::mutex> lock(mu_);
    ensureTable(table);
    catalog_->getTableMetadataMutable(table).setColumnPrimaryKey(column, true);
}

void DatabaseInstance::declareColumn(const std::string& table, const std::string& column, bool nullable) {
    engine_->invalidate();  // cached programs were planned on the old catalog
    std::lock_guard<std::mutex> lock(mu_);
    ensureTable(table);
    auto& meta = catalog_->getTableMetadataMutable(table);
    std::vector<std::string> names = meta.getColumnNames();
    names.push_back(column);
    TableMetadata updated(meta.getName(), names, meta.getSize(), meta.isFactTable());
    for (const auto& n : meta.getColumnNames())
        if (meta.hasColumnStats(n)) updated.setColumnStats(n, meta.getColumnStats(n));
    if (nullable) updated.setColumnNullable(column, true);
    meta = updated;
}

void DatabaseInstance::setColumnStats(const std::string& column, int64_t mn, int64_t mx, uint64_t card) {
    engine_->invalidate();  // cached programs were planned on the old catalog
    std::lock_guard<std::mutex> lock(mu_);
    const std::string t = catalog_->tableOfColumn(column);
    ensureTable(t);
    ColumnStatistics st;
    st.min_value_ = mn;
    st.max_value_ = mx;
    st.cardinality_ = card;
    catalog_->getTableMetadataMutable(t).setColumnStats(column, st);
}

uint64_t DatabaseInstance::resultSize(const std::string& sql) const {
    auto tree = buildOptimizedTree(sql, catalog_.get());
    const OperatorNode* n = tree.get();
    while (n && n->getType() != OperatorType::AGGREGATE)
        n = n->getChildren().empty() ? nullptr : n->getChildren()[0].get();
    return n ? static_cast<const AggregateNode*>(n)->calculateResultSize(*catalog_) : 0;
}

void DatabaseInstance::generateSSB(const GenSpec& spec) {
    engine_->invalidate();  // cached programs were planned on the old catalog
    if (device_ < 0) throw std::runtime_error("generateSSB needs a GPU device");
    for (const auto& t : genTables()) createTable(t, genTableRows(spec, t), t == "LINEORDER");
    for (const auto& t : genTables()) {
        const int64_t rows = genTableRows(spec, t);
        for (const auto& c : genTableColumns(t)) {
            const cq_gen_recipe rc = recipeFor(spec, t, c);
            std::lock_guard<std::mutex> lock(mu_);
            void* d = deviceAlloc(static_cast<size_t>(rows) * 4);
            void* v = nullptr;
            if (rc.null_bp > 0) v = deviceAlloc(static_cast<size_t>((rows + 63) / 64) * 8);
            hipCheck(cq_launch_generate(static_cast<int*>(d), static_cast<unsigned long long*>(v), &rc, 0, rows, ctx_->stream),
                     "generate column");
            DeviceColumn dc;
            dc.data = static_cast<const int*>(d);
            dc.validity = static_cast<const uint64_t*>(v);
            dc.rows = rows;
            columns_[c] = dc;
            auto& meta = catalog_->getTableMetadataMutable(t);
            std::vector<std::string> names = meta.getColumnNames();
            names.push_back(c);
            TableMetadata updated(meta.getName(), names, meta.getSize(), meta.isFactTable());
            if (v) updated.setColumnNullable(c, true);
            meta = updated;
        }
    }
    hipCheck(hipStreamSynchronize(ctx_->stream), "generate sync");
    setPrimaryKey("PART", "p_partkey");
    setPrimaryKey("SUPPLIER", "s_suppkey");
    setPrimaryKey("CUSTOMER", "c_custkey");
    setPrimaryKey("DDATE", "d_datekey");
}

ColumnStatistics DatabaseInstance::computeStats(const std::string& column, bool& nullable) {
    const DeviceColumn& dc = columns_.at(column);
    ColumnStatistics st;
    nullable = dc.validity != nullptr;
    if (dc.rows == 0) return st;
    auto preset = preset_stats_.find(column);
    if (preset != preset_stats_.end()) {
        st.min_value_ = preset->second.min_value_;
        st.max_value_ = preset->second.max_value_;
        st.cardinality_ = preset->second.cardinality_;
        if (dc.rows <= (int64_t(1) << 24) && st.min_value_ >= 10000101 && st.max_value_ <= 99991231) {
            std::vector<int32_t> host(static_cast<size_t>(dc.rows));
            fetchColumn(column, host.data(), nullptr);
            st.is_calendar_ = calendarColumn(host.data(), dc.rows, st.min_value_, st.max_value_);
        }
        return st;
    }
    if (device_ < 0) {
        const int32_t* d = dc.data;
        int64_t mn = d[0], mx = d[0];
        for (int64_t i = 1; i < dc.rows; ++i) mn = std::min<int64_t>(mn, d[i]), mx = std::max<int64_t>(mx, d[i]);
        std::vector<uint64_t> seen(static_cast<size_t>((mx - mn) / 64 + 1), 0);
        uint64_t card = 0;
        for (int64_t i = 0; i < dc.rows; ++i) {
            const uint64_t k = static_cast<uint64_t>(d[i] - mn);
            if (!(seen[k >> 6] & (1ULL << (k & 63)))) seen[k >> 6] |= 1ULL << (k & 63), ++card;
        }
        st.min_value_ = mn;
        st.max_value_ = mx;
        st.cardinality_ = card;
        st.is_calendar_ = calendarColumn(d, dc.rows, mn, mx);
        return st;
    }
    hipStream_t s = ctx_->stream;
    DeviceBuffer& mm = ctx_->scratch("stats_minmax");
    mm.ensure(24);
    auto* m = static_cast<unsigned long long*>(mm.data());
    const unsigned long long init[3] = {~0ull, 0ull, 0ull};
    hipCheck(hipMemcpyAsync(m, init, 24, hipMemcpyHostToDevice, s), "stats init");
    hipCheck(cq_launch_minmax(dc.data, dc.rows, m, s), "minmax");
    unsigned long long h[3];
    hipCheck(hipMemcpyAsync(h, m, 24, hipMemcpyDeviceToHost, s), "stats read");
    hipCheck(hipStreamSynchronize(s), "stats sync");
    st.min_value_ = static_cast<int64_t>(h[0]) - 0x80000000LL;
    st.max_value_ = static_cast<int64_t>(h[1]) - 0x80000000LL;
    const uint64_t range = static_cast<uint64_t>(st.max_value_ - st.min_value_) + 1;
    const long long words = static_cast<long long>((range + 31) / 32);
    DeviceBuffer& bits = ctx_->scratch("stats_bits");
    bits.ensure(static_cast<size_t>(words) * 4);
    hipCheck(cq_launch_fill_u32(static_cast<unsigned int*>(bits.data()), 0u, words, s), "stats fill");
    hipCheck(hipMemcpyAsync(m + 2, init + 2, 8, hipMemcpyHostToDevice, s), "stats init");
    hipCheck(cq_launch_distinct(dc.data, dc.rows, st.min_value_, static_cast<unsigned int*>(bits.data()), words, m + 2, s),
             "distinct");
    hipCheck(hipMemcpyAsync(&h[2], m + 2, 8, hipMemcpyDeviceToHost, s), "stats read");
    hipCheck(hipStreamSynchronize(s), "stats sync");
    st.cardinality_ = h[2];
    // dimension-sized columns of 8-digit values: are they all yyyymmdd dates?
    if (dc.rows <= (int64_t(1) << 24) && st.min_value_ >= 10000101 && st.max_value_ <= 99991231) {
        std::vector<int32_t> host(static_cast<size_t>(dc.rows));
        hipCheck(hipMemcpyAsync(host.data(), dc.data, static_cast<size_t>(dc.rows) * 4, hipMemcpyDeviceToHost, s), "calendar check");
        hipCheck(hipStreamSynchronize(s), "calendar check");
        st.is_calendar_ = calendarColumn(host.data(), dc.rows, st.min_value_, st.max_value_);
    }
    return st;
}

void DatabaseInstance::finalizeCatalog() {
    engine_->invalidate();  // cached programs were planned on the old catalog
    std::lock_guard<std::mutex> lock(mu_);
    for (auto& meta : catalog_->getTablesMetadataMutable()) {
        for (const auto& c : meta.getColumnNames()) {
            bool nul = false;
            ColumnStatistics st = computeStats(c, nul);
            if (st.cardinality_ == meta.getSize() && meta.getSize() > 0) st.is_unique_ = true;
            meta.setColumnStats(c, st);
            if (nul) meta.setColumnNullable(c, true);
        }
    }
}

void DatabaseInstance::loadReferenceFiles(const std::string& dir, const int64_t rows[5], const std::string& stats_cache,
                                          const std::vector<std::string>& only) {
    std::string d = dir;
    if (!d.empty() && d.back() != '/') d += '/';
    auto wanted = [&](const std::string& c) { return only.empty() || std::find(only.begin(), only.end(), c) != only.end(); };
    for (const auto& c : only) (void)storage::tableOf(c);  // unknown names are errors
    std::unique_ptr<storage::StatsCache> cache;
    if (!stats_cache.empty()) cache = std::make_unique<storage::StatsCache>(stats_cache);
    struct Pending {
        std::string column;
        uintmax_t size, mtime;
    };
    std::vector<Pending> pending;
    const auto& tables = storage::referenceTables();
    for (size_t t = 0; t < tables.size(); ++t) {
        const std::string& table = tables[t];
        bool any = false;
        for (const auto& c : storage::referenceColumns(table)) any = any || wanted(c);
        if (!any) continue;
        if (rows[t] < 0) throw std::runtime_error("negative row count for " + table);
        createTable(table, rows[t], table == "LINEORDER");
        for (const auto& c : storage::referenceColumns(table)) {
            if (!wanted(c)) continue;
            const std::string path = d + storage::physicalName(c);
            const std::vector<int32_t> data = storage::readColumnFile(path, rows[t]);
            const std::vector<uint64_t> valid = storage::readValidity(d, c, rows[t]);
            loadColumn(table, c, data.data(), valid.empty() ? nullptr : valid.data());
            if (!cache) continue;
            const uintmax_t size = storage::StatsCache::fileSize(path), mtime = storage::StatsCache::fileMtime(path);
            storage::CachedStats cs;
            if (cache->lookup(c, size, mtime, cs) && cs.cardinality > 0) {
                ColumnStatistics st;
                st.min_value_ = cs.min_val;
                st.max_value_ = cs.max_val;
                st.cardinality_ = static_cast<uint64_t>(cs.cardinality);
                preset_stats_[c] = st;
            } else {
                pending.push_back({c, size, mtime});
            }
        }
        const std::string pk = storage::referencePrimaryKey(table);
        if (!pk.empty() && wanted(pk)) setPrimaryKey(table, pk);
    }
    finalizeCatalog();
    if (!cache) return;
    for (const auto& p : pending) {
        const auto& meta = catalog_->getTableMetadata(storage::tableOf(p.column));
        if (!meta.hasColumnStats(p.column)) continue;
        const ColumnStatistics* st = &meta.getColumnStats(p.column);
        storage::CachedStats cs;
        cs.file_size = p.size;
        cs.mtime = p.mtime;
        cs.min_val = static_cast<int>(st->min_value_);
        cs.max_val = static_cast<int>(st->max_value_);
        cs.cardinality = static_cast<int>(st->cardinality_);
        cache->put(p.column, cs);
    }
    cache->save();
}

void DatabaseInstance::saveReferenceFiles(const std::string& dir, const std::string& validity_suffix) const {
    std::string d = dir;
    if (!d.empty() && d.back() != '/') d += '/';
    for (const auto& table : storage::referenceTables()) {
        if (!catalog_->hasTable(table)) continue;
        for (const auto& c : storage::referenceColumns(table)) {
            auto it = columns_.find(c);
            if (it == columns_.end()) continue;
            const int64_t rows = it->second.rows;
            std::vector<int32_t> data(static_cast<size_t>(rows));
            std::vector<uint64_t> valid(static_cast<size_t>((rows + 63) / 64));
            fetchColumn(c, data.data(), valid.data());
            storage::writeColumnFile(d + storage::physicalName(c), data.data(), rows);
            if (it->second.validity) storage::writeValidity(d, c, validity_suffix, valid.data(), rows);
        }
    }
}

QueryResult DatabaseInstance::executeQuery(const std::string& sql) { return executeQueryRange(sql, 0, -1); }

QueryResult DatabaseInstance::executeQueryRange(const std::string& sql, int64_t row_begin, int64_t row_end) {
    if (device_ < 0) throw std::runtime_error("executeQuery needs a GPU device (host-only database)");
    std::lock_guard<std::mutex> lock(mu_);
    hipCheck(hipSetDevice(device_), "hipSetDevice");
    ctx_->row_begin = row_begin;
    ctx_->row_end = row_end;
    engine_->executeQuery(sql, ctx_.get());
    return ctx_->result;
}

void DatabaseInstance::groupContexts(ExecutionContext& base, std::vector<std::unique_ptr<ExecutionContext>>& pool, size_t n) {
    while (pool.size() < n) {
        auto c = std::make_unique<ExecutionContext>();
        c->device = base.device;
        c->stream = base.stream;
        c->owns_stream = false;
        c->columns = base.columns;
        pool.push_back(std::move(c));
    }
    for (size_t k = 0; k < n; ++k) {
        ExecutionContext& c = *pool[k];
        c.config = base.config;
        c.loaded_device_bytes = base.loaded_device_bytes;
        c.row_begin = 0;
        c.row_end = -1;
    }
}

std::vector<QueryResult> DatabaseInstance::executeGroup(const std::vector<std::string>& sqls, std::vector<std::string>& errs) {
    if (device_ < 0) throw std::runtime_error("executeGroup needs a GPU device (host-only database)");
    std::lock_guard<std::mutex> lock(mu_);
    hipCheck(hipSetDevice(device_), "hipSetDevice");
    groupContexts(*ctx_, group_ctx_, sqls.size());
    std::vector<ExecutionContext*> cs;
    for (size_t k = 0; k < sqls.size(); ++k) cs.push_back(group_ctx_[k].get());
    errs = engine_->executeInterpretedGroup(sqls, cs);
    std::vector<QueryResult> out(sqls.size());
    for (size_t k = 0; k < sqls.size(); ++k)
        if (errs[k].empty()) out[k] = std::move(cs[k]->result);
    return out;
}

QueryProgram DatabaseInstance::prepare(const std::string& sql) const {
    if (ctx_->config.prepared_cache) return engine_->prepareCached(sql, ctx_->config.codegen);
    return engine_->prepare(sql, ctx_->config.codegen);
}

std::string DatabaseInstance::explain(const std::string& sql) const {
    auto tree = buildOptimizedTree(sql, catalog_.get());
    std::ostringstream o;
    printPlanSexpr(o, tree.get());
    return o.str();
}

std::string DatabaseInstance::explainTree(const std::string& sql) const {
    auto tree = buildOptimizedTree(sql, catalog_.get());
    std::ostringstream o;
    printPlanTree(o, tree.get());
    return o.str();
}

std::string DatabaseInstance::generateQueryCode(const std::string& sql) const { return prepare(sql).source; }

ModuleHandle DatabaseInstance::compile(const std::string& sql) {
    QueryProgram p = prepare(sql);
    ModuleHandle m = engine_->moduleFor(p, nullptr);
    if (device_ >= 0) {
        hipCheck(hipSetDevice(device_), "hipSetDevice");
        m->load();  // loaded now, so later executions load nothing
    }
    return m;
}

QueryResult DatabaseInstance::executeQueryInto(const std::string& sql, void* scratch, size_t bytes) {
    if (device_ < 0) throw std::runtime_error("executeQuery needs a GPU device (host-only database)");
    std::lock_guard<std::mutex> lock(mu_);
    hipCheck(hipSetDevice(device_), "hipSetDevice");
    ctx_->resetForQuery();
    ctx_->program = prepare(sql);
    ModuleHandle m = ctx_->program.vm ? nullptr : engine_->cachedModule(ctx_->program);
    executor_->executeInto(m, ctx_.get(), scratch, bytes);
    return ctx_->result;
}

void DatabaseInstance::fetchColumn(const std::string& column, int32_t* out, uint64_t* validity_out) const {
    auto it = columns_.find(column);
    if (it == columns_.end()) throw std::runtime_error("unknown column " + column);
    const DeviceColumn& dc = it->second;
    const size_t words = static_cast<size_t>((dc.rows + 63) / 64);
    if (device_ < 0) {
        std::memcpy(out, dc.data, static_cast<size_t>(dc.rows) * 4);
        if (validity_out) {
            if (dc.validity) std::memcpy(validity_out, dc.validity, words * 8);
            else std::fill(validity_out, validity_out + words, ~0ULL);
        }
        return;
    }
    {
        DeviceMapLock lock;  // null-stream copy: synchronizes with every stream
        hipCheck(hipMemcpy(out, dc.data, static_cast<size_t>(dc.rows) * 4, hipMemcpyDeviceToHost), "fetch column");
    }
    if (validity_out) {
        if (dc.validity) {
            DeviceMapLock lock;  // null-stream copy: synchronizes with every stream
            hipCheck(hipMemcpy(validity_out, dc.validity, words * 8, hipMemcpyDeviceToHost), "fetch validity");
        } else {
            std::fill(validity_out, validity_out + words, ~0ULL);
        }
    }
}

}  // namespace db
