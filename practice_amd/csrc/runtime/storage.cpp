// Reference on-disk column store: readers, writers and the statistics cache.
#include "storage.h"

#include <chrono>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace db {
namespace storage {
namespace fs = std::filesystem;

namespace {
struct Schema {
    std::vector<std::string> tables;
    std::map<std::string, std::vector<std::string>> columns;
    std::map<std::string, std::string> pk;
};
// Column order of include/crystal/utils.h:47-74 (the file index); primary
// keys of database_instance.cpp:104-151.
const Schema& schema() {
    static const Schema s = [] {
        Schema x;
        x.tables = {"LINEORDER", "SUPPLIER", "CUSTOMER", "PART", "DDATE"};
        x.columns["LINEORDER"] = {"lo_orderkey", "lo_linenumber", "lo_custkey", "lo_partkey", "lo_suppkey", "lo_orderdate",
                                  "lo_orderpriority", "lo_shippriority", "lo_quantity", "lo_extendedprice",
                                  "lo_ordtotalprice", "lo_discount", "lo_revenue", "lo_supplycost", "lo_tax",
                                  "lo_commitdate", "lo_shipmode"};
        x.columns["SUPPLIER"] = {"s_suppkey", "s_name", "s_address", "s_city", "s_nation", "s_region", "s_phone"};
        x.columns["CUSTOMER"] = {"c_custkey", "c_name", "c_address", "c_city", "c_nation", "c_region", "c_phone",
                                 "c_mktsegment"};
        x.columns["PART"] = {"p_partkey", "p_name", "p_mfgr", "p_category", "p_brand1", "p_color", "p_type", "p_size",
                             "p_container"};
        x.columns["DDATE"] = {"d_datekey",       "d_date",          "d_dayofweek",       "d_month",
                              "d_year",          "d_yearmonthnum",  "d_yearmonth",       "d_daynuminweek",
                              "d_daynuminmonth", "d_daynuminyear",  "d_sellingseason",   "d_lastdayinweekfl",
                              "d_lastdayinmonthfl", "d_holidayfl", "d_weekdayfl"};
        x.pk = {{"LINEORDER", ""}, {"SUPPLIER", "s_suppkey"}, {"CUSTOMER", "c_custkey"}, {"PART", "p_partkey"},
                {"DDATE", "d_datekey"}};
        return x;
    }();
    return s;
}

uint64_t lastWordMask(int64_t rows) {
    const int64_t rem = rows & 63;
    return rem == 0 ? ~0ull : ((1ull << rem) - 1ull);
}

const char* const kSuffixes[] = {".valid64", ".valid", ".validity", ".valid_bitmap", ".valid8",
                                 ".null64",  ".null",  ".nulls",    ".null_bitmap",  ".null8"};
ValidityKind kindOf(const std::string& suffix) {
    if (suffix == ".valid8") return ValidityKind::ValidBytes;
    if (suffix == ".null8") return ValidityKind::NullBytes;
    if (suffix.compare(0, 5, ".null") == 0) return ValidityKind::NullBits64;
    return ValidityKind::ValidBits64;
}
}  // namespace

const std::vector<std::string>& referenceTables() { return schema().tables; }

const std::vector<std::string>& referenceColumns(const std::string& table) {
    auto it = schema().columns.find(table);
    if (it == schema().columns.end()) throw std::runtime_error("unknown table " + table);
    return it->second;
}

std::string referencePrimaryKey(const std::string& table) {
    auto it = schema().pk.find(table);
    if (it == schema().pk.end()) throw std::runtime_error("unknown table " + table);
    return it->second;
}

std::string tableOf(const std::string& column) {
    for (const auto& t : schema().tables)
        for (const auto& c : schema().columns.at(t))
            if (c == column) return t;
    throw std::runtime_error("column " + column + " is not in the SSB schema");
}

std::string physicalName(const std::string& column) {
    const std::string t = tableOf(column);
    const auto& cols = schema().columns.at(t);
    for (size_t i = 0; i < cols.size(); ++i)
        if (cols[i] == column) return t + std::to_string(i);
    return t + "-1";
}

bool referenceRows(int sf, int64_t rows[5]) {
    if (sf == 10) {
        const int64_t r[5] = {59986217, 20000, 300000, 800000, 2557};
        std::memcpy(rows, r, sizeof(r));
        return true;
    }
    if (sf == 20) {
        const int64_t r[5] = {119994746, 40000, 600000, 1000000, 2556};
        std::memcpy(rows, r, sizeof(r));
        return true;
    }
    return false;
}

std::vector<int32_t> readColumnFile(const std::string& path, int64_t rows) {
    std::vector<int32_t> v(static_cast<size_t>(rows), 0);
    std::ifstream in(path, std::ios::binary);
    if (!in) throw std::runtime_error("Cannot open file: " + path);
    in.read(reinterpret_cast<char*>(v.data()), static_cast<std::streamsize>(rows) * 4);  // short file: rest stays 0
    return v;
}

void writeColumnFile(const std::string& path, const int32_t* data, int64_t rows) {
    std::ofstream out(path, std::ios::binary | std::ios::trunc);
    if (!out) throw std::runtime_error("Cannot create file: " + path);
    out.write(reinterpret_cast<const char*>(data), static_cast<std::streamsize>(rows) * 4);
    if (!out) throw std::runtime_error("Failed to write file: " + path);
}

std::vector<std::pair<std::string, ValidityKind>> validityCandidates(const std::string& dir, const std::string& column) {
    const std::string physical = dir + physicalName(column), logical = dir + column;
    std::vector<std::pair<std::string, ValidityKind>> c;
    for (const char* s : kSuffixes) c.emplace_back(physical + s, kindOf(s));
    if (logical != physical)
        for (const char* s : kSuffixes) c.emplace_back(logical + s, kindOf(s));
    return c;
}

std::vector<uint64_t> readValidity(const std::string& dir, const std::string& column, int64_t rows, std::string* used_path) {
    const size_t words = static_cast<size_t>((rows + 63) / 64);
    if (words == 0) return {};
    for (const auto& [path, kind] : validityCandidates(dir, column)) {
        std::error_code ec;
        if (!fs::exists(path, ec)) continue;
        const uintmax_t size = fs::file_size(path);
        std::ifstream in(path, std::ios::binary);
        if (!in) throw std::runtime_error("Cannot open null bitmap file: " + path);
        std::vector<uint64_t> valid(words, 0);
        if (kind == ValidityKind::ValidBits64 || kind == ValidityKind::NullBits64) {
            const uintmax_t want = static_cast<uintmax_t>(words) * 8;
            if (size != want)
                throw std::runtime_error("Invalid null bitmap size for " + column + " at " + path + ": expected " +
                                         std::to_string(want) + " bytes for uint64 bitset, got " + std::to_string(size));
            in.read(reinterpret_cast<char*>(valid.data()), static_cast<std::streamsize>(want));
            if (!in) throw std::runtime_error("Failed to read null bitmap file: " + path);
            if (kind == ValidityKind::NullBits64)
                for (auto& w : valid) w = ~w;
        } else {
            if (size != static_cast<uintmax_t>(rows))
                throw std::runtime_error("Invalid null bitmap size for " + column + " at " + path + ": expected " +
                                         std::to_string(rows) + " bytes for byte bitmap, got " + std::to_string(size));
            std::vector<unsigned char> bytes(static_cast<size_t>(rows));
            in.read(reinterpret_cast<char*>(bytes.data()), static_cast<std::streamsize>(rows));
            if (!in) throw std::runtime_error("Failed to read null bitmap file: " + path);
            const bool ones_valid = kind == ValidityKind::ValidBytes;
            for (int64_t r = 0; r < rows; ++r)
                if ((bytes[static_cast<size_t>(r)] != 0) == ones_valid) valid[static_cast<size_t>(r >> 6)] |= 1ull << (r & 63);
        }
        valid.back() &= lastWordMask(rows);
        if (used_path) *used_path = path;
        return valid;
    }
    return {};
}

void writeValidity(const std::string& dir, const std::string& column, const std::string& suffix, const uint64_t* valid,
                   int64_t rows) {
    bool known = false;
    for (const char* s : kSuffixes) known = known || suffix == s;
    if (!known) throw std::runtime_error("unknown validity suffix " + suffix);
    const std::string path = dir + physicalName(column) + suffix;
    std::ofstream out(path, std::ios::binary | std::ios::trunc);
    if (!out) throw std::runtime_error("Cannot create file: " + path);
    const ValidityKind kind = kindOf(suffix);
    const size_t words = static_cast<size_t>((rows + 63) / 64);
    if (kind == ValidityKind::ValidBits64 || kind == ValidityKind::NullBits64) {
        std::vector<uint64_t> w(valid, valid + words);
        if (kind == ValidityKind::NullBits64)
            for (auto& x : w) x = ~x;
        if (!w.empty()) w.back() &= lastWordMask(rows);
        out.write(reinterpret_cast<const char*>(w.data()), static_cast<std::streamsize>(words * 8));
    } else {
        std::vector<unsigned char> b(static_cast<size_t>(rows));
        for (int64_t r = 0; r < rows; ++r) {
            const bool v = (valid[r >> 6] >> (r & 63)) & 1ull;
            b[static_cast<size_t>(r)] = (v == (kind == ValidityKind::ValidBytes)) ? 1 : 0;
        }
        out.write(reinterpret_cast<const char*>(b.data()), static_cast<std::streamsize>(rows));
    }
    if (!out) throw std::runtime_error("Failed to write file: " + path);
}

StatsCache::StatsCache(std::string path) : path_(std::move(path)) {
    std::ifstream in(path_);
    if (!in) return;
    std::string line;
    while (std::getline(in, line)) {
        if (line.empty()) continue;
        std::istringstream is(line);
        std::string col;
        CachedStats s;
        if (is >> col >> s.file_size >> s.mtime >> s.min_val >> s.max_val >> s.cardinality) entries_[col] = s;
    }
}

bool StatsCache::lookup(const std::string& column, uintmax_t size, uintmax_t mtime, CachedStats& out) const {
    auto it = entries_.find(column);
    if (it == entries_.end() || it->second.file_size != size || it->second.mtime != mtime) return false;
    out = it->second;
    return true;
}

void StatsCache::put(const std::string& column, const CachedStats& stats) {
    entries_[column] = stats;
    modified_ = true;
}

void StatsCache::save() {
    if (!modified_ || path_.empty()) return;
    const fs::path p(path_);
    if (p.has_parent_path()) fs::create_directories(p.parent_path());
    std::ofstream out(path_, std::ios::trunc);
    if (!out) throw std::runtime_error("Cannot write statistics cache " + path_);
    for (const auto& [col, s] : entries_)
        out << col << " " << s.file_size << " " << s.mtime << " " << s.min_val << " " << s.max_val << " " << s.cardinality
            << "\n";
    modified_ = false;
}

uintmax_t StatsCache::fileMtime(const std::string& path) {
    // seconds of the file clock's time since its epoch, as an unsigned
    // number (catalog_cache.h:32-36 stores the same quantity)
    const auto t = fs::last_write_time(path);
    return static_cast<uintmax_t>(std::chrono::duration_cast<std::chrono::seconds>(t.time_since_epoch()).count());
}

uintmax_t StatsCache::fileSize(const std::string& path) { return fs::file_size(path); }

}  // namespace storage
}  // namespace db
