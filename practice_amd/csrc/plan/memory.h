// Catalog: tables, column statistics, keys, nullability.
//
// Same types and accessors as the reference catalog (include/core/memory.h:9-185):
// ColumnStatistics{min_value_, max_value_, cardinality_, is_unique_,
// is_primary_key_, is_nullable_}, TableMetadata, Catalog.  One addition:
// Catalog::tableOfColumn() resolves a column to its table through the column
// lists, falling back to the reference's first-letter rule (lo_ -> LINEORDER,
// s_ -> SUPPLIER, c_ -> CUSTOMER, p_ -> PART, d_ -> DDATE; crystal/utils.h:38-46)
// so tables outside the SSB schema resolve too.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

enum class LogicalType : uint8_t {
    Int64,
    UInt64,
    Float64,
};

struct ColumnStatistics {
    int64_t min_value_ = 0;
    int64_t max_value_ = 0;
    uint64_t cardinality_ = 0;  // number of distinct values
    bool is_unique_ = false;
    bool is_primary_key_ = false;
    bool is_nullable_ = false;
    // every value is a calendar date written yyyymmdd (month 1-12, day 1-31):
    // key-indexed tables may index such keys by (year, month, day)
    bool is_calendar_ = false;
};

class TableMetadata {
public:
    TableMetadata(std::string name, std::vector<std::string> column_names, uint64_t size, bool is_fact_table)
        : name_(std::move(name)), column_names_(std::move(column_names)), size_(size), is_fact_table_(is_fact_table) {}

    const std::string& getName() const { return name_; }
    const std::vector<std::string>& getColumnNames() const { return column_names_; }
    uint64_t getSize() const { return size_; }
    uint64_t getColumnCount() const { return column_names_.size(); }
    bool isFactTable() const { return is_fact_table_; }
    bool hasColumn(const std::string& col) const {
        for (const auto& c : column_names_)
            if (c == col) return true;
        return false;
    }

    // Keeps key / nullability flags that were set before (memory.h:79-88).
    void setColumnStats(const std::string& col, const ColumnStatistics& stats) {
        auto& dst = column_stats_[col];
        const bool was_unique = dst.is_unique_, was_pk = dst.is_primary_key_, was_nullable = dst.is_nullable_;
        dst = stats;
        dst.is_unique_ = stats.is_unique_ || was_unique || stats.is_primary_key_ || was_pk;
        dst.is_primary_key_ = stats.is_primary_key_ || was_pk;
        dst.is_nullable_ = stats.is_nullable_ || was_nullable;
    }
    const ColumnStatistics& getColumnStats(const std::string& col) const {
        auto it = column_stats_.find(col);
        if (it == column_stats_.end()) throw std::runtime_error("Column stats not found: " + col);
        return it->second;
    }
    bool hasColumnStats(const std::string& col) const { return column_stats_.count(col) > 0; }
    void setColumnUnique(const std::string& col, bool unique = true) {
        auto& s = column_stats_[col];
        s.is_unique_ = unique;
        if (!unique) s.is_primary_key_ = false;
    }
    void setColumnPrimaryKey(const std::string& col, bool pk = true) {
        auto& s = column_stats_[col];
        s.is_primary_key_ = pk;
        if (pk) s.is_unique_ = true;
    }
    void setColumnNullable(const std::string& col, bool nullable = true) { column_stats_[col].is_nullable_ = nullable; }
    bool isColumnUnique(const std::string& col) const {
        auto it = column_stats_.find(col);
        return it != column_stats_.end() && it->second.is_unique_;
    }
    bool isColumnPrimaryKey(const std::string& col) const {
        auto it = column_stats_.find(col);
        return it != column_stats_.end() && it->second.is_primary_key_;
    }
    bool isColumnNullable(const std::string& col) const {
        auto it = column_stats_.find(col);
        return it != column_stats_.end() && it->second.is_nullable_;
    }

private:
    std::string name_;
    std::vector<std::string> column_names_;
    uint64_t size_;
    bool is_fact_table_;
    std::unordered_map<std::string, ColumnStatistics> column_stats_;
};

class Catalog {
public:
    Catalog() = default;
    explicit Catalog(std::vector<TableMetadata> tables) : tables_(std::move(tables)) {}

    void pushTableMetadata(TableMetadata tb) { tables_.push_back(std::move(tb)); }
    const std::vector<TableMetadata>& getTablesMetadata() const { return tables_; }
    std::vector<TableMetadata>& getTablesMetadataMutable() { return tables_; }
    uint64_t getTablesMetadataCount() const { return tables_.size(); }
    bool hasTable(const std::string& name) const {
        for (const auto& t : tables_)
            if (t.getName() == name) return true;
        return false;
    }
    const TableMetadata& getTableMetadata(const std::string& name) const {
        for (const auto& t : tables_)
            if (t.getName() == name) return t;
        throw std::runtime_error("Table not found: " + name);
    }
    TableMetadata& getTableMetadataMutable(const std::string& name) {
        for (auto& t : tables_)
            if (t.getName() == name) return t;
        throw std::runtime_error("Table not found: " + name);
    }

    // Table owning `column`; "" when unknown.
    std::string tableOfColumn(const std::string& column) const {
        for (const auto& t : tables_)
            if (t.hasColumn(column)) return t.getName();
        return prefixTable(column);
    }
    static std::string prefixTable(const std::string& column) {
        if (column.empty()) return "";
        switch (column[0]) {
            case 'l': return "LINEORDER";
            case 's': return "SUPPLIER";
            case 'c': return "CUSTOMER";
            case 'p': return "PART";
            case 'd': return "DDATE";
            default: return "";
        }
    }
    bool columnNullable(const std::string& column) const {
        const std::string t = tableOfColumn(column);
        if (!hasTable(t)) return false;
        return getTableMetadata(t).isColumnNullable(column);
    }

private:
    std::vector<TableMetadata> tables_;
};
