// The reference's on-disk column store, read (and written) natively.
//
//   columns     one raw little-endian int32 file per column, named by the
//               table and the column's position in the table's schema,
//               e.g. lo_revenue -> LINEORDER12 (include/crystal/utils.h:46-96,
//               loadColumn 240-254: exactly `rows` values are read; a short
//               file leaves the rest 0, a missing one is an error)
//   validity    an optional bitmap beside the column, looked up under ten
//               suffixes, first beside the physical name and then beside the
//               column name (utils.h:142-238): u64 valid / null bitsets of
//               exactly ceil(rows/64) words, or one valid / null byte per row;
//               bits past the last row are cleared
//   statistics  a text cache of "column file_size mtime min max cardinality"
//               lines; a line is used while the column file's size and
//               modification time (seconds, std::filesystem clock) match,
//               otherwise the statistics are recomputed and the line
//               replaced (include/core/catalog_cache.h:17-123)
//   row counts  fixed per scale factor: SF 10 and SF 20 (utils.h:11-25)
#pragma once
#include <cstdint>
#include <map>
#include <string>
#include <utility>
#include <vector>

namespace db {
namespace storage {

// Tables in the reference's catalog order (database_instance.cpp:82-151).
const std::vector<std::string>& referenceTables();
// Columns of a table in file-index order; primary key ("" for the fact table).
const std::vector<std::string>& referenceColumns(const std::string& table);
std::string referencePrimaryKey(const std::string& table);
// TABLE<index>; throws for a column outside the schema.
std::string physicalName(const std::string& column);
std::string tableOf(const std::string& column);
// rows[5] in referenceTables() order for SF 10 or 20; false otherwise.
bool referenceRows(int sf, int64_t rows[5]);

std::vector<int32_t> readColumnFile(const std::string& path, int64_t rows);
void writeColumnFile(const std::string& path, const int32_t* data, int64_t rows);

enum class ValidityKind { ValidBits64, NullBits64, ValidBytes, NullBytes };
// Candidate files in lookup order.
std::vector<std::pair<std::string, ValidityKind>> validityCandidates(const std::string& dir, const std::string& column);
// LSB-first valid bits of ceil(rows/64) words, or empty when no file exists.
std::vector<uint64_t> readValidity(const std::string& dir, const std::string& column, int64_t rows,
                                   std::string* used_path = nullptr);
// suffix: one of the candidate suffixes (".valid64", ".null8", ...)
void writeValidity(const std::string& dir, const std::string& column, const std::string& suffix, const uint64_t* valid,
                   int64_t rows);

struct CachedStats {
    uintmax_t file_size = 0;
    uintmax_t mtime = 0;
    int min_val = 0;
    int max_val = 0;
    int cardinality = 0;
};

class StatsCache {
public:
    explicit StatsCache(std::string path);  // reads the file if it exists
    // the entry for `column` if its size and mtime match
    bool lookup(const std::string& column, uintmax_t size, uintmax_t mtime, CachedStats& out) const;
    void put(const std::string& column, const CachedStats& stats);
    void save();  // writes the file (creating its directory) when entries changed
    static uintmax_t fileMtime(const std::string& path);
    static uintmax_t fileSize(const std::string& path);

private:
    std::string path_;
    std::map<std::string, CachedStats> entries_;
    bool modified_ = false;
};

}  // namespace storage
}  // namespace db
