/*
 * crystal_hip.h -- C ABI of the MI355X-native JIT query backend.
 *
 * Drop-in boundary for the reference's JIT path.  Every entry point returns
 * 0 on success and a negative code on failure (no C++ exception crosses this
 * ABI); cq_last_error() returns the message of the calling thread's last
 * failure.  A C++ ICompiler / IExecutor adapter re-throws, so QueryEngine's
 * contract (exceptions out of executeQuery) is unchanged.
 *
 * Interfaces replaced (reference file:line):
 *   cq_compile / cq_module_release
 *       ICompiler::compile(source, query_hash) -> lib path
 *       (include/core/execution.h:291-295; AdaptiveCppCompiler,
 *       src/core/execution.cpp:114-151)
 *   cq_launch / cq_module_kernel_count
 *       IExecutor::execute(lib_path, ctx) + extern "C" execute_query(ctx)
 *       (include/core/execution.h:297-301; DynamicLibraryExecutor,
 *       src/core/execution.cpp:161-192; generated ABI, visitor.cpp:4804)
 *   cq_db_create / cq_db_destroy / cq_db_generate_ssb / cq_db_create_table /
 *   cq_db_load_column / cq_db_set_primary_key / cq_db_finalize /
 *   cq_db_load_reference_files / cq_reference_rows
 *       DatabaseInstance ctor, initCatalog, loadData
 *       (src/db/database_instance.cpp:21-174)
 *   cq_db_execute + cq_result_*
 *       DatabaseInstance::executeQuery -> QueryResult
 *       (src/db/database_instance.cpp:180-439; include/core/execution.h:56-96)
 *   cq_db_explain / cq_db_generate_code
 *       DatabaseInstance::generateQueryCode (database_instance.cpp:441-445)
 *   cq_db_column_stats
 *       CatalogCacheManager::getStatsForColumn (include/core/catalog_cache.h:76-123)
 *   cq_batch_execute
 *       no reference counterpart: the batched shared-scan form of many
 *       queries over resident columns (callers: serving / batch schedulers)
 *   cq_server_*
 *       no reference counterpart: the continuous-batching serving loop
 *       (callers: request front ends)
 */
#ifndef CRYSTAL_HIP_API_H
#define CRYSTAL_HIP_API_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CQ_OK 0
#define CQ_ERR_INVALID (-1)
#define CQ_ERR_COMPILE (-2)
#define CQ_ERR_RUNTIME (-3)
#define CQ_ERR_UNSUPPORTED (-4)
#define CQ_ERR_MEMORY (-5)

typedef struct cq_module cq_module;
typedef struct cq_db cq_db;
typedef struct cq_result cq_result;
typedef struct cq_server cq_server;
#include "cq_query_params.h" /* struct cq_query_params: the kernel argument block cq_launch passes */

const char* cq_last_error(void);
const char* cq_version(void);

/* ---- compile / launch seam ---------------------------------------- */
/* hipRTC-compile HIP source for `arch` ("gfx950"); the device primitive
 * library crystal_hip.h is available as an include. */
int cq_compile(const char* source, size_t length, const char* arch, cq_module** out);
/* Launch kernel `name` of a module with a parameter block on a HIP stream
 * (NULL = default stream).  grid/block in work-groups / work-items.  A
 * module of cq_db_generate_code holds cq_build_<j> (one per join, any grid
 * of 256-thread blocks: grid-stride over the dimension rows) and cq_scan
 * (blocks of "block_threads", any grid: grid-stride over the tiles); the
 * caller zero-fills the hash tables and the result slots first. */
int cq_launch(cq_module* module, const char* name, const struct cq_query_params* params, unsigned grid, unsigned block,
              void* stream);
size_t cq_module_code_size(const cq_module* module);
/* sizeof(struct cq_query_params) as this library was built (binding check) */
size_t cq_query_params_size(void);
void cq_module_release(cq_module* module);

/* ---- database ------------------------------------------------------ */
/* device < 0 creates a host-only database: catalog, planning, code
 * generation and hipRTC compilation work; execution does not. */
int cq_db_create(int device, cq_db** out);
void cq_db_destroy(cq_db* db);
int cq_db_generate_ssb(cq_db* db, int sf, long long fact_rows, unsigned long long seed, const char* nullable);
int cq_db_create_table(cq_db* db, const char* table, long long rows, int is_fact);
int cq_db_load_column(cq_db* db, const char* table, const char* column, const int32_t* data, const uint64_t* validity);
int cq_db_set_primary_key(cq_db* db, const char* table, const char* column);
/* The reference's on-disk column store, replacing DatabaseInstance::initCatalog
 * + loadData (src/db/database_instance.cpp:57-174) with loadColumn /
 * loadValidityBitmapIfExists (include/crystal/utils.h:142-254) and the
 * statistics cache (include/core/catalog_cache.h:17-123):
 *   data_dir     directory of TABLE<i> int32 column files (and validity files)
 *   rows         rows of LINEORDER, SUPPLIER, CUSTOMER, PART, DDATE
 *   stats_cache  "column size mtime min max cardinality" text file; NULL or ""
 *                for none; stale or missing lines are recomputed and saved
 *   columns      comma-separated subset of the schema; NULL or "" for all
 * Finalizes the catalog.  A missing column file fails with "Cannot open
 * file: <path>"; a validity file of the wrong size fails as the reference's. */
int cq_db_load_reference_files(cq_db* db, const char* data_dir, const long long rows[5], const char* stats_cache,
                               const char* columns);
/* The row counts the reference compiles in for SF 10 and 20
 * (include/crystal/utils.h:11-25); CQ_ERR_INVALID for any other sf. */
int cq_reference_rows(int sf, long long rows[5]);
/* Every loaded SSB column written in the reference's format; nullable
 * columns get a validity file with the given suffix (".valid64", ".valid",
 * ".validity", ".valid_bitmap", ".valid8", ".null64", ".null", ".nulls",
 * ".null_bitmap", ".null8").  No reference counterpart (its loader is
 * bench/loader); it lets the reference and this library read one data set. */
int cq_db_save_reference_files(cq_db* db, const char* data_dir, const char* validity_suffix);
int cq_db_finalize(cq_db* db);
/* Catalog-only column and explicit statistics (planning without data). */
int cq_db_declare_column(cq_db* db, const char* table, const char* column, int nullable);
int cq_db_set_column_stats(cq_db* db, const char* column, long long min_value, long long max_value,
                           unsigned long long cardinality);
/* AggregateNode::calculateResultSize of the optimized plan
 * (include/core/operators.h:331-369); -1 on error. */
long long cq_db_result_size(cq_db* db, const char* sql);
int cq_db_column_stats(cq_db* db, const char* column, long long* min_value, long long* max_value,
                       unsigned long long* cardinality, int* nullable, int* unique);
int cq_db_fetch_column(cq_db* db, const char* column, int32_t* out, uint64_t* validity_out);
/* Device address of a resident column and of its LSB-first u64 validity
 * bitmap (NULL when the column has none) and its rows, for binding to
 * cq_query_params slots: ExecutionContext::getBuffer<int>("d_<col>") and
 * getNullBitmap("d_<col>") (include/core/execution.h:163-177). */
int cq_db_column_device(cq_db* db, const char* column, const int32_t** data, const uint64_t** validity, long long* rows);
long long cq_db_table_rows(cq_db* db, const char* table);
/* option names: "kernel_timing", "memory_guard", "lds_budget", "block_threads",
 * "vectors_per_thread", "lds_tables", "lds_aggregation", "compact_tables",
 * "dense_tables" (key-indexed tables), "calendar_tables" (key-indexed tables
 * of yyyymmdd date keys indexed by year, month and day), "stream_tables",
 * "memory_guard_reserve_bytes", "prefetch" (load the next tile while the
 * current one is probed), "defer_probes" (batched mask form: probe tables
 * too large for LDS only for items some query still accepts), "interpret" (1: run plans on the ahead-of-time
 * interpreter instead of hipRTC-compiled kernels; same results, no compile),
 * "vm_row_kernel", "vm_linear" (0: interpreted plans always run their bytecode,
 * also those with a linear form), "vm_grid_percent" (1-100: linear-form scans
 * on that share of a full-device grid, leaving CUs to concurrent work), "host_sort_rows", "prepared_cache", "module_cache_capacity" (compiled modules kept,
 * least recently used evicted) */
int cq_db_set_option(cq_db* db, const char* name, long long value);
long long cq_db_cache_stat(cq_db* db, const char* name); /* "size","hits","misses","evictions","compiles" */
int cq_db_clear_cache(cq_db* db);

int cq_db_execute(cq_db* db, const char* sql, cq_result** out);
/* Several statements run together; no reference counterpart (the
 * reference runs one statement per call), for servers that hold many
 * requests at once.  With the "interpret" option on, statements whose
 * interpreted plans have a linear form over the same fact table share scans
 * of at most 16 statements and 8 distinct fact columns (the columns are read
 * once for all of them); every other statement runs as cq_db_execute runs
 * it.  out[i] receives statement i's result, or NULL when that statement
 * failed; the call then returns CQ_ERR_RUNTIME with the first failure's
 * message. */
int cq_db_execute_group(cq_db* db, const char* const* sqls, int n, cq_result** out);
/* cq_db_execute for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * No reference counterpart.  Calls no hipMalloc, hipFree, hipHostMalloc or
 * other device-allocating routine (the cq_result it returns is host memory,
 * as cq_db_execute's): every scratch buffer of the query -- hash tables,
 * aggregate slots, materialization and sort buffers -- comes from one static
 * device array of 32 MiB per GPU, held until the process ends and shared by
 * every call on that GPU (calls that overlap on one GPU race; one database
 * serializes its own calls).  Aggregate queries only; the plan's module
 * must be compiled beforehand with cq_db_compile (which also loads it).
 * CQ_ERR_MEMORY when the query needs more scratch than the array holds,
 * CQ_ERR_INVALID for a projection or a plan not compiled yet. */
int cq_db_execute_noalloc(cq_db* db, const char* sql, cq_result** out);
/* scan only fact rows [row_begin, row_end) */
int cq_db_execute_range(cq_db* db, const char* sql, long long row_begin, long long row_end, cq_result** out);
/* Copies a NUL-terminated string into buf (truncating); *needed gets the full size. */
int cq_db_explain(cq_db* db, const char* sql, char* buf, size_t len, size_t* needed);
int cq_db_generate_code(cq_db* db, const char* sql, char* buf, size_t len, size_t* needed);
/* JSON description of the query program (layout, signature, LDS / table placement) */
int cq_db_describe(cq_db* db, const char* sql, char* buf, size_t len, size_t* needed);
int cq_db_compile(cq_db* db, const char* sql);

/* ---- results ------------------------------------------------------- */
long long cq_result_rows(const cq_result* r);
int cq_result_cols(const cq_result* r);
int cq_result_type(const cq_result* r, int col); /* 0 int64, 1 uint64, 2 float64 */
int cq_result_nullable(const cq_result* r, int col);
int cq_result_ordered(const cq_result* r);
const unsigned long long* cq_result_column(const cq_result* r, int col);
const unsigned char* cq_result_validity(const cq_result* r, int col);
/* "gpu_execute_ms","scan_kernel_ms","compile_ms","codegen_ms","parse_ms",
 * "host_fetch_ms","engine_ms","cache_hit" */
double cq_result_timing(const cq_result* r, const char* name);
/* A result column's name: its alias, else the group column's name or the
 * aggregate's function (inferResultColumnNames, execution.cpp:540-579). */
const char* cq_result_column_name(const cq_result* r, int col);
/* The text block the reference console prints for a result
 * (src/cli/terminal_app.cpp:263-302): "--- Result ---", "Columns: a | b",
 * the rows ("Row i: x | y", or the bare cell of a one-column result; the
 * last row_limit rows when row_limit > 0), "Rows returned: N", "Rows shown:
 * K" when the limit cuts, "GPU execution time: X.XXX ms", with
 * extended_timing the code generation / compilation / library load /
 * materialization / engine lines, "--- End ---".  Two-call string protocol
 * as cq_db_explain. */
int cq_result_format(const cq_result* r, long long row_limit, int extended_timing, char* buf, size_t len, size_t* needed);
/* The accessors above in bulk, for bindings that read many results (no
 * reference counterpart: QueryResult's fields, execution.h:56-96, are read
 * in place by C++ callers).  cq_result_shape: rows, columns, ordered flag,
 * and per column its type and nullability (arrays of max_cols entries;
 * CQ_ERR_INVALID when the result has more columns).  cq_result_copy: every
 * column into values[c * rows + i] and its validity into valid[c * rows + i]
 * (either may be NULL).  cq_result_timings: out[0..7] = gpu_execute_ms,
 * scan_kernel_ms, compile_ms, codegen_ms, parse_ms, host_fetch_ms,
 * engine_ms, cache_hit. */
int cq_result_shape(const cq_result* r, long long* rows, int* cols, int* ordered, int* types, int* nullable, int max_cols);
int cq_result_copy(const cq_result* r, unsigned long long* values, unsigned char* valid);
int cq_result_timings(const cq_result* r, double* out);
void cq_result_free(cq_result* r);

/* ---- batched shared scan ------------------------------------------- */
/* Executes n queries over the resident columns; outs[i] receives query i's
 * result, equal to what cq_db_execute returns for it alone.  Queries are
 * grouped by plan signature and every group of up to 16 signatures over
 * one fact table shares ONE persistent scan launch that reads each fact
 * column once; the device work of a batch is captured in a HIP graph and
 * replayed for batches of the same shape.  No reference counterpart: the
 * reference runs QueryEngine::executeQuery per query
 * (src/core/execution.cpp:652-725).
 * options: NULL or "key=value,..." with bt (64..1024, multiple of 64,
 * default 1024), vpt (1..4 128-bit loads per column per lane, default 1),
 * graph (0/1), autotune (0/1: time every bt x vpt once per signature set and
 * keep the fastest), max_templates (<= 16), max_queries (per scan), shared
 * (0/1: one probe per template and dimension), lds_agg (0/1: per-query LDS
 * aggregation), lds_kb (16..160, default 160: LDS per scan block for the
 * per-query aggregation regions), masks (0/1: query-mask form), table_mask
 * (bit k: stage the k-th candidate shared table in LDS), table_kb (16..96,
 * default 32: LDS of the staged tables), grid_percent (1..100, default 100:
 * scan grid as a percentage of the blocks that fit on the device at once). */
int cq_batch_execute(cq_db* db, const char* const* sqls, int n, const char* options, cq_result** outs);
/* JSON of the last batch (per shared scan: queries, templates, columns,
 * launch shape, graph replay, scan / device milliseconds, fact bytes). */
int cq_batch_last_report(cq_db* db, char* buf, size_t len, size_t* needed);
/* Generated HIP source of a batch (host-only databases too): part < 0 the
 * shared-scan module of all n queries, part = k query k's template module. */
int cq_db_generate_batch_code(cq_db* db, const char* const* sqls, int n, const char* options, int part, char* buf, size_t len,
                              size_t* needed);

/* ---- serving -------------------------------------------------------- */
/* Continuous-batching server over the database's resident columns (no
 * reference counterpart).  Lanes (own stream each) take every queued
 * request they can admit, up to max_batch, and run them as one shared scan;
 * admission reserves each request's device footprint against the budget;
 * cache misses compile on a background thread.  options: NULL or
 * "key=value,..." with lanes (1..8, default 1), max_batch (128),
 * max_templates (16), promote_after (2: requests of a new signature before
 * it is compiled into the shared scan), compile_threads (2), budget_mb
 * (default: a quarter of free memory), window_us (0), async_compile (1),
 * interpret (1: new signatures run at once on the interpreted tier),
 * cold_lanes (1: lanes for not-yet-compiled signatures), cold_group (1:
 * interpreted requests a cold lane runs in one shared pass), cold_grid_percent
 * (100: cold scans' grid), per_template (1: one shared-scan launch per hot
 * signature in a batch; 0: one module for the whole hot set; 2: one launch
 * of the hot-set module while fewer than cold_backlog (2) cold requests are
 * queued, one launch per signature otherwise), max_hot (64),
 * bt, vpt, graph, lds_kb, grid_percent (as for cq_batch_execute). */
int cq_server_start(cq_db* db, const char* options, cq_server** out);
/* Enqueue a request; returns a ticket >= 0 (negative on error). */
long long cq_server_submit(cq_server* s, const char* sql);
/* Wait for a ticket (timeout_ms < 0: forever); *out owns the result. */
int cq_server_wait(cq_server* s, long long ticket, int timeout_ms, cq_result** out);
/* JSON: completed, p50_ms, p99_ms, max_ms, qps, batches, mean_batch,
 * admission_waits, compiles, rejected, lanes, budget_bytes, in_flight_bytes */
int cq_server_stats(cq_server* s, char* buf, size_t len, size_t* needed);
/* Restart the latency / throughput window. */
int cq_server_reset_stats(cq_server* s);
/* Stops the lanes; requests still queued fail with "server stopped". */
void cq_server_stop(cq_server* s);

#ifdef __cplusplus
}
#endif
#endif
