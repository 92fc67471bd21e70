// Query engine runtime: the compile / execute seam of the JIT path.
//
// Reference interfaces replaced (include/core/execution.h:284-346):
//   ICompiler::compile(source, hash) -> path of an acpp-built .so
//       -> HipRtcCompiler: hiprtcCompileProgram for gfx950, code object kept
//          in process and loaded with hipModuleLoadData (returns a handle)
//   IExecutor::execute(path, ctx) -> dlopen + execute_query(ctx)
//       -> HipModuleExecutor: binds columns / scratch to cq_query_params and
//          launches build, scan and materialization kernels on ctx's stream
//   IQueryCache keyed by std::hash(SQL text)
//       -> PlanSignatureCache keyed by the generated source (constants are
//          kernel arguments), thread-safe, LRU-bounded
//   ExecutionContext (execution.h:131-282) -> column registry, grow-only
//          named scratch (DynamicDeviceBuffer, device_buffer.h:12-118), the
//          query program and the host result
//   QueryEngine::executeQuery (execution.cpp:652-725) incl. the memory guard
//          preflightDeviceMemoryOrThrow / estimateJitTemporaryBytes (298-404)
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <list>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <shared_mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../codegen/hip_codegen.h"
#include "../device/query_params.h"
#include "../kernels/aot_kernels.h"
#include "../plan/memory.h"

namespace db {

void hipCheck(hipError_t e, const char* what);

// HIP calls that synchronize the whole device or change its memory map or
// loaded code objects (hipMalloc, hipFree, hipDeviceSynchronize,
// hipHostMalloc / hipHostFree, hipModuleLoadData / hipModuleUnload,
// hipStreamCreate / Destroy, graph destruction) synchronize implicitly with
// every stream, so one of them issued by any thread while another thread
// captures a HIP graph invalidates that capture.  Such calls hold this lock
// exclusively (DeviceMapLock); a capture holds it shared from begin to
// instantiation (CaptureLock).  A thread that captures must not make such a
// call: DeviceMapLock throws instead of deadlocking.
std::shared_mutex& deviceMapMutex();
bool threadIsCapturing();

class DeviceMapLock {
public:
    DeviceMapLock();
    ~DeviceMapLock();
    DeviceMapLock(const DeviceMapLock&) = delete;
    DeviceMapLock& operator=(const DeviceMapLock&) = delete;
};

class CaptureLock {
public:
    CaptureLock();
    ~CaptureLock();
    CaptureLock(const CaptureLock&) = delete;
    CaptureLock& operator=(const CaptureLock&) = delete;
};

// Grow-only device allocation.  Grows by 1.5x below 1 Mi elements and exactly
// above (device_buffer.h:59-75); frees after synchronizing its stream.
class DeviceBuffer {
public:
    DeviceBuffer() = default;
    ~DeviceBuffer();
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    void ensure(size_t bytes);
    void release();
    void* data() const { return ptr_; }
    size_t capacity() const { return bytes_; }

private:
    void* ptr_ = nullptr;
    size_t bytes_ = 0;
};

struct QueryTimingStats {
    double parse_ms = 0, codegen_ms = 0, compile_ms = 0, library_load_ms = 0;
    double gpu_execute_ms = 0;   // device work of the query (events on the stream)
    double scan_kernel_ms = 0;   // cq_scan alone, when kernel timing is enabled
    double host_fetch_ms = 0, engine_ms = 0;
    bool cache_hit = false;
};

struct ResultColumnDesc {
    LogicalType type = LogicalType::UInt64;
    bool nullable = false;
    std::string name;
};

// Host result: columnar 64-bit cells (int64 / uint64 / double bits) + per-row
// validity (1 = valid).  `ordered` is false for the sparse output path whose
// row order the reference leaves unspecified (database_instance.cpp:299).
struct QueryResult {
    std::vector<ResultColumnDesc> columns;
    size_t row_count = 0;
    bool ordered = true;
    bool dense_result = true;
    std::vector<std::vector<uint64_t>> data;      // [column][row]
    std::vector<std::vector<uint8_t>> validity;   // [column][row]
    QueryTimingStats timing;
};

struct EngineConfig {
    bool memory_guard_enabled = true;
    size_t memory_guard_reserve_bytes = 512ull << 20;
    double memory_guard_reserve_fraction = 0.10;
    bool kernel_timing = false;          // record cq_scan duration with events
    bool vm_row_kernel = false;          // interpreted tier: always the one-row-per-thread scan
    bool vm_linear = true;               // interpreted tier: the linear-form kernel for plans that have one
    // interpreted tier: the linear-form scans' grid as a percentage of a
    // full-device persistent grid (below 100 the scan leaves CUs to
    // concurrent work, such as a server's shared scans)
    int vm_grid_percent = 100;
    long long host_sort_rows = 4096;     // ORDER BY results up to this size are sorted on the host
    bool prepared_cache = false;         // DatabaseInstance::prepare goes through QueryEngine::prepareCached
    CodegenOptions codegen;
    size_t module_cache_capacity = 256;  // compiled modules kept (LRU)
};

// Grow-only pinned host allocation (hipHostMalloc), for asynchronous
// device-to-host copies.
class PinnedBuffer {
public:
    PinnedBuffer() = default;
    ~PinnedBuffer();
    PinnedBuffer(const PinnedBuffer&) = delete;
    PinnedBuffer& operator=(const PinnedBuffer&) = delete;
    void* ensure(size_t bytes);

private:
    void* ptr_ = nullptr;
    size_t bytes_ = 0;
};

struct DeviceColumn {
    const int* data = nullptr;
    const uint64_t* validity = nullptr;  // LSB-first u64 words, null = all valid
    int64_t rows = 0;
};

class CompiledModule {
public:
    CompiledModule(std::string signature, std::vector<char> code);
    ~CompiledModule();
    const std::string& signature() const { return signature_; }
    size_t codeSize() const { return code_.size(); }
    const std::vector<char>& code() const { return code_; }
    // Loads the module on the current device on first use (thread-safe).
    hipFunction_t function(const std::string& name);
    // Loads the code object now (hipModuleLoadData) / whether it is loaded.
    void load();
    bool loaded();

private:
    std::string signature_;
    std::vector<char> code_;
    std::mutex mu_;
    hipModule_t module_ = nullptr;
    std::map<std::string, hipFunction_t> functions_;
};
using ModuleHandle = std::shared_ptr<CompiledModule>;

struct ExecutionContext {
    ExecutionContext() = default;
    ExecutionContext(const ExecutionContext&) = delete;
    ExecutionContext& operator=(const ExecutionContext&) = delete;
    ~ExecutionContext() {
        scratch_.clear();
        if (stream && owns_stream) {
            DeviceMapLock lock;
            (void)hipStreamDestroy(stream);
        }
    }
    int device = 0;
    hipStream_t stream = nullptr;
    bool owns_stream = false;
    EngineConfig config;
    const std::unordered_map<std::string, DeviceColumn>* columns = nullptr;
    size_t loaded_device_bytes = 0;

    // per query
    QueryProgram program;
    QueryResult result;
    QueryTimingStats timing;
    int64_t row_begin = 0, row_end = -1;  // fact rows to scan (-1: all)

    DeviceBuffer& scratch(const std::string& name) { return scratch_[name]; }
    // pinned host staging of result fetches
    PinnedBuffer& pinned() { return pinned_; }
    size_t scratchBytes() const {
        size_t b = 0;
        for (const auto& kv : scratch_) b += kv.second.capacity();
        return b;
    }
    void resetForQuery() {
        result = QueryResult{};
        timing = QueryTimingStats{};
        row_begin = 0;
        row_end = -1;
    }

private:
    std::map<std::string, DeviceBuffer> scratch_;
    PinnedBuffer pinned_;
};

class ICompiler {
public:
    virtual ~ICompiler() = default;
    virtual ModuleHandle compile(const std::string& source_code, const std::string& query_hash) = 0;
};

class IExecutor {
public:
    virtual ~IExecutor() = default;
    virtual void execute(const ModuleHandle& module, ExecutionContext* ctx) = 0;
};

class IQueryCache {
public:
    virtual ~IQueryCache() = default;
    virtual ModuleHandle get(const std::string& key) = 0;
    virtual void put(const std::string& key, ModuleHandle module) = 0;
    virtual size_t size() const = 0;
    virtual void clear() = 0;
};

// hipRTC for --offload-arch=gfx950 with the device primitive library as an
// in-memory header.  Throws std::runtime_error with the compiler log on failure.
class HipRtcCompiler : public ICompiler {
public:
    explicit HipRtcCompiler(std::string arch = "gfx950", std::vector<std::string> extra_options = {});
    ModuleHandle compile(const std::string& source_code, const std::string& query_hash) override;
    size_t compileCount() const { return compiles_.load(); }

private:
    std::string arch_;
    std::vector<std::string> options_;
    std::atomic<size_t> compiles_{0};
};

// Thread-safe LRU map from plan signature to compiled module.
class PlanSignatureCache : public IQueryCache {
public:
    explicit PlanSignatureCache(size_t capacity = 256) : capacity_(capacity) {}
    ModuleHandle get(const std::string& key) override;
    void put(const std::string& key, ModuleHandle module) override;
    size_t size() const override;
    void clear() override;
    size_t hits() const { return hits_; }
    size_t misses() const { return misses_; }
    size_t evictions() const { return evictions_; }
    void setCapacity(size_t capacity);

private:
    void trimLocked(std::vector<ModuleHandle>& evicted);
    size_t capacity_;
    size_t evictions_ = 0;
    mutable std::mutex mu_;
    std::list<std::pair<std::string, ModuleHandle>> lru_;
    std::unordered_map<std::string, std::list<std::pair<std::string, ModuleHandle>>::iterator> index_;
    size_t hits_ = 0, misses_ = 0;
};

// Device state of one query in flight: its parameter block (columns, tables,
// aggregate slots, constants), module, bound scratch and host result.
// Scratch buffers are named under `prefix`, so several queries can be in
// flight in one context (the batched shared scan binds one run per query).
struct QueryRun {
    ModuleHandle module;
    const QueryProgram* program = nullptr;
    std::string prefix;
    cq_query_params P;
    std::vector<hipFunction_t> builds;
    std::vector<hipFunction_t> mht_writes;  // per join: cq_mhtw_<j> of a multi-value table, else null
    hipFunction_t dense_count = nullptr;
    hipFunction_t keys = nullptr;       // cq_keys (group keys decoded after the scan)
    cq_vm_program* vm = nullptr;        // interpreted tier: the program in device memory
    int64_t G = 1;
    size_t slots = 0;
    unsigned long long* res = nullptr;  // visible tuples (after AVG finalization)
    unsigned long long* rv = nullptr;
    bool single_cell = false;
    long long cap_rows = 1, vwords = 1;
    unsigned long long *o = nullptr, *ov = nullptr, *fl = nullptr, *off = nullptr, *bsums = nullptr;
    unsigned long long *sidx = nullptr, *o2 = nullptr, *ov2 = nullptr;  // ORDER BY buffers
    long long rows = 0;              // result rows after LIMIT (fetchLaunch)
    bool host_sort = false;          // ORDER BY of a small result: sorted (and limited) by fetchFinish
    char* fetch_host = nullptr;      // host staging of the result (fetchLaunch)
    QueryResult result;
};

// Sub-allocates the device scratch of many runs from one block (256-byte
// aligned).  With no base it only measures: bind() every run once to size
// the block, allocate it, then bind() again with the base set.
class ScratchArena {
public:
    void measure() { base_ = nullptr, capacity_ = 0, used_ = 0; }
    void assign(void* base, size_t capacity) { base_ = base, capacity_ = capacity, used_ = 0; }
    void* take(size_t bytes);
    size_t used() const { return used_; }

private:
    void* base_ = nullptr;
    size_t capacity_ = 0, used_ = 0;
};

// Executes a module in phases.  bind() does host work only (parameter block,
// scratch allocation, module load), so setup() / scan / shape() enqueue
// kernels and nothing else and can be captured into a HIP graph.
class HipModuleExecutor : public IExecutor {
public:
    void execute(const ModuleHandle& module, ExecutionContext* ctx) override;
    void bind(const ModuleHandle& module, const QueryProgram& program, ExecutionContext* ctx, const std::string& prefix,
              QueryRun& run, ScratchArena* arena);
    // zero tables / aggregate slots, build the dimension tables.  With
    // `device_params` the generated kernels read the block from device memory.
    void setup(QueryRun& run, ExecutionContext* ctx, const cq_query_params* device_params);
    // the build kernels of run's tables (setup() runs them unless the
    // program's tables are shared by a batch template)
    void buildTables(QueryRun& run, ExecutionContext* ctx, const cq_query_params* device_params);
    void scan(QueryRun& run, ExecutionContext* ctx);
    // AVG finalization, present/HAVING flags, compaction; the number of
    // result rows lands in device memory at `row_count` (default: after the offsets)
    void shape(QueryRun& run, ExecutionContext* ctx, const cq_query_params* device_params, unsigned long long* row_count);
    // projection: count pass, scan of the counts, result buffers for the
    // rows found, write pass, columnar rows (returns the row count)
    unsigned long long project(QueryRun& run, ExecutionContext* ctx, hipEvent_t kernel_start, hipEvent_t kernel_end);
    // sort / limit and copy the result to the host (synchronizes);
    // fetchPinned: the same through the context's pinned staging, the
    // result gathered into one range on the device and copied once
    void fetch(QueryRun& run, ExecutionContext* ctx, unsigned long long rows);
    void fetchPinned(QueryRun& run, ExecutionContext* ctx, unsigned long long rows);
    // the same in two steps, so many runs share one synchronization:
    // fetchLaunch enqueues sort / gather and the copies into `host`
    // (fetchBytes(run, rows) bytes, pinned for asynchronous copies),
    // fetchFinish decodes once the stream has reached that point
    size_t fetchBytes(const QueryRun& run, unsigned long long rows) const;
    void fetchLaunch(QueryRun& run, ExecutionContext* ctx, unsigned long long rows, void* host);
    // the same, as device ranges to copy (u64 words; destination words from
    // host_word on, in the host layout fetchFinish decodes at `host`)
    void fetchSegments(QueryRun& run, ExecutionContext* ctx, unsigned long long rows, char* host, unsigned long long host_word,
                       std::vector<cq_copy_seg>& segs);
    void fetchFinish(QueryRun& run);
    void sortHost(QueryRun& run);
    // execute() without device or pinned allocation: every scratch buffer
    // (hash tables, aggregate slots, materialization, sort) is carved out of
    // `scratch` (`bytes` long), the module must be loaded, no events.
    // Aggregate plans only (a projection's size is known after its scan).
    void executeInto(const ModuleHandle& module, ExecutionContext* ctx, void* scratch, size_t bytes);
    // Interpreted plans in their linear form over the same fact rows, in one
    // shared pass (k_lin_multi): each context's program is bound, set up,
    // shaped and fetched as execute() does, one scan serves all of them.
    // The contexts share one stream and device; at most CQ_LIN_BATCH plans
    // reading at most CQ_LIN_COLS distinct fact columns (linearGroupable).
    static bool linearGroupable(const QueryProgram& p);
    void executeLinearGroup(const std::vector<ExecutionContext*>& ctxs);
};

// Persistent grid for a scan kernel: occupancy x CUs, at most one block per tile.
unsigned scanGrid(hipFunction_t f, int block_threads, int device, long long num_tiles);

// Device bytes a query needs beyond the resident columns (hash tables,
// aggregate slots, materialization buffers).
size_t estimateQueryTemporaryBytes(const QueryProgram& program);

class QueryEngine {
public:
    QueryEngine(std::shared_ptr<IQueryCache> cache, std::shared_ptr<ICompiler> compiler, std::shared_ptr<IExecutor> executor,
                std::shared_ptr<const Catalog> catalog);
    // SQL -> optimized plan -> program -> (cached) module -> execution
    void executeQuery(const std::string& sql, ExecutionContext* ctx);
    // Interpreted statements run together: plans with a linear form over
    // the same fact table share passes of at most CQ_LIN_BATCH plans and
    // CQ_LIN_COLS distinct columns, the others run one by one.  The
    // contexts share one stream; ctxs[k] receives statement k's result.
    // Returns per statement its error message ("" when it ran).
    std::vector<std::string> executeInterpretedGroup(const std::vector<std::string>& sqls, const std::vector<ExecutionContext*>& ctxs);
    // plan -> program without executing (codegen text / signature / layout)
    QueryProgram prepare(const std::string& sql, const CodegenOptions& options) const;
    // prepare() for statements that repeat with other integer constants:
    // the SQL with its integer literals taken out is the key of a cached
    // program whose constants are re-bound (lits only; plan_sexpr is left
    // empty).  A statement is cached only when planning it with marker
    // constants shows that every constant reaches the program through
    // `lits` and nothing else depends on the values.
    QueryProgram prepareCached(const std::string& sql, const CodegenOptions& options);
    // compile or fetch the module for a program
    ModuleHandle moduleFor(const QueryProgram& program, QueryTimingStats* timing);
    ModuleHandle moduleForSource(const std::string& source, const std::string& signature, QueryTimingStats* timing);
    // the cached module of a program, or null (never compiles)
    ModuleHandle cachedModule(const QueryProgram& program) { return cache_->get(program.signature); }
    HipModuleExecutor* hipExecutor() { return dynamic_cast<HipModuleExecutor*>(executor_.get()); }
    // a new or changed catalog (tables, columns, statistics) drops every
    // cached program
    void setCatalog(std::shared_ptr<const Catalog> catalog) {
        catalog_ = std::move(catalog);
        invalidate();
    }
    void invalidate() {
        {
            std::lock_guard<std::mutex> a(prepared_mu_);
            prepared_.clear();
        }
        std::lock_guard<std::mutex> b(programs_mu_);
        programs_.clear();
    }
    IQueryCache& cache() { return *cache_; }

private:
    // the program of `sql` into ctx->program (statement cache, parse, code
    // generation), with the parse / codegen times
    void planQuery(const std::string& sql, ExecutionContext* ctx, QueryTimingStats& timing);
    std::shared_ptr<IQueryCache> cache_;
    std::shared_ptr<ICompiler> compiler_;
    std::shared_ptr<IExecutor> executor_;
    std::shared_ptr<const Catalog> catalog_;
    // one compilation per signature at a time; different signatures compile
    // concurrently (hipRTC programs are independent)
    std::mutex compile_mu_;
    std::condition_variable compile_cv_;
    std::set<std::string> compiling_;
    // prepareCached: normalized statement -> program and, per lits slot, the
    // index of the statement's literal it takes (-1: fixed value)
    struct PreparedShape {
        std::shared_ptr<const QueryProgram> program;
        std::vector<int> slot_literal;
        bool cacheable = false;
    };
    std::mutex prepared_mu_;
    std::unordered_map<std::string, PreparedShape> prepared_;
    // executeQuery: statement text (and code generation options) -> its
    // program, so a repeated statement skips parsing and code generation
    // (the reference skips code generation for a statement it has seen,
    // execution.cpp:692-699)
    std::mutex programs_mu_;
    std::unordered_map<std::string, std::shared_ptr<const QueryProgram>> programs_;
};

// Throws when the device cannot hold the query (preflightDeviceMemoryOrThrow).
void preflightDeviceMemoryOrThrow(const ExecutionContext& ctx, const QueryProgram& program);

inline double elapsedMs(std::chrono::steady_clock::time_point a, std::chrono::steady_clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

}  // namespace db
