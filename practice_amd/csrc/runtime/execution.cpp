// Query engine runtime implementation.
#include "execution.h"

#include <hip/hiprtc.h>

#include <algorithm>
#include <cctype>
#include <cstring>
#include <sstream>
#include <stdexcept>

#include "../frontend/sql.h"
#include "../kernels/aot_kernels.h"
#include "../kernels/vm_kernels.h"

extern const char* const kCrystalHipHeaderSource;
extern const char* const kQueryParamsHeaderSource;
extern const char* const kHtLayoutsHeaderSource;
extern const char* const kParamsLayoutHeaderSource;

namespace db {

void hipCheck(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// ------------------------------------------------------------------ device map lock

namespace {
thread_local bool t_capturing = false;
}

std::shared_mutex& deviceMapMutex() {
    static std::shared_mutex mu;
    return mu;
}

bool threadIsCapturing() { return t_capturing; }

DeviceMapLock::DeviceMapLock() {
    if (t_capturing)
        throw std::logic_error("device allocation, free or synchronization inside a HIP graph capture");
    deviceMapMutex().lock();
}

DeviceMapLock::~DeviceMapLock() { deviceMapMutex().unlock(); }

CaptureLock::CaptureLock() {
    if (t_capturing) throw std::logic_error("nested HIP graph capture");
    deviceMapMutex().lock_shared();
    t_capturing = true;
}

CaptureLock::~CaptureLock() {
    t_capturing = false;
    deviceMapMutex().unlock_shared();
}

// ------------------------------------------------------------------ DeviceBuffer

DeviceBuffer::~DeviceBuffer() {
    if (!ptr_) return;
    if (t_capturing) {  // unwinding out of a capture: leak rather than break it
        ptr_ = nullptr;
        return;
    }
    release();
}

void DeviceBuffer::release() {
    if (ptr_) {
        DeviceMapLock lock;
        (void)hipDeviceSynchronize();
        (void)hipFree(ptr_);
    }
    ptr_ = nullptr;
    bytes_ = 0;
}

void DeviceBuffer::ensure(size_t bytes) {
    if (bytes <= bytes_ && ptr_) return;
    size_t want = bytes;
    if (bytes < (8ull << 20)) want = std::max(bytes, bytes_ + bytes_ / 2);
    if (want == 0) want = 256;
    release();
    DeviceMapLock lock;
    hipCheck(hipMalloc(&ptr_, want), "hipMalloc scratch");
    bytes_ = want;
}

PinnedBuffer::~PinnedBuffer() {
    if (ptr_) {
        DeviceMapLock lock;
        (void)hipHostFree(ptr_);
    }
}

void* PinnedBuffer::ensure(size_t bytes) {
    if (bytes <= bytes_ && ptr_) return ptr_;
    DeviceMapLock lock;
    if (ptr_) hipCheck(hipHostFree(ptr_), "hipHostFree");
    ptr_ = nullptr;
    bytes_ = 0;
    const size_t want = std::max<size_t>(bytes, 64 * 1024);
    hipCheck(hipHostMalloc(&ptr_, want, hipHostMallocDefault), "hipHostMalloc");
    bytes_ = want;
    return ptr_;
}

// ------------------------------------------------------------------ CompiledModule

CompiledModule::CompiledModule(std::string signature, std::vector<char> code)
    : signature_(std::move(signature)), code_(std::move(code)) {}

CompiledModule::~CompiledModule() {
    if (!module_) return;
    if (threadIsCapturing()) return;  // last handle dropped while unwinding a capture: keep the code loaded
    DeviceMapLock lock;
    (void)hipModuleUnload(module_);
}

void CompiledModule::load() {
    std::lock_guard<std::mutex> lock(mu_);
    if (module_) return;
    DeviceMapLock dl;
    hipCheck(hipModuleLoadData(&module_, code_.data()), "hipModuleLoadData");
}

bool CompiledModule::loaded() {
    std::lock_guard<std::mutex> lock(mu_);
    return module_ != nullptr;
}

hipFunction_t CompiledModule::function(const std::string& name) {
    std::lock_guard<std::mutex> lock(mu_);
    if (!module_) {
        DeviceMapLock dl;
        hipCheck(hipModuleLoadData(&module_, code_.data()), "hipModuleLoadData");
    }
    auto it = functions_.find(name);
    if (it != functions_.end()) return it->second;
    hipFunction_t f = nullptr;
    hipCheck(hipModuleGetFunction(&f, module_, name.c_str()), ("hipModuleGetFunction " + name).c_str());
    functions_[name] = f;
    return f;
}

// ------------------------------------------------------------------ HipRtcCompiler

HipRtcCompiler::HipRtcCompiler(std::string arch, std::vector<std::string> extra)
    : arch_(std::move(arch)), options_(std::move(extra)) {}

ModuleHandle HipRtcCompiler::compile(const std::string& source, const std::string& query_hash) {
    hiprtcProgram prog = nullptr;
    const char* headers[] = {kCrystalHipHeaderSource, kQueryParamsHeaderSource, kParamsLayoutHeaderSource, kHtLayoutsHeaderSource};
    const char* names[] = {"crystal_hip.h", "query_params.h", "cq_query_params.h", "ht_layouts.h"};
    const std::string file = "q_" + query_hash + ".hip";
    if (hiprtcCreateProgram(&prog, source.c_str(), file.c_str(), 4, headers, names) != HIPRTC_SUCCESS)
        throw std::runtime_error("hiprtcCreateProgram failed");
    std::vector<std::string> opts = {"--offload-arch=" + arch_, "-O3", "-std=c++17", "-fwrapv",
                                     "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
    for (const auto& o : options_) opts.push_back(o);
    std::vector<const char*> copts;
    for (const auto& o : opts) copts.push_back(o.c_str());
    const hiprtcResult rc = hiprtcCompileProgram(prog, static_cast<int>(copts.size()), copts.data());
    size_t log_size = 0;
    hiprtcGetProgramLogSize(prog, &log_size);
    std::string log(log_size, '\0');
    if (log_size) hiprtcGetProgramLog(prog, &log[0]);
    if (rc != HIPRTC_SUCCESS) {
        hiprtcDestroyProgram(&prog);
        throw std::runtime_error("hipRTC compilation failed for " + file + ": " + log);
    }
    size_t code_size = 0;
    hiprtcGetCodeSize(prog, &code_size);
    std::vector<char> code(code_size);
    hiprtcGetCode(prog, code.data());
    hiprtcDestroyProgram(&prog);
    ++compiles_;
    return std::make_shared<CompiledModule>(query_hash, std::move(code));
}

// ------------------------------------------------------------------ PlanSignatureCache

ModuleHandle PlanSignatureCache::get(const std::string& key) {
    std::lock_guard<std::mutex> lock(mu_);
    auto it = index_.find(key);
    if (it == index_.end()) {
        ++misses_;
        return nullptr;
    }
    lru_.splice(lru_.begin(), lru_, it->second);
    ++hits_;
    return it->second->second;
}

void PlanSignatureCache::put(const std::string& key, ModuleHandle module) {
    std::vector<ModuleHandle> evicted;  // unloaded (if last held) after the cache lock is dropped
    {
        std::lock_guard<std::mutex> lock(mu_);
        auto it = index_.find(key);
        if (it != index_.end()) {
            evicted.push_back(std::move(it->second->second));
            it->second->second = std::move(module);
            lru_.splice(lru_.begin(), lru_, it->second);
        } else {
            lru_.emplace_front(key, std::move(module));
            index_[key] = lru_.begin();
            trimLocked(evicted);
        }
    }
}

void PlanSignatureCache::trimLocked(std::vector<ModuleHandle>& evicted) {
    while (lru_.size() > capacity_) {
        evicted.push_back(std::move(lru_.back().second));
        index_.erase(lru_.back().first);
        lru_.pop_back();
        ++evictions_;
    }
}

void PlanSignatureCache::setCapacity(size_t capacity) {
    std::vector<ModuleHandle> evicted;
    std::lock_guard<std::mutex> lock(mu_);
    capacity_ = std::max<size_t>(1, capacity);
    trimLocked(evicted);
}

size_t PlanSignatureCache::size() const {
    std::lock_guard<std::mutex> lock(mu_);
    return lru_.size();
}

void PlanSignatureCache::clear() {
    std::list<std::pair<std::string, ModuleHandle>> dropped;
    std::lock_guard<std::mutex> lock(mu_);
    dropped.swap(lru_);
    index_.clear();
}

// ------------------------------------------------------------------ memory guard

size_t estimateQueryTemporaryBytes(const QueryProgram& p) {
    size_t bytes = 0;
    for (const auto& j : p.joins) bytes += j.ht_words * 4;
    if (p.projection) {
        // per-(tile, vector, wave) counts and offsets; the rows are checked after the count
        const int64_t tile = static_cast<int64_t>(std::max(p.block_threads, 64)) * 4 * std::max(p.vectors_per_thread, 1);
        const size_t n = static_cast<size_t>((p.fact_rows + tile - 1) / tile) * static_cast<size_t>(std::max(p.vectors_per_thread, 1)) *
                         static_cast<size_t>(std::max(p.proj_waves, 1));
        return bytes + 2 * n * 8 + 64;
    }
    const size_t slots = static_cast<size_t>(p.groups) * static_cast<size_t>(p.storage_ts);
    bytes += slots * 8 + ((slots + 63) / 64) * 8;                              // aggregate state
    const size_t vslots = static_cast<size_t>(p.groups) * static_cast<size_t>(p.visible_ts);
    if (p.has_avg) bytes += vslots * 8 + ((vslots + 63) / 64) * 8;            // visible copy
    const size_t rows = static_cast<size_t>(p.groups);
    const size_t words = (rows + 63) / 64;
    bytes += 2 * (static_cast<size_t>(p.visible_ts) * (rows * 8 + words * 8));  // output + sort buffers
    bytes += 3 * rows * 8;                                                      // flags, offsets, indices
    return bytes;
}

void preflightDeviceMemoryOrThrow(const ExecutionContext& ctx, const QueryProgram& p) {
    if (!ctx.config.memory_guard_enabled) return;
    size_t free_b = 0, total = 0;
    if (hipMemGetInfo(&free_b, &total) != hipSuccess || total == 0) return;
    const size_t temp = estimateQueryTemporaryBytes(p);
    const size_t existing = ctx.scratchBytes();
    const size_t required = ctx.loaded_device_bytes + std::max(existing, temp);
    const size_t reserve = std::max(static_cast<size_t>(static_cast<double>(total) * ctx.config.memory_guard_reserve_fraction),
                                    ctx.config.memory_guard_reserve_bytes);
    const size_t budget = total > reserve ? total - reserve : total / 2;
    if (required > budget) {
        std::ostringstream m;
        m << "Insufficient GPU memory for query before JIT execution. Estimated resident requirement: " << required
          << " bytes (loaded columns: " << ctx.loaded_device_bytes << ", query temporaries: " << temp
          << "). Device memory: " << total << ", safety budget: " << budget << ".";
        throw std::runtime_error(m.str());
    }
}

// ------------------------------------------------------------------ HipModuleExecutor

namespace {

int deviceCUs(int device) {
    static int cached[64] = {0};
    if (device >= 0 && device < 64 && cached[device]) return cached[device];
    hipDeviceProp_t prop;
    hipCheck(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties");
    if (device >= 0 && device < 64) cached[device] = prop.multiProcessorCount;
    return prop.multiProcessorCount;
}

void launch(hipFunction_t f, unsigned grid, unsigned block, hipStream_t s, void** args, const char* what) {
    hipCheck(hipModuleLaunchKernel(f, grid, 1, 1, block, 1, 1, 0, s, args, nullptr), what);
}

int aggKind(const std::string& f) {
    if (f == "COUNT") return CQ_AGG_COUNT;
    if (f == "MIN") return CQ_AGG_MIN;
    if (f == "MAX") return CQ_AGG_MAX;
    if (f == "AVG") return CQ_AGG_AVG;
    return CQ_AGG_SUM;
}

int typeCode(LogicalType t) {
    if (t == LogicalType::Int64) return CQ_TYPE_INT64;
    if (t == LogicalType::Float64) return CQ_TYPE_FLOAT64;
    return CQ_TYPE_UINT64;
}

}  // namespace

void* ScratchArena::take(size_t bytes) {
    const size_t off = (used_ + 255) & ~static_cast<size_t>(255);
    used_ = off + std::max<size_t>(bytes, 8);
    if (!base_) return nullptr;  // measuring pass
    if (used_ > capacity_) throw std::runtime_error("scratch arena overflow");
    return static_cast<char*>(base_) + off;
}

void HipModuleExecutor::bind(const ModuleHandle& module, const QueryProgram& p, ExecutionContext* ctx,
                             const std::string& prefix, QueryRun& run, ScratchArena* arena) {
    hipCheck(hipSetDevice(ctx->device), "hipSetDevice");
    if (!ctx->columns) throw std::runtime_error("ExecutionContext has no column registry");
    run.module = module;
    run.program = &p;
    run.prefix = prefix;
    cq_query_params& P = run.P;
    std::memset(&P, 0, sizeof P);
    auto column = [&](const std::string& name) -> const DeviceColumn& {
        auto it = ctx->columns->find(name);
        if (it == ctx->columns->end()) throw std::runtime_error("column not loaded on the device: " + name);
        return it->second;
    };
    for (size_t i = 0; i < p.fact_cols.size(); ++i) {
        const DeviceColumn& c = column(p.fact_cols[i]);
        P.col[i] = c.data;
        P.valid[i] = reinterpret_cast<const unsigned int*>(c.validity);
        if (p.fact_nullable[i] && !c.validity) throw std::runtime_error("nullable column without bitmap: " + p.fact_cols[i]);
    }
    for (size_t i = 0; i < p.dim_cols.size(); ++i) {
        const DeviceColumn& c = column(p.dim_cols[i]);
        P.dcol[i] = c.data;
        P.dvalid[i] = reinterpret_cast<const unsigned int*>(c.validity);
    }
    for (size_t i = 0; i < p.lits.size(); ++i) P.lit[i] = p.lits[i];
    for (size_t i = 0; i < p.flits.size(); ++i) P.flit[i] = p.flits[i];

    // device scratch: the batch's arena when given, else named buffers of ctx
    auto take = [&](const std::string& name, size_t bytes) -> void* {
        if (arena) return arena->take(bytes);
        DeviceBuffer& b = ctx->scratch(prefix + name);
        b.ensure(std::max<size_t>(bytes, 8));
        return b.data();
    };
    auto take64 = [&](const std::string& name, size_t bytes) { return static_cast<unsigned long long*>(take(name, bytes)); };

    run.builds.clear();
    run.mht_writes.clear();
    for (size_t j = 0; j < p.joins.size(); ++j) {
        // shared (template) tables are owned and built by the batch
        P.ht[j] = p.shared_tables ? nullptr
                                  : static_cast<unsigned int*>(take("ht" + std::to_string(j), std::max<size_t>(p.joins[j].ht_words * 4, 16)));
        P.dim_rows[j] = p.joins[j].dim_rows;
        if (!p.vm) run.builds.push_back(module->function("cq_build_" + std::to_string(j)));
        run.mht_writes.push_back(p.joins[j].mht ? module->function("cq_mhtw_" + std::to_string(j)) : nullptr);
    }
    run.dense_count = (p.fast_sparse || p.vm || p.projection) ? nullptr : module->function("cq_dense_count");
    run.keys = p.keys_kernel ? module->function("cq_keys") : nullptr;
    run.vm = p.vm ? static_cast<cq_vm_program*>(take("vm", sizeof(cq_vm_program))) : nullptr;

    // aggregate state
    const int64_t G = p.groups;
    run.G = G;
    const size_t slots = static_cast<size_t>(G) * static_cast<size_t>(p.storage_ts);
    run.slots = slots;
    P.result = take64("result", slots * 8);
    P.result_valid = take64("result_valid", ((slots + 63) / 64) * 8);
    run.res = P.result;
    run.rv = P.result_valid;
    if (p.has_avg) {
        const size_t vslots = static_cast<size_t>(G) * static_cast<size_t>(p.visible_ts);
        run.res = take64("vresult", vslots * 8);
        run.rv = take64("vresult_valid", ((vslots + 63) / 64) * 8);
    }

    // materialization buffers
    const int vts = p.visible_ts;
    run.single_cell = p.fast_sparse && vts <= 1;
    run.cap_rows = std::max<long long>(G, 1);
    run.vwords = (run.cap_rows + 63) / 64;
    run.o = take64("out", static_cast<size_t>(vts) * run.cap_rows * 8);
    run.ov = take64("out_valid", static_cast<size_t>(vts) * run.vwords * 8);
    run.fl = take64("flags", run.cap_rows * 8);
    run.off = take64("offsets", run.cap_rows * 8 + 8);
    run.bsums = take64("block_sums", static_cast<size_t>(cq_scan_block_count(run.cap_rows) + 1) * 8);
    run.sidx = run.o2 = run.ov2 = nullptr;
    if (p.has_sort_node && !p.sort_keys.empty()) {
        run.sidx = take64("sort_idx", static_cast<size_t>(cq_sort_padded(run.cap_rows)) * 8);
        run.o2 = take64("out2", static_cast<size_t>(vts) * run.cap_rows * 8);
        run.ov2 = take64("out2_valid", static_cast<size_t>(vts) * run.vwords * 8);
    }

    const int64_t row_end = ctx->row_end < 0 ? p.fact_rows : std::min<int64_t>(ctx->row_end, p.fact_rows);
    if (ctx->row_begin < 0 || ctx->row_begin % 4 != 0) throw std::runtime_error("scan range must start at a multiple of 4 rows");
    P.row_begin = ctx->row_begin;
    P.row_end = std::max<int64_t>(row_end, ctx->row_begin);
    const int64_t tile = static_cast<int64_t>(p.block_threads) * 4 * p.vectors_per_thread;
    P.num_tiles = (P.row_end - P.row_begin + tile - 1) / tile;
}

void HipModuleExecutor::setup(QueryRun& run, ExecutionContext* ctx, const cq_query_params* device_params) {
    const QueryProgram& p = *run.program;
    cq_query_params& P = run.P;
    hipStream_t s = ctx->stream;
    if (!p.shared_tables)
        for (size_t j = 0; j < p.joins.size(); ++j)
            hipCheck(cq_launch_fill_u32(P.ht[j], 0u, static_cast<long long>(p.joins[j].ht_words), s), "fill hash table");
    const size_t words = (run.slots + 63) / 64;
    hipCheck(cq_launch_fill_u64(P.result, 0ull, static_cast<long long>(run.slots), s), "fill result");
    // scalar COUNT of a non-empty scan is valid even when no row qualifies
    unsigned long long count_bits = 0;
    if (!p.grouped && P.row_end > P.row_begin)
        for (int slot : p.count_slots_scalar)
            if (slot < 64) count_bits |= 1ull << slot;
    if (words > 0) hipCheck(cq_launch_fill_u64(P.result_valid, count_bits, 1, s), "fill result validity");
    if (words > 1) hipCheck(cq_launch_fill_u64(P.result_valid + 1, 0ull, static_cast<long long>(words - 1), s), "fill");
    unsigned int min_mask = 0;
    for (int slot : p.min_slots) min_mask |= 1u << slot;
    if (min_mask) hipCheck(cq_launch_init_min(P.result, run.G, p.storage_ts, min_mask, s), "init MIN slots");
    if (p.vm) {
        // the interpreted tier: program upload, then the generic build kernel
        hipCheck(hipMemcpyAsync(run.vm, p.vm.get(), sizeof(cq_vm_program), hipMemcpyHostToDevice, s), "upload program");
        for (size_t j = 0; j < p.joins.size(); ++j)
            hipCheck(cq_vm_launch_build(&P, run.vm, static_cast<int>(j), p.joins[j].dim_rows, s), "interpreted build");
        return;
    }
    if (!p.shared_tables) buildTables(run, ctx, device_params);
}

void HipModuleExecutor::buildTables(QueryRun& run, ExecutionContext* ctx, const cq_query_params* device_params) {
    const QueryProgram& p = *run.program;
    for (size_t j = 0; j < p.joins.size(); ++j) {
        const JoinProgram& jp = p.joins[j];
        const long long rows = jp.dim_rows;
        const unsigned grid = static_cast<unsigned>(std::min<long long>(std::max<long long>((rows + 255) / 256, 1), 8192));
        const cq_query_params* dp = device_params;
        void* by_value[] = {&run.P};
        void* by_pointer[] = {&dp};
        launch(run.builds[j], grid, 256, ctx->stream, device_params ? by_pointer : by_value, "launch build");
        if (!jp.mht) continue;
        // multi-value table: first positions = exclusive scan of the slot
        // counts, then the row ids and directory entries
        auto* T = reinterpret_cast<unsigned long long*>(run.P.ht[j]);
        unsigned long long* cnt = T + jp.mht_cnt / 2;
        unsigned long long* off = T + jp.mht_off / 2;
        unsigned long long* sums = T + jp.mht_sums / 2;
        const long long nb = cq_scan_block_count(jp.len);
        hipCheck(cq_launch_exclusive_scan(cnt, off, jp.len, sums, sums + nb + 1, ctx->stream), "scan multi-value counts");
        launch(run.mht_writes[j], grid, 256, ctx->stream, device_params ? by_pointer : by_value, "launch multi-value write");
    }
}

unsigned scanGrid(hipFunction_t f, int block_threads, int device, long long num_tiles) {
    int per_cu = 0;
    if (hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, block_threads, 0) != hipSuccess || per_cu < 1) per_cu = 1;
    const int64_t max_blocks = static_cast<int64_t>(per_cu) * deviceCUs(device);
    return static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>(num_tiles, max_blocks)));
}

void HipModuleExecutor::scan(QueryRun& run, ExecutionContext* ctx) {
    const QueryProgram& p = *run.program;
    if (p.vm) {
        const long long rows = run.P.row_end - run.P.row_begin;
        if (rows <= 0) return;
        if (!ctx->config.vm_row_kernel && ctx->config.vm_linear && cq_lin_supported(p.vm.get()) &&
            run.P.row_begin % cq_lin_row_align() == 0) {
            hipCheck(cq_lin_launch_scan(&run.P, run.vm, p.vm.get(), static_cast<unsigned>(p.vm_lds_bytes), ctx->device,
                                        ctx->config.vm_grid_percent, ctx->stream),
                     "interpreted scan (linear form)");
            return;
        }
        if (!ctx->config.vm_row_kernel && cq_vm_scan4_supported(p.vm.get()) && run.P.row_begin % 4 == 0) {
            hipCheck(cq_vm_launch_scan4(&run.P, run.vm, p.vm->nregs, static_cast<unsigned>(p.vm_lds_bytes), ctx->device, ctx->stream),
                     "interpreted scan");
            return;
        }
        const long long bt = cq_vm_block_threads();
        const unsigned grid = static_cast<unsigned>(std::min<long long>((rows + bt - 1) / bt, static_cast<long long>(deviceCUs(ctx->device)) * 8));
        hipCheck(cq_vm_launch_scan(&run.P, run.vm, static_cast<unsigned>(p.vm_lds_bytes), grid, ctx->stream), "interpreted scan");
        return;
    }
    hipFunction_t f = run.module->function("cq_scan");
    const unsigned grid = scanGrid(f, p.block_threads, ctx->device, run.P.num_tiles);
    void* args[] = {&run.P};
    if (run.P.num_tiles > 0) launch(f, grid, static_cast<unsigned>(p.block_threads), ctx->stream, args, "launch scan");
}

void HipModuleExecutor::shape(QueryRun& run, ExecutionContext* ctx, const cq_query_params* device_params,
                              unsigned long long* row_count) {
    const QueryProgram& p = *run.program;
    cq_query_params& P = run.P;
    hipStream_t s = ctx->stream;
    const int64_t G = run.G;
    const int vts = p.visible_ts;
    if (run.keys) {
        const cq_query_params* dp = device_params;
        void* by_value[] = {&P};
        void* by_pointer[] = {&dp};
        launch(run.keys, static_cast<unsigned>((G + 255) / 256), 256, s, device_params ? by_pointer : by_value, "launch keys");
    }
    if (p.has_avg) {
        const size_t vslots = static_cast<size_t>(G) * static_cast<size_t>(vts);
        hipCheck(cq_launch_fill_u64(run.res, 0ull, static_cast<long long>(vslots), s), "fill");
        hipCheck(cq_launch_fill_u64(run.rv, 0ull, static_cast<long long>((vslots + 63) / 64), s), "fill");
        cq_finalize_desc d;
        std::memset(&d, 0, sizeof d);
        d.ngroups = p.ngroups;
        d.naggs = p.naggs;
        d.sts = p.storage_ts;
        d.vts = vts;
        for (int i = 0; i < p.naggs && i < 32; ++i) {
            d.kind[i] = aggKind(p.agg_funcs[static_cast<size_t>(i)]);
            d.hidden[i] = p.hidden_slot[static_cast<size_t>(i)];
        }
        hipCheck(cq_launch_finalize(P.result, P.result_valid, G, &d, run.res, run.rv, s), "finalize AVG");
    }
    if (run.single_cell) return;
    hipCheck(cq_launch_fill_u64(run.ov, 0ull, static_cast<long long>(vts) * run.vwords, s), "fill out validity");
    if (p.fast_sparse) {
        hipCheck(cq_launch_sparse_count(run.res, G, vts, run.fl, s), "sparse count");
    } else {
        long long g = G;
        const unsigned long long* R = run.res;
        const unsigned long long* RV = run.rv;
        unsigned long long* counts = run.fl;
        const cq_query_params* dp = device_params;
        void* by_value[] = {&P, &R, &RV, &counts, &g};
        void* by_pointer[] = {&dp, &R, &RV, &counts, &g};
        if (p.vm) hipCheck(cq_vm_launch_dense_count(&P, run.vm, R, RV, counts, g, s), "interpreted dense count");
        else
            launch(run.dense_count, static_cast<unsigned>((G + 255) / 256), 256, s, device_params ? by_pointer : by_value,
                   "launch dense count");
    }
    hipCheck(cq_launch_exclusive_scan(run.fl, run.off, G, run.bsums, row_count ? row_count : run.off + run.cap_rows, s), "scan");
    unsigned int nmask = 0;
    for (int c = 0; c < vts; ++c)
        if (p.out[static_cast<size_t>(c)].nullable) nmask |= 1u << c;
    hipCheck(cq_launch_write_rows(run.res, run.rv, G, vts, run.fl, run.off, nmask, p.fast_sparse ? 0 : 1, run.o, run.ov,
                                  run.cap_rows, run.vwords, s),
             "write rows");
}

unsigned long long HipModuleExecutor::project(QueryRun& run, ExecutionContext* ctx, hipEvent_t kernel_start, hipEvent_t kernel_end) {
    const QueryProgram& p = *run.program;
    cq_query_params& P = run.P;
    hipStream_t s = ctx->stream;
    auto take64 = [&](const std::string& name, size_t bytes) {
        DeviceBuffer& b = ctx->scratch(name);
        b.ensure(std::max<size_t>(bytes, 8));
        return static_cast<unsigned long long*>(b.data());
    };
    // one count per (tile, vector of 4 rows per lane, wave) in row order
    const long long n = std::max<long long>(P.num_tiles * p.vectors_per_thread * p.proj_waves, 1);
    P.tile_count = take64("proj_counts", static_cast<size_t>(n) * 8);
    P.tile_off = take64("proj_offsets", static_cast<size_t>(n) * 8 + 8);
    unsigned long long* bs = take64("proj_block_sums", static_cast<size_t>(cq_scan_block_count(n) + 1) * 8);
    hipCheck(cq_launch_fill_u64(P.tile_count, 0ull, n, s), "fill projection counts");
    hipFunction_t fc = run.module->function("cq_pcount");
    hipFunction_t fw = run.module->function("cq_pwrite");
    const unsigned grid = scanGrid(fc, p.block_threads, ctx->device, P.num_tiles);
    void* args[] = {&P};
    if (kernel_start) hipCheck(hipEventRecord(kernel_start, s), "hipEventRecord");
    if (P.num_tiles > 0) launch(fc, grid, static_cast<unsigned>(p.block_threads), s, args, "launch projection count");
    hipCheck(cq_launch_exclusive_scan(P.tile_count, P.tile_off, n, bs, P.tile_off + n, s), "scan projection counts");
    unsigned long long rows = 0;
    hipCheck(hipMemcpyAsync(&rows, P.tile_off + n, 8, hipMemcpyDeviceToHost, s), "copy projection rows");
    hipCheck(hipStreamSynchronize(s), "sync");
    // the exact result must fit beside the resident data (the reference's
    // check after its count pass)
    const int vts = p.visible_ts;
    const size_t words = static_cast<size_t>((rows + 63) / 64);
    const size_t need = static_cast<size_t>(rows) * static_cast<size_t>(vts) * 8 * 3 + words * 8 * static_cast<size_t>(vts) * 3 +
                        static_cast<size_t>(rows) * 8 * 3;
    if (ctx->config.memory_guard_enabled) {
        size_t free_b = 0, total = 0;
        if (hipMemGetInfo(&free_b, &total) == hipSuccess && need > free_b)
            throw std::runtime_error("Insufficient GPU memory after projection count: exact materialized result does not fit device memory");
    }
    // result rows in the aggregate-state layout, then the shared shaping path
    run.G = static_cast<int64_t>(rows);
    run.slots = static_cast<size_t>(rows) * static_cast<size_t>(p.storage_ts);
    P.result = take64("result", run.slots * 8);
    P.result_valid = take64("result_valid", ((run.slots + 63) / 64) * 8);
    run.res = P.result;
    run.rv = P.result_valid;
    run.single_cell = false;
    run.cap_rows = std::max<long long>(static_cast<long long>(rows), 1);
    run.vwords = (run.cap_rows + 63) / 64;
    run.o = take64("out", static_cast<size_t>(vts) * run.cap_rows * 8);
    run.ov = take64("out_valid", static_cast<size_t>(vts) * run.vwords * 8);
    run.fl = take64("flags", run.cap_rows * 8);
    run.off = take64("offsets", run.cap_rows * 8 + 8);
    run.bsums = take64("block_sums", static_cast<size_t>(cq_scan_block_count(run.cap_rows) + 1) * 8);
    if (p.has_sort_node && !p.sort_keys.empty()) {
        run.sidx = take64("sort_idx", static_cast<size_t>(cq_sort_padded(run.cap_rows)) * 8);
        run.o2 = take64("out2", static_cast<size_t>(vts) * run.cap_rows * 8);
        run.ov2 = take64("out2_valid", static_cast<size_t>(vts) * run.vwords * 8);
    }
    hipCheck(cq_launch_fill_u64(P.result_valid, 0ull, static_cast<long long>((run.slots + 63) / 64), s), "fill result validity");
    if (P.num_tiles > 0 && rows > 0) launch(fw, grid, static_cast<unsigned>(p.block_threads), s, args, "launch projection write");
    if (kernel_end) hipCheck(hipEventRecord(kernel_end, s), "hipEventRecord");
    // every row is present: offsets are the row indices
    hipCheck(cq_launch_fill_u64(run.ov, 0ull, static_cast<long long>(vts) * run.vwords, s), "fill out validity");
    if (rows > 0) {
        hipCheck(cq_launch_fill_u64(run.fl, 1ull, static_cast<long long>(rows), s), "fill projection flags");
        hipCheck(cq_launch_exclusive_scan(run.fl, run.off, static_cast<long long>(rows), run.bsums, run.off + run.cap_rows, s), "scan");
    }
    unsigned int nmask = 0;
    for (int c = 0; c < vts; ++c)
        if (p.out[static_cast<size_t>(c)].nullable) nmask |= 1u << c;
    if (rows > 0)
        hipCheck(cq_launch_write_rows(run.res, run.rv, static_cast<long long>(rows), vts, run.fl, run.off, nmask, 1, run.o, run.ov,
                                      run.cap_rows, run.vwords, s),
                 "write rows");
    return rows;
}

size_t HipModuleExecutor::fetchBytes(const QueryRun& run, unsigned long long rows) const {
    const QueryProgram& p = *run.program;
    if (run.single_cell) return 8;
    const size_t r = static_cast<size_t>(std::min<unsigned long long>(rows, static_cast<unsigned long long>(run.cap_rows)));
    return static_cast<size_t>(p.visible_ts) * (r * 8 + static_cast<size_t>(run.vwords) * 8);
}

void HipModuleExecutor::fetchSegments(QueryRun& run, ExecutionContext* ctx, unsigned long long rows_found, char* host,
                                      unsigned long long host_word, std::vector<cq_copy_seg>& segs) {
    const QueryProgram& p = *run.program;
    hipStream_t s = ctx->stream;
    const int vts = p.visible_ts;
    run.fetch_host = host;
    if (run.single_cell) {
        run.rows = 1;
        segs.push_back(cq_copy_seg{run.res, host_word, 1});
        return;
    }
    long long rows = static_cast<long long>(rows_found);
    unsigned long long* o = run.o;
    unsigned long long* ov = run.ov;
    const long long cap_rows = run.cap_rows, vwords = run.vwords;
    // small results are sorted on the host after the copy (same order as the
    // device sort: a total order with the group index as last key), which
    // saves the sort launches of every such query
    run.host_sort = p.has_sort_node && !p.sort_keys.empty() && rows > 1 && rows <= ctx->config.host_sort_rows;
    if (p.has_sort_node && !p.sort_keys.empty() && rows > 1 && !run.host_sort) {
        if (!run.sidx) {
            DeviceBuffer& ib = ctx->scratch(run.prefix + "sort_idx");
            DeviceBuffer& o2 = ctx->scratch(run.prefix + "out2");
            DeviceBuffer& ov2 = ctx->scratch(run.prefix + "out2_valid");
            ib.ensure(static_cast<size_t>(cq_sort_padded(cap_rows)) * 8);
            o2.ensure(static_cast<size_t>(vts) * cap_rows * 8);
            ov2.ensure(static_cast<size_t>(vts) * vwords * 8);
            run.sidx = static_cast<unsigned long long*>(ib.data());
            run.o2 = static_cast<unsigned long long*>(o2.data());
            run.ov2 = static_cast<unsigned long long*>(ov2.data());
        }
        cq_sort_desc d;
        std::memset(&d, 0, sizeof d);
        d.nkeys = static_cast<int>(std::min<size_t>(p.sort_keys.size(), 16));
        for (int k = 0; k < d.nkeys; ++k) {
            const int c = static_cast<int>(p.sort_keys[static_cast<size_t>(k)].column_index);
            d.col[k] = c;
            d.type[k] = typeCode(p.out[static_cast<size_t>(c)].type);
            d.nullable[k] = p.out[static_cast<size_t>(c)].nullable ? 1 : 0;
            d.desc[k] = p.sort_keys[static_cast<size_t>(k)].descending ? 1 : 0;
        }
        hipCheck(cq_launch_sort(run.sidx, rows, &d, o, ov, cap_rows, vwords, s), "sort");
        hipCheck(cq_launch_fill_u64(run.ov2, 0ull, static_cast<long long>(vts) * vwords, s), "fill");
        hipCheck(cq_launch_gather(o, ov, run.sidx, rows, vts, cap_rows, vwords, run.o2, run.ov2, s), "gather");
        o = run.o2;
        ov = run.ov2;
    }
    if (!run.host_sort && p.has_sort_node && p.has_limit && static_cast<unsigned long long>(rows) > p.limit)
        rows = static_cast<long long>(p.limit);
    run.rows = rows;
    // host layout: [column][rows] values, then [column][vwords] validity words
    for (int c = 0; c < vts && rows; ++c)
        segs.push_back(cq_copy_seg{o + static_cast<size_t>(c) * cap_rows, host_word + static_cast<unsigned long long>(c) * rows,
                                   static_cast<unsigned long long>(rows)});
    if (vts)
        segs.push_back(cq_copy_seg{ov, host_word + static_cast<unsigned long long>(vts) * rows,
                                   static_cast<unsigned long long>(vts) * vwords});
}

void HipModuleExecutor::fetchLaunch(QueryRun& run, ExecutionContext* ctx, unsigned long long rows_found, void* host) {
    std::vector<cq_copy_seg> segs;
    fetchSegments(run, ctx, rows_found, static_cast<char*>(host), 0, segs);
    char* h = static_cast<char*>(host);
    for (const auto& sg : segs)
        hipCheck(hipMemcpyAsync(h + sg.dst_word * 8, sg.src, sg.words * 8, hipMemcpyDeviceToHost, ctx->stream), "fetch");
}

void HipModuleExecutor::fetchFinish(QueryRun& run) {
    const QueryProgram& p = *run.program;
    const int vts = p.visible_ts;
    QueryResult& out = run.result;
    out.columns.clear();
    for (const auto& oc : p.out) out.columns.push_back(ResultColumnDesc{oc.type, oc.nullable, oc.name});
    out.data.assign(static_cast<size_t>(vts), {});
    out.validity.assign(static_cast<size_t>(vts), {});
    out.ordered = (!p.fast_sparse || run.single_cell) && !(p.rows_unordered && !(p.has_sort_node && !p.sort_keys.empty()));
    out.dense_result = !p.fast_sparse;
    const char* h = run.fetch_host;
    if (run.single_cell) {
        out.row_count = 1;
        out.data[0].resize(1);
        std::memcpy(out.data[0].data(), h, 8);
        out.validity[0].assign(1, 1);
        return;
    }
    const long long rows = run.rows;
    out.row_count = static_cast<size_t>(rows);
    const auto* vbits = reinterpret_cast<const uint64_t*>(h + static_cast<size_t>(vts) * rows * 8);
    for (int c = 0; c < vts; ++c) {
        auto& dv = out.data[static_cast<size_t>(c)];
        dv.resize(static_cast<size_t>(rows));
        if (rows) std::memcpy(dv.data(), h + static_cast<size_t>(c) * rows * 8, static_cast<size_t>(rows) * 8);
        auto& vv = out.validity[static_cast<size_t>(c)];
        vv.resize(static_cast<size_t>(rows));
        const uint64_t* w = vbits + static_cast<size_t>(c) * run.vwords;
        for (long long r = 0; r < rows; ++r) vv[static_cast<size_t>(r)] = static_cast<uint8_t>((w[r >> 6] >> (r & 63)) & 1ull);
    }
    if (run.host_sort) sortHost(run);
}

// ORDER BY [LIMIT] of a fetched result on the host: the device sort's order
// (k_bitonic_*: per key, a NULL sorts after values ascending, the type's
// comparison, DESC flips it, the group index breaks ties)
void HipModuleExecutor::sortHost(QueryRun& run) {
    const QueryProgram& p = *run.program;
    QueryResult& out = run.result;
    const size_t rows = out.row_count;
    std::vector<size_t> idx(rows);
    for (size_t i = 0; i < rows; ++i) idx[i] = i;
    const size_t nk = std::min<size_t>(p.sort_keys.size(), 16);
    auto before = [&](size_t l, size_t r) {  // row l sorts before row r
        for (size_t k = 0; k < nk; ++k) {
            const size_t c = p.sort_keys[k].column_index;
            const bool desc = p.sort_keys[k].descending;
            if (p.out[c].nullable) {
                const uint8_t lv = out.validity[c][l], rv = out.validity[c][r];
                if (lv != rv) return lv != 0;  // valid first
                if (!lv) continue;
            }
            const uint64_t a = out.data[c][l], b = out.data[c][r];
            bool lt, gt;
            switch (typeCode(p.out[c].type)) {
                case CQ_TYPE_INT64:
                    lt = static_cast<long long>(a) < static_cast<long long>(b);
                    gt = static_cast<long long>(b) < static_cast<long long>(a);
                    break;
                case CQ_TYPE_FLOAT64: {
                    double x, y;
                    std::memcpy(&x, &a, 8);
                    std::memcpy(&y, &b, 8);
                    lt = x < y;
                    gt = y < x;
                    break;
                }
                default:
                    lt = a < b;
                    gt = b < a;
                    break;
            }
            if (lt) return !desc;
            if (gt) return desc;
        }
        return l < r;
    };
    std::sort(idx.begin(), idx.end(), before);
    size_t keep = rows;
    if (p.has_limit && keep > p.limit) keep = static_cast<size_t>(p.limit);
    for (size_t c = 0; c < out.data.size(); ++c) {
        std::vector<uint64_t> d(keep);
        std::vector<uint8_t> v(keep);
        for (size_t i = 0; i < keep; ++i) {
            d[i] = out.data[c][idx[i]];
            v[i] = out.validity[c][idx[i]];
        }
        out.data[c].swap(d);
        out.validity[c].swap(v);
    }
    out.row_count = keep;
    run.rows = static_cast<long long>(keep);
}

void HipModuleExecutor::fetch(QueryRun& run, ExecutionContext* ctx, unsigned long long rows_found) {
    std::vector<char> host(fetchBytes(run, rows_found) + 8);
    fetchLaunch(run, ctx, rows_found, host.data());
    hipCheck(hipStreamSynchronize(ctx->stream), "sync");
    fetchFinish(run);
}

void HipModuleExecutor::fetchPinned(QueryRun& run, ExecutionContext* ctx, unsigned long long rows_found) {
    hipStream_t s = ctx->stream;
    const size_t bytes = fetchBytes(run, rows_found);
    std::vector<cq_copy_seg> segs;
    fetchSegments(run, ctx, rows_found, nullptr, 0, segs);
    // the result's columns and validity words gathered on the device into
    // one range, then one copy into pinned memory (not one copy per column
    // into pageable memory)
    const size_t seg_at = (bytes + 255) & ~size_t(255);
    const size_t seg_bytes = segs.size() * sizeof(cq_copy_seg);
    char* host = static_cast<char*>(ctx->pinned().ensure(seg_at + seg_bytes + 8));
    if (segs.size() == 1) {
        hipCheck(hipMemcpyAsync(host + segs[0].dst_word * 8, segs[0].src, segs[0].words * 8, hipMemcpyDeviceToHost, s), "fetch");
    } else if (!segs.empty()) {
        std::memcpy(host + seg_at, segs.data(), seg_bytes);
        DeviceBuffer& stage = ctx->scratch(run.prefix + "fetch_stage");
        stage.ensure(seg_at + seg_bytes + 8);
        char* d = static_cast<char*>(stage.data());
        auto* dseg = reinterpret_cast<cq_copy_seg*>(d + seg_at);
        hipCheck(hipMemcpyAsync(dseg, host + seg_at, seg_bytes, hipMemcpyHostToDevice, s), "upload fetch segments");
        hipCheck(cq_launch_copy_segments(dseg, static_cast<int>(segs.size()), reinterpret_cast<unsigned long long*>(d), s),
                 "gather result");
        hipCheck(hipMemcpyAsync(host, d, bytes, hipMemcpyDeviceToHost, s), "fetch");
    }
    run.fetch_host = host;
    hipCheck(hipStreamSynchronize(s), "sync");
    fetchFinish(run);
}

void HipModuleExecutor::execute(const ModuleHandle& module, ExecutionContext* ctx) {
    hipStream_t s = ctx->stream;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ks = nullptr, ke = nullptr;
    hipCheck(hipEventCreate(&ev0), "hipEventCreate");
    hipCheck(hipEventCreate(&ev1), "hipEventCreate");
    if (ctx->config.kernel_timing) {
        hipCheck(hipEventCreate(&ks), "hipEventCreate");
        hipCheck(hipEventCreate(&ke), "hipEventCreate");
    }
    QueryRun run;
    bind(module, ctx->program, ctx, "", run, nullptr);
    hipCheck(hipEventRecord(ev0, s), "hipEventRecord");
    setup(run, ctx, nullptr);
    unsigned long long rows = 0;
    if (ctx->program.projection) {
        rows = project(run, ctx, ks, ke);
    } else {
        if (ks) hipCheck(hipEventRecord(ks, s), "hipEventRecord");
        scan(run, ctx);
        if (ke) hipCheck(hipEventRecord(ke, s), "hipEventRecord");
        shape(run, ctx, nullptr, nullptr);
    }
    if (!run.single_cell && !ctx->program.projection) {
        auto* pr = static_cast<unsigned long long*>(ctx->pinned().ensure(64));
        hipCheck(hipMemcpyAsync(pr, run.off + run.cap_rows, 8, hipMemcpyDeviceToHost, s), "copy row count");
        hipCheck(hipStreamSynchronize(s), "sync");
        rows = *pr;
    }
    const auto f0 = std::chrono::steady_clock::now();
    fetchPinned(run, ctx, rows);
    hipCheck(hipEventRecord(ev1, s), "hipEventRecord");
    hipCheck(hipEventSynchronize(ev1), "sync");
    ctx->result = std::move(run.result);
    const auto f1 = std::chrono::steady_clock::now();
    float ms = 0;
    (void)hipEventElapsedTime(&ms, ev0, ev1);
    ctx->timing.gpu_execute_ms = ms;
    if (ks && ke) {
        float kms = 0;
        (void)hipEventElapsedTime(&kms, ks, ke);
        ctx->timing.scan_kernel_ms = kms;
    }
    ctx->timing.host_fetch_ms = elapsedMs(f0, f1);
    (void)hipEventDestroy(ev0);
    (void)hipEventDestroy(ev1);
    if (ks) (void)hipEventDestroy(ks);
    if (ke) (void)hipEventDestroy(ke);
}

bool HipModuleExecutor::linearGroupable(const QueryProgram& p) {
    return p.vm && !p.projection && cq_lin_supported(p.vm.get());
}

void HipModuleExecutor::executeLinearGroup(const std::vector<ExecutionContext*>& ctxs) {
    const size_t n = ctxs.size();
    if (n == 0) return;
    if (n > static_cast<size_t>(CQ_LIN_BATCH)) throw std::logic_error("linear group of more than CQ_LIN_BATCH plans");
    ExecutionContext* c0 = ctxs.front();
    hipStream_t s = c0->stream;
    for (ExecutionContext* c : ctxs) {
        if (c->stream != s || c->device != c0->device) throw std::logic_error("linear group: contexts on different streams");
        if (!linearGroupable(c->program)) throw std::logic_error("linear group: a plan without a linear form");
    }
    hipCheck(hipSetDevice(c0->device), "hipSetDevice");
    std::vector<QueryRun> runs(n);
    // union of the plans' columns (by device address); every plan's linear
    // form re-indexed to the union positions
    std::vector<const int*> ucol;
    std::vector<const unsigned int*> uvalid;
    for (size_t k = 0; k < n; ++k) {
        ExecutionContext* c = ctxs[k];
        bind(nullptr, c->program, c, "", runs[k], nullptr);
        const cq_query_params& P = runs[k].P;
        if (P.row_begin != runs[0].P.row_begin || P.row_end != runs[0].P.row_end)
            throw std::logic_error("linear group: plans over different rows");
        if (P.row_begin % cq_lin_row_align() != 0) throw std::logic_error("linear group: unaligned first row");
        auto prog = std::make_shared<cq_vm_program>(*c->program.vm);
        cq_lin_program& L = prog->lin;
        int map[CQ_LIN_COLS];
        for (int c2 = 0; c2 < L.ncols; ++c2) {
            const int* col = P.col[L.slot[c2]];
            const unsigned int* vb = L.nullable[c2] ? P.valid[L.slot[c2]] : nullptr;
            size_t u = 0;
            while (u < ucol.size() && ucol[u] != col) ++u;
            if (u == ucol.size()) {
                if (ucol.size() == static_cast<size_t>(CQ_LIN_COLS)) throw std::logic_error("linear group: more than CQ_LIN_COLS columns");
                ucol.push_back(col);
                uvalid.push_back(vb);
            } else if (uvalid[u] != vb) {
                throw std::logic_error("linear group: a column read with and without its validity");
            }
            map[c2] = static_cast<int>(u);
        }
        for (int t = 0; t < L.nterms; ++t) L.term[t].col = map[L.term[t].col];
        if (L.join >= 0) L.join_col = map[L.join_col];
        for (int g = 0; g < L.ngroups; ++g) {
            if (L.key[g].kind == CQ_LIN_COL) L.key[g].col = map[L.key[g].col];
            if (L.key[g].valid_col >= 0) L.key[g].valid_col = map[L.key[g].valid_col];
        }
        for (int i = 0; i < L.naggs; ++i) {
            cq_lin_agg& A = L.agg[i];
            for (int v = 0; v < A.nvalid; ++v) A.valid_col[v] = map[A.valid_col[v]];
            for (int st = 0; st < A.nsteps; ++st)
                if (A.step[st].kind == CQ_LIN_COL) A.step[st].idx = map[A.step[st].idx];
        }
        c->program.vm = prog;  // setup() uploads this copy
    }
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipCheck(hipEventCreate(&ev0), "hipEventCreate");
    hipCheck(hipEventCreate(&ev1), "hipEventCreate");
    hipCheck(hipEventRecord(ev0, s), "hipEventRecord");
    for (size_t k = 0; k < n; ++k) setup(runs[k], ctxs[k], nullptr);

    // LDS after the staged columns: group regions first (they replace
    // global atomics), then join tables, within a budget that keeps two
    // blocks per CU
    const int nu = static_cast<int>(ucol.size());
    const unsigned stage = cq_lin_batch_stage_bytes(nu);
    const unsigned limit = 80u * 1024u > stage ? 80u * 1024u - stage : 0u;
    cq_lin_batch hb;
    std::memset(&hb, 0, sizeof hb);
    hb.nprog = static_cast<int>(n);
    hb.ncols = nu;
    hb.row_begin = runs[0].P.row_begin;
    hb.row_end = runs[0].P.row_end;
    for (int u = 0; u < nu; ++u) {
        hb.col[u] = ucol[static_cast<size_t>(u)];
        hb.valid[u] = uvalid[static_cast<size_t>(u)];
    }
    unsigned rest = 0;
    for (size_t k = 0; k < n; ++k) {
        const QueryProgram& p = ctxs[k]->program;
        hb.tab_off[k] = hb.grp_off[k] = CQ_LIN_NONE;
        const unsigned gb = p.vm->ngroups > 0 ? static_cast<unsigned>((p.vm_lds_bytes + 7) & ~7ull) : 0u;
        if (gb > 0 && rest + gb <= limit) {
            hb.grp_off[k] = rest;
            rest += gb;
        }
    }
    for (size_t k = 0; k < n; ++k) {
        const cq_lin_program& L = ctxs[k]->program.vm->lin;
        const unsigned tb = L.join >= 0 && L.join_lds_words > 0 ? static_cast<unsigned>((L.join_lds_words + 1) & ~1) * 4u : 0u;
        if (tb > 0 && rest + tb <= limit) {
            hb.tab_off[k] = rest;
            rest += tb;
        }
    }
    // programs, parameter blocks and the descriptor -> device, as arrays
    const size_t vbytes = n * sizeof(cq_vm_program), pbytes = n * sizeof(cq_query_params);
    const size_t poff = (vbytes + 255) & ~static_cast<size_t>(255), boff = (poff + pbytes + 255) & ~static_cast<size_t>(255);
    DeviceBuffer& db = c0->scratch("lin_group");
    db.ensure(boff + sizeof(cq_lin_batch));
    char* base = static_cast<char*>(db.data());
    std::vector<char> host(boff + sizeof(cq_lin_batch));
    for (size_t k = 0; k < n; ++k) {
        std::memcpy(host.data() + k * sizeof(cq_vm_program), ctxs[k]->program.vm.get(), sizeof(cq_vm_program));
        std::memcpy(host.data() + poff + k * sizeof(cq_query_params), &runs[k].P, sizeof(cq_query_params));
    }
    std::memcpy(host.data() + boff, &hb, sizeof hb);
    hipCheck(hipMemcpyAsync(base, host.data(), host.size(), hipMemcpyHostToDevice, s), "upload linear group");
    hipCheck(cq_lin_launch_batch(reinterpret_cast<const cq_lin_batch*>(base + boff), reinterpret_cast<const cq_vm_program*>(base),
                                 reinterpret_cast<const cq_query_params*>(base + poff), nu, rest, hb.row_end - hb.row_begin,
                                 c0->device, c0->config.vm_grid_percent, s),
             "linear group scan");
    for (size_t k = 0; k < n; ++k) shape(runs[k], ctxs[k], nullptr, nullptr);
    // row counts of every plan in one copy each, one synchronization
    auto* counts = static_cast<unsigned long long*>(c0->pinned().ensure(8 * n));
    for (size_t k = 0; k < n; ++k) {
        counts[k] = 0;
        if (!runs[k].single_cell)
            hipCheck(hipMemcpyAsync(&counts[k], runs[k].off + runs[k].cap_rows, 8, hipMemcpyDeviceToHost, s), "copy row count");
    }
    hipCheck(hipStreamSynchronize(s), "sync");
    std::vector<unsigned long long> rows(counts, counts + n);
    for (size_t k = 0; k < n; ++k) {
        const auto f0 = std::chrono::steady_clock::now();
        fetchPinned(runs[k], ctxs[k], rows[k]);
        ctxs[k]->result = std::move(runs[k].result);
        ctxs[k]->timing.host_fetch_ms = elapsedMs(f0, std::chrono::steady_clock::now());
    }
    hipCheck(hipEventRecord(ev1, s), "hipEventRecord");
    hipCheck(hipEventSynchronize(ev1), "sync");
    float ms = 0;
    (void)hipEventElapsedTime(&ms, ev0, ev1);
    for (ExecutionContext* c : ctxs) c->timing.gpu_execute_ms = ms;
    (void)hipEventDestroy(ev0);
    (void)hipEventDestroy(ev1);
}

void HipModuleExecutor::executeInto(const ModuleHandle& module, ExecutionContext* ctx, void* scratch, size_t bytes) {
    const QueryProgram& p = ctx->program;
    if (p.projection)
        throw std::invalid_argument("the allocation-free execute takes aggregate queries only (a projection's result size is known "
                                    "only after its scan)");
    if (!p.vm && (!module || !module->loaded()))
        throw std::invalid_argument("the allocation-free execute needs the query's module compiled and loaded on the device "
                                    "(cq_db_compile first)");
    hipStream_t s = ctx->stream;
    QueryRun run;
    ScratchArena arena;
    arena.measure();
    bind(module, p, ctx, "", run, &arena);
    if (arena.used() > bytes)
        throw std::length_error("the allocation-free execute needs " + std::to_string(arena.used()) +
                                " bytes of scratch, over its static " + std::to_string(bytes) + " bytes");
    arena.assign(scratch, bytes);
    bind(module, p, ctx, "", run, &arena);
    setup(run, ctx, nullptr);
    scan(run, ctx);
    shape(run, ctx, nullptr, nullptr);
    unsigned long long rows = 0;
    if (!run.single_cell) {
        hipCheck(hipMemcpyAsync(&rows, run.off + run.cap_rows, 8, hipMemcpyDeviceToHost, s), "copy row count");
        hipCheck(hipStreamSynchronize(s), "sync");
    }
    fetch(run, ctx, rows);
    ctx->result = std::move(run.result);
}

// ------------------------------------------------------------------ QueryEngine

QueryEngine::QueryEngine(std::shared_ptr<IQueryCache> cache, std::shared_ptr<ICompiler> compiler,
                         std::shared_ptr<IExecutor> executor, std::shared_ptr<const Catalog> catalog)
    : cache_(std::move(cache)), compiler_(std::move(compiler)), executor_(std::move(executor)), catalog_(std::move(catalog)) {}

QueryProgram QueryEngine::prepare(const std::string& sql, const CodegenOptions& options) const {
    auto tree = buildOptimizedTree(sql, catalog_.get());
    return generateQueryProgram(tree.get(), *catalog_, options);
}

namespace {

// Integer literals of a statement (digit runs not attached to an
// identifier, not part of a decimal number), and the statement with each
// replaced by '?'.  Empty `values` and no '?' when the text has none.
bool splitLiterals(const std::string& sql, std::string& shape, std::vector<long long>& values,
                   std::vector<std::pair<size_t, size_t>>& spans) {
    shape.clear();
    values.clear();
    spans.clear();
    size_t i = 0;
    const size_t n = sql.size();
    bool in_string = false;
    while (i < n) {
        const char c = sql[i];
        if (c == '\'') in_string = !in_string;
        const bool ident_before = i > 0 && (std::isalnum(static_cast<unsigned char>(sql[i - 1])) || sql[i - 1] == '_' || sql[i - 1] == '.');
        if (!in_string && std::isdigit(static_cast<unsigned char>(c)) && !ident_before) {
            size_t j = i;
            while (j < n && std::isdigit(static_cast<unsigned char>(sql[j]))) ++j;
            if (j < n && (sql[j] == '.' || std::isalpha(static_cast<unsigned char>(sql[j])) || sql[j] == '_')) return false;
            if (j - i > 18) return false;
            values.push_back(std::stoll(sql.substr(i, j - i)));
            spans.emplace_back(i, j);
            shape += '?';
            i = j;
            continue;
        }
        shape += c;
        ++i;
    }
    return !in_string;
}

std::string optionsKey(const CodegenOptions& o) {
    std::ostringstream k;
    k << o.block_threads << ',' << o.vectors_per_thread << ',' << o.lds_budget << ',' << o.allow_lds_tables << o.allow_lds_aggregation
      << o.allow_compact_tables << o.allow_dense_tables << o.calendar_tables << o.split_probes << o.buffer_tables << o.hoist_gathers << o.defer_probes << o.prefetch << o.stream_large_tables << o.batch_body << o.batch_shared_tables << o.batch_masks
      << o.vm;
    return k.str();
}

}  // namespace

QueryProgram QueryEngine::prepareCached(const std::string& sql, const CodegenOptions& options) {
    std::string shape;
    std::vector<long long> values;
    std::vector<std::pair<size_t, size_t>> spans;
    if (!splitLiterals(sql, shape, values, spans) || values.empty()) return prepare(sql, options);
    const std::string key = optionsKey(options) + "|" + shape;
    {
        std::lock_guard<std::mutex> lock(prepared_mu_);
        auto it = prepared_.find(key);
        if (it != prepared_.end()) {
            if (!it->second.cacheable) return prepare(sql, options);
            QueryProgram p = *it->second.program;
            for (size_t k = 0; k < p.lits.size(); ++k) {
                const int li = it->second.slot_literal[k];
                if (li >= 0) p.lits[k] = values[static_cast<size_t>(li)];
            }
            return p;
        }
    }
    QueryProgram p = prepare(sql, options);
    // plan again with marker constants: distinct values of the same C
    // literal class (int / long long) as each original
    PreparedShape entry;
    try {
        std::string marked;
        std::vector<long long> markers;
        size_t at = 0;
        for (size_t k = 0; k < spans.size(); ++k) {
            const long long v = values[k];
            const long long m = v <= 2147483647ll ? 2000000011ll - 7919ll * static_cast<long long>(k)
                                                  : 9000000000000000000ll - 7919ll * static_cast<long long>(k);
            markers.push_back(m);
            marked += sql.substr(at, spans[k].first - at) + std::to_string(m);
            at = spans[k].second;
        }
        marked += sql.substr(at);
        const QueryProgram q = prepare(marked, options);
        bool ok = q.signature == p.signature && q.has_limit == p.has_limit && q.limit == p.limit && q.lits.size() == p.lits.size() &&
                  q.flits == p.flits && static_cast<bool>(q.vm) == static_cast<bool>(p.vm);
        if (ok && p.vm) ok = std::memcmp(p.vm.get(), q.vm.get(), sizeof(cq_vm_program)) == 0;
        std::vector<int> slot(q.lits.size(), -1);
        for (size_t k = 0; ok && k < q.lits.size(); ++k) {
            int found = -1;
            for (size_t m = 0; m < markers.size(); ++m)
                if (q.lits[k] == markers[m]) found = static_cast<int>(m);
            if (found < 0) {
                ok = q.lits[k] == p.lits[k];  // a constant the statement does not vary
            } else {
                ok = p.lits[k] == values[static_cast<size_t>(found)];
                slot[k] = found;
            }
        }
        entry.cacheable = ok;
        entry.slot_literal = std::move(slot);
        if (ok) {
            auto cached = std::make_shared<QueryProgram>(p);
            cached->plan_sexpr.clear();
            entry.program = std::move(cached);
        }
    } catch (const std::exception&) {
        entry.cacheable = false;
    }
    std::lock_guard<std::mutex> lock(prepared_mu_);
    if (prepared_.size() >= 4096) prepared_.clear();
    prepared_[key] = std::move(entry);
    return p;
}

ModuleHandle QueryEngine::moduleFor(const QueryProgram& program, QueryTimingStats* timing) {
    return moduleForSource(program.source, program.signature, timing);
}

ModuleHandle QueryEngine::moduleForSource(const std::string& source, const std::string& signature, QueryTimingStats* timing) {
    ModuleHandle m = cache_->get(signature);
    if (m) {
        if (timing) timing->cache_hit = true;
        return m;
    }
    {
        std::unique_lock<std::mutex> lock(compile_mu_);
        compile_cv_.wait(lock, [&] { return compiling_.count(signature) == 0; });
        m = cache_->get(signature);
        if (m) {
            if (timing) timing->cache_hit = true;
            return m;
        }
        compiling_.insert(signature);
    }
    const auto c0 = std::chrono::steady_clock::now();
    try {
        m = compiler_->compile(source, signature);
    } catch (...) {
        std::lock_guard<std::mutex> lock(compile_mu_);
        compiling_.erase(signature);
        compile_cv_.notify_all();
        throw;
    }
    if (timing) timing->compile_ms += elapsedMs(c0, std::chrono::steady_clock::now());
    std::lock_guard<std::mutex> lock(compile_mu_);
    cache_->put(signature, m);
    compiling_.erase(signature);
    compile_cv_.notify_all();
    return m;
}

void QueryEngine::planQuery(const std::string& sql, ExecutionContext* ctx, QueryTimingStats& timing) {
    const auto t0 = std::chrono::steady_clock::now();
    const int64_t row_begin = ctx->row_begin, row_end = ctx->row_end;
    ctx->resetForQuery();
    ctx->row_begin = row_begin;
    ctx->row_end = row_end;
    const std::string key = optionsKey(ctx->config.codegen) + "|" + sql;
    std::shared_ptr<const QueryProgram> known;
    {
        std::lock_guard<std::mutex> lock(programs_mu_);
        auto it = programs_.find(key);
        if (it != programs_.end()) known = it->second;
    }
    if (known) {
        ctx->program = *known;
        return;
    }
    auto tree = buildOptimizedTree(sql, catalog_.get());
    const auto t1 = std::chrono::steady_clock::now();
    try {
        ctx->program = generateQueryProgram(tree.get(), *catalog_, ctx->config.codegen);
    } catch (const std::runtime_error& e) {
        // plans the interpreter does not take (OR-key joins) run compiled
        if (!ctx->config.codegen.vm || std::string(e.what()).rfind("interpreted tier:", 0) != 0) throw;
        CodegenOptions co = ctx->config.codegen;
        co.vm = false;
        ctx->program = generateQueryProgram(tree.get(), *catalog_, co);
    }
    const auto t2 = std::chrono::steady_clock::now();
    timing.parse_ms = elapsedMs(t0, t1);
    timing.codegen_ms = elapsedMs(t1, t2);
    std::lock_guard<std::mutex> lock(programs_mu_);
    if (programs_.size() >= 1024) programs_.clear();
    programs_[key] = std::make_shared<const QueryProgram>(ctx->program);
}

void QueryEngine::executeQuery(const std::string& sql, ExecutionContext* ctx) {
    const auto t0 = std::chrono::steady_clock::now();
    QueryTimingStats timing;
    planQuery(sql, ctx, timing);
    preflightDeviceMemoryOrThrow(*ctx, ctx->program);
    ModuleHandle m = ctx->program.vm ? nullptr : moduleFor(ctx->program, &timing);
    executor_->execute(m, ctx);
    timing.gpu_execute_ms = ctx->timing.gpu_execute_ms;
    timing.scan_kernel_ms = ctx->timing.scan_kernel_ms;
    timing.host_fetch_ms = ctx->timing.host_fetch_ms;
    timing.engine_ms = elapsedMs(t0, std::chrono::steady_clock::now());
    ctx->timing = timing;
    ctx->result.timing = timing;
}

std::vector<std::string> QueryEngine::executeInterpretedGroup(const std::vector<std::string>& sqls,
                                                              const std::vector<ExecutionContext*>& ctxs) {
    const size_t n = sqls.size();
    if (ctxs.size() != n) throw std::invalid_argument("executeInterpretedGroup: one context per statement");
    std::vector<std::string> errs(n);
    std::vector<QueryTimingStats> timing(n);
    std::vector<std::chrono::steady_clock::time_point> start(n);
    HipModuleExecutor* hip = hipExecutor();
    auto finishOne = [&](size_t k) {
        ExecutionContext* c = ctxs[k];
        timing[k].gpu_execute_ms = c->timing.gpu_execute_ms;
        timing[k].scan_kernel_ms = c->timing.scan_kernel_ms;
        timing[k].host_fetch_ms = c->timing.host_fetch_ms;
        timing[k].engine_ms = elapsedMs(start[k], std::chrono::steady_clock::now());
        c->timing = timing[k];
        c->result.timing = timing[k];
    };
    auto runAlone = [&](size_t k) {
        try {
            ExecutionContext* c = ctxs[k];
            ModuleHandle m = c->program.vm ? nullptr : moduleFor(c->program, &timing[k]);
            executor_->execute(m, c);
            finishOne(k);
        } catch (const std::exception& e) {
            errs[k] = e.what();
        }
    };
    // plans; the ones with a linear form over the full fact range group by
    // fact table and union columns, in statement order
    std::vector<size_t> alone;
    struct Group {
        std::string table;
        std::vector<size_t> members;
        std::vector<std::string> cols;
    };
    std::vector<Group> groups;
    for (size_t k = 0; k < n; ++k) {
        start[k] = std::chrono::steady_clock::now();
        ExecutionContext* c = ctxs[k];
        try {
            planQuery(sqls[k], c, timing[k]);
            preflightDeviceMemoryOrThrow(*c, c->program);
        } catch (const std::exception& e) {
            errs[k] = e.what();
            continue;
        }
        const QueryProgram& p = c->program;
        if (!hip || !HipModuleExecutor::linearGroupable(p) || c->row_begin != 0 || c->row_end >= 0) {
            alone.push_back(k);
            continue;
        }
        std::vector<std::string> cols;
        for (int i = 0; i < p.vm->lin.ncols; ++i) cols.push_back(p.fact_cols[static_cast<size_t>(p.vm->lin.slot[i])]);
        bool placed = false;
        for (Group& g : groups) {
            if (g.table != p.fact_table || g.members.size() >= static_cast<size_t>(CQ_LIN_BATCH)) continue;
            std::vector<std::string> u = g.cols;
            for (const auto& col : cols)
                if (std::find(u.begin(), u.end(), col) == u.end()) u.push_back(col);
            if (u.size() > static_cast<size_t>(CQ_LIN_COLS)) continue;
            g.cols = u;
            g.members.push_back(k);
            placed = true;
            break;
        }
        if (!placed) {
            Group g;
            g.table = p.fact_table;
            g.members.push_back(k);
            for (const auto& col : cols)
                if (std::find(g.cols.begin(), g.cols.end(), col) == g.cols.end()) g.cols.push_back(col);
            groups.push_back(std::move(g));
        }
    }
    for (const Group& g : groups) {
        if (g.members.size() == 1) {
            runAlone(g.members.front());
            continue;
        }
        std::vector<ExecutionContext*> gc;
        for (size_t k : g.members) gc.push_back(ctxs[k]);
        try {
            hip->executeLinearGroup(gc);
            for (size_t k : g.members) finishOne(k);
        } catch (const std::exception&) {
            // the shared pass failed: each plan on its own, so one plan's
            // error does not become every member's
            for (size_t k : g.members) {
                errs[k].clear();
                try {
                    planQuery(sqls[k], ctxs[k], timing[k]);  // the statement's own linear form again
                } catch (const std::exception& e) {
                    errs[k] = e.what();
                    continue;
                }
                runAlone(k);
            }
        }
    }
    for (size_t k : alone) runAlone(k);
    return errs;
}

}  // namespace db
