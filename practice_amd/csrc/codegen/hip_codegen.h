// HIP code generator for the JIT query path.
//
// Replaces the reference JITOperatorVisitor (src/core/visitor.cpp,
// include/core/visitor.h:200-356), which walks the optimized plan in
// produce/consume order and emits SYCL text.  This visitor walks the same plan
// types and emits one hipRTC module per query shape:
//   cq_build_<j>   one dimension hash-table build per join
//                  (emitPHTBuildKernel, visitor.cpp:4206-4290)
//   cq_scan        the fused fact kernel: 128-bit column loads, filters,
//                  probes, perfect-hash GROUP BY with LDS-privatized u64
//                  aggregates (produceTableScan + consume*, 1657-1751,
//                  1823-1893, 4556-4731, 2987-3196)
//   cq_dense_count present && HAVING flags per group
//                  (emitAggregateDenseColumnarMaterialization, 2589-2697)
// Predicate constants are kernel arguments, so the emitted source is the plan
// signature: queries that differ only in constants share one compiled module.
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "../device/vm_program.h"
#include "../plan/operators.h"

namespace db {

struct CodegenOptions {
    int block_threads = 0;            // 0: chosen from the LDS footprint
    int vectors_per_thread = 2;       // 128-bit loads per column per lane per tile
    uint64_t lds_budget = 160 * 1024;  // bytes of LDS a scan block may use (gfx950: 160 KiB per CU)
    bool allow_lds_tables = true;     // stage compact hash tables into LDS
    bool allow_lds_aggregation = true;
    bool allow_compact_tables = true;
    bool allow_dense_tables = true;   // key-indexed tables where the keys allow
    bool calendar_tables = true;      // key-indexed tables of yyyymmdd keys indexed by (year, month, day)
    // load tile t + stride while tile t is processed (scan kernels whose
    // prefetched columns take at most 48 registers per lane)
    bool prefetch = true;
    bool stream_large_tables = false; // streaming loads for probes of HBM tables beyond L1 (measured 2x slower on C3: off)
    bool split_probes = true;         // probes as locate / load / decode loops (all of a thread's loads in flight at once)
    bool buffer_tables = true;        // key-indexed tables in HBM read through a buffer resource (32-bit offsets, range-checked)
    // issue the loads of an HBM table's probe right after the fact filters,
    // ahead of LDS probes that remove no rows, so the L2 round trip overlaps
    // their work
    bool hoist_gathers = true;
    // batched mask form: probe tables too large for LDS after the other
    // joins' masks, only for items some query still accepts
    bool defer_probes = true;
    // Also emit the per-tile pipeline body for the batched shared scan
    // (global tables, per-tile scalar reductions, global GROUP BY atomics).
    bool batch_body = false;
    // batched templates share one probe per dimension (packed attributes)
    bool batch_shared_tables = true;
    // shared templates whose filters reduce to small lookup tables evaluate
    // all their queries at once through per-row query bit masks
    bool batch_masks = true;
    // interpreted tier: register code for the ahead-of-time interpreter
    // instead of HIP source
    bool vm = false;
};

struct JoinProgram {
    std::string dim_table, pk_col, fk_col, payload_col;
    // OR-key join (lo_orderdate = d_datekey OR lo_commitdate = d_datekey):
    // every fact key alternative in condition order (fk_col is the first)
    std::vector<std::string> fk_alts;
    // projection of this dimension's columns: the payload is the build row's
    // index and the columns are gathered by it (with their validity)
    bool rowid = false;
    int64_t dim_rows = 0;
    int pk_dslot = -1, payload_dslot = -1;
    bool pair = false;        // PHT_2 (carries a payload) vs PHT_1
    int64_t kmin = 0, len = 0;
    bool compact = false;     // compact tagged layout vs reference int layout
    int eb = 0, tb = 0, jlo = 0;
    int64_t pmin = 0;
    // key-indexed layout (chip::ht_probe_dense): candidate and choice
    bool dense_ok = false;    // keys unique, non-null, within the reference slot range
    bool dense = false;       // chosen
    int64_t dmin = 0, dspan = 0;
    int dbits = 0;            // exact entry width (0: all keys present, no payload)
    int doff = 1;             // 1: code 0 marks an absent key; 0: every key present
    int db = 0;               // chosen width (exact in LDS, power of two in HBM)
    int zmode = 0, zx = 0;    // probe key 0 vs. the reference's empty-slot match (chip::ht_probe_dense)
    int cal = 0;              // calendar index from this first year (keys yyyymmdd; chip::dense_index), 0: key - dmin
    // batched templates: one table per template, every dimension row
    // inserted with its packed attributes (payload + build-filter columns);
    // each query applies its own build filter to the attributes
    bool shared = false;
    std::vector<std::string> s_cols;
    std::vector<int64_t> s_min;
    std::vector<int> s_bits, s_off;
    int s_present = 0;        // 1: bit 0 marks an inserted key
    int s_payload = -1;       // attribute index of the payload column
    std::string s_key;        // layout key: tables with equal keys hold equal entries
    // the shared probe's loops with placeholders (@E@ entry, @M@ / @Z@ lane
    // masks, @X@ entry index, @FK@ the key of item i, @FV@ its validity bit
    // or empty, @TAB@ the table), so a batch can run one probe for several
    // templates with the same layout and fact key
    std::string s_probe;
    // (the shared probe reads the table through `cq_sht<j>`, which the batch
    // kernel points at LDS or at the table in HBM)
    bool lds = false;
    uint64_t ht_words = 0;    // u32 words of the table
    // multi-value table of a non-unique build key (MHT, visitor.cpp:3946-3970,
    // 4295-4464): one 16-byte directory entry {key, first, count, 0} per slot
    // (key - kmin) % len, u64 counts and offsets, the prefix scan's block
    // sums, u32 write cursors and the build row ids grouped by slot; the
    // payload is always the build row id.  Word offsets in the table:
    bool mht = false;
    uint64_t mht_cnt = 0, mht_off = 0, mht_sums = 0, mht_pos = 0, mht_rows = 0;
};

struct OutputColumn {
    LogicalType type = LogicalType::UInt64;
    bool nullable = false;
    std::string name;
};

struct QueryProgram {
    // projection (non-aggregate SELECT): cq_pcount counts each wave's rows per
    // tile, cq_pwrite writes them at the scanned offsets (row order); the
    // output columns are the visible slots of `out`
    bool projection = false;
    int proj_waves = 0;
    // rows of a projection through a multi-value join come out grouped by
    // scan position, in table order within one key (unspecified order)
    bool rows_unordered = false;
    std::string source;
    std::string signature;    // hex hash of `source`
    std::string plan_sexpr;

    std::string fact_table;
    int64_t fact_rows = 0;
    std::vector<std::string> fact_cols;  // scan slot -> column
    std::vector<bool> fact_nullable;
    std::vector<std::string> dim_cols;   // build slot -> column
    std::vector<bool> dim_nullable;
    std::vector<JoinProgram> joins;
    std::vector<int64_t> lits;
    std::vector<double> flits;

    // aggregate layout (operators.h storage/visible tuples)
    int ngroups = 0, naggs = 0, visible_ts = 0, storage_ts = 0;
    int64_t groups = 1;
    std::vector<std::string> agg_funcs;
    std::vector<int> hidden_slot;        // per aggregate: AVG count slot or -1
    std::vector<int> min_slots;          // storage slots initialised to ~0
    bool grouped = false;
    bool fast_sparse = false;            // aggregateFastSparseOutputEligible
    bool lds_aggregation = false;
    bool has_avg = false;
    bool has_having = false;
    std::vector<OutputColumn> out;

    bool has_sort_node = false;
    std::vector<SortKeyDef> sort_keys;
    bool has_limit = false;
    uint64_t limit = 0;

    int block_threads = 256;
    int vectors_per_thread = 2;
    uint64_t scan_lds_bytes = 0;

    // algorithmic bytes of one scan (fact columns + bitmaps + dimension columns)
    uint64_t algorithmic_bytes = 0;

    // batched shared scan: pipeline + aggregation code of one tile, reading
    // c<slot>[i] / v<slot> registers, the flag mask `fl` and the query's
    // parameter block `P`; `continue` leaves the query's body
    std::string batch_body;
    std::vector<int> count_slots_scalar;  // scalar COUNT slots (validity set whenever rows are scanned)
    // shared-probe form: template-level code run once per tile before the
    // per-query bodies (probes of the template's shared tables), and
    // per-query LDS aggregation (init / merge code, bytes per query)
    bool shared_tables = false;
    bool batch_bool_flags = false;  // the body keeps row flags in `bool fl[IPT]` (else `u32 fl`)
    // query-mask form (CodegenOptions::batch_masks): per launch, LDS tables
    // map a filter column's value / a join's filter attributes to the bit
    // mask of the template's queries that accept it (built by
    // batch_mask_build), and batch_mask_body handles a tile for all queries
    // of the template (up to 64).  "@T@" in both stands for the template's
    // index in the launch.
    bool batch_mask = false;
    std::vector<std::pair<std::string, int>> batch_mask_tables;  // LDS u64 tables: name, entries
    std::string batch_mask_build, batch_mask_body;
    // joins whose probe batch_mask_body_deferred makes itself, after the
    // other joins' masks and only for items some query still accepts
    // (tables too large for LDS, so each skipped item saves an L2 gather);
    // a launch of this template alone uses that body and emits no shared
    // probe part for them (empty: no such join)
    std::vector<int> batch_mask_deferred;
    std::string batch_mask_body_deferred;
    std::string batch_shared;
    std::vector<std::string> batch_shared_parts;  // batch_shared per join (declarations + loops)
    // (scope of the batch kernel: cq_hbits = log2 slots of the query's LDS
    // hash table).  batch_hash_slot_bytes > 0 marks a hash region whose size
    // the launch chooses (stride = slot bytes << hbits, at least
    // batch_lds_stride).
    std::string batch_lds_init, batch_lds_flush;
    uint64_t batch_lds_stride = 0;
    uint64_t batch_hash_slot_bytes = 0;
    bool keys_kernel = false;             // cq_keys writes group keys after the scan
    // interpreted tier (CodegenOptions::vm): the program and the dynamic LDS
    // of its group aggregates (0: global atomics)
    std::shared_ptr<cq_vm_program> vm;
    uint64_t vm_lds_bytes = 0;
};

// One hipRTC module with `cq_batch_scan`: the union of the templates' fact
// columns is loaded once per tile and each template's body runs for each of
// its queries (a runtime loop over parameter blocks).  Templates must have
// been generated with CodegenOptions::batch_body.  `union_cols` receives the
// union slot order and `union_nullable` which of them carry validity.
// `table_lds_bytes` (optional) receives the static LDS of the shared
// dimension tables staged into LDS (at most table_lds_limit).
// Placement: the shared tables that fit LDS are candidates, ranked by the
// templates that probe them, then by size; candidate k is staged into LDS
// only when bit k of `table_mask` is set (and it still fits), else it is
// probed in HBM.  `lds_candidates` (optional) receives the candidate count.
// `table_lds_limit`: the static LDS the staged tables and query-mask
// tables may take (default kBatchTableLdsBytes, at most
// kBatchTableLdsMax); the per-query regions get what the block has left.
constexpr unsigned kBatchTableLdsBytes = 32 * 1024;
constexpr unsigned kBatchTableLdsMax = 96 * 1024;
std::string generateBatchScanSource(const std::vector<const QueryProgram*>& templates, int block_threads,
                                    int vectors_per_thread, std::vector<std::string>& union_cols,
                                    std::vector<bool>& union_nullable, unsigned* table_lds_bytes = nullptr,
                                    unsigned table_mask = ~0u, int* lds_candidates = nullptr,
                                    unsigned table_lds_limit = kBatchTableLdsBytes);

// Visitor entry points as in the reference: AggregateNode / SortLimitNode
// roots start generation (visitor.cpp:1589-1621); projections are rejected.
class HipOperatorVisitor : public OperatorVisitor {
public:
    HipOperatorVisitor(const Catalog& catalog, CodegenOptions options = {});
    void visit(const TableScanNode& node) override;
    void visit(const FilterNode& node) override;
    void visit(const HashJoinNode& node) override;
    void visit(const AggregateNode& node) override;
    void visit(const ProjectionNode& node) override;
    void visit(const SortLimitNode& node) override;
    QueryProgram takeProgram();

private:
    const Catalog& catalog_;
    CodegenOptions options_;
    QueryProgram program_;
    bool done_ = false;
};

QueryProgram generateQueryProgram(const OperatorNode* root, const Catalog& catalog, const CodegenOptions& options = {});

// 64-bit FNV-1a of a string, hex encoded (module cache key).
std::string signatureOf(const std::string& text);

}  // namespace db
