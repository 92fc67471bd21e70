/*
 * cq_query_params.h -- kernel argument block of every generated module.
 *
 * Shared by the host runtime, every generated (hipRTC) or ahead-of-time
 * kernel and any caller that launches a module itself through cq_launch.
 * Plain C layout: the host fills it and passes it by value (kernarg segment,
 * < 4 KB), the device reads it.
 *
 * Replaces the reference's ExecutionContext lookups inside the generated
 * execute_query() (ctx->getBuffer<int>("d_<col>"), getNullBitmap,
 * getResultPointer, getScratchIntBuffer; include/core/execution.h:163-212):
 * the code generator assigns each column a slot and the caller binds device
 * pointers to slots.  cq_db_describe lists the slots of a plan: "fact_cols"
 * in col[] order, "dim_cols" in dcol[] order, "joins" in ht[] order with the
 * table's "bytes", "lits" in lit[] order, "groups" x "storage_ts" u64 result
 * slots, "block_threads" x 4 x "vectors_per_thread" rows per tile.
 */
#ifndef CQ_QUERY_PARAMS_H
#define CQ_QUERY_PARAMS_H

#define CQ_MAX_FACT_COLS 16
#define CQ_MAX_DIM_COLS 24
#define CQ_MAX_JOINS 8
#define CQ_MAX_LITS 48
#define CQ_MAX_FLITS 8

struct cq_query_params {
    // fact (probe-side) columns by slot, int32
    const int* col[CQ_MAX_FACT_COLS];
    // LSB-first validity bitmaps of the fact columns viewed as 32-bit words
    // (bit r of the u64 bitmap is bit r%32 of word r/32); null = no bitmap
    const unsigned int* valid[CQ_MAX_FACT_COLS];
    // dimension (build-side) columns by slot
    const int* dcol[CQ_MAX_DIM_COLS];
    const unsigned int* dvalid[CQ_MAX_DIM_COLS];
    // hash tables, one per join (layout chosen by the code generator)
    unsigned int* ht[CQ_MAX_JOINS];
    long long dim_rows[CQ_MAX_JOINS];
    // aggregate state: u64 slots [group * storage_tuple + slot] and a bitmap
    // with one validity bit per slot (reference d_result / d_result_validity)
    unsigned long long* result;
    unsigned long long* result_valid;
    unsigned long long* tile_count;   // projection: rows per (tile, vector, wave)
    unsigned long long* tile_off;     // projection: exclusive scan of tile_count
    // predicate constants (runtime parameters, so one compiled module serves
    // every query of the same plan signature)
    long long lit[CQ_MAX_LITS];
    double flit[CQ_MAX_FLITS];
    // fact rows [row_begin, row_end) scanned by this launch; row_begin % 4 == 0
    long long row_begin;
    long long row_end;
    long long num_tiles;
};

#define CQ_MAX_BATCH_COLS 24
#define CQ_MAX_TEMPLATES 16

// Kernel arguments of the batched shared scan: the union of fact columns is
// loaded once per tile; template t runs its pipeline for queries
// [tq_begin[t], tq_end[t]) of the parameter-block array `qp` (device memory).
struct cq_batch_params {
    const int* col[CQ_MAX_BATCH_COLS];
    const unsigned int* valid[CQ_MAX_BATCH_COLS];
    const struct cq_query_params* qp;
    int tq_begin[CQ_MAX_TEMPLATES];
    int tq_end[CQ_MAX_TEMPLATES];
    // byte offset of each template's per-query LDS aggregation regions in
    // the dynamic shared memory of the launch, the bytes of one query's
    // region, and log2 of the slots of a per-query LDS hash table (templates
    // whose region size is chosen at launch)
    unsigned int lds_off[CQ_MAX_TEMPLATES];
    unsigned int lds_stride[CQ_MAX_TEMPLATES];
    unsigned int hbits[CQ_MAX_TEMPLATES];
    long long row_begin;
    long long row_end;
    long long num_tiles;
};

#endif
