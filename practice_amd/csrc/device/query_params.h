// Device-side name of the kernel argument block (include/cq_query_params.h).
#ifndef CQ_DEVICE_QUERY_PARAMS_H
#define CQ_DEVICE_QUERY_PARAMS_H
#include "cq_query_params.h"

// Result shaping of the queries of a batched shared scan, one launch for
// all of them: per query, where its aggregate state and output columns are
// (device addresses), read by the generated cq_keys_b / cq_dense_count_b and
// by the ahead-of-time batch shaping kernels.
struct cq_shape_desc {
    unsigned long long* res;        // visible tuples, G x vts
    unsigned long long* rv;         // their validity bits
    unsigned long long* fl;         // per-group present flag (0 / 1)
    unsigned long long* o;          // output columns, vts x cap_rows
    unsigned long long* ov;         // output validity, vts x vwords
    unsigned long long* rows;       // result row count
    long long G, cap_rows, vwords;
    int vts;
    int sparse;                     // present = any non-zero cell (no HAVING, no dense count)
    int use_validity;
    unsigned nullable_mask;
};

// Aggregate-state initialization of one query of a batch.
struct cq_init_desc {
    unsigned long long* res;        // storage slots (slots words)
    unsigned long long* rv;         // slot validity ((slots + 63) / 64 words)
    long long slots, G;
    int sts;
    unsigned min_mask;              // storage slots that start at ~0 (MIN)
    unsigned long long count_bits;  // validity word 0 (scalar COUNT of a non-empty scan)
};

#endif
