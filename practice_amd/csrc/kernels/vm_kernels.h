// Host launchers of the interpreted tier (vm_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "../device/query_params.h"
#include "../device/vm_program.h"

// V: the program in device memory
hipError_t cq_vm_launch_build(const cq_query_params* P, const cq_vm_program* V, int join, long long dim_rows, hipStream_t s);
// lds_group_bytes: dynamic LDS for the group aggregates (0: global atomics)
hipError_t cq_vm_launch_scan(const cq_query_params* P, const cq_vm_program* V, unsigned lds_group_bytes, unsigned grid,
                             hipStream_t s);
hipError_t cq_vm_launch_dense_count(const cq_query_params* P, const cq_vm_program* V, const unsigned long long* R,
                                    const unsigned long long* RV, unsigned long long* counts, long long G, hipStream_t s);
int cq_vm_block_threads();
// 4 rows per thread: programs with at most 16 registers whose scan uses no
// build / HAVING instructions (V: host copy)
bool cq_vm_scan4_supported(const cq_vm_program* V);
// row_begin must be a multiple of 4; nregs: the program's register count;
// the grid fills the device
hipError_t cq_vm_launch_scan4(const cq_query_params* P, const cq_vm_program* V, int nregs, unsigned lds_group_bytes, int device,
                              hipStream_t s);
// the plan's linear form (cq_lin_program) when it has one (V: host copy)
bool cq_lin_supported(const cq_vm_program* V);
// row_begin must be a multiple of cq_lin_row_align(); the grid fills the
// device (V: the program in device memory, host: its host copy, which
// picks the kernel instantiation)
int cq_lin_row_align();
// grid_percent: the grid as a percentage of a full-device persistent grid
hipError_t cq_lin_launch_scan(const cq_query_params* P, const cq_vm_program* V, const cq_vm_program* host, unsigned lds_group_bytes,
                              int device, int grid_percent, hipStream_t s);
// Several linear-form programs over the same rows in one pass (B: the
// descriptor, V / P: B.nprog programs, whose linear forms index B's union
// columns, and their parameter blocks, all in device memory).  Dynamic LDS:
// cq_lin_batch_stage_bytes(ncols) of staged columns, then rest_bytes of
// join tables and group regions at the descriptor's offsets.  rows:
// B.row_end - B.row_begin (row_begin a multiple of cq_lin_row_align()).
unsigned cq_lin_batch_stage_bytes(int ncols);
hipError_t cq_lin_launch_batch(const cq_lin_batch* B, const cq_vm_program* V, const cq_query_params* P, int ncols, unsigned rest_bytes,
                               long long rows, int device, int grid_percent, hipStream_t s);
