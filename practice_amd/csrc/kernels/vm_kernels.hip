// Interpreted tier: ahead-of-time kernels that execute the register code of
// device/vm_program.h, so a plan runs without waiting for hipRTC.
//
//   cq_vm_build        one dimension table: build-filter program per row,
//                      insert in the layout the code generator chose
//   cq_vm_scan         the fact pipeline (filters, probes) and aggregation,
//                      one row per thread per tile; the register file lives
//                      in LDS ([register][thread], conflict-free), group
//                      aggregates in LDS when the domain decodes and fits
//   cq_vm_dense_count  present && HAVING per group slot
//
// Every instruction is uniform across the block (the program is read with
// scalar loads), so interpretation costs one switch per instruction per wave.
#include "vm_kernels.h"

#include <algorithm>

#include "../device/crystal_hip.h"

namespace {

using chip::i64;
using chip::u32;
using chip::u64;

constexpr int VM_BT = 128;

__device__ __forceinline__ bool truthy(u64 v, int t) {
    return t == CQ_T_F64 ? __builtin_bit_cast(double, v) != 0.0 : v != 0ull;
}
__device__ __forceinline__ u64 canon(i64 v, int t) {  // keep the C value of type t in 64 bits
    switch (t) {
        case CQ_T_I32: return (u64)(i64)(int)(u32)(u64)v;
        case CQ_T_U32: return (u64)(u32)(u64)v;
        default: return (u64)v;
    }
}
__device__ __forceinline__ double as_f64(u64 v, int t) {
    switch (t) {
        case CQ_T_I32: return (double)(int)(i64)v;
        case CQ_T_U32: return (double)(u32)v;
        case CQ_T_I64: return (double)(i64)v;
        case CQ_T_U64: return (double)v;
        default: return __builtin_bit_cast(double, v);
    }
}
__device__ __forceinline__ u64 conv(u64 v, int to, int from) {
    if (from == CQ_T_F64) {
        const double d = __builtin_bit_cast(double, v);
        switch (to) {
            case CQ_T_I32: return (u64)(i64)(int)d;
            case CQ_T_U32: return (u64)(u32)d;
            case CQ_T_I64: return (u64)(i64)d;
            case CQ_T_U64: return (u64)d;
            default: return v;
        }
    }
    if (to == CQ_T_F64) return __builtin_bit_cast(u64, as_f64(v, from));
    return canon((i64)v, to);  // integer registers already hold the value modulo 2^64
}
__device__ __forceinline__ u64 arith(int op, u64 a, u64 b, int t) {
    if (t == CQ_T_F64) {
        const double x = __builtin_bit_cast(double, a), y = __builtin_bit_cast(double, b);
        double r = 0;
        switch (op) {
            case CQ_OP_ADD: r = x + y; break;
            case CQ_OP_SUB: r = x - y; break;
            case CQ_OP_MUL: r = x * y; break;
            default: r = x / y; break;
        }
        return __builtin_bit_cast(u64, r);
    }
    switch (op) {
        case CQ_OP_ADD: return canon((i64)(a + b), t);
        case CQ_OP_SUB: return canon((i64)(a - b), t);
        case CQ_OP_MUL: return canon((i64)(a * b), t);
        case CQ_OP_DIV:
            switch (t) {
                case CQ_T_I32: return canon(chip::div_((int)(i64)a, (int)(i64)b), t);
                case CQ_T_U32: return canon((i64)chip::div_((u64)(u32)a, (u64)(u32)b), t);
                case CQ_T_I64: return (u64)chip::div_((i64)a, (i64)b);
                default: return chip::div_(a, b);
            }
        default:  // MOD (x % 0 and x % -1 are 0 here)
            switch (t) {
                case CQ_T_I32: { const int x = (int)(i64)a, y = (int)(i64)b; return canon((y == 0 || y == -1) ? 0 : x % y, t); }
                case CQ_T_U32: { const u32 x = (u32)a, y = (u32)b; return y ? (u64)(x % y) : 0ull; }
                case CQ_T_I64: { const i64 x = (i64)a, y = (i64)b; return (y == 0 || y == -1) ? 0ull : (u64)(x % y); }
                default: return b ? a % b : 0ull;
            }
    }
}
__device__ __forceinline__ int compare(int op, u64 a, u64 b, int t) {
    int lt, eq;
    if (t == CQ_T_F64) {
        const double x = __builtin_bit_cast(double, a), y = __builtin_bit_cast(double, b);
        switch (op) {
            case CQ_OP_EQ: return x == y;
            case CQ_OP_NE: return x != y;
            case CQ_OP_LT: return x < y;
            case CQ_OP_LE: return x <= y;
            case CQ_OP_GT: return x > y;
            default: return x >= y;
        }
    }
    if (t == CQ_T_I32 || t == CQ_T_I64) lt = (i64)a < (i64)b;
    else lt = a < b;
    eq = a == b;
    switch (op) {
        case CQ_OP_EQ: return eq;
        case CQ_OP_NE: return !eq;
        case CQ_OP_LT: return lt;
        case CQ_OP_LE: return lt || eq;
        case CQ_OP_GT: return !lt && !eq;
        default: return !lt;
    }
}

// Row context of one interpreted program.
struct RowCtx {
    const cq_query_params* P;
    long long row;                 // fact row (scan) or dimension row (build)
    const u64* R;                  // result slots of the group (HAVING)
    const u64* RV;
    long long slot_base;
    bool in;                       // row exists (loads are guarded)
};

__device__ __forceinline__ bool probe(const cq_vm_join& J, const u32* ht, int fk, int& payload) {
    switch (J.layout) {
        case CQ_L_COMPACT: return chip::ht_probe_compact(ht, fk, J.kmin, J.len, J.eb, J.tb, J.jlo, J.pmin, payload);
        case CQ_L_DENSE:
            if ((J.db & (J.db - 1)) == 0)
                return chip::ht_probe_dense<true>(ht, fk, J.dmin, J.dspan, J.db, J.pbase, J.doff != 0, J.zmode, J.zx, payload, J.cal);
            return chip::ht_probe_dense<false>(ht, fk, J.dmin, J.dspan, J.db, J.pbase, J.doff != 0, J.zmode, J.zx, payload, J.cal);
        default: return chip::ht_probe_ref((const int*)ht, fk, J.kmin, J.len, J.payload_dslot >= 0, payload);
    }
}

// Runs code[begin, end) for this thread's row; registers are RF[r * VM_BT].
// Returns whether the row survived the filters and probes.
__device__ bool run(const cq_vm_program* __restrict__ V, int begin, int end, u64* RF, const RowCtx& c, bool alive) {
    for (int pc = begin; pc < end; ++pc) {
        const cq_vm_instr in = V->code[pc];
        u64* d = RF + (int)in.d * VM_BT;
        const u64 a = RF[(int)in.a * VM_BT], b = RF[(int)in.b * VM_BT];
        switch (in.op) {
            case CQ_OP_COL: *d = c.in ? (u64)(i64)c.P->col[in.imm][c.row] : 0ull; break;
            case CQ_OP_VALID: *d = (c.in && c.P->valid[in.imm]) ? (u64)chip::bit32(c.P->valid[in.imm], c.row) : (c.in ? 1ull : 0ull); break;
            case CQ_OP_DCOL: *d = c.in ? (u64)(i64)c.P->dcol[in.imm][c.row] : 0ull; break;
            case CQ_OP_DVALID: *d = (c.in && c.P->dvalid[in.imm]) ? (u64)chip::bit32(c.P->dvalid[in.imm], c.row) : (c.in ? 1ull : 0ull); break;
            case CQ_OP_LIT: *d = (u64)c.P->lit[in.imm]; break;
            case CQ_OP_FLIT: *d = __builtin_bit_cast(u64, c.P->flit[in.imm]); break;
            case CQ_OP_IMM: *d = canon((i64)(((u64)in.imm2 << 32) | (u64)(u32)in.imm), in.t); break;
            case CQ_OP_SLOT: *d = c.in ? c.R[c.slot_base + in.imm] : 0ull; break;
            case CQ_OP_SLOTV: *d = c.in ? (u64)chip::bit64(c.RV, c.slot_base + in.imm) : 0ull; break;
            case CQ_OP_CONV: *d = conv(a, in.t, in.t2); break;
            case CQ_OP_ADD: case CQ_OP_SUB: case CQ_OP_MUL: case CQ_OP_DIV: case CQ_OP_MOD: *d = arith(in.op, a, b, in.t); break;
            case CQ_OP_EQ: case CQ_OP_NE: case CQ_OP_LT: case CQ_OP_LE: case CQ_OP_GT: case CQ_OP_GE:
                *d = (u64)compare(in.op, a, b, in.t);
                break;
            case CQ_OP_NOT: *d = truthy(a, in.t) ? 0ull : 1ull; break;
            case CQ_OP_AND: *d = (truthy(a, in.t) && truthy(b, in.t2)) ? 1ull : 0ull; break;
            case CQ_OP_OR: *d = (truthy(a, in.t) || truthy(b, in.t2)) ? 1ull : 0ull; break;
            case CQ_OP_SEL: *d = truthy(RF[(int)in.c * VM_BT], in.t2) ? a : b; break;
            case CQ_OP_FILTER: alive = alive && truthy(a, in.t); break;
            case CQ_OP_PROBE: {
                const cq_vm_join& J = V->joins[in.imm];
                int pay = 0;
                if (alive) alive = probe(J, c.P->ht[in.imm], (int)(i64)a, pay);
                *d = (u64)(i64)pay;
                break;
            }
            case CQ_OP_EXIT:
                if (__ballot(alive) == 0ull) return false;
                break;
            case CQ_OP_FCMP:
                if (alive) {
                    const int v = c.P->col[in.imm][c.row];
                    alive = compare(in.t, (u64)(i64)v, (u64)(i64)(int)c.P->lit[in.c], CQ_T_I32) != 0;
                }
                break;
            default: break;
        }
    }
    return alive;
}

__global__ __launch_bounds__(VM_BT) void k_vm_build(const cq_query_params P, const cq_vm_program* __restrict__ V, int j) {
    __shared__ u64 RF_all[CQ_VM_MAX_REGS * VM_BT];
    const cq_vm_join& J = V->joins[j];
    const long long n = P.dim_rows[j];
    const int tid = threadIdx.x;
    u64* RF = RF_all + tid;
    for (long long base = (long long)blockIdx.x * VM_BT; base < n; base += (long long)gridDim.x * VM_BT) {
        RowCtx c{&P, base + tid, nullptr, nullptr, 0, base + tid < n};
        bool alive = run(V, J.build_begin, J.build_end, RF, c, c.in);
        if (!alive) continue;
        if (J.pk_nullable && !chip::bit32(P.dvalid[J.pk_dslot], c.row)) continue;
        const int k = P.dcol[J.pk_dslot][c.row];
        const int pay = J.payload_dslot >= 0 ? P.dcol[J.payload_dslot][c.row] : 0;
        switch (J.layout) {
            case CQ_L_COMPACT: chip::ht_insert_compact(P.ht[j], k, J.kmin, J.len, J.eb, J.tb, J.jlo, J.pmin, pay); break;
            case CQ_L_DENSE: chip::ht_insert_dense(P.ht[j], k, J.dmin, J.dspan, J.db, J.pbase, pay, J.cal); break;
            default: chip::ht_insert_ref((int*)P.ht[j], k, J.kmin, J.len, J.payload_dslot >= 0, pay); break;
        }
    }
}

// scalar slots accumulated per thread in registers by k_vm_scan
constexpr int VM_RACC = 4;
// function of scalar storage slot k: aggregate k's, or SUM for a hidden count
__device__ __forceinline__ int slot_fn(const cq_vm_program* V, int k) {
    return k < V->naggs ? V->agg_fn[k] : CQ_VM_SUM;
}
__device__ __forceinline__ bool slot_is_min(const cq_vm_program* V, int k) { return k < V->naggs && V->agg_fn[k] == CQ_VM_MIN; }
__device__ __forceinline__ u64 combine(u64 acc, u64 x, int fn) {
    if (fn == CQ_VM_MIN) return x < acc ? x : acc;
    if (fn == CQ_VM_MAX) return x > acc ? x : acc;
    return acc + x;
}

__device__ __forceinline__ u64 wave_reduce(u64 v, int fn) {
    return fn == CQ_VM_MIN ? chip::wave_min(v) : (fn == CQ_VM_MAX ? chip::wave_max(v) : chip::wave_sum(v));
}
__device__ __forceinline__ void atomic_fn(u64* p, u64 v, int fn) {
    if (fn == CQ_VM_MIN) atomicMin(p, v);
    else if (fn == CQ_VM_MAX) atomicMax(p, v);
    else atomicAdd(p, v);
}
__device__ __forceinline__ void set_bit_once(u64* w, long long b) {
    u64* x = &w[b >> 6];
    const u64 m = 1ull << (b & 63);
    if (!(*x & m)) atomicOr(x, m);
}

// lds_groups: aggregate slots of every group in LDS (dynamic shared memory,
// G * nas u64 + G u32 flag words: bit 31 hit, bit s validity of slot s)
__global__ __launch_bounds__(VM_BT) void k_vm_scan(const cq_query_params P, const cq_vm_program* __restrict__ V, int lds_groups) {
    __shared__ u64 RF_all[CQ_VM_MAX_REGS * VM_BT];
    __shared__ u64 s_acc[CQ_VM_MAX_AGGS * 2];
    __shared__ u32 s_seen[CQ_VM_MAX_AGGS * 2];
    extern __shared__ u64 s_dyn[];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    u64* RF = RF_all + tid;
    const int ng = V->ngroups, sts = V->storage_ts, nas = sts - ng;
    const long long G = V->groups;
    u64* s_agg = s_dyn;
    u32* s_flag = (u32*)(s_dyn + (lds_groups ? G * nas : 0));
    if (ng == 0) {
        for (int k = tid; k < sts; k += VM_BT) {
            bool is_min = false;
            for (int i = 0; i < V->naggs; ++i) is_min = is_min || (i == k && V->agg_fn[i] == CQ_VM_MIN);
            s_acc[k] = is_min ? ~0ull : 0ull;
            s_seen[k] = 0u;
        }
    } else if (lds_groups) {
        for (long long k = tid; k < G * nas; k += VM_BT) {
            const int s = (int)(k % nas);
            bool is_min = false;
            for (int i = 0; i < V->naggs; ++i) is_min = is_min || (ng + i == ng + s && V->agg_fn[i] == CQ_VM_MIN);
            s_agg[k] = is_min ? ~0ull : 0ull;
        }
        for (long long k = tid; k < G; k += VM_BT) s_flag[k] = 0u;
    }
    __syncthreads();
    // scalar slots 0..VM_RACC-1 accumulate per thread in registers across
    // rows (one wave reduction at the end); later slots per row and wave
    u64 racc[VM_RACC];
    u32 rseen = 0u;
#pragma unroll
    for (int k = 0; k < VM_RACC; ++k) racc[k] = slot_is_min(V, k) ? ~0ull : 0ull;
    const long long nrows = P.row_end - P.row_begin;
    for (long long base = (long long)blockIdx.x * VM_BT; base < nrows; base += (long long)gridDim.x * VM_BT) {
        RowCtx c{&P, P.row_begin + base + tid, nullptr, nullptr, 0, base + tid < nrows};
        const bool alive = run(V, V->scan_begin, V->scan_end, RF, c, c.in);
        if (ng == 0) {
            for (int i = 0; i < V->naggs; ++i) {
                const int fn = V->agg_fn[i];
                const bool v = alive && (V->agg_valid[i] < 0 || truthy(RF[V->agg_valid[i] * VM_BT], CQ_T_U32));
                const u64 x = fn == CQ_VM_COUNT ? 1ull : (V->agg_val[i] >= 0 ? RF[V->agg_val[i] * VM_BT] : 1ull);
                if (i < VM_RACC) {
#pragma unroll
                    for (int k = 0; k < VM_RACC; ++k)
                        if (k == i && v) racc[k] = combine(racc[k], x, fn);
                    if (v) rseen |= 1u << i;
                } else {
                    const u64 id = fn == CQ_VM_MIN ? ~0ull : 0ull;
                    const u64 w = wave_reduce(v ? x : id, fn);
                    const bool any = __ballot(v) != 0ull;
                    if (lane == 0 && any) {
                        atomic_fn(&s_acc[i], w, fn);
                        s_seen[i] = 1u;
                    }
                }
                const int hs = V->agg_hidden[i];
                if (hs >= 0 && hs < VM_RACC) {
#pragma unroll
                    for (int k = 0; k < VM_RACC; ++k)
                        if (k == hs && v) racc[k] += 1ull;
                } else if (hs >= 0) {
                    const u64 cnt = wave_reduce(v ? 1ull : 0ull, CQ_VM_SUM);
                    if (lane == 0 && cnt) atomicAdd(&s_acc[hs], cnt);
                }
            }
            continue;
        }
        if (!alive) continue;
        const int h = (int)(i64)RF[V->group_reg * VM_BT];
        if ((unsigned)h >= (unsigned long long)G) continue;
        if (lds_groups) {
            u32 want = 0x80000000u;
            for (int g = 0; g < ng; ++g)
                if (!V->fast_sparse && (V->key_valid[g] < 0 || truthy(RF[V->key_valid[g] * VM_BT], CQ_T_U32))) want |= 1u << g;
            for (int i = 0; i < V->naggs; ++i) {
                const int fn = V->agg_fn[i];
                const bool v = V->agg_valid[i] < 0 || truthy(RF[V->agg_valid[i] * VM_BT], CQ_T_U32);
                if (v) {
                    const u64 x = fn == CQ_VM_COUNT ? 1ull : RF[V->agg_val[i] * VM_BT];
                    atomic_fn(&s_agg[(long long)h * nas + i], x, fn);
                    if (V->agg_hidden[i] >= 0) atomicAdd(&s_agg[(long long)h * nas + (V->agg_hidden[i] - ng)], 1ull);
                    else if (!V->fast_sparse) want |= 1u << (ng + i);
                }
                if (fn == CQ_VM_COUNT && !V->fast_sparse) want |= 1u << (ng + i);
            }
            if ((s_flag[h] & want) != want) atomicOr(&s_flag[h], want);
        } else {
            const long long out = (long long)h * sts;
            for (int g = 0; g < ng; ++g) {
                P.result[out + g] = RF[V->key_val[g] * VM_BT];
                if (!V->fast_sparse && (V->key_valid[g] < 0 || truthy(RF[V->key_valid[g] * VM_BT], CQ_T_U32)))
                    set_bit_once(P.result_valid, out + g);
            }
            for (int i = 0; i < V->naggs; ++i) {
                const int fn = V->agg_fn[i];
                const int slot = ng + i;
                const bool v = V->agg_valid[i] < 0 || truthy(RF[V->agg_valid[i] * VM_BT], CQ_T_U32);
                if (v) {
                    const u64 x = fn == CQ_VM_COUNT ? 1ull : RF[V->agg_val[i] * VM_BT];
                    atomic_fn(&P.result[out + slot], x, fn);
                    if (V->agg_hidden[i] >= 0) atomicAdd(&P.result[out + V->agg_hidden[i]], 1ull);
                    else if (fn != CQ_VM_COUNT && !V->fast_sparse) set_bit_once(P.result_valid, out + slot);
                }
                if (fn == CQ_VM_COUNT && !V->fast_sparse) set_bit_once(P.result_valid, out + slot);
            }
        }
    }
    if (ng == 0) {
#pragma unroll
        for (int k = 0; k < VM_RACC; ++k) {
            if (k >= sts) break;
            const int fn = slot_fn(V, k);
            const u64 w = wave_reduce(racc[k], fn);
            const bool any = __ballot((rseen >> k) & 1u) != 0ull;
            const bool nonid = fn == CQ_VM_MIN ? w != ~0ull : w != 0ull;
            if (lane == 0 && nonid) atomic_fn(&s_acc[k], w, fn);
            if (lane == 0 && any) s_seen[k] = 1u;
        }
    }
    __syncthreads();
    if (ng == 0) {
        for (int k = tid; k < sts; k += VM_BT) {
            int fn = CQ_VM_SUM;
            bool hidden = false, scalar_valid_bit = false;
            for (int i = 0; i < V->naggs; ++i) {
                if (i == k) {
                    fn = V->agg_fn[i];
                    scalar_valid_bit = fn != CQ_VM_COUNT && V->agg_hidden[i] < 0;
                }
                hidden = hidden || V->agg_hidden[i] == k;
            }
            const u64 x = s_acc[k];
            if (fn == CQ_VM_MIN) { if (x != ~0ull) atomicMin(&P.result[k], x); }
            else if (fn == CQ_VM_MAX) { if (x) atomicMax(&P.result[k], x); }
            else if (x) atomicAdd(&P.result[k], x);
            if (!hidden && scalar_valid_bit && s_seen[k]) chip::set_bit64(P.result_valid, k);
        }
        return;
    }
    if (!lds_groups) return;
    for (long long g = tid; g < G; g += VM_BT) {
        const u32 f = s_flag[g];
        if (!(f & 0x80000000u)) continue;
        const long long base = g * sts;
        long long rem = g;
        long long d[CQ_VM_MAX_AGGS];
        for (int k = ng - 1; k >= 0; --k) {
            d[k] = rem % V->key_card[k];
            rem /= V->key_card[k];
        }
        for (int k = 0; k < ng; ++k) {
            const long long v = V->key_nullable[k] ? (d[k] == 0 ? 0ll : V->key_min[k] + d[k] - 1) : V->key_min[k] + d[k];
            P.result[base + k] = (u64)v;
        }
        for (int s = 0; s < nas; ++s) {
            int fn = CQ_VM_SUM;
            for (int i = 0; i < V->naggs; ++i)
                if (i == s) fn = V->agg_fn[i];
            const u64 x = s_agg[g * nas + s];
            if (fn == CQ_VM_MIN) { if (x != ~0ull) atomicMin(&P.result[base + ng + s], x); }
            else if (fn == CQ_VM_MAX) { if (x) atomicMax(&P.result[base + ng + s], x); }
            else if (x) atomicAdd(&P.result[base + ng + s], x);
        }
        if (!V->fast_sparse)
            for (int s = 0; s < sts; ++s)
                if ((f >> s) & 1u) chip::set_bit64(P.result_valid, base + s);
    }
}

__global__ __launch_bounds__(VM_BT) void k_vm_dense_count(const cq_query_params P, const cq_vm_program* __restrict__ V,
                                                          const u64* __restrict__ R, const u64* __restrict__ RV, u64* counts,
                                                          long long G) {
    __shared__ u64 RF_all[CQ_VM_MAX_REGS * VM_BT];
    const int tid = threadIdx.x;
    u64* RF = RF_all + tid;
    const long long g = (long long)blockIdx.x * VM_BT + tid;
    const int vts = V->visible_ts;
    const long long base = g * vts;
    RowCtx c{&P, g, R, RV, base, g < G};
    bool present = V->ngroups == 0;
    if (c.in)
        for (int s = 0; s < vts && !present; ++s) present = chip::bit64(RV, base + s) != 0u || R[base + s] != 0ull;
    bool keep = present;
    if (V->having_reg >= 0) {
        run(V, V->having_begin, V->having_end, RF, c, c.in);
        keep = keep && truthy(RF[V->having_reg * VM_BT], V->having_type);
    }
    if (c.in) counts[g] = keep ? 1ull : 0ull;
}

// ---------------------------------------------------------------- 4-row scan
// The same interpreter with 4 consecutive rows per thread (one 16-byte load
// per column): decoding and dispatch (switches on the wave-uniform opcode
// and types, resolved to straight-line code per case) are paid once per
// instruction for 4 rows, and the 4 register accesses of an operand are
// independent LDS operations.  The register file is dynamic LDS sized to
// the program ([register][row][thread] u64, conflict-free), so occupancy
// follows the program's register count.  Programs with more than VM4_REGS
// registers, more scalar slots than VM_RACC, or instructions that only
// build / HAVING programs use run on k_vm_scan.
constexpr int VM4_BT = 128;
constexpr int VM4_REGS = 16;

// register r of row k of this thread (RFt = RF + tid)
struct VmReg {
    u64* base;
    __device__ __forceinline__ u64& operator[](int r) const { return base[r * (4 * VM4_BT)]; }
};

#define VM4_EACH(stmt)                                   \
    {                                                    \
        _Pragma("unroll") for (int k = 0; k < 4; ++k) {  \
            const VmReg R{RFt + k * VM4_BT};             \
            stmt;                                        \
        }                                                \
    }

// dispatch on a wave-uniform type to code with that type as a constant
#define VM4_TYPES(tv, M)                      \
    switch (tv) {                             \
        case CQ_T_I32: M(CQ_T_I32) break;     \
        case CQ_T_U32: M(CQ_T_U32) break;     \
        case CQ_T_I64: M(CQ_T_I64) break;     \
        case CQ_T_U64: M(CQ_T_U64) break;     \
        default: M(CQ_T_F64) break;           \
    }
#define VM4_ARITH_ADD(T) VM4_EACH(R[d] = arith(CQ_OP_ADD, R[a], R[b], T))
#define VM4_ARITH_SUB(T) VM4_EACH(R[d] = arith(CQ_OP_SUB, R[a], R[b], T))
#define VM4_ARITH_MUL(T) VM4_EACH(R[d] = arith(CQ_OP_MUL, R[a], R[b], T))
#define VM4_ARITH_DIV(T) VM4_EACH(R[d] = arith(CQ_OP_DIV, R[a], R[b], T))
#define VM4_ARITH_MOD(T) VM4_EACH(R[d] = arith(CQ_OP_MOD, R[a], R[b], T))
#define VM4_CMP_EQ(T) VM4_EACH(R[d] = (u64)compare(CQ_OP_EQ, R[a], R[b], T))
#define VM4_CMP_NE(T) VM4_EACH(R[d] = (u64)compare(CQ_OP_NE, R[a], R[b], T))
#define VM4_CMP_LT(T) VM4_EACH(R[d] = (u64)compare(CQ_OP_LT, R[a], R[b], T))
#define VM4_CMP_LE(T) VM4_EACH(R[d] = (u64)compare(CQ_OP_LE, R[a], R[b], T))
#define VM4_CMP_GT(T) VM4_EACH(R[d] = (u64)compare(CQ_OP_GT, R[a], R[b], T))
#define VM4_CMP_GE(T) VM4_EACH(R[d] = (u64)compare(CQ_OP_GE, R[a], R[b], T))
#define VM4_CONV_FROM_F64(T) VM4_EACH(R[d] = conv(R[a], T, CQ_T_F64))
#define VM4_CONV_TO_F64(T) VM4_EACH(R[d] = conv(R[a], CQ_T_F64, T))

// one row's aggregation after the pipeline (k_vm_scan's epilogue per row)
template <typename Regs>
__device__ __forceinline__ void vm4_aggregate_row(const cq_query_params& P, const cq_vm_program* __restrict__ V, const Regs& R,
                                                  bool alive, int lds_groups, u64* s_agg, u32* s_flag, u64 (&racc)[VM_RACC],
                                                  u32& rseen) {
    const int ng = V->ngroups, sts = V->storage_ts, nas = sts - ng;
    const long long G = V->groups;
    if (!alive) return;
    if (ng == 0) {
        for (int i = 0; i < V->naggs && i < VM_RACC; ++i) {
            const int fn = V->agg_fn[i];
            const bool v = V->agg_valid[i] < 0 || truthy(R[V->agg_valid[i]], CQ_T_U32);
            if (!v) continue;
            const u64 x = fn == CQ_VM_COUNT ? 1ull : (V->agg_val[i] >= 0 ? R[V->agg_val[i]] : 1ull);
#pragma unroll
            for (int k = 0; k < VM_RACC; ++k)
                if (k == i) racc[k] = combine(racc[k], x, fn);
            rseen |= 1u << i;
            const int hs = V->agg_hidden[i];
#pragma unroll
            for (int k = 0; k < VM_RACC; ++k)
                if (k == hs) racc[k] += 1ull;
        }
        return;
    }
    const int h = (int)(i64)R[V->group_reg];
    if ((unsigned)h >= (unsigned long long)G) return;
    if (lds_groups) {
        u32 want = 0x80000000u;
        for (int g = 0; g < ng; ++g)
            if (!V->fast_sparse && (V->key_valid[g] < 0 || truthy(R[V->key_valid[g]], CQ_T_U32))) want |= 1u << g;
        for (int i = 0; i < V->naggs; ++i) {
            const int fn = V->agg_fn[i];
            const bool v = V->agg_valid[i] < 0 || truthy(R[V->agg_valid[i]], CQ_T_U32);
            if (v) {
                const u64 x = fn == CQ_VM_COUNT ? 1ull : R[V->agg_val[i]];
                atomic_fn(&s_agg[(long long)h * nas + i], x, fn);
                if (V->agg_hidden[i] >= 0) atomicAdd(&s_agg[(long long)h * nas + (V->agg_hidden[i] - ng)], 1ull);
                else if (!V->fast_sparse) want |= 1u << (ng + i);
            }
            if (fn == CQ_VM_COUNT && !V->fast_sparse) want |= 1u << (ng + i);
        }
        if ((s_flag[h] & want) != want) atomicOr(&s_flag[h], want);
        return;
    }
    const long long out = (long long)h * sts;
    for (int g = 0; g < ng; ++g) {
        P.result[out + g] = R[V->key_val[g]];
        if (!V->fast_sparse && (V->key_valid[g] < 0 || truthy(R[V->key_valid[g]], CQ_T_U32)))
            set_bit_once(P.result_valid, out + g);
    }
    for (int i = 0; i < V->naggs; ++i) {
        const int fn = V->agg_fn[i];
        const int slot = ng + i;
        const bool v = V->agg_valid[i] < 0 || truthy(R[V->agg_valid[i]], CQ_T_U32);
        if (v) {
            const u64 x = fn == CQ_VM_COUNT ? 1ull : R[V->agg_val[i]];
            atomic_fn(&P.result[out + slot], x, fn);
            if (V->agg_hidden[i] >= 0) atomicAdd(&P.result[out + V->agg_hidden[i]], 1ull);
            else if (fn != CQ_VM_COUNT && !V->fast_sparse) set_bit_once(P.result_valid, out + slot);
        }
        if (fn == CQ_VM_COUNT && !V->fast_sparse) set_bit_once(P.result_valid, out + slot);
    }
}

// Block-level aggregation state shared by the 4-row kernels: scalar slots
// in s_acc / s_seen, or G groups of nas slots in s_agg with a flag word per
// group in s_flag (bit 31 hit, bit s validity of slot s).
template <int BT>
__device__ __forceinline__ void agg_init(const cq_vm_program* __restrict__ V, int lds_groups, u64* s_acc, u32* s_seen, u64* s_agg,
                                         u32* s_flag) {
    const int tid = threadIdx.x;
    const int ng = V->ngroups, sts = V->storage_ts, nas = sts - ng;
    const long long G = V->groups;
    if (ng == 0) {
        for (int k = tid; k < sts; k += BT) {
            s_acc[k] = slot_is_min(V, k) ? ~0ull : 0ull;
            s_seen[k] = 0u;
        }
    } else if (lds_groups) {
        for (long long k = tid; k < G * nas; k += BT) s_agg[k] = slot_is_min(V, (int)(k % nas)) ? ~0ull : 0ull;
        for (long long k = tid; k < G; k += BT) s_flag[k] = 0u;
    }
    __syncthreads();
}

// A block's scalar slots (s_acc / s_seen, filled by its waves) into the
// result: one atomic per slot, validity of the slots that saw an input.
template <int BT>
__device__ __forceinline__ void flush_scalar_slots(const cq_query_params& P, const cq_vm_program* __restrict__ V, const u64* s_acc,
                                                   const u32* s_seen) {
    const int sts = V->storage_ts;
    for (int k = threadIdx.x; k < sts; k += BT) {
        const int fn = slot_fn(V, k);
        bool hidden = false;
        for (int i = 0; i < V->naggs; ++i) hidden = hidden || V->agg_hidden[i] == k;
        const bool scalar_valid_bit = k < V->naggs && fn != CQ_VM_COUNT && V->agg_hidden[k] < 0;
        const u64 x = s_acc[k];
        if (fn == CQ_VM_MIN) { if (x != ~0ull) atomicMin(&P.result[k], x); }
        else if (fn == CQ_VM_MAX) { if (x) atomicMax(&P.result[k], x); }
        else if (x) atomicAdd(&P.result[k], x);
        if (!hidden && scalar_valid_bit && s_seen[k]) chip::set_bit64(P.result_valid, k);
    }
}

// A block's LDS groups into the result: keys decoded from the group index,
// one atomic per touched slot, validity bits from the flag words.
template <int BT>
__device__ __forceinline__ void flush_lds_groups(const cq_query_params& P, const cq_vm_program* __restrict__ V, const u64* s_agg,
                                                 const u32* s_flag) {
    const int ng = V->ngroups, sts = V->storage_ts, nas = sts - ng;
    const long long G = V->groups;
    for (long long g = threadIdx.x; g < G; g += BT) {
        const u32 f = s_flag[g];
        if (!(f & 0x80000000u)) continue;
        const long long base = g * sts;
        long long rem = g;
        long long dg[CQ_VM_MAX_AGGS];
        for (int k = ng - 1; k >= 0; --k) {
            dg[k] = rem % V->key_card[k];
            rem /= V->key_card[k];
        }
        for (int k = 0; k < ng; ++k) {
            const long long v = V->key_nullable[k] ? (dg[k] == 0 ? 0ll : V->key_min[k] + dg[k] - 1) : V->key_min[k] + dg[k];
            P.result[base + k] = (u64)v;
        }
        for (int s2 = 0; s2 < nas; ++s2) {
            const int fn = s2 < V->naggs ? V->agg_fn[s2] : CQ_VM_SUM;
            const u64 x = s_agg[g * nas + s2];
            if (fn == CQ_VM_MIN) { if (x != ~0ull) atomicMin(&P.result[base + ng + s2], x); }
            else if (fn == CQ_VM_MAX) { if (x) atomicMax(&P.result[base + ng + s2], x); }
            else if (x) atomicAdd(&P.result[base + ng + s2], x);
        }
        if (!V->fast_sparse)
            for (int s2 = 0; s2 < sts; ++s2)
                if ((f >> s2) & 1u) chip::set_bit64(P.result_valid, base + s2);
    }
}

// Merge of a block: per-thread scalar slots racc (wave reduction, then one
// atomic per slot and block), or the LDS groups with their keys decoded
// from the group index.
template <int BT, int NR>
__device__ __forceinline__ void agg_flush(const cq_query_params& P, const cq_vm_program* __restrict__ V, int lds_groups, u64* s_acc,
                                          u32* s_seen, const u64* s_agg, const u32* s_flag, const u64 (&racc)[NR], u32 rseen) {
    const int lane = threadIdx.x & 63;
    const int ng = V->ngroups, sts = V->storage_ts;
    if (ng == 0) {
#pragma unroll
        for (int k = 0; k < NR; ++k) {
            if (k >= sts) break;
            const int fn = slot_fn(V, k);
            const u64 w = wave_reduce(racc[k], fn);
            const bool any = __ballot((rseen >> k) & 1u) != 0ull;
            const bool nonid = fn == CQ_VM_MIN ? w != ~0ull : w != 0ull;
            if (lane == 0 && nonid) atomic_fn(&s_acc[k], w, fn);
            if (lane == 0 && any) s_seen[k] = 1u;
        }
    }
    __syncthreads();
    if (ng == 0) {
        flush_scalar_slots<BT>(P, V, s_acc, s_seen);
        return;
    }
    if (lds_groups) flush_lds_groups<BT>(P, V, s_agg, s_flag);
}

// dynamic LDS: register file (nregs x 4 x VM4_BT u64), then the group
// aggregates when lds_groups (G * nas u64 + G u32 flag words)
__global__ __launch_bounds__(VM4_BT) void k_vm_scan4(const cq_query_params P, const cq_vm_program* __restrict__ V, int lds_groups) {
    __shared__ u64 s_acc[CQ_VM_MAX_AGGS * 2];
    __shared__ u32 s_seen[CQ_VM_MAX_AGGS * 2];
    extern __shared__ u64 s_dyn[];
    const int tid = threadIdx.x;
    const int ng = V->ngroups, sts = V->storage_ts, nas = sts - ng;
    const long long G = V->groups;
    u64* RFt = s_dyn + tid;
    u64* s_agg = s_dyn + (long long)V->nregs * 4 * VM4_BT;
    u32* s_flag = (u32*)(s_agg + (lds_groups ? G * nas : 0));
    agg_init<VM4_BT>(V, lds_groups, s_acc, s_seen, s_agg, s_flag);
    u64 racc[VM_RACC];
    u32 rseen = 0u;
#pragma unroll
    for (int k = 0; k < VM_RACC; ++k) racc[k] = slot_is_min(V, k) ? ~0ull : 0ull;
    for (int r = 0; r < V->nregs; ++r) VM4_EACH(R[r] = 0ull)
    const long long nrows = P.row_end - P.row_begin;
    const int code_end = V->scan_end;
    for (long long base = (long long)blockIdx.x * (VM4_BT * 4); base < nrows; base += (long long)gridDim.x * (VM4_BT * 4)) {
        const long long off = base + (long long)tid * 4;  // first of this thread's 4 rows
        const long long row0 = P.row_begin + off;          // multiple of 4
        const bool full = off + 4 <= nrows;
        bool in[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) in[k] = off + k < nrows;
        bool alive[4] = {in[0], in[1], in[2], in[3]};
        for (int pc = V->scan_begin; pc < code_end; ++pc) {
            const cq_vm_instr ins = V->code[pc];
            const int d = ins.d, a = ins.a, b = ins.b, t = ins.t, t2 = ins.t2;
            switch (ins.op) {
                case CQ_OP_COL: {
                    const int* col = P.col[ins.imm];
                    if (full) {
                        const chip::v4i x = *(const chip::v4i*)(col + row0);
                        RFt[(d * 4 + 0) * VM4_BT] = (u64)(i64)x.x;
                        RFt[(d * 4 + 1) * VM4_BT] = (u64)(i64)x.y;
                        RFt[(d * 4 + 2) * VM4_BT] = (u64)(i64)x.z;
                        RFt[(d * 4 + 3) * VM4_BT] = (u64)(i64)x.w;
                    } else {
                        VM4_EACH(R[d] = in[k] ? (u64)(i64)col[row0 + k] : 0ull)
                    }
                    break;
                }
                case CQ_OP_VALID: {
                    const u32* vb = P.valid[ins.imm];
                    // 4 rows from a multiple of 4: bits of one 32-bit word
                    const u32 w = (vb && off < nrows) ? vb[row0 >> 5] >> (row0 & 31) : 0xfu;
                    VM4_EACH(R[d] = in[k] ? (u64)((w >> k) & 1u) : 0ull)
                    break;
                }
                case CQ_OP_LIT: VM4_EACH(R[d] = (u64)P.lit[ins.imm]) break;
                case CQ_OP_FLIT: VM4_EACH(R[d] = __builtin_bit_cast(u64, P.flit[ins.imm])) break;
                case CQ_OP_IMM: {
                    const u64 v = canon((i64)(((u64)ins.imm2 << 32) | (u64)(u32)ins.imm), t);
                    VM4_EACH(R[d] = v)
                    break;
                }
                // every op / type case below is straight-line code for its
                // constant op and types (the helpers fold), so no lane
                // evaluates another case's arithmetic
                case CQ_OP_CONV:
                    if (t2 == CQ_T_F64) {
                        VM4_TYPES(t, VM4_CONV_FROM_F64)
                    } else if (t == CQ_T_F64) {
                        VM4_TYPES(t2, VM4_CONV_TO_F64)
                    } else if (t == CQ_T_I32) {
                        VM4_EACH(R[d] = canon((i64)R[a], CQ_T_I32))
                    } else if (t == CQ_T_U32) {
                        VM4_EACH(R[d] = canon((i64)R[a], CQ_T_U32))
                    } else {
                        VM4_EACH(R[d] = R[a])
                    }
                    break;
                case CQ_OP_ADD: VM4_TYPES(t, VM4_ARITH_ADD) break;
                case CQ_OP_SUB: VM4_TYPES(t, VM4_ARITH_SUB) break;
                case CQ_OP_MUL: VM4_TYPES(t, VM4_ARITH_MUL) break;
                case CQ_OP_DIV: VM4_TYPES(t, VM4_ARITH_DIV) break;
                case CQ_OP_MOD: VM4_TYPES(t, VM4_ARITH_MOD) break;
                case CQ_OP_EQ: VM4_TYPES(t, VM4_CMP_EQ) break;
                case CQ_OP_NE: VM4_TYPES(t, VM4_CMP_NE) break;
                case CQ_OP_LT: VM4_TYPES(t, VM4_CMP_LT) break;
                case CQ_OP_LE: VM4_TYPES(t, VM4_CMP_LE) break;
                case CQ_OP_GT: VM4_TYPES(t, VM4_CMP_GT) break;
                case CQ_OP_GE: VM4_TYPES(t, VM4_CMP_GE) break;
                case CQ_OP_NOT:
                    if (t == CQ_T_F64) VM4_EACH(R[d] = truthy(R[a], CQ_T_F64) ? 0ull : 1ull)
                    else VM4_EACH(R[d] = R[a] != 0ull ? 0ull : 1ull)
                    break;
                case CQ_OP_AND:
                    if (t != CQ_T_F64 && t2 != CQ_T_F64) VM4_EACH(R[d] = (R[a] != 0ull && R[b] != 0ull) ? 1ull : 0ull)
                    else VM4_EACH(R[d] = (truthy(R[a], t) && truthy(R[b], t2)) ? 1ull : 0ull)
                    break;
                case CQ_OP_OR:
                    if (t != CQ_T_F64 && t2 != CQ_T_F64) VM4_EACH(R[d] = (R[a] != 0ull || R[b] != 0ull) ? 1ull : 0ull)
                    else VM4_EACH(R[d] = (truthy(R[a], t) || truthy(R[b], t2)) ? 1ull : 0ull)
                    break;
                case CQ_OP_SEL: {
                    const int c = ins.c;
                    if (t2 != CQ_T_F64) VM4_EACH(R[d] = R[c] != 0ull ? R[a] : R[b])
                    else VM4_EACH(R[d] = truthy(R[c], CQ_T_F64) ? R[a] : R[b])
                    break;
                }
                case CQ_OP_FILTER:
                    if (t != CQ_T_F64) VM4_EACH(alive[k] = alive[k] && R[a] != 0ull)
                    else VM4_EACH(alive[k] = alive[k] && truthy(R[a], CQ_T_F64))
                    break;
                case CQ_OP_PROBE: {
                    const cq_vm_join& J = V->joins[ins.imm];
                    const u32* ht = P.ht[ins.imm];
                    switch (J.layout) {
                        case CQ_L_DENSE:
                            if ((J.db & (J.db - 1)) == 0)
                                VM4_EACH({
                                    int pay = 0;
                                    if (alive[k])
                                        alive[k] = chip::ht_probe_dense<true>(ht, (int)(i64)R[a], J.dmin, J.dspan, J.db, J.pbase,
                                                                              J.doff != 0, J.zmode, J.zx, pay, J.cal);
                                    R[d] = (u64)(i64)pay;
                                })
                            else
                                VM4_EACH({
                                    int pay = 0;
                                    if (alive[k])
                                        alive[k] = chip::ht_probe_dense<false>(ht, (int)(i64)R[a], J.dmin, J.dspan, J.db, J.pbase,
                                                                               J.doff != 0, J.zmode, J.zx, pay, J.cal);
                                    R[d] = (u64)(i64)pay;
                                })
                            break;
                        case CQ_L_COMPACT:
                            VM4_EACH({
                                int pay = 0;
                                if (alive[k]) alive[k] = chip::ht_probe_compact(ht, (int)(i64)R[a], J.kmin, J.len, J.eb, J.tb, J.jlo, J.pmin, pay);
                                R[d] = (u64)(i64)pay;
                            })
                            break;
                        default:
                            VM4_EACH({
                                int pay = 0;
                                if (alive[k]) alive[k] = chip::ht_probe_ref((const int*)ht, (int)(i64)R[a], J.kmin, J.len, J.payload_dslot >= 0, pay);
                                R[d] = (u64)(i64)pay;
                            })
                            break;
                    }
                    break;
                }
                case CQ_OP_EXIT:
                    if (__ballot(alive[0] || alive[1] || alive[2] || alive[3]) == 0ull) pc = code_end;
                    break;
                case CQ_OP_FCMP: {
                    const int* col = P.col[ins.imm];
                    const int l = (int)P.lit[ins.c];
                    int x[4];
                    if (full) {
                        const chip::v4i v = *(const chip::v4i*)(col + row0);
                        x[0] = v.x;
                        x[1] = v.y;
                        x[2] = v.z;
                        x[3] = v.w;
                    } else {
#pragma unroll
                        for (int k = 0; k < 4; ++k) x[k] = in[k] ? col[row0 + k] : 0;
                    }
                    switch (t) {
#define VM4_FCMP(OPC, EXPR)                                                   \
    case OPC:                                                                  \
        _Pragma("unroll") for (int k = 0; k < 4; ++k) alive[k] = alive[k] && (EXPR); \
        break;
                        VM4_FCMP(CQ_OP_EQ, x[k] == l)
                        VM4_FCMP(CQ_OP_NE, x[k] != l)
                        VM4_FCMP(CQ_OP_LT, x[k] < l)
                        VM4_FCMP(CQ_OP_LE, x[k] <= l)
                        VM4_FCMP(CQ_OP_GT, x[k] > l)
                        default:
#pragma unroll
                            for (int k = 0; k < 4; ++k) alive[k] = alive[k] && (x[k] >= l);
                            break;
                    }
                    break;
                }
                default: break;  // build / HAVING instructions never reach this kernel
            }
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) vm4_aggregate_row(P, V, VmReg{RFt + k * VM4_BT}, alive[k], lds_groups, s_agg, s_flag, racc, rseen);
    }
    agg_flush<VM4_BT>(P, V, lds_groups, s_acc, s_seen, s_agg, s_flag, racc, rseen);
}

// ---------------------------------------------------------------- linear form
// k_lin_scan evaluates cq_lin_program (a plan's filters, one key-indexed
// probe, group index and aggregate chains) directly, with no register file
// and no per-instruction dispatch.  A wave owns tiles of 512 rows; it
// stages every column of the plan for its tile into its own LDS region with
// direct global-to-LDS loads (global_load_lds_dwordx4: 1 KiB per
// instruction, all of the tile's columns in flight at once, no registers
// held while they fly), then every filter term, the probe key, the group
// keys and the chain operands read their column from LDS at a computed
// address.  Thread `lane` holds rows t0 + k * 256 + lane * 4 + j (k < 2,
// j < 4) of the tile starting at t0, as bits 4k + j of its row masks.  It
// produces the values k_vm_scan4 produces from the same program's bytecode.
constexpr int LIN_BT = 256;
constexpr int LIN_R = 8;                    // rows per thread
constexpr int LIN_TILE = 64 * LIN_R;        // rows per wave tile
constexpr int LIN_COL_INTS = LIN_TILE;      // LDS ints per column per wave
constexpr u32 LIN_ALL = (1u << LIN_R) - 1u;

// one wave tile: its first row, whether all its rows lie in the range
// (wave-uniform), this lane's live-row mask, the range's last row
struct LinRows {
    long long t0, last;
    bool full;
    u32 inm;
    int lane;
};
__device__ __forceinline__ long long lin_row(const LinRows& q, int bit) { return q.t0 + (bit >> 2) * 256 + q.lane * 4 + (bit & 3); }

// stage the tile's rows of a column into the wave's LDS column `wl`
__device__ __forceinline__ void lin_stage(const int* __restrict__ col, int* wl, const LinRows& q) {
    if (q.full) {
        __builtin_amdgcn_global_load_lds((const void*)(col + q.t0 + q.lane * 4), (__attribute__((address_space(3))) void*)wl, 16, 0, 0);
        __builtin_amdgcn_global_load_lds((const void*)(col + q.t0 + 256 + q.lane * 4),
                                         (__attribute__((address_space(3))) void*)(wl + 256), 16, 0, 0);
    } else {
        // the range's last, partial tile: rows past the end read the last row
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            chip::v4i v;
            long long r = q.t0 + k * 256 + q.lane * 4;
            v.x = col[r < q.last ? r : q.last];
            v.y = col[r + 1 < q.last ? r + 1 : q.last];
            v.z = col[r + 2 < q.last ? r + 2 : q.last];
            v.w = col[r + 3 < q.last ? r + 3 : q.last];
            *(chip::v4i*)(wl + k * 256 + q.lane * 4) = v;
        }
    }
}
// this lane's LIN_R values of a staged column
__device__ __forceinline__ void lin_read(const int* wl, const LinRows& q, int (&x)[LIN_R]) {
    const chip::v4i a = *(const chip::v4i*)(wl + q.lane * 4);
    const chip::v4i b = *(const chip::v4i*)(wl + 256 + q.lane * 4);
    x[0] = a.x, x[1] = a.y, x[2] = a.z, x[3] = a.w;
    x[4] = b.x, x[5] = b.y, x[6] = b.z, x[7] = b.w;
}
// validity bits of this lane's rows; a column without a bitmap is valid
// everywhere
__device__ __forceinline__ u32 lin_valid(const u32* vb, const LinRows& q) {
    if (!vb) return q.inm;
    u32 m = 0u;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        long long r = q.t0 + k * 256 + q.lane * 4;  // a multiple of 4: 4 bits of one word
        r = r < q.last ? r : q.last & ~3ll;
        m |= ((vb[r >> 5] >> (r & 31)) & 0xfu) << (4 * k);
    }
    return m & q.inm;
}

// validity bits of column position p from the tile's packed bits
__device__ __forceinline__ u32 lin_bits(u64 vbits, int p) { return (u32)(vbits >> (LIN_R * p)) & LIN_ALL; }

template <bool ALIGNED>
__device__ __forceinline__ u32 lin_probe(const cq_vm_join& J, const u32* ht, const int (&fk)[LIN_R], u32 alive, int (&pay)[LIN_R]) {
#pragma unroll
    for (int r = 0; r < LIN_R; ++r) {
        int pv = 0;
        if ((alive >> r) & 1u) {
            if (!chip::ht_probe_dense<ALIGNED>(ht, fk[r], J.dmin, J.dspan, J.db, J.pbase, J.doff != 0, J.zmode, J.zx, pv, J.cal))
                alive &= ~(1u << r);
        }
        pay[r] = pv;
    }
    return alive;
}

// one aggregate's input for this lane's rows: the left-deep chain in u64
// (column and payload operands stay int until each step converts them)
__device__ __forceinline__ void lin_chain(const cq_query_params& P, const cq_lin_agg& A, int ns, const int* wcols, const LinRows& q,
                                          const int (&pay)[LIN_R], u64 (&x)[LIN_R]) {
#pragma unroll
    for (int r = 0; r < LIN_R; ++r) x[r] = 1ull;
    for (int s = 0; s < ns; ++s) {
        const cq_lin_step& st = A.step[s];
        const int op = s == 0 ? 0 : st.op;
        if (st.kind == CQ_LIN_LIT) {
            const u64 l = st.i32 ? (u64)(i64)(int)P.lit[st.idx] : (u64)P.lit[st.idx];
            switch (op) {
                case CQ_OP_ADD:
#pragma unroll
                    for (int r = 0; r < LIN_R; ++r) x[r] += l;
                    break;
                case CQ_OP_SUB:
#pragma unroll
                    for (int r = 0; r < LIN_R; ++r) x[r] -= l;
                    break;
                case CQ_OP_MUL:
#pragma unroll
                    for (int r = 0; r < LIN_R; ++r) x[r] *= l;
                    break;
                default:
#pragma unroll
                    for (int r = 0; r < LIN_R; ++r) x[r] = l;
                    break;
            }
            continue;
        }
        int v[LIN_R];
        if (st.kind == CQ_LIN_COL) lin_read(wcols + st.idx * LIN_COL_INTS, q, v);
        else
#pragma unroll
            for (int r = 0; r < LIN_R; ++r) v[r] = pay[r];
        switch (op) {
            case CQ_OP_ADD:
#pragma unroll
                for (int r = 0; r < LIN_R; ++r) x[r] += (u64)(i64)v[r];
                break;
            case CQ_OP_SUB:
#pragma unroll
                for (int r = 0; r < LIN_R; ++r) x[r] -= (u64)(i64)v[r];
                break;
            case CQ_OP_MUL:
#pragma unroll
                for (int r = 0; r < LIN_R; ++r) x[r] *= (u64)(i64)v[r];
                break;
            default:
#pragma unroll
                for (int r = 0; r < LIN_R; ++r) x[r] = (u64)(i64)v[r];
                break;
        }
    }
}

// per-row atomics of one aggregate into group slots base[h * stride]
// (rows of vm only; one branch on the function).  Callers pass an LDS or a
// global pointer, never one that may be either, so the atomics are ds_* or
// global_* instructions, not flat ones.
template <typename Index>
__device__ __forceinline__ void lin_group_atomics(u64* base, Index stride, const int (&h)[LIN_R], const u64 (&x)[LIN_R], u32 vm, int fn) {
    switch (fn) {
        case CQ_VM_MIN:
#pragma unroll
            for (int r = 0; r < LIN_R; ++r)
                if ((vm >> r) & 1u) atomicMin(&base[(Index)h[r] * stride], x[r]);
            break;
        case CQ_VM_MAX:
#pragma unroll
            for (int r = 0; r < LIN_R; ++r)
                if ((vm >> r) & 1u) atomicMax(&base[(Index)h[r] * stride], x[r]);
            break;
        default:
#pragma unroll
            for (int r = 0; r < LIN_R; ++r)
                if ((vm >> r) & 1u) atomicAdd(&base[(Index)h[r] * stride], x[r]);
            break;
    }
}

// the rows of this lane that pass program L's fact filters
__device__ __forceinline__ u32 lin_filters(const cq_query_params& P, const cq_lin_program& L, const int* wcols, const LinRows& q,
                                           u32 alive) {
    const int nterms = L.nterms;
    for (int t = 0; t < nterms; ++t) {
        int x[LIN_R];
        lin_read(wcols + L.term[t].col * LIN_COL_INTS, q, x);
        const int l = (int)P.lit[L.term[t].lit];
        u32 m = 0u;
        switch (L.term[t].op) {
#define LIN_CMP(OPC, OPR)                                                              \
    case OPC:                                                                           \
        _Pragma("unroll") for (int r = 0; r < LIN_R; ++r) m |= (u32)(x[r] OPR l) << r; \
        break;
            LIN_CMP(CQ_OP_EQ, ==)
            LIN_CMP(CQ_OP_NE, !=)
            LIN_CMP(CQ_OP_LT, <)
            LIN_CMP(CQ_OP_LE, <=)
            LIN_CMP(CQ_OP_GT, >)
            LIN_CMP(CQ_OP_GE, >=)
#undef LIN_CMP
            default: break;
        }
        alive &= m;
    }
    return alive;
}

// program L's probe of table `tab` (its LDS copy or the HBM table): the
// rows with a match and their payloads
__device__ __forceinline__ u32 lin_join(const cq_vm_program* V, const cq_lin_program& L, const u32* tab, const int* wcols, u64 vbits,
                                        const LinRows& q, u32 alive, int (&pay)[LIN_R]) {
    const cq_vm_join& J = V->joins[L.join];
    int fk[LIN_R];
    lin_read(wcols + L.join_col * LIN_COL_INTS, q, fk);
    alive &= lin_bits(vbits, L.join_col);
    const bool al = (J.db & (J.db - 1)) == 0;
    return al ? lin_probe<true>(J, tab, fk, alive, pay) : lin_probe<false>(J, tab, fk, alive, pay);
}

// scalar aggregates: this lane's rows of the tile folded into racc slot by
// slot (AVG counts into their hidden slots); returns the aggregates that
// saw a valid input
template <int NR>
__device__ __forceinline__ u32 lin_scalar_fold(const cq_query_params& P, const cq_vm_program* V, const cq_lin_program& L,
                                               const int* wcols, u64 vbits, const LinRows& q, const int (&pay)[LIN_R], u32 alive,
                                               u64 (&racc)[NR]) {
    u32 seen = 0u;
    const int naggs = L.naggs;
#pragma unroll
    for (int i = 0; i < CQ_LIN_AGGS; ++i) {
        if (i >= naggs) break;
        const cq_lin_agg& A = L.agg[i];
        const int fn = V->agg_fn[i];
        const int hs = V->agg_hidden[i];
        u32 vm = alive;
        for (int c = 0; c < A.nvalid; ++c) vm &= lin_bits(vbits, A.valid_col[c]);
        u64 part;
        if (fn == CQ_VM_COUNT) {
            part = (u64)__builtin_popcount(vm);
        } else {
            u64 x[LIN_R];
            lin_chain(P, A, A.nsteps, wcols, q, pay, x);
            if (fn == CQ_VM_MIN) {
                part = ~0ull;
#pragma unroll
                for (int r = 0; r < LIN_R; ++r) part = (((vm >> r) & 1u) && x[r] < part) ? x[r] : part;
            } else if (fn == CQ_VM_MAX) {
                part = 0ull;
#pragma unroll
                for (int r = 0; r < LIN_R; ++r) part = (((vm >> r) & 1u) && x[r] > part) ? x[r] : part;
            } else {
                part = 0ull;
#pragma unroll
                for (int r = 0; r < LIN_R; ++r) part += ((vm >> r) & 1u) ? x[r] : 0ull;
            }
        }
        if (fn == CQ_VM_MIN) racc[i] = part < racc[i] ? part : racc[i];
        else if (fn == CQ_VM_MAX) racc[i] = part > racc[i] ? part : racc[i];
        else racc[i] += part;
        seen |= vm ? 1u << i : 0u;
        const u64 cnt = (u64)__builtin_popcount(vm);
#pragma unroll
        for (int k = 0; k < NR; ++k) racc[k] += k == hs ? cnt : 0ull;
    }
    return seen;
}

// GROUP BY: group index of every live row (Horner over dense domains, no
// modulo, in 32 bits: the index is the low word of the u64 Horner value,
// and add, sub and mul commute with taking the low word), then the
// aggregate atomics into LDS (lds_groups) or global group slots, with the
// keys written and validity bits set in global memory in the latter case
__device__ __forceinline__ void lin_group_tile(const cq_query_params& P, const cq_vm_program* V, const cq_lin_program& L,
                                               const int* wcols, u64 vbits, const LinRows& q, const int (&pay)[LIN_R], u32 alive,
                                               int lds_groups, u64* s_agg, u32* s_flag) {
    const int ng = V->ngroups, sts = V->storage_ts, nas = sts - ng;
    const long long G = V->groups;
    const bool fast_sparse = V->fast_sparse != 0;
    const int naggs = L.naggs;
    int h[LIN_R];
    u32 keyok[CQ_LIN_KEYS] = {0u, 0u};
    {
        u32 hv[LIN_R];
#pragma unroll
        for (int r = 0; r < LIN_R; ++r) hv[r] = 0u;
#pragma unroll
        for (int g = 0; g < CQ_LIN_KEYS; ++g) {
            if (g >= ng) break;
            const cq_lin_key& K = L.key[g];
            int kv[LIN_R];
            if (K.kind == CQ_LIN_COL) lin_read(wcols + K.col * LIN_COL_INTS, q, kv);
            else
#pragma unroll
                for (int r = 0; r < LIN_R; ++r) kv[r] = pay[r];
            keyok[g] = K.valid_col >= 0 ? lin_bits(vbits, K.valid_col) : q.inm;
            const u32 mn = (u32)V->key_min[g], card = (u32)V->key_card[g];
            const bool nl = V->key_nullable[g] != 0;
#pragma unroll
            for (int r = 0; r < LIN_R; ++r) {
                const u32 v = (u32)kv[r];
                const u32 term = nl ? (((keyok[g] >> r) & 1u) ? v - mn + 1u : 0u) : v - mn;
                hv[r] = g == 0 ? term : hv[r] * card + term;
            }
        }
#pragma unroll
        for (int r = 0; r < LIN_R; ++r) {
            h[r] = (int)hv[r];
            if ((unsigned)h[r] >= (unsigned long long)G) alive &= ~(1u << r);
        }
    }
    if (__ballot(alive != 0u) == 0ull) return;
    if (!lds_groups) {
        // group rows in global memory: keys and their validity
#pragma unroll
        for (int g = 0; g < CQ_LIN_KEYS; ++g) {
            if (g >= ng) break;
            const cq_lin_key& K = L.key[g];
            int kv[LIN_R];
            if (K.kind == CQ_LIN_COL) lin_read(wcols + K.col * LIN_COL_INTS, q, kv);
            else
#pragma unroll
                for (int r = 0; r < LIN_R; ++r) kv[r] = pay[r];
#pragma unroll
            for (int r = 0; r < LIN_R; ++r) {
                if (!((alive >> r) & 1u)) continue;
                const long long out = (long long)h[r] * sts;
                P.result[out + g] = (u64)(i64)kv[r];
                if (!fast_sparse && ((keyok[g] >> r) & 1u)) set_bit_once(P.result_valid, out + g);
            }
        }
    }
    // slots that turn valid, per aggregate: COUNT on every hit, others
    // on a valid input (AVG keeps its count instead)
    u32 slotbits[CQ_LIN_AGGS] = {0u, 0u, 0u, 0u};
#pragma unroll
    for (int i = 0; i < CQ_LIN_AGGS; ++i) {
        if (i >= naggs) break;
        const cq_lin_agg& A = L.agg[i];
        const int fn = V->agg_fn[i];
        const int hs = V->agg_hidden[i];
        u32 vm = alive;
        for (int c = 0; c < A.nvalid; ++c) vm &= lin_bits(vbits, A.valid_col[c]);
        u64 x[LIN_R];
        lin_chain(P, A, fn == CQ_VM_COUNT ? 0 : A.nsteps, wcols, q, pay, x);
        u64 one[LIN_R];
#pragma unroll
        for (int r = 0; r < LIN_R; ++r) one[r] = 1ull;
        if (lds_groups) {
            lin_group_atomics<int>(s_agg + i, nas, h, x, vm, fn);
            if (hs >= 0) lin_group_atomics<int>(s_agg + (hs - ng), nas, h, one, vm, CQ_VM_SUM);
        } else {
            lin_group_atomics<long long>(P.result + ng + i, (long long)sts, h, x, vm, fn);
            if (hs >= 0) lin_group_atomics<long long>(P.result + hs, (long long)sts, h, one, vm, CQ_VM_SUM);
        }
        slotbits[i] = fast_sparse ? 0u : (fn == CQ_VM_COUNT ? alive : (hs >= 0 ? 0u : vm));
        if (!lds_groups && slotbits[i]) {
#pragma unroll
            for (int r = 0; r < LIN_R; ++r)
                if ((slotbits[i] >> r) & 1u) set_bit_once(P.result_valid, (long long)h[r] * sts + ng + i);
        }
    }
    if (lds_groups) {
#pragma unroll
        for (int r = 0; r < LIN_R; ++r) {
            if (!((alive >> r) & 1u)) continue;
            u32 want = 0x80000000u;
            if (!fast_sparse) {
#pragma unroll
                for (int g = 0; g < CQ_LIN_KEYS; ++g)
                    if (g < ng) want |= ((keyok[g] >> r) & 1u) << g;
#pragma unroll
                for (int k = 0; k < CQ_LIN_AGGS; ++k) want |= ((slotbits[k] >> r) & 1u) << (ng + k);
            }
            if ((s_flag[h[r]] & want) != want) atomicOr(&s_flag[h[r]], want);
        }
    }
}

// the wave tile starting at row tile * LIN_TILE of [row_begin, row_end)
__device__ __forceinline__ void lin_tile_rows(LinRows& q, long long row_begin, long long row_end, long long tile) {
    q.t0 = row_begin + tile * LIN_TILE;
    q.full = q.t0 + LIN_TILE <= row_end;
    q.inm = LIN_ALL;
    if (!q.full) {
        q.inm = 0u;
#pragma unroll
        for (int b = 0; b < LIN_R; ++b) q.inm |= (u32)(lin_row(q, b) <= q.last) << b;
    }
}

// GROUPED: GROUP BY plans (the scalar accumulators exist only without).
// Dynamic LDS: per wave ncols staged columns (LIN_COL_INTS ints each), the
// join table when it is probed from LDS (join_lds_words, rounded up to an
// even count), then the group aggregates when lds_groups (G * nas u64 +
// G u32 flag words).
template <bool GROUPED>
__global__ __launch_bounds__(LIN_BT, GROUPED ? 3 : 4) void k_lin_scan(const cq_query_params P, const cq_vm_program* __restrict__ V, int lds_groups) {
    __shared__ u64 s_acc[CQ_VM_MAX_AGGS * 2];
    __shared__ u32 s_seen[CQ_VM_MAX_AGGS * 2];
    extern __shared__ u64 s_dyn[];
    constexpr int NR = GROUPED ? 1 : CQ_LIN_RACC;  // scalar slots kept per thread
    const cq_lin_program& L = V->lin;
    const int tid = threadIdx.x;
    const long long G = V->groups;
    const int nas = V->storage_ts - V->ngroups;
    const int nc = L.ncols;
    int* wcols = (int*)s_dyn + (tid >> 6) * nc * LIN_COL_INTS;  // this wave's staged columns
    u32* s_table = (u32*)(s_dyn + (long long)(LIN_BT / 64) * nc * LIN_COL_INTS / 2);
    const int tw = L.join >= 0 ? L.join_lds_words : 0;
    u64* s_agg = (u64*)(s_table + ((tw + 1) & ~1));
    u32* s_flag = (u32*)(s_agg + (lds_groups ? G * nas : 0));
    const u32* ht = L.join >= 0 ? P.ht[L.join] : nullptr;
    // agg_init's barrier orders the copy before the probes
    for (int k = tid; k < tw; k += LIN_BT) s_table[k] = ht[k];
    agg_init<LIN_BT>(V, lds_groups, s_acc, s_seen, s_agg, s_flag);
    u64 racc[NR];
    u32 rseen = 0u;
#pragma unroll
    for (int k = 0; k < NR; ++k) racc[k] = slot_is_min(V, k) ? ~0ull : 0ull;
    const long long nrows = P.row_end - P.row_begin;
    const long long ntiles = (nrows + LIN_TILE - 1) / LIN_TILE;
    const long long waves = (long long)gridDim.x * (LIN_BT / 64);
    LinRows q;
    q.lane = tid & 63;
    q.last = P.row_end - 1;
    for (long long tile = (long long)blockIdx.x * (LIN_BT / 64) + (tid >> 6); tile < ntiles; tile += waves) {
        lin_tile_rows(q, P.row_begin, P.row_end, tile);
        // every column of the plan into LDS and the validity bits of its
        // nullable columns (LIN_R bits per position), all in flight at once
        __builtin_amdgcn_s_waitcnt(0);  // the previous tile's LDS reads are done
        u64 vbits = 0ull;
        for (int c = 0; c < nc; ++c) {
            lin_stage(P.col[L.slot[c]], wcols + c * LIN_COL_INTS, q);
            vbits |= (u64)(L.nullable[c] ? lin_valid(P.valid[L.slot[c]], q) : q.inm) << (LIN_R * c);
        }
        __builtin_amdgcn_s_waitcnt(0);  // the staged columns have landed
        u32 alive = lin_filters(P, L, wcols, q, q.inm);
        if (__ballot(alive != 0u) == 0ull) continue;
        int pay[LIN_R];
#pragma unroll
        for (int r = 0; r < LIN_R; ++r) pay[r] = 0;
        if (L.join >= 0) {
            if (tw > 0) alive = lin_join(V, L, s_table, wcols, vbits, q, alive, pay);
            else alive = lin_join(V, L, ht, wcols, vbits, q, alive, pay);
            if (__ballot(alive != 0u) == 0ull) continue;
        }
        if (!GROUPED) rseen |= lin_scalar_fold<NR>(P, V, L, wcols, vbits, q, pay, alive, racc);
        else lin_group_tile(P, V, L, wcols, vbits, q, pay, alive, lds_groups, s_agg, s_flag);
    }
    agg_flush<LIN_BT>(P, V, lds_groups, s_acc, s_seen, s_agg, s_flag, racc, rseen);
}

// Several linear-form programs in one pass (cq_lin_batch): each wave tile
// of the union columns is staged once and every program evaluates it.
// Scalar programs fold each tile into a wave reduction and one LDS atomic
// per slot (the program count is not known at compile time, so no
// per-thread accumulators); grouped programs aggregate into their LDS
// region or with global atomics.  Static LDS: the scalar slots of every
// program; dynamic: the staged columns, then the programs' join tables and
// group regions at the descriptor's offsets.  The programs and parameter
// blocks come as restrict arrays, so their uniform fields are read with
// scalar loads.
__global__ __launch_bounds__(LIN_BT, 2) void k_lin_multi(const cq_lin_batch* __restrict__ Bp, const cq_vm_program* __restrict__ Vs,
                                                         const cq_query_params* __restrict__ Ps) {
    __shared__ u64 s_acc[CQ_LIN_BATCH * CQ_LIN_RACC];
    __shared__ u32 s_seen[CQ_LIN_BATCH * CQ_LIN_RACC];
    extern __shared__ u64 s_dyn[];
    const cq_lin_batch& B = *Bp;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int nc = B.ncols, np = B.nprog;
    int* wcols = (int*)s_dyn + (tid >> 6) * nc * LIN_COL_INTS;
    char* s_rest = (char*)(s_dyn + (long long)(LIN_BT / 64) * nc * LIN_COL_INTS / 2);
    for (int k = 0; k < np; ++k) {
        const cq_vm_program* V = Vs + k;
        const cq_query_params& P = Ps[k];
        const cq_lin_program& L = V->lin;
        if (B.tab_off[k] != CQ_LIN_NONE) {
            u32* t = (u32*)(s_rest + B.tab_off[k]);
            const u32* ht = P.ht[L.join];
            for (int w = tid; w < L.join_lds_words; w += LIN_BT) t[w] = ht[w];
        }
        const int ng = V->ngroups, sts = V->storage_ts, nas = sts - ng;
        if (ng == 0) {
            for (int s = tid; s < CQ_LIN_RACC; s += LIN_BT) {
                s_acc[k * CQ_LIN_RACC + s] = slot_is_min(V, s) ? ~0ull : 0ull;
                s_seen[k * CQ_LIN_RACC + s] = 0u;
            }
        } else if (B.grp_off[k] != CQ_LIN_NONE) {
            const long long G = V->groups;
            u64* s_agg = (u64*)(s_rest + B.grp_off[k]);
            u32* s_flag = (u32*)(s_agg + G * nas);
            for (long long w = tid; w < G * nas; w += LIN_BT) s_agg[w] = slot_is_min(V, (int)(w % nas)) ? ~0ull : 0ull;
            for (long long w = tid; w < G; w += LIN_BT) s_flag[w] = 0u;
        }
    }
    __syncthreads();
    const long long nrows = B.row_end - B.row_begin;
    const long long ntiles = (nrows + LIN_TILE - 1) / LIN_TILE;
    const long long waves = (long long)gridDim.x * (LIN_BT / 64);
    LinRows q;
    q.lane = lane;
    q.last = B.row_end - 1;
    for (long long tile = (long long)blockIdx.x * (LIN_BT / 64) + (tid >> 6); tile < ntiles; tile += waves) {
        lin_tile_rows(q, B.row_begin, B.row_end, tile);
        __builtin_amdgcn_s_waitcnt(0);  // the previous tile's LDS reads are done
        u64 vbits = 0ull;
        for (int c = 0; c < nc; ++c) {
            lin_stage(B.col[c], wcols + c * LIN_COL_INTS, q);
            vbits |= (u64)lin_valid(B.valid[c], q) << (LIN_R * c);
        }
        __builtin_amdgcn_s_waitcnt(0);  // the staged columns have landed
        for (int k = 0; k < np; ++k) {
            const cq_vm_program* V = Vs + k;
            const cq_query_params& P = Ps[k];
            const cq_lin_program& L = V->lin;
            u32 alive = lin_filters(P, L, wcols, q, q.inm);
            if (__ballot(alive != 0u) == 0ull) continue;
            int pay[LIN_R];
#pragma unroll
            for (int r = 0; r < LIN_R; ++r) pay[r] = 0;
            if (L.join >= 0) {
                if (B.tab_off[k] != CQ_LIN_NONE) alive = lin_join(V, L, (const u32*)(s_rest + B.tab_off[k]), wcols, vbits, q, alive, pay);
                else alive = lin_join(V, L, P.ht[L.join], wcols, vbits, q, alive, pay);
                if (__ballot(alive != 0u) == 0ull) continue;
            }
            const int ng = V->ngroups, sts = V->storage_ts;
            if (ng == 0) {
                u64 racc[CQ_LIN_RACC];
#pragma unroll
                for (int s = 0; s < CQ_LIN_RACC; ++s) racc[s] = slot_is_min(V, s) ? ~0ull : 0ull;
                const u32 seen = lin_scalar_fold<CQ_LIN_RACC>(P, V, L, wcols, vbits, q, pay, alive, racc);
#pragma unroll
                for (int s = 0; s < CQ_LIN_RACC; ++s) {
                    if (s >= sts) break;
                    const int fn = slot_fn(V, s);
                    const u64 w = wave_reduce(racc[s], fn);
                    const bool nonid = fn == CQ_VM_MIN ? w != ~0ull : w != 0ull;
                    const bool any = __ballot((seen >> s) & 1u) != 0ull;
                    if (lane == 0 && nonid) atomic_fn(&s_acc[k * CQ_LIN_RACC + s], w, fn);
                    if (lane == 0 && any) s_seen[k * CQ_LIN_RACC + s] = 1u;
                }
            } else {
                const bool lg = B.grp_off[k] != CQ_LIN_NONE;
                u64* s_agg = lg ? (u64*)(s_rest + B.grp_off[k]) : nullptr;
                u32* s_flag = lg ? (u32*)(s_agg + V->groups * (sts - ng)) : nullptr;
                lin_group_tile(P, V, L, wcols, vbits, q, pay, alive, lg ? 1 : 0, s_agg, s_flag);
            }
        }
    }
    __syncthreads();
    for (int k = 0; k < np; ++k) {
        const cq_vm_program* V = Vs + k;
        const cq_query_params& P = Ps[k];
        if (V->ngroups == 0) {
            flush_scalar_slots<LIN_BT>(P, V, s_acc + k * CQ_LIN_RACC, s_seen + k * CQ_LIN_RACC);
        } else if (B.grp_off[k] != CQ_LIN_NONE) {
            const u64* s_agg = (const u64*)(s_rest + B.grp_off[k]);
            flush_lds_groups<LIN_BT>(P, V, s_agg, (const u32*)(s_agg + V->groups * (V->storage_ts - V->ngroups)));
        }
    }
}

// lds_bytes: staged columns and (group_bytes of them) LDS group aggregates
// blocks of a persistent grid: occupancy x CUs, scaled by grid_percent,
// at most one per `rows_per_block` rows
unsigned lin_grid(int per_cu, int cus, long long rows, long long rows_per_block, int grid_percent) {
    const long long full = std::max<long long>(1, (long long)per_cu * cus * std::max(1, std::min(100, grid_percent)) / 100);
    const long long tiles = (rows + rows_per_block - 1) / rows_per_block;
    return (unsigned)std::max<long long>(1, std::min<long long>(tiles, full));
}

template <bool GROUPED>
hipError_t lin_launch(const cq_query_params* P, const cq_vm_program* V, unsigned lds_bytes, unsigned group_bytes, int device,
                      int grid_percent, hipStream_t s) {
    const long long rows = P->row_end - P->row_begin;
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(k_lin_scan<GROUPED>), LIN_BT, lds_bytes) !=
            hipSuccess ||
        per_cu < 1)
        per_cu = 1;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cus < 1) cus = 1;
    const unsigned grid = lin_grid(per_cu, cus, rows, (long long)LIN_BT * LIN_R, grid_percent);  // blocks' worth of wave tiles
    hipLaunchKernelGGL(k_lin_scan<GROUPED>, dim3(grid), dim3(LIN_BT), lds_bytes, s, *P, V, group_bytes ? 1 : 0);
    return hipGetLastError();
}

}  // namespace

hipError_t cq_vm_launch_build(const cq_query_params* P, const cq_vm_program* V, int join, long long dim_rows, hipStream_t s) {
    const unsigned grid = (unsigned)std::min<long long>(std::max<long long>((dim_rows + VM_BT - 1) / VM_BT, 1), 8192);
    hipLaunchKernelGGL(k_vm_build, dim3(grid), dim3(VM_BT), 0, s, *P, V, join);
    return hipGetLastError();
}

hipError_t cq_vm_launch_scan(const cq_query_params* P, const cq_vm_program* V, unsigned lds_group_bytes, unsigned grid,
                             hipStream_t s) {
    hipLaunchKernelGGL(k_vm_scan, dim3(grid), dim3(VM_BT), lds_group_bytes, s, *P, V, lds_group_bytes ? 1 : 0);
    return hipGetLastError();
}

hipError_t cq_vm_launch_dense_count(const cq_query_params* P, const cq_vm_program* V, const unsigned long long* R,
                                    const unsigned long long* RV, unsigned long long* counts, long long G, hipStream_t s) {
    const unsigned grid = (unsigned)std::max<long long>((G + VM_BT - 1) / VM_BT, 1);
    hipLaunchKernelGGL(k_vm_dense_count, dim3(grid), dim3(VM_BT), 0, s, *P, V, R, RV, counts, G);
    return hipGetLastError();
}

int cq_vm_block_threads() { return VM_BT; }

bool cq_vm_scan4_supported(const cq_vm_program* V) {
    if (V->nregs > VM4_REGS || V->nregs < 1 || V->scan_begin < 0 || V->scan_end > V->ncode) return false;
    // scalar aggregates: every slot in the per-thread registers
    if (V->ngroups == 0 && V->storage_ts > VM_RACC) return false;
    for (int pc = V->scan_begin; pc < V->scan_end; ++pc) {
        const int op = V->code[pc].op;
        if (op == CQ_OP_DCOL || op == CQ_OP_DVALID || op == CQ_OP_SLOT || op == CQ_OP_SLOTV) return false;
    }
    return true;
}

int cq_lin_row_align() { return LIN_R; }

bool cq_lin_supported(const cq_vm_program* V) {
    const cq_lin_program& L = V->lin;
    if (!L.enabled || L.ncols < 1 || L.ncols > CQ_LIN_COLS || L.nterms > CQ_LIN_TERMS || L.naggs != V->naggs || L.naggs > CQ_LIN_AGGS)
        return false;
    if (L.ngroups != V->ngroups || L.ngroups > CQ_LIN_KEYS || (V->ngroups == 0 && V->storage_ts > CQ_LIN_RACC)) return false;
    if (L.nfilter_cols < 0 || L.nfilter_cols > L.ncols || L.join_col < 0 || L.join_col >= L.ncols) return false;
    if (L.join_lds_words < 0 || L.join_lds_words > CQ_LIN_LDS_TABLE_WORDS) return false;
    if (L.join >= 0 && (L.join >= V->njoins || V->joins[L.join].layout != CQ_L_DENSE)) return false;
    return true;
}

hipError_t cq_lin_launch_scan(const cq_query_params* P, const cq_vm_program* V, const cq_vm_program* host, unsigned lds_group_bytes,
                              int device, int grid_percent, hipStream_t s) {
    const long long rows = P->row_end - P->row_begin;
    if (rows <= 0) return hipSuccess;
    if (!cq_lin_supported(host)) return hipErrorInvalidValue;
    const unsigned stage = (unsigned)(LIN_BT / 64) * (unsigned)host->lin.ncols * LIN_COL_INTS * 4u;
    const unsigned table = host->lin.join >= 0 ? (unsigned)((host->lin.join_lds_words + 1) & ~1) * 4u : 0u;
    if (host->ngroups > 0) return lin_launch<true>(P, V, stage + table + lds_group_bytes, lds_group_bytes, device, grid_percent, s);
    return lin_launch<false>(P, V, stage + table, 0u, device, grid_percent, s);
}

unsigned cq_lin_batch_stage_bytes(int ncols) { return (unsigned)(LIN_BT / 64) * (unsigned)ncols * LIN_COL_INTS * 4u; }

hipError_t cq_lin_launch_batch(const cq_lin_batch* B, const cq_vm_program* V, const cq_query_params* P, int ncols, unsigned rest_bytes,
                               long long rows, int device, int grid_percent, hipStream_t s) {
    if (rows <= 0) return hipSuccess;
    if (ncols < 1 || ncols > CQ_LIN_COLS) return hipErrorInvalidValue;
    const unsigned lds = cq_lin_batch_stage_bytes(ncols) + rest_bytes;
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(k_lin_multi), LIN_BT, lds) != hipSuccess ||
        per_cu < 1)
        per_cu = 1;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cus < 1) cus = 1;
    const unsigned grid = lin_grid(per_cu, cus, rows, (long long)LIN_BT * LIN_R, grid_percent);
    hipLaunchKernelGGL(k_lin_multi, dim3(grid), dim3(LIN_BT), lds, s, B, V, P);
    return hipGetLastError();
}

hipError_t cq_vm_launch_scan4(const cq_query_params* P, const cq_vm_program* V, int nregs, unsigned lds_group_bytes, int device,
                              hipStream_t s) {
    const long long rows = P->row_end - P->row_begin;
    if (rows <= 0) return hipSuccess;
    const unsigned lds = (unsigned)nregs * 4u * VM4_BT * 8u + lds_group_bytes;
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(k_vm_scan4), VM4_BT, lds) !=
            hipSuccess ||
        per_cu < 1)
        per_cu = 1;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cus < 1) cus = 1;
    const long long tiles = (rows + VM4_BT * 4 - 1) / (VM4_BT * 4);
    const unsigned grid = (unsigned)std::max<long long>(1, std::min<long long>(tiles, (long long)per_cu * cus));
    hipLaunchKernelGGL(k_vm_scan4, dim3(grid), dim3(VM4_BT), lds, s, *P, V, lds_group_bytes ? 1 : 0);
    return hipGetLastError();
}
