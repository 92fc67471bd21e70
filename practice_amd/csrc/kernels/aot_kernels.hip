// Ahead-of-time kernels of the crystal-hip runtime (compiled by hipcc for
// gfx950 into libcrystal_hip.so).  Query-specific work lives in the hipRTC
// modules; these kernels are data-layout generic:
//   - synthetic star-schema generator (same streams as the oracle generator)
//   - catalog statistics (min / max / exact distinct count)
//   - fills and MIN-slot initialisation (no hipMemset inside graphs)
//   - AVG finalisation (finalize_aggregate_slots, visitor.cpp:2505-2586)
//   - exclusive prefix scan of per-group flags
//   - dense columnar write (AggregateDenseWrite, visitor.cpp:2589-2697)
//   - sparse compaction (CompactSparseAggregateResult,
//     database_instance.cpp:286-307), written in group order
//   - ORDER BY: bitonic sort of row indices with the reference comparator
//     (emitRightBeforeLeftComparison, visitor.cpp:307-360) and a gather
#include <hip/hip_runtime.h>

#include "../device/crystal_hip.h"
#include "aot_kernels.h"

using namespace chip;

namespace {

__device__ __forceinline__ u64 splitmix64(u64 x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
__device__ __forceinline__ i64 uni(u64 s, i64 row, i64 lo, i64 hi) {
    const u64 h = splitmix64(s + (u64)row * 0x9E3779B97F4A7C15ull);
    return lo + (i64)__umul64hi(h, (u64)(hi - lo + 1));
}
__device__ __forceinline__ bool leap(int y) { return (y % 4 == 0 && y % 100 != 0) || y % 400 == 0; }
__device__ void civil(i64 d, int& y, int& m, int& dom, int& doy) {
    int yy = 1992;
    for (;;) {
        const int len = leap(yy) ? 366 : 365;
        if (d < len) break;
        d -= len;
        ++yy;
    }
    doy = (int)d + 1;
    int mm = 0;
    for (;;) {
        int len = 31;
        if (mm == 1) len = leap(yy) ? 29 : 28;
        else if (mm == 3 || mm == 5 || mm == 8 || mm == 10) len = 30;
        if (d < len) break;
        d -= len;
        ++mm;
    }
    y = yy;
    m = mm + 1;
    dom = (int)d + 1;
}
__device__ __forceinline__ int datekey(i64 d) {
    int y, m, dom, doy;
    civil(d, y, m, dom, doy);
    return y * 10000 + m * 100 + dom;
}

__device__ int gen_cell(const cq_gen_recipe& rc, i64 r) {
    int y, m, dom, doy;
    switch (rc.kind) {
        case CQ_GEN_KEY: return (int)(r + 1);
        case CQ_GEN_UNIFORM: return (int)uni(rc.s0, r, rc.lo, rc.hi);
        case CQ_GEN_DATE_FK: return datekey(uni(rc.s0, r, rc.lo, rc.hi));
        case CQ_GEN_PCAT: return (int)uni(rc.s0, r, 0, 24);
        case CQ_GEN_PMFGR: return (int)(uni(rc.s0, r, 0, 24) / 5);
        case CQ_GEN_PBRAND: return (int)(uni(rc.s0, r, 0, 24) * 40 + uni(rc.s1, r, 0, 39));
        case CQ_GEN_NATION: return (int)uni(rc.s0, r, 0, 24);
        case CQ_GEN_REGION: return (int)(uni(rc.s0, r, 0, 24) / 5);
        case CQ_GEN_CITY: return (int)(uni(rc.s0, r, 0, 24) * 10 + uni(rc.s1, r, 0, 9));
        case CQ_GEN_ORDERKEY: return (int)(r / 4 + 1);
        case CQ_GEN_LINENUMBER: return (int)(r % 4 + 1);
        default: break;
    }
    civil(r, y, m, dom, doy);
    switch (rc.kind) {
        case CQ_GEN_DKEY: return y * 10000 + m * 100 + dom;
        case CQ_GEN_DYEAR: return y;
        case CQ_GEN_DMONTH: return m;
        case CQ_GEN_DYM: return y * 100 + m;
        case CQ_GEN_DDOW: return (int)((r + 3) % 7) + 1;
        case CQ_GEN_DDOM: return dom;
        case CQ_GEN_DDOY: return doy;
        case CQ_GEN_DWEEK: return (doy - 1) / 7 + 1;
        default: return 0;
    }
}

// one thread per row; a wave covers 64 rows = one validity word (start % 64 == 0)
__global__ __launch_bounds__(256) void k_generate(int* __restrict__ out, u64* __restrict__ valid, cq_gen_recipe rc,
                                                  i64 start, i64 count) {
    const i64 stride = (i64)gridDim.x * 256;
    const i64 rounded = (count + 63) & ~63ll;
    for (i64 i = (i64)blockIdx.x * 256 + threadIdx.x; i < rounded; i += stride) {
        const bool in = i < count;
        const i64 r = start + i;
        bool ok = true;
        int v = 0;
        if (in) {
            v = gen_cell(rc, r);
            if (rc.null_bp > 0) ok = uni(rc.vseed, r, 0, 9999) >= rc.null_bp;
            out[i] = ok ? v : 0;
        }
        if (valid) {
            const u64 w = __ballot(in && ok);
            if ((threadIdx.x & 63) == 0) valid[i >> 6] = w;
        }
    }
}

// min / max over int32 values (order-preserving u64 transform)
__global__ __launch_bounds__(256) void k_minmax(const int* __restrict__ col, i64 n, u64* __restrict__ mm) {
    u64 lo = ~0ull, hi = 0;
    for (i64 i = (i64)blockIdx.x * 256 + threadIdx.x; i < n; i += (i64)gridDim.x * 256) {
        const u64 t = (u64)((i64)col[i] + 0x80000000ll);
        lo = t < lo ? t : lo;
        hi = t > hi ? t : hi;
    }
    lo = wave_min(lo);
    hi = wave_max(hi);
    if ((threadIdx.x & 63) == 0) {
        atomicMin(&mm[0], lo);
        atomicMax(&mm[1], hi);
    }
}

__global__ __launch_bounds__(256) void k_mark(const int* __restrict__ col, i64 n, i64 mn, u32* __restrict__ bits) {
    for (i64 i = (i64)blockIdx.x * 256 + threadIdx.x; i < n; i += (i64)gridDim.x * 256) {
        const u64 k = (u64)((i64)col[i] - mn);
        const u32 m = 1u << (k & 31);
        if (!(bits[k >> 5] & m)) atomicOr(&bits[k >> 5], m);
    }
}

__global__ __launch_bounds__(256) void k_popcount(const u32* __restrict__ bits, i64 words, u64* __restrict__ out) {
    u64 c = 0;
    for (i64 i = (i64)blockIdx.x * 256 + threadIdx.x; i < words; i += (i64)gridDim.x * 256) c += __popc(bits[i]);
    c = wave_sum(c);
    if ((threadIdx.x & 63) == 0) atomicAdd(out, c);
}

__global__ __launch_bounds__(256) void k_fill_u64(u64* p, u64 v, i64 n) {
    for (i64 i = (i64)blockIdx.x * 256 + threadIdx.x; i < n; i += (i64)gridDim.x * 256) p[i] = v;
}

__global__ __launch_bounds__(256) void k_fill_u32(u32* p, u32 v, i64 n) {
    for (i64 i = (i64)blockIdx.x * 256 + threadIdx.x; i < n; i += (i64)gridDim.x * 256) p[i] = v;
}

__global__ __launch_bounds__(256) void k_init_min(u64* res, i64 groups, int sts, u32 mask) {
    for (i64 g = (i64)blockIdx.x * 256 + threadIdx.x; g < groups; g += (i64)gridDim.x * 256)
        for (int s = 0; s < sts && s < 32; ++s)
            if ((mask >> s) & 1u) res[g * sts + s] = ~0ull;
}

// storage tuple -> visible tuple (AVG = sum / count as double bits)
__global__ __launch_bounds__(256) void k_finalize(const u64* __restrict__ res, const u64* __restrict__ rv, i64 groups,
                                                  cq_finalize_desc d, u64* __restrict__ vres, u64* __restrict__ vvalid) {
    for (i64 g = (i64)blockIdx.x * 256 + threadIdx.x; g < groups; g += (i64)gridDim.x * 256) {
        const i64 src = g * d.sts, dst = g * d.vts;
        for (int c = 0; c < d.ngroups; ++c) {
            vres[dst + c] = res[src + c];
            if (bit64(rv, src + c)) set_bit64(vvalid, dst + c);
        }
        for (int a = 0; a < d.naggs; ++a) {
            const int c = d.ngroups + a;
            const int kind = d.kind[a];
            if (kind == CQ_AGG_COUNT) {
                vres[dst + c] = res[src + c];
                if (d.ngroups == 0 || res[src + c] != 0ull) set_bit64(vvalid, dst + c);
            } else if (d.hidden[a] >= 0) {
                const u64 cnt = res[src + d.hidden[a]];
                if (cnt != 0ull) {
                    const double avg = (double)res[src + c] / (double)cnt;
                    vres[dst + c] = double_to_bits(avg);
                    set_bit64(vvalid, dst + c);
                }
            } else {
                const u64 v = res[src + c];
                vres[dst + c] = (kind == CQ_AGG_MIN && v == ~0ull) ? 0ull : v;
                if (bit64(rv, src + c)) set_bit64(vvalid, dst + c);
            }
        }
    }
}

// present = any non-zero cell (sparse output)
__global__ __launch_bounds__(256) void k_sparse_count(const u64* __restrict__ res, i64 groups, int vts, u64* __restrict__ flags) {
    for (i64 g = (i64)blockIdx.x * 256 + threadIdx.x; g < groups; g += (i64)gridDim.x * 256) {
        bool present = false;
        for (int c = 0; c < vts && !present; ++c) present = res[g * vts + c] != 0ull;
        flags[g] = present ? 1ull : 0ull;
    }
}

// ---- exclusive scan of u64 flags, three passes (block 1024 threads x 4 items)
constexpr int SCAN_BT = 1024, SCAN_IPT = 4, SCAN_TILE = SCAN_BT * SCAN_IPT;

__device__ u64 block_exclusive(u64 v, u64* s_warp, u64& total) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    u64 x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const u64 t = __shfl_up(x, o);
        if (lane >= o) x += t;
    }
    if (lane == 63) s_warp[wid] = x;
    __syncthreads();
    if (wid == 0) {
        u64 w = lane < SCAN_BT / 64 ? s_warp[lane] : 0ull;
        u64 y = w;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const u64 t = __shfl_up(y, o);
            if (lane >= o) y += t;
        }
        if (lane < SCAN_BT / 64) s_warp[lane] = y - w;
        if (lane == SCAN_BT / 64 - 1) s_warp[SCAN_BT / 64] = y;
    }
    __syncthreads();
    total = s_warp[SCAN_BT / 64];
    const u64 r = s_warp[wid] + x - v;
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(SCAN_BT) void k_scan_local(const u64* __restrict__ in, u64* __restrict__ out, i64 n,
                                                       u64* __restrict__ block_sums) {
    __shared__ u64 s_warp[SCAN_BT / 64 + 1];
    const i64 base = (i64)blockIdx.x * SCAN_TILE + (i64)threadIdx.x * SCAN_IPT;
    u64 v[SCAN_IPT], sum = 0;
#pragma unroll
    for (int k = 0; k < SCAN_IPT; ++k) {
        v[k] = base + k < n ? in[base + k] : 0ull;
        sum += v[k];
    }
    u64 total;
    u64 run = block_exclusive(sum, s_warp, total);
#pragma unroll
    for (int k = 0; k < SCAN_IPT; ++k) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
    if (threadIdx.x == 0) block_sums[blockIdx.x] = total;
}

// single block: exclusive scan of block sums in place (nblocks <= 1024 * 4 * 8)
__global__ __launch_bounds__(SCAN_BT) void k_scan_sums(u64* __restrict__ sums, i64 nblocks, u64* __restrict__ grand) {
    __shared__ u64 s_warp[SCAN_BT / 64 + 1];
    u64 carry = 0;
    for (i64 off = 0; off < nblocks; off += SCAN_TILE) {
        const i64 base = off + (i64)threadIdx.x * SCAN_IPT;
        u64 v[SCAN_IPT], sum = 0;
#pragma unroll
        for (int k = 0; k < SCAN_IPT; ++k) {
            v[k] = base + k < nblocks ? sums[base + k] : 0ull;
            sum += v[k];
        }
        u64 total;
        u64 run = carry + block_exclusive(sum, s_warp, total);
#pragma unroll
        for (int k = 0; k < SCAN_IPT; ++k) {
            if (base + k < nblocks) sums[base + k] = run;
            run += v[k];
        }
        carry += total;
    }
    if (threadIdx.x == 0) *grand = carry;
}

__global__ __launch_bounds__(SCAN_BT) void k_scan_add(u64* __restrict__ out, i64 n, const u64* __restrict__ sums) {
    const i64 base = (i64)blockIdx.x * SCAN_TILE + (i64)threadIdx.x * SCAN_IPT;
    const u64 add = sums[blockIdx.x];
#pragma unroll
    for (int k = 0; k < SCAN_IPT; ++k)
        if (base + k < n) out[base + k] += add;
}

// dense / sparse write: rows of present groups, in group order.  `offsets`
// is the exclusive scan of the 0/1 `flags`, so the present groups of one
// wave land on consecutive rows, which span at most two 64-bit validity
// words: each column's validity bits are OR-reduced across the wave and
// written with two atomics per wave (not one per cell on a shared word).
__device__ __forceinline__ u64 wave_or64(u64 v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v |= __shfl_xor(v, o);
    return v;
}
__global__ __launch_bounds__(256) void k_write_rows(const u64* __restrict__ res, const u64* __restrict__ rv, i64 groups,
                                                    int vts, const u64* __restrict__ flags, const u64* __restrict__ offsets,
                                                    u32 nullable_mask, int use_validity, u64* __restrict__ out,
                                                    u64* __restrict__ out_valid, i64 cap_rows, i64 valid_words) {
    const int lane = threadIdx.x & 63;
    // wave-uniform trip count: every lane takes part in the reductions
    for (i64 g0 = (i64)blockIdx.x * 256 + (threadIdx.x & ~63); g0 < groups; g0 += (i64)gridDim.x * 256) {
        const i64 g = g0 + lane;
        const bool present = g < groups && flags[g] != 0ull;
        const u64 pm = __ballot(present);
        if (pm == 0ull) continue;
        const i64 row = present ? (i64)offsets[g] : 0;
        const int first = __ffsll((unsigned long long)pm) - 1;
        const i64 base = __shfl(row, first) & ~(i64)63;
        const i64 off = row - base;
        for (int c = 0; c < vts; ++c) {
            bool valid = false;
            if (present) {
                out[(i64)c * cap_rows + row] = res[g * vts + c];
                valid = !use_validity || !((nullable_mask >> c) & 1u) || bit64(rv, g * vts + c);
            }
            u64* ov = out_valid + (i64)c * valid_words;
            u64 lo = 0ull, hi = 0ull;
            if (valid) {
                if (off < 64) lo = 1ull << off;
                else if (off < 128) hi = 1ull << (off - 64);
                else set_bit64(ov, row);  // not reached when flags are 0/1
            }
            lo = wave_or64(lo);
            hi = wave_or64(hi);
            if (lane == 0) {
                if (lo) atomicOr(&ov[base >> 6], lo);
                if (hi) atomicOr(&ov[(base >> 6) + 1], hi);
            }
        }
    }
}

// ---- batched shared scans: one launch for all queries of a batch
// (blockIdx.y / blockIdx.x = query)
__global__ __launch_bounds__(256) void k_batch_init(const cq_init_desc* __restrict__ D) {
    const cq_init_desc d = D[blockIdx.y];
    for (i64 i = (i64)blockIdx.x * 256 + threadIdx.x; i < d.slots; i += (i64)gridDim.x * 256) {
        u64 v = 0ull;
        if (d.min_mask) {
            const int s = (int)(i % d.sts);
            if (s < 32 && i < d.G * d.sts && ((d.min_mask >> s) & 1u)) v = ~0ull;
        }
        d.res[i] = v;
    }
    const i64 words = (d.slots + 63) / 64;
    for (i64 w = (i64)blockIdx.x * 256 + threadIdx.x; w < words; w += (i64)gridDim.x * 256) d.rv[w] = w == 0 ? d.count_bits : 0ull;
}

// Present groups -> rows in group order, as the dense count / exclusive scan
// / row writer sequence of a single query, for groups <= 16 x 1024: each
// thread owns a contiguous run of groups, a block-wide exclusive scan gives
// its first row, output validity words are assembled in LDS.
constexpr int BSHAPE_BT = 1024;
__global__ __launch_bounds__(BSHAPE_BT) void k_batch_shape(const cq_shape_desc* __restrict__ D) {
    const cq_shape_desc d = D[blockIdx.x];
    if (d.G < 0) return;
    __shared__ u64 s_valid[kBatchShapeMaxCols * (kBatchShapeMaxGroups / 64)];
    __shared__ u32 s_wave[BSHAPE_BT / 64];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int per = (int)((d.G + BSHAPE_BT - 1) / BSHAPE_BT);
    const i64 g0 = (i64)tid * per;
    const int vts = d.vts;
    const int vw = (int)d.vwords;
    for (int i = tid; i < vts * vw; i += BSHAPE_BT) s_valid[i] = 0ull;
    u32 pm = 0u;
    u32 cnt = 0u;
    for (int k = 0; k < per; ++k) {
        const i64 g = g0 + k;
        if (g >= d.G) break;
        bool p = false;
        if (d.sparse) {
            for (int c = 0; c < vts && !p; ++c) p = d.res[g * vts + c] != 0ull;
        } else {
            p = d.fl[g] != 0ull;
        }
        if (p) {
            pm |= 1u << k;
            ++cnt;
        }
    }
    // block-wide exclusive scan of cnt
    u32 x = cnt;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const u32 t = __shfl_up(x, o);
        if (lane >= o) x += t;
    }
    if (lane == 63) s_wave[wid] = x;
    __syncthreads();
    if (wid == 0) {
        u32 w = lane < BSHAPE_BT / 64 ? s_wave[lane] : 0u;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const u32 t = __shfl_up(w, o);
            if (lane >= o) w += t;
        }
        if (lane < BSHAPE_BT / 64) s_wave[lane] = w;  // inclusive totals
    }
    __syncthreads();
    const u32 before = (wid ? s_wave[wid - 1] : 0u) + x - cnt;
    const u32 total = s_wave[BSHAPE_BT / 64 - 1];
    i64 row = before;
    for (int k = 0; k < per; ++k) {
        if (!((pm >> k) & 1u)) continue;
        const i64 g = g0 + k;
        for (int c = 0; c < vts; ++c) {
            d.o[(i64)c * d.cap_rows + row] = d.res[g * vts + c];
            const bool valid = !d.use_validity || !((d.nullable_mask >> c) & 1u) || bit64(d.rv, g * vts + c);
            if (valid) atomicOr(&s_valid[c * vw + (int)(row >> 6)], 1ull << (row & 63));
        }
        ++row;
    }
    __syncthreads();
    for (int i = tid; i < vts * vw; i += BSHAPE_BT) d.ov[(i64)(i / vw) * d.vwords + (i % vw)] = s_valid[i];
    if (tid == 0) *d.rows = total;
}

// ---- ORDER BY
__device__ bool right_before_left(u64 l, u64 r, const cq_sort_desc& d, const u64* cols, const u64* valid, i64 cap_rows,
                                  i64 valid_words) {
    const u64 INV = ~0ull;
    if (r == l) return false;
    if (r == INV) return false;
    if (l == INV) return true;
    for (int k = 0; k < d.nkeys; ++k) {
        const int c = d.col[k];
        if (d.nullable[k]) {
            const u32 lv = bit64(valid + (i64)c * valid_words, (i64)l);
            const u32 rvb = bit64(valid + (i64)c * valid_words, (i64)r);
            if (lv != rvb) return rvb != 0u;
            if (!lv) continue;
        }
        const u64 a = cols[(i64)c * cap_rows + (i64)l], b = cols[(i64)c * cap_rows + (i64)r];
        bool lt, gt;
        if (d.type[k] == CQ_TYPE_INT64) {
            lt = (i64)b < (i64)a;
            gt = (i64)a < (i64)b;
        } else if (d.type[k] == CQ_TYPE_FLOAT64) {
            lt = bits_to_double(b) < bits_to_double(a);
            gt = bits_to_double(a) < bits_to_double(b);
        } else {
            lt = b < a;
            gt = a < b;
        }
        if (lt) return !d.desc[k];
        if (gt) return d.desc[k] != 0;
    }
    return r < l;
}

__global__ __launch_bounds__(256) void k_sort_init(u64* idx, i64 n, i64 padded) {
    for (i64 i = (i64)blockIdx.x * 256 + threadIdx.x; i < padded; i += (i64)gridDim.x * 256) idx[i] = i < n ? (u64)i : ~0ull;
}

__global__ __launch_bounds__(256) void k_bitonic_step(u64* idx, i64 padded, i64 k, i64 j, cq_sort_desc d, const u64* cols,
                                                      const u64* valid, i64 cap_rows, i64 valid_words) {
    for (i64 i = (i64)blockIdx.x * 256 + threadIdx.x; i < padded; i += (i64)gridDim.x * 256) {
        const i64 ixj = i ^ j;
        if (ixj <= i) continue;
        const u64 l = idx[i], r = idx[ixj];
        const bool rbl = right_before_left(l, r, d, cols, valid, cap_rows, valid_words);
        const bool lbr = (l != r) && !rbl;
        const bool asc = (i & k) == 0;
        if (asc ? rbl : lbr) {
            idx[i] = r;
            idx[ixj] = l;
        }
    }
}

constexpr int LOCAL_SORT = 4096;
__global__ __launch_bounds__(1024) void k_bitonic_local(u64* idx, i64 padded, cq_sort_desc d, const u64* cols,
                                                        const u64* valid, i64 cap_rows, i64 valid_words) {
    __shared__ u64 s[LOCAL_SORT];
    for (i64 i = threadIdx.x; i < padded; i += 1024) s[i] = idx[i];
    __syncthreads();
    for (i64 k = 2; k <= padded; k <<= 1) {
        for (i64 j = k >> 1; j > 0; j >>= 1) {
            for (i64 i = threadIdx.x; i < padded; i += 1024) {
                const i64 ixj = i ^ j;
                if (ixj > i) {
                    const u64 l = s[i], r = s[ixj];
                    const bool rbl = right_before_left(l, r, d, cols, valid, cap_rows, valid_words);
                    const bool lbr = (l != r) && !rbl;
                    if (((i & k) == 0) ? rbl : lbr) {
                        s[i] = r;
                        s[ixj] = l;
                    }
                }
            }
            __syncthreads();
        }
    }
    for (i64 i = threadIdx.x; i < padded; i += 1024) idx[i] = s[i];
}

// one block per segment (segments are a few KB: result columns of a query)
__global__ __launch_bounds__(256) void k_copy_segments(const cq_copy_seg* __restrict__ segs, u64* __restrict__ dst) {
    const cq_copy_seg sg = segs[blockIdx.x];
    u64* out = dst + sg.dst_word;
    for (u64 k = threadIdx.x; k < sg.words; k += 256) out[k] = sg.src[k];
}

__global__ __launch_bounds__(256) void k_gather(const u64* __restrict__ cols, const u64* __restrict__ valid,
                                                const u64* __restrict__ idx, i64 rows, int vts, i64 cap_rows,
                                                i64 valid_words, u64* __restrict__ out, u64* __restrict__ out_valid) {
    // a wave writes rows 64k .. 64k + 63: it owns their validity word, which
    // is one ballot (plain store, no atomics on a shared word)
    for (i64 r0 = (i64)blockIdx.x * 256 + (threadIdx.x & ~63); r0 < rows; r0 += (i64)gridDim.x * 256) {
        const i64 r = r0 + (threadIdx.x & 63);
        const bool in = r < rows;
        const i64 src = in ? (i64)idx[r] : 0;
        for (int c = 0; c < vts; ++c) {
            if (in) out[(i64)c * cap_rows + r] = cols[(i64)c * cap_rows + src];
            const u64 w = __ballot(in && bit64(valid + (i64)c * valid_words, src));
            if ((threadIdx.x & 63) == 0) out_valid[(i64)c * valid_words + (r0 >> 6)] = w;
        }
    }
}

int grid_for(i64 n, int per_block = 256) {
    i64 g = (n + per_block - 1) / per_block;
    if (g < 1) g = 1;
    if (g > 16384) g = 16384;
    return (int)g;
}

}  // namespace

// ------------------------------------------------------------------ host launchers

hipError_t cq_launch_generate(int* out, unsigned long long* valid, const cq_gen_recipe* rc, long long start, long long count,
                              hipStream_t s) {
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_generate, dim3(grid_for(count)), dim3(256), 0, s, out, valid, *rc, start, count);
    return hipGetLastError();
}

hipError_t cq_launch_minmax(const int* col, long long n, unsigned long long* mm, hipStream_t s) {
    hipLaunchKernelGGL(k_minmax, dim3(grid_for(n)), dim3(256), 0, s, col, n, mm);
    return hipGetLastError();
}

hipError_t cq_launch_distinct(const int* col, long long n, long long mn, unsigned int* bits, long long words,
                              unsigned long long* out, hipStream_t s) {
    hipLaunchKernelGGL(k_mark, dim3(grid_for(n)), dim3(256), 0, s, col, n, mn, bits);
    hipLaunchKernelGGL(k_popcount, dim3(grid_for(words)), dim3(256), 0, s, bits, words, out);
    return hipGetLastError();
}

hipError_t cq_launch_fill_u64(unsigned long long* p, unsigned long long v, long long n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(n)), dim3(256), 0, s, p, v, n);
    return hipGetLastError();
}

hipError_t cq_launch_fill_u32(unsigned int* p, unsigned int v, long long n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n)), dim3(256), 0, s, p, v, n);
    return hipGetLastError();
}

hipError_t cq_launch_init_min(unsigned long long* res, long long groups, int sts, unsigned int mask, hipStream_t s) {
    if (!mask) return hipSuccess;
    hipLaunchKernelGGL(k_init_min, dim3(grid_for(groups)), dim3(256), 0, s, res, groups, sts, mask);
    return hipGetLastError();
}

hipError_t cq_launch_finalize(const unsigned long long* res, const unsigned long long* rv, long long groups,
                              const cq_finalize_desc* d, unsigned long long* vres, unsigned long long* vvalid, hipStream_t s) {
    hipLaunchKernelGGL(k_finalize, dim3(grid_for(groups)), dim3(256), 0, s, res, rv, groups, *d, vres, vvalid);
    return hipGetLastError();
}

hipError_t cq_launch_sparse_count(const unsigned long long* res, long long groups, int vts, unsigned long long* flags,
                                  hipStream_t s) {
    hipLaunchKernelGGL(k_sparse_count, dim3(grid_for(groups)), dim3(256), 0, s, res, groups, vts, flags);
    return hipGetLastError();
}

long long cq_scan_block_count(long long n) { return (n + SCAN_TILE - 1) / SCAN_TILE; }

hipError_t cq_launch_exclusive_scan(const unsigned long long* in, unsigned long long* out, long long n,
                                    unsigned long long* block_sums, unsigned long long* grand, hipStream_t s) {
    const long long nb = cq_scan_block_count(n);
    if (nb < 1) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_scan_local, dim3((unsigned)nb), dim3(SCAN_BT), 0, s, in, out, n, block_sums);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(SCAN_BT), 0, s, block_sums, nb, grand);
    hipLaunchKernelGGL(k_scan_add, dim3((unsigned)nb), dim3(SCAN_BT), 0, s, out, n, block_sums);
    return hipGetLastError();
}

hipError_t cq_launch_write_rows(const unsigned long long* res, const unsigned long long* rv, long long groups, int vts,
                                const unsigned long long* flags, const unsigned long long* offsets, unsigned int nullable_mask,
                                int use_validity, unsigned long long* out, unsigned long long* out_valid, long long cap_rows,
                                long long valid_words, hipStream_t s) {
    hipLaunchKernelGGL(k_write_rows, dim3(grid_for(groups)), dim3(256), 0, s, res, rv, groups, vts, flags, offsets,
                       nullable_mask, use_validity, out, out_valid, cap_rows, valid_words);
    return hipGetLastError();
}

hipError_t cq_launch_sort(unsigned long long* idx, long long rows, const cq_sort_desc* d, const unsigned long long* cols,
                          const unsigned long long* valid, long long cap_rows, long long valid_words, hipStream_t s) {
    long long padded = 1;
    while (padded < rows) padded <<= 1;
    hipLaunchKernelGGL(k_sort_init, dim3(grid_for(padded)), dim3(256), 0, s, idx, rows, padded);
    if (padded <= LOCAL_SORT) {
        hipLaunchKernelGGL(k_bitonic_local, dim3(1), dim3(1024), 0, s, idx, padded, *d, cols, valid, cap_rows, valid_words);
    } else {
        for (long long k = 2; k <= padded; k <<= 1)
            for (long long j = k >> 1; j > 0; j >>= 1)
                hipLaunchKernelGGL(k_bitonic_step, dim3(grid_for(padded)), dim3(256), 0, s, idx, padded, k, j, *d, cols, valid,
                                   cap_rows, valid_words);
    }
    return hipGetLastError();
}

long long cq_sort_padded(long long rows) {
    long long padded = 1;
    while (padded < rows) padded <<= 1;
    return padded;
}

hipError_t cq_launch_copy_segments(const cq_copy_seg* segs, int n, unsigned long long* dst, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_copy_segments, dim3(n), dim3(256), 0, s, segs, dst);
    return hipGetLastError();
}

hipError_t cq_launch_gather(const unsigned long long* cols, const unsigned long long* valid, const unsigned long long* idx,
                            long long rows, int vts, long long cap_rows, long long valid_words, unsigned long long* out,
                            unsigned long long* out_valid, hipStream_t s) {
    if (rows <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gather, dim3(grid_for(rows)), dim3(256), 0, s, cols, valid, idx, rows, vts, cap_rows, valid_words,
                       out, out_valid);
    return hipGetLastError();
}

hipError_t cq_launch_batch_init(const cq_init_desc* d, int n, long long max_slots, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    long long bx = (max_slots + 255) / 256;
    if (bx < 1) bx = 1;
    if (bx > 64) bx = 64;
    hipLaunchKernelGGL(k_batch_init, dim3((unsigned)bx, (unsigned)n), dim3(256), 0, s, d);
    return hipGetLastError();
}

hipError_t cq_launch_batch_shape(const cq_shape_desc* d, int n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_batch_shape, dim3((unsigned)n), dim3(BSHAPE_BT), 0, s, d);
    return hipGetLastError();
}
