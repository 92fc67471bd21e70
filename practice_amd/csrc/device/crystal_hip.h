// crystal_hip.h -- device primitives for the generated query kernels (CDNA4,
// wave64).  Compiled by hipRTC together with every generated module and by
// hipcc into the ahead-of-time kernels; self-contained (no standard headers).
//
// Counterparts of the reference primitives:
//   include/crystal/load.h   striped BlockLoad of 128 x 4 items  ->  each lane
//                            owns 4 consecutive rows per 128-bit load, a wave
//                            streams 1 KiB per column per instruction
//   include/crystal/pred.h   int flag arrays                    ->  one flag bit
//                            per item in a 32-bit lane mask, wave-uniform
//                            early exit through __ballot
//   include/crystal/join.h   HASH(k,len,min) = (k-min) % len, CAS 0->key
//                            -> same slot function; either the reference
//                            interleaved int layout or a compact tagged layout
//                            (see ht_probe_compact) or a key-indexed layout
//                            (ht_probe_dense) that keep the exact
//                            `ht[slot] == key` semantics in a few bits per key
//   include/core/inline_math.h  db::safe_* promotion rules, u64 atomics
//                            ->  safe_* below, u64 atomics / wave reductions
#ifndef CRYSTAL_HIP_H
#define CRYSTAL_HIP_H

#include "query_params.h"
#include "ht_layouts.h"

namespace chip {

typedef unsigned long long u64;
typedef long long i64;
typedef unsigned int u32;
typedef int v4i __attribute__((ext_vector_type(4)));
typedef unsigned int v4u __attribute__((ext_vector_type(4)));

// ------------------------------------------------------------------ arithmetic
// Integer division with defined behaviour: x / 0 == 0, INT_MIN / -1 wraps.
// (The reference leaves both undefined.)
__device__ __forceinline__ int div_(int a, int b) {
    return b == 0 ? 0 : (b == -1 ? (int)(0u - (u32)a) : a / b);
}
__device__ __forceinline__ i64 div_(i64 a, i64 b) {
    return b == 0 ? 0 : (b == -1 ? (i64)(0ull - (u64)a) : a / b);
}
__device__ __forceinline__ u64 div_(u64 a, u64 b) { return b == 0 ? 0 : a / b; }
__device__ __forceinline__ double div_(double a, double b) { return a / b; }

// a / b in the C type of the expression `a / b`
template <class A, class B>
__device__ __forceinline__ auto cdiv(A a, B b) -> decltype(a / b) {
    typedef decltype(a / b) T;
    return div_((T)a, (T)b);
}

// db::safe_* promote int32 to int64 (and float to double) before the usual
// arithmetic conversions (inline_math.h:9-45).
__device__ __forceinline__ i64 prom(int v) { return v; }
__device__ __forceinline__ i64 prom(i64 v) { return v; }
__device__ __forceinline__ u64 prom(u64 v) { return v; }
__device__ __forceinline__ double prom(double v) { return v; }

template <class A, class B> __device__ __forceinline__ auto safe_add(A a, B b) { return prom(a) + prom(b); }
template <class A, class B> __device__ __forceinline__ auto safe_sub(A a, B b) { return prom(a) - prom(b); }
template <class A, class B> __device__ __forceinline__ auto safe_mul(A a, B b) { return prom(a) * prom(b); }
template <class A, class B> __device__ __forceinline__ auto safe_div(A a, B b) { return cdiv(prom(a), prom(b)); }
template <class A, class B> __device__ __forceinline__ int safe_lt(A a, B b) { return prom(a) < prom(b) ? 1 : 0; }
template <class A, class B> __device__ __forceinline__ int safe_gt(A a, B b) { return prom(a) > prom(b) ? 1 : 0; }
template <class A, class B> __device__ __forceinline__ int safe_lte(A a, B b) { return prom(a) <= prom(b) ? 1 : 0; }
template <class A, class B> __device__ __forceinline__ int safe_gte(A a, B b) { return prom(a) >= prom(b) ? 1 : 0; }
template <class A, class B> __device__ __forceinline__ int safe_eq(A a, B b) { return prom(a) == prom(b) ? 1 : 0; }
template <class A, class B> __device__ __forceinline__ int safe_neq(A a, B b) { return prom(a) != prom(b) ? 1 : 0; }
// int32 against int32: the int64 promotion cannot change the outcome, so the
// comparison runs in 32 bits
__device__ __forceinline__ int safe_lt(int a, int b) { return a < b ? 1 : 0; }
__device__ __forceinline__ int safe_gt(int a, int b) { return a > b ? 1 : 0; }
__device__ __forceinline__ int safe_lte(int a, int b) { return a <= b ? 1 : 0; }
__device__ __forceinline__ int safe_gte(int a, int b) { return a >= b ? 1 : 0; }
__device__ __forceinline__ int safe_eq(int a, int b) { return a == b ? 1 : 0; }
__device__ __forceinline__ int safe_neq(int a, int b) { return a != b ? 1 : 0; }

__device__ __forceinline__ double bits_to_double(u64 v) { return __builtin_bit_cast(double, v); }
__device__ __forceinline__ u64 double_to_bits(double v) { return __builtin_bit_cast(u64, v); }

// ------------------------------------------------------------------ bitmaps
__device__ __forceinline__ u32 bit32(const u32* w, i64 row) { return (w[row >> 5] >> (row & 31)) & 1u; }
__device__ __forceinline__ u32 bit64(const u64* w, i64 idx) { return (u32)((w[idx >> 6] >> (idx & 63)) & 1ull); }
__device__ __forceinline__ void set_bit64(u64* w, i64 idx) { atomicOr(&w[idx >> 6], 1ull << (idx & 63)); }

// ------------------------------------------------------------------ hash tables
// Layouts and their probe / insert primitives: ht_layouts.h

// any item of a lane still alive
template <int N>
__device__ __forceinline__ bool any_of(const bool (&f)[N]) {
    bool a = false;
#pragma unroll
    for (int i = 0; i < N; ++i) a = a | f[i];
    return a;
}

// ------------------------------------------------------------------ multi-value tables
// Slot of a build or probe key in a multi-value table: (key - kmin) % len in
// int32 arithmetic (HASH, join.h:3); a key below the minimum has no slot.
__device__ __forceinline__ bool mht_slot(int key, int kmin, long long len, long long& slot) {
    const int x = (int)((u32)key - (u32)kmin);
    if (x < 0) return false;
    slot = (long long)x % len;
    return true;
}
// ProbeMultiHT (join.h:278-299): the directory entry of the key's slot
// matches when it holds that key and a non-zero count; returns the count and
// the first payload index.  An empty entry ({0, 0, 0, 0}) never matches.
__device__ __forceinline__ u32 mht_find(const u32* __restrict__ ht, int key, int kmin, long long len, u32& first) {
    long long s;
    if (!mht_slot(key, kmin, len, s)) return 0u;
    const v4u d = *(const v4u*)(ht + 4 * s);
    if ((int)d.x != key || d.z == 0u) return 0u;
    first = d.y;
    return d.z;
}

// ------------------------------------------------------------------ wave reductions
__device__ __forceinline__ u64 wave_sum(u64 v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}
__device__ __forceinline__ u64 wave_min(u64 v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const u64 t = __shfl_xor(v, o);
        v = t < v ? t : v;
    }
    return v;
}
__device__ __forceinline__ u64 wave_max(u64 v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const u64 t = __shfl_xor(v, o);
        v = t > v ? t : v;
    }
    return v;
}

// exclusive prefix sum over the 64 lanes of a wave
__device__ __forceinline__ u32 wave_exclusive(u32 v) {
    const int lane = threadIdx.x & 63;
    u32 x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const u32 t = __shfl_up(x, o);
        if (lane >= o) x += t;
    }
    return x - v;
}

}  // namespace chip

#endif
