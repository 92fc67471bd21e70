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

namespace chip {

typedef unsigned long long u64;
typedef long long i64;
typedef unsigned int u32;
typedef int v4i __attribute__((ext_vector_type(4)));

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

__device__ __forceinline__ double bits_to_double(u64 v) { return __builtin_bit_cast(double, v); }
__device__ __forceinline__ u64 double_to_bits(double v) { return __builtin_bit_cast(u64, v); }

// ------------------------------------------------------------------ bitmaps
__device__ __forceinline__ u32 bit32(const u32* w, i64 row) { return (w[row >> 5] >> (row & 31)) & 1u; }
__device__ __forceinline__ u32 bit64(const u64* w, i64 idx) { return (u32)((w[idx >> 6] >> (idx & 63)) & 1ull); }
__device__ __forceinline__ void set_bit64(u64* w, i64 idx) { atomicOr(&w[idx >> 6], 1ull << (idx & 63)); }

// ------------------------------------------------------------------ hash tables
// Reference slot: (k - kmin) % len in int arithmetic.  A negative k - kmin
// (an out-of-bounds access in the reference) never matches here.
__device__ __forceinline__ int ht_slot(int k, int kmin, int len, int& x) {
    x = (int)((u32)k - (u32)kmin);
    return x < len ? x : x % len;
}

// Compact tagged layout.  Entry of `eb` bits per slot:
//   bit 0            occupied
//   bits 1..tb       tag t: the key stored in slot s is kmin + s + (jlo+t)*len
//   bits tb+1..      payload - pmin
// Valid when the keys inserted are unique and span at most `len` consecutive
// values starting at or above kmin (then no two keys share a slot).  An empty
// slot reads as key 0 / payload 0, exactly like the zero-filled reference table.
__device__ __forceinline__ bool ht_probe_compact(const u32* w, int fk, int kmin, int len, int eb, int tb, int jlo,
                                                 int pmin, int& payload) {
    int x;
    const int s = ht_slot(fk, kmin, len, x);
    if (x < 0) return false;
    const u64 bit = (u64)(u32)s * (u32)eb;
    const u32 mask = eb >= 32 ? 0xffffffffu : ((1u << eb) - 1u);
    const u32 e = (w[bit >> 5] >> (bit & 31)) & mask;
    const bool occ = (e & 1u) != 0;
    const int tag = tb ? (int)((e >> 1) & ((1u << tb) - 1u)) : 0;
    const int key = occ ? (int)((u32)kmin + (u32)s + (u32)((jlo + tag) * len)) : 0;
    payload = occ ? pmin + (int)(e >> (1 + tb)) : 0;
    return key == fk;
}

__device__ __forceinline__ void ht_insert_compact(u32* w, int k, int kmin, int len, int eb, int tb, int jlo, int pmin,
                                                  int payload) {
    int x;
    const int s = ht_slot(k, kmin, len, x);
    if (x < 0) return;
    const u32 tag = (u32)(x / len - jlo);
    const u32 e = 1u | (tag << 1) | ((u32)(payload - pmin) << (1 + tb));
    const u64 bit = (u64)(u32)s * (u32)eb;
    atomicOr(&w[bit >> 5], e << (bit & 31));
}

// Key-indexed layout.  Entry x = fk - dmin of `db` bits holds
// payload - pbase; with `off` set, code 0 marks a key that was not inserted
// (filtered out on the build side, or no such row) and pbase = pmin - 1;
// without it every key of the span is present and pbase = pmin.  Entries may
// straddle two words (tables carry one padding word).  Same validity
// condition as the compact layout (unique keys spanning at most `len`
// consecutive values at or above the reference kmin: then ht[slot] == fk holds
// exactly for the inserted keys), and dmin..dmin+dspan-1 covers every key.
// db == 0: every key present, no payload: the probe is the range test alone.
__device__ __forceinline__ u32 dense_entry(const u32* w, u32 x, int db) {
    const u32 b = x * (u32)db;
    const u32 lo = w[b >> 5] >> (b & 31);
    const u32 mask = (1u << db) - 1u;
    if ((b & 31) + (u32)db <= 32u) return lo & mask;
    return (lo | (w[(b >> 5) + 1] << (32u - (b & 31)))) & mask;
}
// db a power of two: an entry never straddles (one load per probe)
__device__ __forceinline__ u32 dense_entry_aligned(const u32* w, u32 x, int db) {
    const u32 b = x * (u32)db;
    return (w[b >> 5] >> (b & 31)) & ((1u << db) - 1u);
}
template <bool ALIGNED>
__device__ __forceinline__ bool ht_probe_dense(const u32* w, int fk, int dmin, int dspan, int db, int pbase, bool off,
                                               int& payload) {
    const u32 x = (u32)fk - (u32)dmin;
    if (x >= (u32)dspan) return false;
    if (db == 0) return true;
    const u32 e = ALIGNED ? dense_entry_aligned(w, x, db) : dense_entry(w, x, db);
    payload = pbase + (int)e;
    return !off || e != 0u;
}
__device__ __forceinline__ void ht_insert_dense(u32* w, int k, int dmin, int dspan, int db, int pbase, int payload) {
    const u32 x = (u32)k - (u32)dmin;
    if (x >= (u32)dspan || db == 0) return;
    const u32 c = (u32)(payload - pbase);
    const u32 b = x * (u32)db;
    atomicOr(&w[b >> 5], c << (b & 31));
    if ((b & 31) + (u32)db > 32u) atomicOr(&w[(b >> 5) + 1], c >> (32u - (b & 31)));
}

// Reference layout: int[2*len]; PHT_1 keys in ht[s], PHT_2 [key, payload] pairs.
__device__ __forceinline__ bool ht_probe_ref(const int* ht, int fk, int kmin, int len, bool pair, int& payload) {
    int x;
    const int s = ht_slot(fk, kmin, len, x);
    if (x < 0) return false;
    if (!pair) return ht[s] == fk;
    if (ht[2 * s] != fk) return false;
    payload = ht[2 * s + 1];
    return true;
}

__device__ __forceinline__ void ht_insert_ref(int* ht, int k, int kmin, int len, bool pair, int payload) {
    int x;
    const int s = ht_slot(k, kmin, len, x);
    if (x < 0) return;
    if (!pair) {
        atomicCAS(&ht[s], 0, k);
        return;
    }
    atomicCAS(&ht[2 * s], 0, k);
    ht[2 * s + 1] = payload;  // unconditional, as BlockBuildSelectivePHT_2 (join.h:203-205)
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

}  // namespace chip

#endif
