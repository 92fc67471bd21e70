// ht_layouts.h -- dimension hash-table layouts of the generated query
// kernels (build: insert, scan: probe).  Part of the device primitive
// library (crystal_hip.h includes it); written so that it also compiles as
// plain C++ on the host, given atomicOr / atomicCAS and empty __device__ /
// __forceinline__, which is how the layouts are tested against the
// reference table semantics on the CPU.
//
// Reference (include/crystal/join.h): HASH(k, len, min) = (k - min) % len,
// build CAS 0 -> key into a zero-filled int table (PHT_2 stores the payload
// beside the key, unconditionally), probe matches iff ht[slot] == fk.
#ifndef CQ_HT_LAYOUTS_H
#define CQ_HT_LAYOUTS_H

namespace chip {

typedef unsigned int u32;
typedef unsigned long long u64;

// A table word.  STREAM: a table much larger than L1 that is probed at
// random; its lines are loaded non-temporally so that they do not push the
// small tables probed beside it out of L1.
template <bool STREAM>
__device__ __forceinline__ u32 ht_word(const u32* w, u32 i) {
#if defined(__clang__)
    if (STREAM) return __builtin_nontemporal_load(w + i);
#endif
    return w[i];
}

// low 32 bits of the product of the low 24 bits of a and b (the masks let
// the compiler use the full-rate 24-bit multiply)
__device__ __forceinline__ u32 mul24(u32 a, u32 b) { return (a & 0xffffffu) * (b & 0xffffffu); }

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
template <bool STREAM = false>
__device__ __forceinline__ bool ht_probe_compact(const u32* w, int fk, int kmin, int len, int eb, int tb, int jlo,
                                                 int pmin, int& payload) {
    int x;
    const int s = ht_slot(fk, kmin, len, x);
    if (x < 0) return false;
    const u64 bit = (u64)(u32)s * (u32)eb;
    const u32 mask = eb >= 32 ? 0xffffffffu : ((1u << eb) - 1u);
    const u32 e = (ht_word<STREAM>(w, (u32)(bit >> 5)) >> (bit & 31)) & mask;
    const bool occ = (e & 1u) != 0;
    const int tag = tb ? (int)((e >> 1) & ((1u << tb) - 1u)) : 0;
    const int key = occ ? (int)((u32)kmin + (u32)s + (u32)((jlo + tag) * len)) : 0;
    payload = occ ? pmin + (int)(e >> (1 + tb)) : 0;
    return key == fk;
}

// The same probe in three steps, so that a thread can issue the loads of
// all its rows before it waits for the first (a wave then has one table
// round trip per tile instead of one per row):
//   compact_locate  slot of fk (a dead or unmatched row reads word 0)
//   ht_word         the entry's word
//   compact_decode  match and payload, as ht_probe_compact
__device__ __forceinline__ u32 compact_locate(int fk, bool live, int kmin, int len, int eb, int& s) {
    int x;
    s = ht_slot(fk, kmin, len, x);
    if (!live || x < 0) s = -1;
    return s < 0 ? 0u : (u32)(((u64)(u32)s * (u32)eb) >> 5);
}
__device__ __forceinline__ bool compact_decode(u32 word, int s, int fk, int kmin, int len, int eb, int tb, int jlo, int pmin,
                                               int& payload) {
    if (s < 0) return false;
    const u64 bit = (u64)(u32)s * (u32)eb;
    const u32 mask = eb >= 32 ? 0xffffffffu : ((1u << eb) - 1u);
    const u32 e = (word >> (bit & 31)) & mask;
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

// the low `db` bits (db = 32: all of them; a shift by 32 is undefined)
__device__ __forceinline__ u32 width_mask(int db) { return db >= 32 ? 0xffffffffu : (1u << db) - 1u; }

// Key-indexed layout.  Entry x = fk - dmin of `db` bits holds
// payload - pbase; with `off` set, code 0 marks a key that was not inserted
// (filtered out on the build side, or no such row) and pbase = pmin - 1;
// without it every key of the span is present and pbase = pmin.  Entries may
// straddle two words (tables carry one padding word).  Same validity
// condition as the compact layout (unique keys spanning at most `len`
// consecutive values at or above the reference kmin: then ht[slot] == fk holds
// exactly for the inserted keys), and dmin..dmin+dspan-1 covers every key.
// db == 0: every key present and a single payload value (pbase) or none:
// the probe is the range test alone (no table).
template <bool STREAM = false>
__device__ __forceinline__ u32 dense_entry(const u32* w, u32 x, int db) {
    const u32 b = x * (u32)db;
    const u32 lo = ht_word<STREAM>(w, b >> 5) >> (b & 31);
    const u32 mask = width_mask(db);
    if ((b & 31) + (u32)db <= 32u) return lo & mask;
    return (lo | (ht_word<STREAM>(w, (b >> 5) + 1) << (32u - (b & 31)))) & mask;
}
// db a power of two: an entry never straddles (one load per probe)
template <bool STREAM = false>
__device__ __forceinline__ u32 dense_entry_aligned(const u32* w, u32 x, int db) {
    const u32 b = x * (u32)db;
    return (ht_word<STREAM>(w, b >> 5) >> (b & 31)) & width_mask(db);
}
// Entry index of key k: k - dmin, or for calendar keys (cal = first year,
// every table key a yyyymmdd date) ((year - cal) * 12 + month - 1) * 31 +
// day - 1, which packs 7 years of dates into 2,604 entries instead of the
// 61,131 of the numeric range.  An index >= dspan: no entry (also every key
// that is not such a date, since no table key is one).
__device__ __forceinline__ u32 dense_index(int k, int dmin, int dspan, int cal) {
    if (!cal) return (u32)k - (u32)dmin;
    // Branch-free (selects only).  u = (y - cal) * 10000 + m * 100 + d must
    // lie below years * 10000 (negative keys and earlier years wrap above).
    // With at most 13 years, u < 130,000 and y - cal = floor((u + 0.5) / 10^4)
    // in single precision is exact: (u + 0.5) / 10^4 stays at least 5e-5 from
    // an integer, far beyond the float rounding (< 2e-6 here); the month is
    // md * 10486 >> 20, exact for md < 43,690.  All full-rate 24-bit / fp32
    // operations instead of 32-bit division.
    const u32 years = (u32)dspan / 372u;
    const u32 u = (u32)k - (u32)cal * 10000u;
    const bool inr = u < years * 10000u;
    const u32 yo = years <= 13u ? (u32)(((float)u + 0.5f) * 1.0e-4f) : u / 10000u;
    // every factor below 2^24 whenever the key is in range (out of range,
    // the products are ignored): 24-bit multiplies
    const u32 md = u - mul24(yo, 10000u);
    const u32 m = mul24(md, 10486u) >> 20;
    const u32 d = md - mul24(m, 100u);
    const bool ok = inr && m - 1u < 12u && d - 1u < 31u;
    const u32 x = mul24(mul24(yo, 12u) + m - 1u, 31u) + (d - 1u);
    return ok ? x : (u32)dspan;
}

// Probe key 0 in the zero-filled reference table matches an EMPTY slot
// (key 0 == empty key, payload 0).  `zmode` reproduces that (fixed per join
// by the code generator from the reference slot of key 0):
//   0  nothing special (key 0 falls below the reference kmin: no match)
//   1  no key shares key 0's slot: key 0 always matches, payload 0
//   2  key zx + dmin != 0 shares the slot: key 0 matches iff it is absent
//   3  key 0 itself owns the slot: key 0 always matches, payload if inserted
template <bool ALIGNED, bool STREAM = false>
__device__ __forceinline__ bool ht_probe_dense(const u32* w, int fk, int dmin, int dspan, int db, int pbase, bool off,
                                               int zmode, int zx, int& payload, int cal = 0) {
    if (zmode != 0 && fk == 0) {
        if (zmode == 1) {
            payload = 0;
            return true;
        }
        const u32 e0 = db == 0 ? 0u : (ALIGNED ? dense_entry_aligned<STREAM>(w, (u32)zx, db) : dense_entry<STREAM>(w, (u32)zx, db));
        const bool in0 = db == 0 || !off || e0 != 0u;
        payload = (zmode == 3 && in0) ? pbase + (int)e0 : 0;
        return zmode == 3 || !in0;
    }
    const u32 x = dense_index(fk, dmin, dspan, cal);
    if (x >= (u32)dspan) return false;
    if (db == 0) {  // every key present, one payload value (or none)
        payload = pbase;
        return true;
    }
    const u32 e = ALIGNED ? dense_entry_aligned<STREAM>(w, x, db) : dense_entry<STREAM>(w, x, db);
    payload = pbase + (int)e;
    return !off || e != 0u;
}
// The key-indexed probe in three steps (see compact_locate):
//   dense_locate  entry x to read and what decides the match (kind):
//                 0 no match, 1 the entry (present code), 2 key 0 matches
//                 iff entry zx is absent, 3 key 0 matches (payload if
//                 present), 4 key 0 matches with payload 0
//   dense_words   the entry's word (and the next one for exact widths)
//   dense_decode  match and payload, as ht_probe_dense
//                 (selects only: the generated code passes zmode as a
//                 literal, so with zmode 0 the key-0 cases vanish)
__device__ __forceinline__ u32 dense_locate(int fk, bool live, int dmin, int dspan, int zmode, int zx, int cal, u32& x) {
    const u32 y = dense_index(fk, dmin, dspan, cal);
    const bool in = y < (u32)dspan;
    u32 kind = in ? 1u : 0u;
    u32 at = in ? y : 0u;
    if (zmode != 0) {
        const bool z = fk == 0;
        kind = z ? (zmode == 1 ? 4u : (u32)zmode) : kind;
        at = z ? (zmode == 1 ? 0u : (u32)zx) : at;
    }
    x = live ? at : 0u;  // a dead item reads entry 0 (in the table, one shared line)
    return live ? kind : 0u;
}
#if defined(__HIP_DEVICE_COMPILE__)
// A key-indexed table in HBM read through a buffer resource: the per-lane
// address is a 32-bit byte offset (no 64-bit address arithmetic per lane)
// and reads beyond `bytes` return 0.  dword 3 of the descriptor: the gfx9
// family's raw 32-bit buffer format.
typedef __amdgpu_buffer_rsrc_t table_rsrc;
__device__ __forceinline__ table_rsrc make_table_rsrc(const u32* w, u32 bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<u32*>(w), (short)0, (int)bytes, 0x00020000);
}
// entry words as dense_words<ALIGNED> (byte-wide entries: the byte itself)
template <bool ALIGNED>
__device__ __forceinline__ void dense_words_rsrc(table_rsrc r, u32 x, int db, u32& lo, u32& hi) {
    const u32 b = x * (u32)db;
    if (db == 8) {
        lo = (u32)(unsigned char)__builtin_amdgcn_raw_buffer_load_b8(r, x, 0, 0) << (b & 31);
        hi = 0u;
        return;
    }
    lo = db == 0 ? 0u : (u32)__builtin_amdgcn_raw_buffer_load_b32(r, (b >> 5) * 4u, 0, 0);
    hi = (ALIGNED || db == 0) ? 0u : (u32)__builtin_amdgcn_raw_buffer_load_b32(r, ((b >> 5) + 1u) * 4u, 0, 0);
}
#endif

template <bool ALIGNED, bool STREAM = false>
__device__ __forceinline__ void dense_words(const u32* w, u32 x, int db, u32& lo, u32& hi) {
    const u32 b = x * (u32)db;
    lo = db == 0 ? 0u : ht_word<STREAM>(w, b >> 5);
    hi = (ALIGNED || db == 0) ? 0u : ht_word<STREAM>(w, (b >> 5) + 1);  // the padding word keeps this in the table
}
// (selects only; ZMODE as in dense_locate: with 0, kind is 0 or 1)
template <bool ALIGNED, int ZMODE = -1>
__device__ __forceinline__ bool dense_decode(u32 kind, u32 x, u32 lo, u32 hi, int db, int pbase, bool off, int& payload) {
    const u32 b = x * (u32)db;
    const u32 mask = width_mask(db);
    const u32 e = db == 0 ? 0u : (ALIGNED ? (lo >> (b & 31)) & mask : (u32)((((u64)hi << 32) | lo) >> (b & 31)) & mask);
    const bool in = db == 0 || !off || e != 0u;
    const int pe = pbase + (int)e;
    if (ZMODE == 0) {
        payload = pe;
        return kind != 0u && in;
    }
    // kind 1: the entry; 2: key 0 matches iff entry zx is absent; 3: key 0
    // matches, payload if present; 4: key 0 matches, payload 0
    payload = (kind == 1u || (kind == 3u && in)) ? pe : 0;
    return kind == 1u ? in : (kind == 2u ? !in : kind >= 3u);
}

__device__ __forceinline__ void ht_insert_dense(u32* w, int k, int dmin, int dspan, int db, int pbase, int payload, int cal = 0) {
    const u32 x = dense_index(k, dmin, dspan, cal);
    if (x >= (u32)dspan || db == 0) return;
    const u32 c = (u32)(payload - pbase);
    // Byte-, halfword- and word-wide entries own their bytes (the layout
    // is used only for unique build keys, so one row writes each entry):
    // plain stores, which neighbouring lanes with consecutive keys coalesce,
    // instead of atomic ORs that serialize on a shared word.
    if (db == 8) {
        reinterpret_cast<unsigned char*>(w)[x] = (unsigned char)c;
        return;
    }
    if (db == 16) {
        reinterpret_cast<unsigned short*>(w)[x] = (unsigned short)c;
        return;
    }
    if (db == 32) {
        w[x] = c;
        return;
    }
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

}  // namespace chip

#endif
