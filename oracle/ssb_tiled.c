/*
 * ssb_tiled.c -- tiled CPU restatement of the kernels the reference code
 * generator emits for the three plain benchmark queries, used as the timed
 * CPU baseline.
 *
 * TEST INFRASTRUCTURE.  Used only by tests/ (checked against the plan
 * interpreter in ssb_oracle.c) and by the cpu_baseline leg of bench.py.  The
 * product never links it.
 *
 * The reference's CPU path is AdaptiveCpp's OpenMP backend running the SYCL
 * text its visitor emits.  That toolchain is absent here, so this file
 * restates what those kernels do, in the same shape:
 *   - 512-row tiles of 128 work-items x 4 items, item i of work-item t being
 *     row tile_offset + t + 128 i  (BlockLoad, include/crystal/load.h:6-36);
 *   - a flag per item, cleared by the fast-path predicates
 *     (BlockPredA*, include/crystal/pred.h:70-275), the universal 3VL filter
 *     (visitor.cpp:837-991) and failed probes;
 *   - reference hash tables int[2*len], slot (key - min) % len, key 0 =
 *     empty, CAS 0 -> key then a plain payload store
 *     (BlockBuildSelectivePHT_1/_2, include/crystal/join.h:124-253; lengths
 *     as computeBuildInfo, visitor.cpp:3881-3896);
 *   - probes that compare the slot's key (BlockProbeAndPHT_1/_2,
 *     join.h:7-120);
 *   - scalar aggregates: per-work-group partial sums, one relaxed u64 atomic
 *     per tile (consumeAggregateVector, visitor.cpp:3017-3111);
 *   - grouped aggregates: Horner perfect hash over catalog (min, distinct
 *     count), +1 and card + 1 for a catalog-nullable column
 *     (generatePerfectHash, visitor.cpp:2056-2127), group columns stored
 *     plainly and one relaxed u64 atomic add per surviving row into
 *     d_result[hash * tuple + slot] plus a fetch_or of the slot's validity
 *     bit (visitor.cpp:3135-3180, inline_math.h:112-158).
 * A tile's work-items are run as loops over the 512 items, stage by stage,
 * the way a CPU backend runs a work-group; tiles are spread over OpenMP
 * threads.  The timed region is the dimension builds plus the fact scan, as
 * the reference's generated execute_query runs them.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#define TILE 512
#define BT 128

typedef struct {
    const int32_t* key;
    const int32_t* payload;   /* NULL: PHT_1 (keys only) */
    const int32_t* filt;      /* build filter column (equality), NULL: none */
    int32_t filt_value;
    int64_t rows;
    int64_t len;              /* hash table slots */
    int32_t kmin;
} tl_dim;

static void tl_build(const tl_dim* d, int32_t* ht, int threads) {
    memset(ht, 0, (size_t)(2 * d->len) * sizeof(int32_t));
    const int64_t tiles = (d->rows + TILE - 1) / TILE;
#pragma omp parallel for num_threads(threads) schedule(static)
    for (int64_t tile = 0; tile < tiles; ++tile) {
        const int64_t off = tile * TILE;
        const int n = (int)(d->rows - off < TILE ? d->rows - off : TILE);
        for (int r = 0; r < n; ++r) {
            const int64_t row = off + r;
            if (d->filt && d->filt[row] != d->filt_value) continue;
            const int32_t k = d->key[row];
            const int64_t x = (int64_t)k - d->kmin;
            if (x < 0) continue;
            const int64_t s = x % d->len;
            int32_t expect = 0;
            __atomic_compare_exchange_n(&ht[2 * s], &expect, k, 0, __ATOMIC_RELAXED, __ATOMIC_RELAXED);
            if (d->payload) ht[2 * s + 1] = d->payload[row];
        }
    }
}

/* probe one tile's keys; clears flags of misses, returns payloads */
static inline void tl_probe(const int32_t* ht, int64_t len, int32_t kmin, const int32_t* keys, uint8_t* flags,
                            int32_t* payload, int n) {
    for (int i = 0; i < n; ++i) {
        if (!flags[i]) continue;
        const int64_t x = (int64_t)keys[i] - kmin;
        if (x < 0) { flags[i] = 0; continue; }
        const int64_t s = x % len;
        if (ht[2 * s] != keys[i]) { flags[i] = 0; continue; }
        if (payload) payload[i] = ht[2 * s + 1];
    }
}

/* C1: sum(lo_extendedprice * lo_discount) over fact x ddate where
 * d_year = year and dlo <= lo_discount <= dhi and lo_quantity < qlt. */
int tl_c1(const int32_t* lo_orderdate, const int32_t* lo_discount, const int32_t* lo_quantity,
          const int32_t* lo_extendedprice, int64_t n, const int32_t* d_datekey, const int32_t* d_year, int64_t dn,
          int32_t year, int32_t dlo, int32_t dhi, int32_t qlt, int threads, uint64_t* out) {
    tl_dim dd = {d_datekey, NULL, d_year, year, dn, 61131, 19920101};
    int32_t* ht = (int32_t*)malloc((size_t)(2 * dd.len) * sizeof(int32_t));
    if (!ht) return -1;
    tl_build(&dd, ht, threads);
    uint64_t total = 0;
    const int64_t tiles = (n + TILE - 1) / TILE;
#pragma omp parallel for num_threads(threads) schedule(static)
    for (int64_t tile = 0; tile < tiles; ++tile) {
        const int64_t off = tile * TILE;
        const int m = (int)(n - off < TILE ? n - off : TILE);
        uint8_t flags[TILE];
        int32_t it[TILE];
        for (int i = 0; i < m; ++i) flags[i] = (uint8_t)(lo_discount[off + i] >= dlo && lo_discount[off + i] <= dhi);
        for (int i = 0; i < m; ++i) flags[i] &= (uint8_t)(lo_quantity[off + i] < qlt);
        for (int i = 0; i < m; ++i) it[i] = lo_orderdate[off + i];
        tl_probe(ht, dd.len, dd.kmin, it, flags, NULL, m);
        uint64_t part = 0;
        for (int i = 0; i < m; ++i)
            if (flags[i]) part += (uint64_t)(int64_t)lo_extendedprice[off + i] * (uint64_t)(int64_t)lo_discount[off + i];
        __atomic_fetch_add(&total, part, __ATOMIC_RELAXED);
    }
    free(ht);
    *out = total;
    return 0;
}

/* C2: d_year, p_brand1, sum(lo_revenue) over fact x part x supplier x ddate
 * where p_category = cat and s_region = reg; result state d_result[h * 3 + k]
 * (k = 0 d_year, 1 p_brand1, 2 sum) over `groups` slots, validity bits in
 * `valid` (groups bits).  yr_min / br_min / br_card are catalog statistics. */
int tl_c2(const int32_t* lo_orderdate, const int32_t* lo_partkey, const int32_t* lo_suppkey, const int32_t* lo_revenue,
          int64_t n, const int32_t* p_partkey, const int32_t* p_category, const int32_t* p_brand1, int64_t pn,
          const int32_t* s_suppkey, const int32_t* s_region, int64_t sn, const int32_t* d_datekey, const int32_t* d_year,
          int64_t dn, int32_t cat, int32_t reg, int32_t yr_min, int64_t yr_card, int32_t br_min, int64_t br_card,
          int threads, uint64_t* d_result, uint64_t* valid) {
    tl_dim pd = {p_partkey, p_brand1, p_category, cat, pn, pn, 0};
    tl_dim sd = {s_suppkey, NULL, s_region, reg, sn, sn, 0};
    tl_dim dd = {d_datekey, d_year, NULL, 0, dn, 61131, 19920101};
    int32_t* pht = (int32_t*)malloc((size_t)(2 * pd.len) * sizeof(int32_t));
    int32_t* sht = (int32_t*)malloc((size_t)(2 * sd.len) * sizeof(int32_t));
    int32_t* dht = (int32_t*)malloc((size_t)(2 * dd.len) * sizeof(int32_t));
    if (!pht || !sht || !dht) { free(pht); free(sht); free(dht); return -1; }
    const int64_t groups = yr_card * br_card;
    memset(d_result, 0, (size_t)groups * 3 * sizeof(uint64_t));
    memset(valid, 0, (size_t)((groups + 63) / 64) * sizeof(uint64_t));
    tl_build(&sd, sht, threads);
    tl_build(&pd, pht, threads);
    tl_build(&dd, dht, threads);
    const int64_t tiles = (n + TILE - 1) / TILE;
#pragma omp parallel for num_threads(threads) schedule(static)
    for (int64_t tile = 0; tile < tiles; ++tile) {
        const int64_t off = tile * TILE;
        const int m = (int)(n - off < TILE ? n - off : TILE);
        uint8_t flags[TILE];
        int32_t it[TILE], brand[TILE], year[TILE];
        memset(flags, 1, (size_t)m);
        for (int i = 0; i < m; ++i) it[i] = lo_suppkey[off + i];
        tl_probe(sht, sd.len, sd.kmin, it, flags, NULL, m);
        for (int i = 0; i < m; ++i) it[i] = lo_partkey[off + i];
        tl_probe(pht, pd.len, pd.kmin, it, flags, brand, m);
        for (int i = 0; i < m; ++i) it[i] = lo_orderdate[off + i];
        tl_probe(dht, dd.len, dd.kmin, it, flags, year, m);
        for (int i = 0; i < m; ++i) {
            if (!flags[i]) continue;
            const int64_t h = (((int64_t)year[i] - yr_min) * br_card + ((int64_t)brand[i] - br_min)) % groups;
            d_result[h * 3 + 0] = (uint64_t)(int64_t)year[i];
            d_result[h * 3 + 1] = (uint64_t)(int64_t)brand[i];
            __atomic_fetch_add(&d_result[h * 3 + 2], (uint64_t)(int64_t)lo_revenue[off + i], __ATOMIC_RELAXED);
            if (!(__atomic_load_n(&valid[h >> 6], __ATOMIC_RELAXED) & (1ull << (h & 63))))
                __atomic_fetch_or(&valid[h >> 6], 1ull << (h & 63), __ATOMIC_RELAXED);
        }
    }
    free(pht); free(sht); free(dht);
    return 0;
}

/* C3: c_nation, s_nation, d_year, sum(lo_revenue) over fact x customer x
 * supplier x ddate where (lo_revenue > rev_gt or lo_revenue is null); state
 * d_result[h * 4 + k] (c_nation, s_nation, d_year, sum).  The catalog makes
 * s_nation nullable, so its hash term is s_nation - min + 1 with card + 1
 * (the payload register carries no validity: a NULL nation groups under its
 * raw value).  rev_valid: LSB-first u64 bitmap of lo_revenue, NULL = all
 * valid.  `valid` holds two bitmaps of `groups` bits each (words rounded up):
 * the group slots hit, then the slots whose SUM saw a valid value. */
int tl_c3(const int32_t* lo_custkey, const int32_t* lo_suppkey, const int32_t* lo_orderdate, const int32_t* lo_revenue,
          const uint64_t* rev_valid, int64_t n, const int32_t* c_custkey, const int32_t* c_nation, int64_t cn,
          const int32_t* s_suppkey, const int32_t* s_nation, int64_t sn, const int32_t* d_datekey, const int32_t* d_year,
          int64_t dn, int32_t rev_gt, int32_t cn_min, int64_t cn_card, int32_t sn_min, int64_t sn_card, int sn_nullable,
          int32_t yr_min, int64_t yr_card, int threads, uint64_t* d_result, uint64_t* valid) {
    tl_dim cd = {c_custkey, c_nation, NULL, 0, cn, cn, 0};
    tl_dim sd = {s_suppkey, s_nation, NULL, 0, sn, sn, 0};
    tl_dim dd = {d_datekey, d_year, NULL, 0, dn, 61131, 19920101};
    int32_t* cht = (int32_t*)malloc((size_t)(2 * cd.len) * sizeof(int32_t));
    int32_t* sht = (int32_t*)malloc((size_t)(2 * sd.len) * sizeof(int32_t));
    int32_t* dht = (int32_t*)malloc((size_t)(2 * dd.len) * sizeof(int32_t));
    if (!cht || !sht || !dht) { free(cht); free(sht); free(dht); return -1; }
    const int64_t sc = sn_card + (sn_nullable ? 1 : 0);
    const int64_t groups = cn_card * sc * yr_card;
    const int64_t words = (groups + 63) / 64;
    uint64_t* sum_valid = valid + words;
    memset(d_result, 0, (size_t)groups * 4 * sizeof(uint64_t));
    memset(valid, 0, (size_t)(2 * words) * sizeof(uint64_t));
    tl_build(&cd, cht, threads);
    tl_build(&sd, sht, threads);
    tl_build(&dd, dht, threads);
    const int64_t tiles = (n + TILE - 1) / TILE;
#pragma omp parallel for num_threads(threads) schedule(static)
    for (int64_t tile = 0; tile < tiles; ++tile) {
        const int64_t off = tile * TILE;
        const int m = (int)(n - off < TILE ? n - off : TILE);
        uint8_t flags[TILE];
        int32_t it[TILE], cnat[TILE], snat[TILE], year[TILE];
        for (int i = 0; i < m; ++i) {
            const int64_t row = off + i;
            const int v = rev_valid ? (int)((rev_valid[row >> 6] >> (row & 63)) & 1) : 1;
            flags[i] = (uint8_t)((v && lo_revenue[row] > rev_gt) || !v);
        }
        for (int i = 0; i < m; ++i) it[i] = lo_custkey[off + i];
        tl_probe(cht, cd.len, cd.kmin, it, flags, cnat, m);
        for (int i = 0; i < m; ++i) it[i] = lo_suppkey[off + i];
        tl_probe(sht, sd.len, sd.kmin, it, flags, snat, m);
        for (int i = 0; i < m; ++i) it[i] = lo_orderdate[off + i];
        tl_probe(dht, dd.len, dd.kmin, it, flags, year, m);
        for (int i = 0; i < m; ++i) {
            if (!flags[i]) continue;
            const int64_t row = off + i;
            const int64_t st = (int64_t)snat[i] - sn_min + (sn_nullable ? 1 : 0);
            const int64_t h = ((((int64_t)cnat[i] - cn_min) * sc + st) * yr_card + ((int64_t)year[i] - yr_min)) % groups;
            d_result[h * 4 + 0] = (uint64_t)(int64_t)cnat[i];
            d_result[h * 4 + 1] = (uint64_t)(int64_t)snat[i];
            d_result[h * 4 + 2] = (uint64_t)(int64_t)year[i];
            /* SUM over a NULL revenue adds nothing (the aggregate is guarded
             * by the value's validity) */
            const int v = rev_valid ? (int)((rev_valid[row >> 6] >> (row & 63)) & 1) : 1;
            if (v) {
                __atomic_fetch_add(&d_result[h * 4 + 3], (uint64_t)(int64_t)lo_revenue[row], __ATOMIC_RELAXED);
                if (!(__atomic_load_n(&sum_valid[h >> 6], __ATOMIC_RELAXED) & (1ull << (h & 63))))
                    __atomic_fetch_or(&sum_valid[h >> 6], 1ull << (h & 63), __ATOMIC_RELAXED);
            }
            if (!(__atomic_load_n(&valid[h >> 6], __ATOMIC_RELAXED) & (1ull << (h & 63))))
                __atomic_fetch_or(&valid[h >> 6], 1ull << (h & 63), __ATOMIC_RELAXED);
        }
    }
    free(cht); free(sht); free(dht);
    return 0;
}

int tl_max_threads(void) { return omp_get_max_threads(); }
