/*
 * ssb_gen.c -- seeded synthetic star schema (SSB shape).  TEST
 * INFRASTRUCTURE: the oracle side of the generator; the product ships its
 * own GPU generator and the tests require both to agree bit for bit.
 *
 * Every value is a pure function of (seed, table, column, row): a
 * counter-based splitmix64 stream per column, so any row range can be
 * produced independently (GPU threads, CPU chunks, test samples).
 *
 * Domains follow the reference's dictionary encodings
 * (bench/loader/convert.py:53-93: nation/region indices, city = nation*10+d,
 * p_mfgr = category/5, p_brand1 = category*40 + b) and the sizes of the SSB
 * generator (SUPPLIER 2,000*SF, CUSTOMER 30,000*SF, PART
 * 200,000*(1+floor(log2 SF)), DDATE 2,556 days from 1992-01-01).
 *
 * Parity: pinned by tests/test_generator.py (closed-form domain checks,
 * exact date calendar, PK density) and by the GPU generator test.
 */
#include "ssb_oracle.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define GAMMA 0x9E3779B97F4A7C15ULL

uint64_t orc_splitmix64(uint64_t x) {
    x += GAMMA;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

uint64_t orc_name_hash(const char* s) {
    uint64_t h = 0xCBF29CE484222325ULL;
    for (; *s; ++s) {
        h ^= (uint8_t)*s;
        h *= 0x100000001B3ULL;
    }
    return h;
}

static uint64_t stream_seed(uint64_t seed, const char* table, const char* column, const char* suffix) {
    char buf[160];
    snprintf(buf, sizeof buf, "%s.%s%s", table, column, suffix ? suffix : "");
    return orc_splitmix64(orc_name_hash(buf) ^ seed);
}

/* value of stream `s` at `row`, uniform over [lo, hi] */
static int64_t uni(uint64_t s, int64_t row, int64_t lo, int64_t hi) {
    uint64_t h = orc_splitmix64(s + (uint64_t)row * GAMMA);
    uint64_t range = (uint64_t)(hi - lo + 1);
    return lo + (int64_t)(((unsigned __int128)h * range) >> 64);
}

static int is_leap(int y) { return (y % 4 == 0 && y % 100 != 0) || y % 400 == 0; }

/* day d (0-based from 1992-01-01) -> year, month (1-12), day of month, day of year (1-based) */
static void civil(int64_t d, int* y, int* m, int* dom, int* doy) {
    static const int mdays[12] = {31, 28, 31, 30, 31, 30, 31, 31, 30, 31, 30, 31};
    int yy = 1992;
    for (;;) {
        int len = is_leap(yy) ? 366 : 365;
        if (d < len) break;
        d -= len;
        ++yy;
    }
    *doy = (int)d + 1;
    int mm = 0;
    for (;;) {
        int len = mdays[mm] + (mm == 1 && is_leap(yy) ? 1 : 0);
        if (d < len) break;
        d -= len;
        ++mm;
    }
    *y = yy;
    *m = mm + 1;
    *dom = (int)d + 1;
}

int32_t orc_datekey(int64_t d) {
    int y, m, dom, doy;
    civil(d, &y, &m, &dom, &doy);
    return y * 10000 + m * 100 + dom;
}

#define DATE_ROWS 2556

static int floor_log2(int v) {
    int r = 0;
    while (v > 1) { v >>= 1; ++r; }
    return r;
}

int64_t orc_table_rows(const orc_genspec* spec, const char* table) {
    int sf = spec->sf < 1 ? 1 : spec->sf;
    if (!strcmp(table, "LINEORDER")) return spec->fact_rows > 0 ? spec->fact_rows : 6000000LL * sf;
    if (!strcmp(table, "PART")) return 200000LL * (1 + floor_log2(sf));
    if (!strcmp(table, "SUPPLIER")) return 2000LL * sf;
    if (!strcmp(table, "CUSTOMER")) return 30000LL * sf;
    if (!strcmp(table, "DDATE")) return DATE_ROWS;
    return -1;
}

static const char* LO_COLS[] = {"lo_custkey", "lo_partkey", "lo_suppkey", "lo_orderdate", "lo_commitdate",
                                "lo_quantity", "lo_extendedprice", "lo_discount", "lo_revenue",
                                "lo_supplycost", "lo_tax", "lo_orderkey", "lo_linenumber"};
static const char* P_COLS[] = {"p_partkey", "p_mfgr", "p_category", "p_brand1", "p_size"};
static const char* S_COLS[] = {"s_suppkey", "s_city", "s_nation", "s_region"};
static const char* C_COLS[] = {"c_custkey", "c_city", "c_nation", "c_region", "c_mktsegment"};
static const char* D_COLS[] = {"d_datekey", "d_year", "d_month", "d_yearmonthnum", "d_daynuminweek",
                               "d_daynuminmonth", "d_daynuminyear", "d_weeknuminyear"};

int orc_table_columns(const char* table, const char** names, int max_names) {
    const char** src = NULL;
    int n = 0;
    if (!strcmp(table, "LINEORDER")) { src = LO_COLS; n = (int)(sizeof LO_COLS / sizeof *LO_COLS); }
    else if (!strcmp(table, "PART")) { src = P_COLS; n = (int)(sizeof P_COLS / sizeof *P_COLS); }
    else if (!strcmp(table, "SUPPLIER")) { src = S_COLS; n = (int)(sizeof S_COLS / sizeof *S_COLS); }
    else if (!strcmp(table, "CUSTOMER")) { src = C_COLS; n = (int)(sizeof C_COLS / sizeof *C_COLS); }
    else if (!strcmp(table, "DDATE")) { src = D_COLS; n = (int)(sizeof D_COLS / sizeof *D_COLS); }
    else return -1;
    for (int i = 0; i < n && i < max_names; ++i) names[i] = src[i];
    return n;
}

int orc_column_null_bp(const orc_genspec* spec, const char* table, const char* column) {
    if (!spec->nullable || !*spec->nullable) return 0;
    char want[160];
    snprintf(want, sizeof want, "%s.%s:", table, column);
    size_t wl = strlen(want);
    const char* p = spec->nullable;
    while (*p) {
        while (*p == ',' || *p == ' ') ++p;
        if (!strncmp(p, want, wl)) return atoi(p + wl);
        while (*p && *p != ',') ++p;
    }
    return 0;
}

/* Column recipe: resolved once per call so the per-row loop does no string work. */
enum { K_KEY, K_UNI, K_DATEFK, K_PCAT, K_PMFGR, K_PBRAND, K_NATION, K_REGION, K_CITY,
       K_DKEY, K_DYEAR, K_DMONTH, K_DYM, K_DDOW, K_DDOM, K_DDOY, K_DWEEK, K_ORDERKEY, K_LINENUMBER };
typedef struct { int kind; uint64_t s0, s1; int64_t lo, hi; } recipe;

static int resolve(const orc_genspec* sp, const char* t, const char* c, recipe* rc) {
    const uint64_t seed = sp->seed;
    memset(rc, 0, sizeof *rc);
#define S(col) stream_seed(seed, t, col, NULL)
#define UNI(l, h) do { rc->kind = K_UNI; rc->s0 = S(c); rc->lo = (l); rc->hi = (h); return 0; } while (0)
    if (!strcmp(t, "LINEORDER")) {
        if (!strcmp(c, "lo_custkey")) UNI(1, orc_table_rows(sp, "CUSTOMER"));
        if (!strcmp(c, "lo_partkey")) UNI(1, orc_table_rows(sp, "PART"));
        if (!strcmp(c, "lo_suppkey")) UNI(1, orc_table_rows(sp, "SUPPLIER"));
        if (!strcmp(c, "lo_orderdate") || !strcmp(c, "lo_commitdate")) {
            rc->kind = K_DATEFK; rc->s0 = S(c); rc->lo = 0; rc->hi = DATE_ROWS - 1; return 0;
        }
        if (!strcmp(c, "lo_quantity")) UNI(1, 50);
        if (!strcmp(c, "lo_extendedprice")) UNI(90000, 10500000);
        if (!strcmp(c, "lo_discount")) UNI(0, 10);
        if (!strcmp(c, "lo_revenue")) UNI(100000, 10000000);
        if (!strcmp(c, "lo_supplycost")) UNI(60000, 600000);
        if (!strcmp(c, "lo_tax")) UNI(0, 8);
        /* orders of 4 lines: order key row / 4 + 1, line number row % 4 + 1 */
        if (!strcmp(c, "lo_orderkey")) { rc->kind = K_ORDERKEY; return 0; }
        if (!strcmp(c, "lo_linenumber")) { rc->kind = K_LINENUMBER; return 0; }
        return -1;
    }
    if (!strcmp(t, "PART")) {
        rc->s0 = S("p_category");
        if (!strcmp(c, "p_partkey")) { rc->kind = K_KEY; return 0; }
        if (!strcmp(c, "p_category")) { rc->kind = K_PCAT; return 0; }
        if (!strcmp(c, "p_mfgr")) { rc->kind = K_PMFGR; return 0; }
        if (!strcmp(c, "p_brand1")) { rc->kind = K_PBRAND; rc->s1 = S(c); return 0; }
        if (!strcmp(c, "p_size")) UNI(1, 50);
        return -1;
    }
    if (!strcmp(t, "SUPPLIER") || !strcmp(t, "CUSTOMER")) {
        const int supp = t[0] == 'S';
        const char* col = c + 2;
        if (strncmp(c, supp ? "s_" : "c_", 2) != 0) return -1;
        rc->s0 = S(supp ? "s_nation" : "c_nation");
        if (!strcmp(col, supp ? "suppkey" : "custkey")) { rc->kind = K_KEY; return 0; }
        if (!strcmp(col, "nation")) { rc->kind = K_NATION; return 0; }
        if (!strcmp(col, "region")) { rc->kind = K_REGION; return 0; }
        if (!strcmp(col, "city")) { rc->kind = K_CITY; rc->s1 = S(c); return 0; }
        if (!supp && !strcmp(col, "mktsegment")) UNI(0, 4);
        return -1;
    }
    if (!strcmp(t, "DDATE")) {
        if (!strcmp(c, "d_datekey")) { rc->kind = K_DKEY; return 0; }
        if (!strcmp(c, "d_year")) { rc->kind = K_DYEAR; return 0; }
        if (!strcmp(c, "d_month")) { rc->kind = K_DMONTH; return 0; }
        if (!strcmp(c, "d_yearmonthnum")) { rc->kind = K_DYM; return 0; }
        if (!strcmp(c, "d_daynuminweek")) { rc->kind = K_DDOW; return 0; }
        if (!strcmp(c, "d_daynuminmonth")) { rc->kind = K_DDOM; return 0; }
        if (!strcmp(c, "d_daynuminyear")) { rc->kind = K_DDOY; return 0; }
        if (!strcmp(c, "d_weeknuminyear")) { rc->kind = K_DWEEK; return 0; }
        return -1;
    }
#undef UNI
#undef S
    return -1;
}

static int32_t cell(const recipe* rc, int64_t r) {
    int y, m, dom, doy;
    switch (rc->kind) {
        case K_KEY: return (int32_t)(r + 1);
        case K_UNI: return (int32_t)uni(rc->s0, r, rc->lo, rc->hi);
        case K_DATEFK: return orc_datekey(uni(rc->s0, r, rc->lo, rc->hi));
        case K_PCAT: return (int32_t)uni(rc->s0, r, 0, 24);
        case K_PMFGR: return (int32_t)(uni(rc->s0, r, 0, 24) / 5);
        case K_PBRAND: return (int32_t)(uni(rc->s0, r, 0, 24) * 40 + uni(rc->s1, r, 0, 39));
        case K_NATION: return (int32_t)uni(rc->s0, r, 0, 24);
        case K_REGION: return (int32_t)(uni(rc->s0, r, 0, 24) / 5);
        case K_CITY: return (int32_t)(uni(rc->s0, r, 0, 24) * 10 + uni(rc->s1, r, 0, 9));
        case K_ORDERKEY: return (int32_t)(r / 4 + 1);
        case K_LINENUMBER: return (int32_t)(r % 4 + 1);
        default: break;
    }
    civil(r, &y, &m, &dom, &doy);
    switch (rc->kind) {
        case K_DKEY: return y * 10000 + m * 100 + dom;
        case K_DYEAR: return y;
        case K_DMONTH: return m;
        case K_DYM: return y * 100 + m;
        case K_DDOW: return (int32_t)((r + 3) % 7) + 1;
        case K_DDOM: return dom;
        case K_DDOY: return doy;
        case K_DWEEK: return (doy - 1) / 7 + 1;
        default: return 0;
    }
}

int orc_gen_column(const orc_genspec* spec, const char* table, const char* column,
                   int64_t start, int64_t count, int32_t* out, uint64_t* validity) {
    recipe rc;
    if (resolve(spec, table, column, &rc) != 0) return -1;
    const int bp = orc_column_null_bp(spec, table, column);
    const uint64_t vseed = stream_seed(spec->seed, table, column, "#valid");
    if (validity && start % 64 != 0) return -1;
    /* counter-based cells: any row range is generated independently, so
     * 64-row groups (one validity word each) are spread over the threads */
    const int64_t groups = (count + 63) / 64;
#pragma omp parallel for schedule(static)
    for (int64_t w = 0; w < groups; ++w) {
        uint64_t word = 0;
        const int64_t end = (w + 1) * 64 < count ? (w + 1) * 64 : count;
        for (int64_t i = w * 64; i < end; ++i) {
            const int64_t r = start + i;
            const int32_t v = cell(&rc, r);
            int valid = 1;
            if (bp > 0) valid = uni(vseed, r, 0, 9999) >= bp;
            out[i] = valid ? v : 0;
            if (valid) word |= 1ULL << (i & 63);
        }
        if (validity) validity[w] = word;
    }
    return 0;
}
