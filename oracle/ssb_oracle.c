/*
 * ssb_oracle.c -- CPU restatement of the reference JIT query path.
 *
 * TEST INFRASTRUCTURE.  Used only by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py, as the checker.  The product never links it.
 *
 * The interpreter walks an optimized plan (S-expression printed by the
 * product's front end) and computes, row by row, the device state that the
 * kernels emitted by the reference code generator leave behind
 * (d_result u64 slots + d_result_validity bits), then shapes the result the
 * way the reference host fetch does.  Functions and the reference code they
 * restate:
 *
 *   build_info()       computeBuildInfo        src/core/visitor.cpp:3864-4016
 *   build_filter()     emitBuildFilter         src/core/visitor.cpp:4047-4201
 *   build_tables()     emitPHTBuildKernel      src/core/visitor.cpp:4206-4290
 *                      BlockBuildSelectivePHT  include/crystal/join.h:124-253
 *   probe()            consumeHashJoinVector   src/core/visitor.cpp:4556-4731
 *                      BlockProbeAndPHT_1/_2   include/crystal/join.h:7-120
 *   build_tables() /   MHT choice (non-unique key) visitor.cpp:3946-3970,
 *   mht_probe()        emitMHTBuildKernels     visitor.cpp:4295-4464,
 *                      BlockBuildMHT_Count/_Write, ProbeMultiHT join.h:276-364,
 *                      the row-expanding probe visitor.cpp:4671-4730
 *   fast_filter()      consumeFilterVector     src/core/visitor.cpp:1823-1893
 *                      BlockPred* (int items)  include/crystal/pred.h:70-275
 *   vexpr()/vvalid()   expressionValueNoEmit / expressionValidExpr
 *                                              src/core/visitor.cpp:837-991
 *   safe_binop()       db::safe_*              include/core/inline_math.h:15-94
 *   mexpr()            translateMathExprPush   src/core/visitor.cpp:1947-2018
 *   group_hash()       generatePerfectHash     src/core/visitor.cpp:2056-2127
 *   aggregate_row()    consumeAggregateVector  src/core/visitor.cpp:2987-3196
 *   finalize_avg()     emitAggregateFinalizationIfNeeded  visitor.cpp:2505-2586
 *   shape_result()     emitAggregateDenseColumnarMaterialization 2589-2697,
 *                      emitAggregateSmallFusedColumnarMaterialization 2774-2885,
 *                      emitSortLimitPostExecution 524-581 (comparator 307-360),
 *                      DatabaseInstance::executeQuery src/db/database_instance.cpp:180-439
 *   having()           aggregateTupleExprValue/Valid  visitor.cpp:2277-2450
 *   finalize stats     CatalogCacheManager::getStatsForColumn
 *                                              include/core/catalog_cache.h:76-123
 *
 * Parity pins (tests/test_oracle.py):
 *   - reference golden fixture test/nullable_data (nullable_numbers.csv,
 *     expected.tsv) for COUNT/SUM/AVG/IS NULL/CASE/GROUP BY/ORDER BY;
 *   - known-answer vectors produced by the reference's own pred.h / load.h /
 *     utils.h compiled on the host (oracle/_ref, tests/golden/ref_vectors.json);
 *   - the reference's SQL three-valued-logic truth table (test_sql_3vl.cpp:24-31);
 *   - closed forms: brute-force nested-loop joins and Python evaluation of the
 *     star-join queries on small instances; u64 sums compared exactly.
 * Parity unpinned: hash-table collision behaviour (the reference's first-CAS
 * winner is nondeterministic, and so is the key an MHT directory slot keeps
 * when two keys share it); the synthetic data never produces collisions.
 * MHT rows within one directory slot are kept in build-row order here (the
 * reference's order is that of its atomic write counters), so only the
 * order of projected rows can differ.
 *
 * Deliberate, documented definitions where the reference is undefined:
 *   - integer division by zero yields 0;  INT_MIN / -1 wraps;
 *   - int32 arithmetic wraps (two's complement) instead of being UB;
 *   - a probe or build key below the hash-table minimum (negative slot,
 *     an out-of-bounds access in the reference) never matches / is skipped;
 *   - every column reference reads that column's value (the reference's
 *     register bookkeeping can alias two columns onto `items`: a fast-path
 *     column stays mapped to a register that a later load reuses,
 *     src/core/visitor.cpp:1765-1776).
 */
#include "ssb_oracle.h"

#include <ctype.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ */
/* errors                                                               */
/* ------------------------------------------------------------------ */

static __thread char g_err[512];
#define FAIL(...) do { snprintf(g_err, sizeof g_err, __VA_ARGS__); return -1; } while (0)
#define FAILP(...) do { snprintf(g_err, sizeof g_err, __VA_ARGS__); return NULL; } while (0)

/* ------------------------------------------------------------------ */
/* catalog                                                              */
/* ------------------------------------------------------------------ */

#define MAXT 16
#define MAXC 64

typedef struct {
    char name[64];
    const int32_t* data;
    const uint64_t* validity;
    int64_t min_v, max_v;
    uint64_t card;
    int pk, unique;
} ocol;

typedef struct {
    char name[64];
    int64_t rows;
    int is_fact;
    int ncols;
    ocol cols[MAXC];
} otable;

struct orc_db {
    int ntables;
    otable tables[MAXT];
};

orc_db* orc_db_create(void) { return (orc_db*)calloc(1, sizeof(orc_db)); }
void orc_db_destroy(orc_db* db) { free(db); }

static otable* find_table(orc_db* db, const char* name) {
    for (int i = 0; i < db->ntables; ++i)
        if (!strcmp(db->tables[i].name, name)) return &db->tables[i];
    return NULL;
}

/* column name -> (table, column); column names are unique across tables */
static ocol* find_col(orc_db* db, const char* name, otable** tab) {
    for (int i = 0; i < db->ntables; ++i)
        for (int j = 0; j < db->tables[i].ncols; ++j)
            if (!strcmp(db->tables[i].cols[j].name, name)) {
                if (tab) *tab = &db->tables[i];
                return &db->tables[i].cols[j];
            }
    return NULL;
}

int orc_db_add_table(orc_db* db, const char* name, int64_t rows, int is_fact) {
    if (db->ntables >= MAXT) return -1;
    otable* t = &db->tables[db->ntables++];
    memset(t, 0, sizeof *t);
    snprintf(t->name, sizeof t->name, "%s", name);
    t->rows = rows;
    t->is_fact = is_fact;
    return 0;
}

int orc_db_add_column(orc_db* db, const char* table, const char* column, const int32_t* data,
                      const uint64_t* validity) {
    otable* t = find_table(db, table);
    if (!t || t->ncols >= MAXC) return -1;
    ocol* c = &t->cols[t->ncols++];
    memset(c, 0, sizeof *c);
    snprintf(c->name, sizeof c->name, "%s", column);
    c->data = data;
    c->validity = validity;
    return 0;
}

int orc_db_set_primary_key(orc_db* db, const char* table, const char* column) {
    otable* t = find_table(db, table);
    if (!t) return -1;
    for (int j = 0; j < t->ncols; ++j)
        if (!strcmp(t->cols[j].name, column)) { t->cols[j].pk = 1; t->cols[j].unique = 1; return 0; }
    return -1;
}

/* min / max / exact distinct count over the raw values (NULL slots included,
 * as the reference's stats scan reads the raw column file). */
int orc_db_finalize(orc_db* db) {
    for (int i = 0; i < db->ntables; ++i) {
        otable* t = &db->tables[i];
        for (int j = 0; j < t->ncols; ++j) {
            ocol* c = &t->cols[j];
            if (t->rows == 0) { c->min_v = c->max_v = 0; c->card = 0; continue; }
            int64_t mn = c->data[0], mx = c->data[0];
            for (int64_t r = 1; r < t->rows; ++r) {
                if (c->data[r] < mn) mn = c->data[r];
                if (c->data[r] > mx) mx = c->data[r];
            }
            c->min_v = mn;
            c->max_v = mx;
            uint64_t range = (uint64_t)(mx - mn) + 1;
            uint8_t* seen = (uint8_t*)calloc((size_t)((range + 7) / 8), 1);
            if (!seen) return -1;
            uint64_t card = 0;
            for (int64_t r = 0; r < t->rows; ++r) {
                uint64_t k = (uint64_t)(c->data[r] - mn);
                if (!(seen[k >> 3] & (1u << (k & 7)))) { seen[k >> 3] |= (uint8_t)(1u << (k & 7)); ++card; }
            }
            free(seen);
            c->card = card;
            if (card == (uint64_t)t->rows) c->unique = 1;
        }
    }
    return 0;
}

int orc_db_column_stats(orc_db* db, const char* column, int64_t* min_v, int64_t* max_v, uint64_t* card,
                        int* nullable, int* unique) {
    ocol* c = find_col(db, column, NULL);
    if (!c) return -1;
    *min_v = c->min_v;
    *max_v = c->max_v;
    *card = c->card;
    *nullable = c->validity != NULL;
    *unique = c->unique || c->pk;
    return 0;
}

/* ------------------------------------------------------------------ */
/* S-expression parsing                                                 */
/* ------------------------------------------------------------------ */

typedef enum {
    E_COL, E_SLOT, E_INT, E_FLOAT, E_NULL, E_STAR, E_AND, E_OR, E_NOT, E_EQ, E_NE, E_LT, E_LE, E_GT,
    E_GE, E_ADD, E_SUB, E_MUL, E_DIV, E_ISNULL, E_ISNOTNULL, E_CASE
} ekind;

typedef struct expr {
    ekind k;
    char name[64];
    int64_t ival;
    double fval;
    struct expr *a, *b, *c;
    /* resolved */
    int src;      /* 0 fact column, 1 join payload, 2 dimension column at the join's row id */
    int colidx;   /* fact column index or join index */
    void* dcol;   /* src 2: the dimension column (ocol*) */
} expr;

typedef enum { P_SCAN, P_FILTER, P_JOIN, P_AGG, P_SORT, P_PROJECT } pkind;

typedef struct plan {
    pkind k;
    char table[64];
    expr* pred;
    struct plan* child[2];
    int nchild;
    struct plan* parent;
    /* aggregate */
    int ngroups;
    expr* groups[16];
    int naggs;
    char aggfn[16][8];
    expr* aggs[16];
    expr* having;
    /* sort */
    int nkeys;
    int keycol[16];
    int keydesc[16];
    int64_t limit; /* -1: none */
    /* projection */
    int nexprs;
    expr* exprs[64];
} plan;

typedef struct {
    const char* s;
    /* arena */
    void* blocks[4096];
    int nblocks;
} parser;

static void* palloc(parser* p, size_t n) {
    if (p->nblocks >= 4096) return NULL;
    void* m = calloc(1, n);
    p->blocks[p->nblocks++] = m;
    return m;
}

static void pskip(parser* p) { while (*p->s && isspace((unsigned char)*p->s)) ++p->s; }

static int ptoken(parser* p, char* out, size_t n) {
    pskip(p);
    size_t i = 0;
    while (*p->s && !isspace((unsigned char)*p->s) && *p->s != '(' && *p->s != ')') {
        if (i + 1 < n) out[i++] = *p->s;
        ++p->s;
    }
    out[i] = 0;
    return i > 0 ? 0 : -1;
}

static int pexpect(parser* p, char ch) {
    pskip(p);
    if (*p->s != ch) return -1;
    ++p->s;
    return 0;
}

static int ppeek(parser* p, char ch) {
    pskip(p);
    return *p->s == ch;
}

static expr* parse_expr(parser* p);

static expr* parse_expr_or_nil(parser* p) {
    pskip(p);
    if (p->s[0] == '(' ) {
        const char* save = p->s;
        ++p->s;
        pskip(p);
        if (*p->s == ')') { ++p->s; return NULL; }
        p->s = save;
    }
    return parse_expr(p);
}

static const struct { const char* n; ekind k; int arity; } OPS[] = {
    {"and", E_AND, 2}, {"or", E_OR, 2}, {"not", E_NOT, 1}, {"eq", E_EQ, 2}, {"ne", E_NE, 2},
    {"lt", E_LT, 2}, {"le", E_LE, 2}, {"gt", E_GT, 2}, {"ge", E_GE, 2}, {"add", E_ADD, 2},
    {"sub", E_SUB, 2}, {"mul", E_MUL, 2}, {"div", E_DIV, 2}, {"isnull", E_ISNULL, 1},
    {"isnotnull", E_ISNOTNULL, 1}, {"case", E_CASE, 3},
};

static expr* parse_expr(parser* p) {
    char tok[128];
    if (pexpect(p, '(') || ptoken(p, tok, sizeof tok)) FAILP("expected expression");
    expr* e = (expr*)palloc(p, sizeof(expr));
    if (!e) FAILP("out of memory");
    if (!strcmp(tok, "col")) {
        e->k = E_COL;
        if (ptoken(p, e->name, sizeof e->name)) FAILP("col without name");
    } else if (!strcmp(tok, "slot")) {
        e->k = E_SLOT;
        char n[32];
        if (ptoken(p, n, sizeof n)) FAILP("slot without index");
        e->ival = atoll(n);
    } else if (!strcmp(tok, "int")) {
        e->k = E_INT;
        char n[64];
        if (ptoken(p, n, sizeof n)) FAILP("int without value");
        e->ival = strtoll(n, NULL, 10);
    } else if (!strcmp(tok, "float")) {
        e->k = E_FLOAT;
        char n[64];
        if (ptoken(p, n, sizeof n)) FAILP("float without value");
        e->fval = strtod(n, NULL);
    } else if (!strcmp(tok, "null")) {
        e->k = E_NULL;
    } else if (!strcmp(tok, "star")) {
        e->k = E_STAR;
    } else {
        int found = 0;
        for (size_t i = 0; i < sizeof OPS / sizeof *OPS; ++i) {
            if (!strcmp(tok, OPS[i].n)) {
                e->k = OPS[i].k;
                found = 1;
                if (!(e->a = parse_expr(p))) return NULL;
                if (OPS[i].arity >= 2 && !(e->b = parse_expr(p))) return NULL;
                if (OPS[i].arity >= 3 && !(e->c = parse_expr(p))) return NULL;
                break;
            }
        }
        if (!found) FAILP("unknown expression '%s'", tok);
    }
    if (pexpect(p, ')')) FAILP("expected ')' after expression '%s'", tok);
    return e;
}

static plan* parse_plan(parser* p) {
    char tok[64];
    if (pexpect(p, '(') || ptoken(p, tok, sizeof tok)) FAILP("expected plan node");
    plan* n = (plan*)palloc(p, sizeof(plan));
    if (!n) FAILP("out of memory");
    n->limit = -1;
    if (!strcmp(tok, "scan")) {
        n->k = P_SCAN;
        if (ptoken(p, n->table, sizeof n->table)) FAILP("scan without table");
    } else if (!strcmp(tok, "filter")) {
        n->k = P_FILTER;
        if (!(n->pred = parse_expr(p))) return NULL;
        if (!(n->child[0] = parse_plan(p))) return NULL;
        n->nchild = 1;
    } else if (!strcmp(tok, "join")) {
        n->k = P_JOIN;
        n->pred = parse_expr_or_nil(p);
        if (!n->pred && g_err[0]) return NULL;
        if (!(n->child[0] = parse_plan(p))) return NULL;
        if (!(n->child[1] = parse_plan(p))) return NULL;
        n->nchild = 2;
    } else if (!strcmp(tok, "agg")) {
        n->k = P_AGG;
        char t2[32];
        if (pexpect(p, '(') || ptoken(p, t2, sizeof t2) || strcmp(t2, "groups")) FAILP("agg: expected (groups");
        while (!ppeek(p, ')')) {
            if (n->ngroups >= 16) FAILP("too many groups");
            if (!(n->groups[n->ngroups++] = parse_expr(p))) return NULL;
        }
        pexpect(p, ')');
        if (pexpect(p, '(') || ptoken(p, t2, sizeof t2) || strcmp(t2, "aggs")) FAILP("agg: expected (aggs");
        while (!ppeek(p, ')')) {
            if (n->naggs >= 16) FAILP("too many aggregates");
            if (pexpect(p, '(') || ptoken(p, n->aggfn[n->naggs], 8)) FAILP("agg: bad aggregate");
            if (!(n->aggs[n->naggs] = parse_expr(p))) return NULL;
            if (pexpect(p, ')')) FAILP("agg: expected ')'");
            ++n->naggs;
        }
        pexpect(p, ')');
        if (pexpect(p, '(') || ptoken(p, t2, sizeof t2) || strcmp(t2, "having")) FAILP("agg: expected (having");
        n->having = parse_expr_or_nil(p);
        if (!n->having && g_err[0]) return NULL;
        if (pexpect(p, ')')) FAILP("agg: expected ')' after having");
        if (!(n->child[0] = parse_plan(p))) return NULL;
        n->nchild = 1;
    } else if (!strcmp(tok, "sort")) {
        n->k = P_SORT;
        char t2[32];
        if (pexpect(p, '(') || ptoken(p, t2, sizeof t2) || strcmp(t2, "keys")) FAILP("sort: expected (keys");
        while (!ppeek(p, ')')) {
            char a[32], d[16];
            if (n->nkeys >= 16) FAILP("too many sort keys");
            if (pexpect(p, '(') || ptoken(p, a, sizeof a) || ptoken(p, d, sizeof d) || pexpect(p, ')'))
                FAILP("sort: bad key");
            n->keycol[n->nkeys] = atoi(a);
            n->keydesc[n->nkeys] = !strcmp(d, "desc");
            ++n->nkeys;
        }
        pexpect(p, ')');
        if (pexpect(p, '(') || ptoken(p, t2, sizeof t2) || strcmp(t2, "limit")) FAILP("sort: expected (limit");
        char lim[32];
        if (ptoken(p, lim, sizeof lim)) FAILP("sort: limit value");
        n->limit = atoll(lim);
        if (pexpect(p, ')')) FAILP("sort: expected ')'");
        if (!(n->child[0] = parse_plan(p))) return NULL;
        n->nchild = 1;
    } else if (!strcmp(tok, "project")) {
        n->k = P_PROJECT;
        char t2[32];
        if (pexpect(p, '(') || ptoken(p, t2, sizeof t2) || strcmp(t2, "exprs")) FAILP("project: expected (exprs");
        while (!ppeek(p, ')')) {
            if (n->nexprs >= 64) FAILP("too many projection expressions");
            if (!(n->exprs[n->nexprs++] = parse_expr(p))) return NULL;
        }
        pexpect(p, ')');
        if (!(n->child[0] = parse_plan(p))) return NULL;
        n->nchild = 1;
    } else {
        FAILP("unknown plan node '%s'", tok);
    }
    if (pexpect(p, ')')) FAILP("expected ')' after plan node '%s'", tok);
    for (int i = 0; i < n->nchild; ++i) n->child[i]->parent = n;
    return n;
}

static void parser_free(parser* p) {
    for (int i = 0; i < p->nblocks; ++i) free(p->blocks[i]);
    p->nblocks = 0;
}

/* ------------------------------------------------------------------ */
/* C value model                                                        */
/* ------------------------------------------------------------------ */

enum { T_I32 = 0, T_I64 = 1, T_U64 = 2, T_F64 = 3 };

typedef struct {
    int t;
    int64_t i;   /* T_I32 / T_I64 */
    uint64_t u;  /* T_U64 */
    double f;    /* T_F64 */
} cval;

static cval mk_i32(int32_t v) { cval r = {T_I32, v, 0, 0}; return r; }
static cval mk_i64(int64_t v) { cval r = {T_I64, v, 0, 0}; return r; }
static cval mk_u64(uint64_t v) { cval r = {T_U64, 0, v, 0}; return r; }
static cval mk_f64(double v) { cval r = {T_F64, 0, 0, v}; return r; }

static cval conv(cval v, int t) {
    if (v.t == t) return v;
    switch (t) {
        case T_I32:
            if (v.t == T_I64) return mk_i32((int32_t)(uint32_t)(uint64_t)v.i);
            if (v.t == T_U64) return mk_i32((int32_t)(uint32_t)v.u);
            return mk_i32((int32_t)v.f);
        case T_I64:
            if (v.t == T_I32) return mk_i64(v.i);
            if (v.t == T_U64) return mk_i64((int64_t)v.u);
            return mk_i64((int64_t)v.f);
        case T_U64:
            if (v.t == T_I32 || v.t == T_I64) return mk_u64((uint64_t)v.i);
            return mk_u64((uint64_t)v.f);
        default:
            if (v.t == T_U64) return mk_f64((double)v.u);
            return mk_f64((double)v.i);
    }
}

static int truthy(cval v) {
    if (v.t == T_U64) return v.u != 0;
    if (v.t == T_F64) return v.f != 0.0;
    return v.i != 0;
}

/* C usual arithmetic conversions (int < long < unsigned long long < double) */
static int common_type(int a, int b) { return a > b ? a : b; }

static cval arith(int op, cval a, cval b) {
    int t = common_type(a.t, b.t);
    a = conv(a, t);
    b = conv(b, t);
    switch (t) {
        case T_I32: {
            uint32_t x = (uint32_t)a.i, y = (uint32_t)b.i;
            int32_t sx = (int32_t)x, sy = (int32_t)y;
            switch (op) {
                case E_ADD: return mk_i32((int32_t)(x + y));
                case E_SUB: return mk_i32((int32_t)(x - y));
                case E_MUL: return mk_i32((int32_t)(x * y));
                default:
                    if (sy == 0) return mk_i32(0);
                    if (sy == -1) return mk_i32((int32_t)(0u - x));
                    return mk_i32(sx / sy);
            }
        }
        case T_I64: {
            uint64_t x = (uint64_t)a.i, y = (uint64_t)b.i;
            switch (op) {
                case E_ADD: return mk_i64((int64_t)(x + y));
                case E_SUB: return mk_i64((int64_t)(x - y));
                case E_MUL: return mk_i64((int64_t)(x * y));
                default:
                    if (b.i == 0) return mk_i64(0);
                    if (b.i == -1) return mk_i64((int64_t)(0ULL - x));
                    return mk_i64(a.i / b.i);
            }
        }
        case T_U64:
            switch (op) {
                case E_ADD: return mk_u64(a.u + b.u);
                case E_SUB: return mk_u64(a.u - b.u);
                case E_MUL: return mk_u64(a.u * b.u);
                default: return mk_u64(b.u == 0 ? 0 : a.u / b.u);
            }
        default:
            switch (op) {
                case E_ADD: return mk_f64(a.f + b.f);
                case E_SUB: return mk_f64(a.f - b.f);
                case E_MUL: return mk_f64(a.f * b.f);
                default: return mk_f64(a.f / b.f);
            }
    }
}

static int compare(int op, cval a, cval b) {
    int t = common_type(a.t, b.t);
    a = conv(a, t);
    b = conv(b, t);
    int lt, eq;
    if (t == T_U64) { lt = a.u < b.u; eq = a.u == b.u; }
    else if (t == T_F64) { lt = a.f < b.f; eq = a.f == b.f; }
    else { lt = a.i < b.i; eq = a.i == b.i; }
    switch (op) {
        case E_EQ: return eq;
        case E_NE: return !eq;
        case E_LT: return lt;
        case E_LE: return lt || eq;
        case E_GT: return !lt && !eq;
        default: return !lt;
    }
}

/* db::safe_*: int32 -> int64 promotion, then the common type. */
static cval promote(cval v) { return v.t == T_I32 ? conv(v, T_I64) : v; }

/* Type of the C literal std::to_string(int64) prints: int when the magnitude
 * fits (the minus sign is a separate unary operator), else long. */
static cval int_literal(int64_t v) {
    if (v >= -2147483647LL && v <= 2147483647LL) return mk_i32((int32_t)v);
    return mk_i64(v);
}

/* std::to_string(double) prints "%f": the generated code sees that decimal. */
static cval float_literal(double v) {
    char buf[512];
    snprintf(buf, sizeof buf, "%f", v);
    return mk_f64(strtod(buf, NULL));
}

/* ------------------------------------------------------------------ */
/* query analysis                                                       */
/* ------------------------------------------------------------------ */

#define MAXJ 8
#define MAXOPS 32

typedef struct {
    otable* dim;
    ocol* pk;
    ocol* fk;          /* fact column (the first key alternative) */
    int fk_idx;        /* index in fact table */
    ocol* fks[MAXJ];   /* OR-key join: every fact key alternative, in condition order */
    int nfk;
    int rowid;         /* payload is the dimension row index (projection of its columns) */
    int mht;           /* non-unique build key: multi-value table, one pipeline pass per build row */
    int32_t* mkey;     /* MHT directory: key, offset and count per slot (len each) */
    int32_t* moff;
    int32_t* mcnt;
    int32_t* mrow;     /* MHT payload: build row ids grouped by slot */
    ocol* payload;     /* NULL for PHT_1 */
    const expr* build_pred;
    int32_t kmin;
    int64_t len;
    int32_t* ht;       /* 2*len ints, reference layout */
} joininfo;

typedef struct {
    int kind;          /* 0 filter, 1 join */
    const expr* pred;
    int fast;
    int join;          /* join index */
} pipeop;

typedef struct {
    orc_db* db;
    otable* fact;
    const plan* agg;   /* the aggregate, or the projection */
    const plan* sort;
    const plan* proj;  /* projection root (else NULL) */
    int nexprs;
    expr* exprs[64];   /* projection output expressions (a star expanded) */
    expr star_cols[64];
    int nops;
    pipeop ops[MAXOPS];
    int njoins;
    joininfo joins[MAXJ];
    /* aggregate layout */
    int ngroups, naggs;
    int visible_ts, storage_ts;
    int hidden_slot[16];
    int64_t G;         /* group slots (perfect hash domain) */
    int64_t gmin[16];
    uint64_t gcard[16];
    int gnullable[16];
    int fast_sparse;
    int out_type[32];
    int out_nullable[32];
} query;

static int is_cmp(ekind k) { return k >= E_EQ && k <= E_GE; }

/* resolve column references of the probe pipeline: fact column or payload */
static int resolve_expr(query* q, expr* e, int njoins_visible) {
    if (!e) return 0;
    if (e->k == E_COL) {
        otable* t = NULL;
        ocol* c = find_col(q->db, e->name, &t);
        if (!c) FAIL("unknown column '%s'", e->name);
        if (t == q->fact) {
            e->src = 0;
            e->colidx = (int)(c - t->cols);
            return 0;
        }
        for (int j = 0; j < njoins_visible; ++j) {
            if (q->joins[j].rowid && q->joins[j].dim == t) { e->src = 2; e->colidx = j; e->dcol = c; return 0; }
            if (q->joins[j].payload == c) { e->src = 1; e->colidx = j; return 0; }
        }
        FAIL("column '%s' of dimension %s is not available in the probe pipeline "
             "(only one payload column per joined dimension)", e->name, t->name);
    }
    if (resolve_expr(q, e->a, njoins_visible)) return -1;
    if (resolve_expr(q, e->b, njoins_visible)) return -1;
    if (resolve_expr(q, e->c, njoins_visible)) return -1;
    return 0;
}

static void collect_cols(const expr* e, const char** out, int* n, int max) {
    if (!e) return;
    if (e->k == E_COL) {
        for (int i = 0; i < *n; ++i)
            if (!strcmp(out[i], e->name)) return;
        if (*n < max) out[(*n)++] = e->name;
        return;
    }
    collect_cols(e->a, out, n, max);
    collect_cols(e->b, out, n, max);
    collect_cols(e->c, out, n, max);
}

/* validateFilterFastPathPredicate: AND-chains of col op int literal */
static int fast_path_ok(const expr* e) {
    if (!e) return 1;
    if (e->k == E_AND) return fast_path_ok(e->a) && fast_path_ok(e->b);
    if (is_cmp(e->k)) return e->a->k == E_COL && e->b->k == E_INT;
    return 0;
}

static int unsupported_filter(const expr* e, char* why, size_t n) {
    if (!e) return 0;
    if (e->k == E_AND) return unsupported_filter(e->a, why, n) || unsupported_filter(e->b, why, n);
    if (e->k == E_OR || e->k == E_NOT || e->k == E_ISNULL || e->k == E_ISNOTNULL) return 0;
    if (!is_cmp(e->k)) { snprintf(why, n, "unsupported filter predicate node"); return 1; }
    if (e->a->k == E_COL && (e->b->k == E_INT || e->b->k == E_COL)) return 0;
    int ak = e->a->k, bk = e->b->k;
    if (ak == E_ADD || ak == E_SUB || ak == E_MUL || ak == E_DIV || bk == E_ADD || bk == E_SUB ||
        bk == E_MUL || bk == E_DIV) return 0;
    if (ak == E_FLOAT || bk == E_FLOAT) { snprintf(why, n, "unsupported float literal in fast-path predicate"); return 1; }
    snprintf(why, n, "unsupported comparison shape in fast-path predicate");
    return 1;
}

static int build_pred_ok(const expr* e) {
    if (!e) return 1;
    if (e->k == E_AND || e->k == E_OR) return build_pred_ok(e->a) && build_pred_ok(e->b);
    if (!is_cmp(e->k)) return 0;
    int lc = e->a->k == E_COL && (e->b->k == E_INT || e->b->k == E_FLOAT);
    int rc = e->b->k == E_COL && (e->a->k == E_INT || e->a->k == E_FLOAT);
    return lc || rc;
}

static const plan* dim_scan(const plan* n) {
    if (n->k == P_SCAN) return n;
    if (n->k == P_FILTER && n->child[0]->k == P_SCAN) return n->child[0];
    return NULL;
}

static int extract_keys(query* q, joininfo* ji, const expr* e) {
    if (!e) return 0;
    if (e->k == E_EQ && e->a->k == E_COL && e->b->k == E_COL) {
        otable* ta = NULL;
        ocol* ca = find_col(q->db, e->a->name, &ta);
        otable* tb = NULL;
        ocol* cb = find_col(q->db, e->b->name, &tb);
        if (!ca || !cb) FAIL("join key column not found");
        /* computeBuildInfo's extractKeys (visitor.cpp:3921-3945): every key
         * pair adds a fact key alternative of the one dimension key */
        ocol* pk = ta == ji->dim ? ca : cb;
        ocol* fk = ta == ji->dim ? cb : ca;
        if (ji->pk && ji->pk != pk) FAIL("OR-key join over two dimension keys (%s, %s)", ji->pk->name, pk->name);
        if (ji->nfk >= MAXJ) FAIL("too many join key alternatives");
        ji->pk = pk;
        ji->fks[ji->nfk++] = fk;
        if (!ji->fk) ji->fk = fk;
        return 0;
    }
    if (e->k == E_AND || e->k == E_OR) {
        if (extract_keys(q, ji, e->a)) return -1;
        return extract_keys(q, ji, e->b);
    }
    return 0;
}

static int build_info(query* q, const plan* join, const plan* build_side, joininfo* ji) {
    const plan* scan = dim_scan(build_side);
    ji->dim = find_table(q->db, scan->table);
    if (!ji->dim) FAIL("unknown table %s", scan->table);
    ji->build_pred = build_side->k == P_FILTER ? build_side->pred : NULL;
    if (ji->build_pred && !build_pred_ok(ji->build_pred))
        FAIL("Unsupported build-side filter predicate for block codegen");
    const char* t = ji->dim->name;
    if (!strcmp(t, "PART") || !strcmp(t, "SUPPLIER") || !strcmp(t, "CUSTOMER")) {
        ji->len = ji->dim->rows;
        ji->kmin = 0;
    } else if (!strcmp(t, "DDATE")) {
        ji->len = 61131;
        ji->kmin = 19920101;
    } else {
        ji->len = 0;
        for (int i = 0; i < ji->dim->ncols; ++i) {
            if (strstr(ji->dim->cols[i].name, "key") || strstr(ji->dim->cols[i].name, "Key")) {
                int64_t range = ji->dim->cols[i].max_v - ji->dim->cols[i].min_v + 1;
                if (range > 0) { ji->kmin = (int32_t)ji->dim->cols[i].min_v; ji->len = range; }
                break;
            }
        }
        if (ji->len == 0) ji->len = ji->dim->rows;
    }
    if (extract_keys(q, ji, join->pred)) return -1;
    if (!ji->pk) FAIL("hash join without an equi-key condition");
    for (int k = 0; k < ji->nfk; ++k) {
        otable* ft = NULL;
        find_col(q->db, ji->fks[k]->name, &ft);
        if (ft != q->fact) FAIL("join key %s is not a fact-table column", ji->fks[k]->name);
    }
    ji->fk_idx = (int)(ji->fk - q->fact->cols);
    /* MHT only when the catalog proves the key non-unique (visitor.cpp:3951-3961);
     * it always carries the build row id (3963-3970) */
    ji->mht = !(ji->pk->pk || ji->pk->unique || ji->pk->card == (uint64_t)ji->dim->rows) &&
              ji->pk->card > 0 && ji->pk->card < (uint64_t)ji->dim->rows;
    if (ji->mht && ji->nfk > 1) FAIL("OR-key join on a non-unique build key");
    /* payload: first column of this dimension needed above the join, aggregate
     * columns (aggregates, then group-bys) before filter columns */
    const char* aggc[64];
    int nagg = 0;
    const char* fltc[64];
    int nflt = 0;
    for (const plan* cur = join->parent; cur; cur = cur->parent) {
        if (cur->k == P_AGG) {
            for (int i = 0; i < cur->naggs; ++i) collect_cols(cur->aggs[i], aggc, &nagg, 64);
            for (int i = 0; i < cur->ngroups; ++i) collect_cols(cur->groups[i], aggc, &nagg, 64);
        } else if (cur->k == P_FILTER) {
            collect_cols(cur->pred, fltc, &nflt, 64);
        }
    }
    ji->payload = NULL;
    ji->rowid = 0;
    /* projectionRequiresRowIdPayload (visitor.cpp:3300-3317): a projection
     * above that reads this dimension gathers its columns by row id */
    for (const plan* cur = join->parent; cur; cur = cur->parent) {
        if (cur->k != P_PROJECT) continue;
        for (int i = 0; i < cur->nexprs; ++i) {
            const char* cols[64];
            int nc = 0;
            if (cur->exprs[i]->k == E_STAR) ji->rowid = 1;
            collect_cols(cur->exprs[i], cols, &nc, 64);
            for (int c = 0; c < nc; ++c) {
                otable* ct = NULL;
                find_col(q->db, cols[c], &ct);
                if (ct == ji->dim) ji->rowid = 1;
            }
        }
    }
    if (ji->mht) ji->rowid = 1;
    for (int pass = 0; pass < 2 && !ji->payload && !ji->rowid; ++pass) {
        const char** lst = pass == 0 ? aggc : fltc;
        int n = pass == 0 ? nagg : nflt;
        for (int i = 0; i < n; ++i) {
            otable* ct = NULL;
            ocol* c = find_col(q->db, lst[i], &ct);
            if (ct == ji->dim && c != ji->pk) { ji->payload = c; break; }
        }
    }
    return 0;
}

static int is_nullable_col(orc_db* db, const char* name) {
    ocol* c = find_col(db, name, NULL);
    return c && c->validity;
}

/* resultColumnNullableForExpr (visitor.cpp:128-170) */
static int expr_nullable(orc_db* db, const expr* e) {
    if (!e) return 1;
    switch (e->k) {
        case E_NULL: return 1;
        case E_STAR: case E_INT: case E_FLOAT: case E_ISNULL: case E_ISNOTNULL: return 0;
        case E_COL: return is_nullable_col(db, e->name);
        case E_NOT: return expr_nullable(db, e->a);
        case E_CASE: return expr_nullable(db, e->b) || expr_nullable(db, e->c);
        case E_SLOT: return 1;
        default: return expr_nullable(db, e->a) || expr_nullable(db, e->b);
    }
}

/* resultLogicalTypeForExpr (visitor.cpp:77-107) */
static int expr_out_type(const expr* e) {
    if (!e) return ORC_T_UINT64;
    switch (e->k) {
        case E_EQ: case E_NE: case E_LT: case E_LE: case E_GT: case E_GE: case E_AND: case E_OR:
        case E_NOT: case E_ISNULL: case E_ISNOTNULL: case E_NULL:
            return ORC_T_UINT64;
        case E_FLOAT: return ORC_T_FLOAT64;
        case E_CASE: {
            int a = expr_out_type(e->b), b = expr_out_type(e->c);
            if (a == ORC_T_FLOAT64 || b == ORC_T_FLOAT64) return ORC_T_FLOAT64;
            if (a == ORC_T_INT64 || b == ORC_T_INT64) return ORC_T_INT64;
            return ORC_T_UINT64;
        }
        default: return ORC_T_INT64;
    }
}

static void case_range(const expr* e, int64_t* lo, int64_t* hi, int* any, int* may_null, int* bad) {
    if (!e) { *may_null = 1; return; }
    if (e->k == E_INT) {
        if (!*any || e->ival < *lo) *lo = e->ival;
        if (!*any || e->ival > *hi) *hi = e->ival;
        *any = 1;
        return;
    }
    if (e->k == E_NULL) { *may_null = 1; return; }
    if (e->k == E_CASE) {
        case_range(e->b, lo, hi, any, may_null, bad);
        case_range(e->c, lo, hi, any, may_null, bad);
        return;
    }
    *bad = 1;
}

static int fast_sparse_eligible(query* q) {
    const plan* a = q->agg;
    for (int i = 0; i < a->naggs; ++i)
        if (!strcmp(a->aggfn[i], "AVG")) return 0;
    if (a->having) return 0;
    for (int c = 0; c < q->visible_ts; ++c)
        if (q->out_nullable[c]) return 0;
    for (int i = 0; i < a->naggs; ++i)
        if (strcmp(a->aggfn[i], "COUNT") && strcmp(a->aggfn[i], "SUM")) return 0;
    if (a->ngroups == 0) return 1;
    for (int g = 0; g < a->ngroups; ++g) {
        const expr* e = a->groups[g];
        if (e->k == E_COL) {
            ocol* c = find_col(q->db, e->name, NULL);
            if (c && c->min_v != 0) return 1;
        }
    }
    for (int i = 0; i < a->naggs; ++i)
        if (!strcmp(a->aggfn[i], "COUNT")) return 1;
    return 0;
}

/* Projection (consumeProjectionVector, visitor.cpp:3466-3548): one output
 * row per row leaving the pipeline, each expression cast to its result type
 * (resultLogicalTypeForExpr), validity from the expression's operands. */
static int analyze_projection(query* q) {
    const plan* pr = q->proj;
    q->nexprs = 0;
    for (int i = 0; i < pr->nexprs; ++i) {
        expr* e = pr->exprs[i];
        if (e->k != E_STAR) {
            if (q->nexprs >= 64) FAIL("too many projection columns");
            q->exprs[q->nexprs++] = e;
            continue;
        }
        /* every column of every scanned table, tables in plan order (depth
         * first, children in order; collectTableScans, visitor.cpp:3426-3432) */
        const plan* stack[64];
        const otable* tabs[16];
        int ns = 0, nt = 0;
        stack[ns++] = pr->child[0];
        while (ns > 0) {
            const plan* c = stack[--ns];
            if (c->k == P_SCAN) {
                const otable* t = find_table(q->db, c->table);
                if (!t) FAIL("unknown table %s", c->table);
                if (nt < 16) tabs[nt++] = t;
                continue;
            }
            for (int k = c->nchild - 1; k >= 0 && ns < 64; --k) stack[ns++] = c->child[k];
        }
        for (int k = 0; k < nt; ++k)
            for (int c = 0; c < tabs[k]->ncols; ++c) {
                if (q->nexprs >= 64) FAIL("too many projection columns");
                expr* x = &q->star_cols[q->nexprs];
                memset(x, 0, sizeof *x);
                x->k = E_COL;
                snprintf(x->name, sizeof x->name, "%s", tabs[k]->cols[c].name);
                q->exprs[q->nexprs++] = x;
            }
    }
    if (q->nexprs == 0) FAIL("projection without columns");
    for (int i = 0; i < q->nexprs; ++i)
        if (resolve_expr(q, q->exprs[i], q->njoins)) return -1;
    q->ngroups = 0;
    q->naggs = 0;
    q->visible_ts = q->storage_ts = q->nexprs;
    for (int c = 0; c < q->nexprs; ++c) {
        q->out_type[c] = expr_out_type(q->exprs[c]);
        q->out_nullable[c] = expr_nullable(q->db, q->exprs[c]);
    }
    q->G = 0;
    q->fast_sparse = 0;
    return 0;
}

static int analyze(query* q, const plan* root) {
    const plan* n = root;
    q->sort = NULL;
    if (n->k == P_SORT) { q->sort = n; n = n->child[0]; }
    q->proj = NULL;
    if (n->k == P_PROJECT) q->proj = n;
    else if (n->k != P_AGG) FAIL("plan root must be an aggregate or a projection");
    q->agg = n;
    /* collect the probe chain from the aggregate down to the fact scan */
    const plan* chain[MAXOPS];
    const plan* chain_build[MAXOPS];
    int nchain = 0;
    const plan* cur = n->child[0];
    while (cur->k != P_SCAN) {
        if (nchain >= MAXOPS) FAIL("plan too deep");
        if (cur->k == P_FILTER) { chain_build[nchain] = NULL; chain[nchain++] = cur; cur = cur->child[0]; continue; }
        if (cur->k == P_JOIN) {
            const plan* build = NULL;
            const plan* probe = NULL;
            for (int side = 0; side < 2 && !build; ++side) {
                const plan* s = dim_scan(cur->child[side]);
                if (s) {
                    otable* t = find_table(q->db, s->table);
                    if (t && !t->is_fact) { build = cur->child[side]; probe = cur->child[1 - side]; }
                }
            }
            if (!build) FAIL("hash join without a dimension build side");
            chain_build[nchain] = build;
            chain[nchain++] = cur;
            cur = probe;
            continue;
        }
        FAIL("unsupported plan node below the aggregate");
    }
    q->fact = find_table(q->db, cur->table);
    if (!q->fact) FAIL("unknown table %s", cur->table);
    if (!q->fact->is_fact) FAIL("the probe pipeline must scan the fact table");
    /* joins in execution order (innermost first) */
    q->nops = 0;
    q->njoins = 0;
    for (int i = nchain - 1; i >= 0; --i) {
        const plan* p = chain[i];
        pipeop* op = &q->ops[q->nops++];
        memset(op, 0, sizeof *op);
        if (p->k == P_FILTER) {
            char why[128];
            if (unsupported_filter(p->pred, why, sizeof why))
                FAIL("Filter predicate is not translatable before JIT C++ compilation: %s", why);
            op->kind = 0;
            op->pred = p->pred;
            op->fast = fast_path_ok(p->pred);
        } else {
            if (q->njoins >= MAXJ) FAIL("too many joins");
            joininfo* ji = &q->joins[q->njoins];
            memset(ji, 0, sizeof *ji);
            if (build_info(q, p, chain_build[i], ji)) return -1;
            for (int j = 0; j < q->njoins; ++j)
                if (q->joins[j].dim == ji->dim) FAIL("dimension %s joined twice", ji->dim->name);
            op->kind = 1;
            op->join = q->njoins++;
        }
    }
    /* resolve column references with the payloads visible at each point */
    int visible = 0;
    for (int i = 0; i < q->nops; ++i) {
        pipeop* op = &q->ops[i];
        if (op->kind == 1) { ++visible; continue; }
        if (resolve_expr(q, (expr*)op->pred, visible)) return -1;
        if (op->fast) {
            /* a fast-path predicate reads its column from the fact table */
            const char* cols[16];
            int nc = 0;
            collect_cols(op->pred, cols, &nc, 16);
            for (int c = 0; c < nc; ++c) {
                otable* t = NULL;
                find_col(q->db, cols[c], &t);
                if (t != q->fact) FAIL("fast-path filter on dimension column %s above its join", cols[c]);
            }
        }
    }
    if (q->proj) return analyze_projection(q);
    const plan* a = q->agg;
    for (int g = 0; g < a->ngroups; ++g)
        if (resolve_expr(q, a->groups[g], q->njoins)) return -1;
    for (int i = 0; i < a->naggs; ++i)
        if (resolve_expr(q, a->aggs[i], q->njoins)) return -1;
    /* layout */
    q->ngroups = a->ngroups;
    q->naggs = a->naggs;
    q->visible_ts = a->ngroups + a->naggs;
    int hidden = 0;
    for (int i = 0; i < a->naggs; ++i) {
        q->hidden_slot[i] = -1;
        if (!strcmp(a->aggfn[i], "AVG")) q->hidden_slot[i] = q->visible_ts + hidden++;
        if (a->aggs[i]->k == E_STAR && strcmp(a->aggfn[i], "COUNT"))
            FAIL("Aggregate function %s does not accept '*'", a->aggfn[i]);
    }
    q->storage_ts = q->visible_ts + hidden;
    if (q->storage_ts == 0) q->storage_ts = 1;
    for (int g = 0; g < a->ngroups; ++g) {
        q->out_type[g] = expr_out_type(a->groups[g]);
        q->out_nullable[g] = expr_nullable(q->db, a->groups[g]);
    }
    for (int i = 0; i < a->naggs; ++i) {
        int c = a->ngroups + i;
        if (!strcmp(a->aggfn[i], "AVG")) q->out_type[c] = ORC_T_FLOAT64;
        else q->out_type[c] = ORC_T_UINT64;
        q->out_nullable[c] = !strcmp(a->aggfn[i], "COUNT") ? 0 : expr_nullable(q->db, a->aggs[i]);
    }
    /* perfect group hash domain */
    q->G = 1;
    for (int g = 0; g < a->ngroups; ++g) {
        const expr* e = a->groups[g];
        int64_t mn = 0;
        uint64_t card = 1;
        int nullable = 0;
        if (e->k == E_COL) {
            ocol* c = find_col(q->db, e->name, NULL);
            mn = c->min_v;
            card = c->card;
            nullable = c->validity != NULL;
        } else if (e->k == E_CASE) {
            int64_t lo = 0, hi = 0;
            int any = 0, mnull = 0, bad = 0;
            case_range(e, &lo, &hi, &any, &mnull, &bad);
            if (bad) FAIL("GROUP BY CASE currently requires integer THEN/ELSE result expressions");
            if (!any && !mnull) FAIL("GROUP BY CASE has no integer output values");
            mn = any ? lo : 0;
            card = any ? (uint64_t)(hi - lo + 1) : 1;
            nullable = 1; /* validity of a CASE is never the literal 1 here */
        } else {
            FAIL("GROUP BY expression is not supported by perfect hash codegen");
        }
        if (nullable) card += 1;
        q->gmin[g] = mn;
        q->gcard[g] = card;
        q->gnullable[g] = nullable;
        q->G *= (int64_t)card;
    }
    if (a->ngroups == 0) q->G = 1;
    q->fast_sparse = fast_sparse_eligible(q) && !q->sort && !a->having;
    return 0;
}

/* ------------------------------------------------------------------ */
/* row evaluation                                                       */
/* ------------------------------------------------------------------ */

typedef struct {
    const query* q;
    int64_t row;
    int32_t payload[MAXJ];
} rowctx;

static int32_t col_value(const rowctx* rc, const expr* e) {
    if (e->src == 2) return ((const ocol*)e->dcol)->data[rc->payload[e->colidx]];
    if (e->src == 1) return rc->payload[e->colidx];
    return rc->q->fact->cols[e->colidx].data[rc->row];
}

static int col_valid(const rowctx* rc, const expr* e) {
    if (e->src == 2) {  /* gathered by row id with its validity (itemColumnValidExpr, visitor.cpp:3570-3582) */
        const ocol* c = (const ocol*)e->dcol;
        const int64_t r = rc->payload[e->colidx];
        return !c->validity || (int)((c->validity[r >> 6] >> (r & 63)) & 1ULL);
    }
    if (e->src == 1) return 1; /* payload registers carry no validity */
    const uint64_t* v = rc->q->fact->cols[e->colidx].validity;
    if (!v) return 1;
    return (int)((v[rc->row >> 6] >> (rc->row & 63)) & 1ULL);
}

static cval vexpr(const rowctx* rc, const expr* e);
static int vvalid(const rowctx* rc, const expr* e);

static cval safe_binop(int op, cval a, cval b) {
    a = promote(a);
    b = promote(b);
    if (op == E_ADD || op == E_SUB || op == E_MUL || op == E_DIV) return arith(op, a, b);
    return mk_i32(compare(op, a, b));
}

/* C conditional operator: both arms converted to their common type */
static cval cond_pick(int take_then, cval t, cval f) {
    int ty = common_type(t.t, f.t);
    return conv(take_then ? t : f, ty);
}

static cval vexpr(const rowctx* rc, const expr* e) {
    switch (e->k) {
        case E_COL: return mk_i32(col_value(rc, e));
        case E_INT: return int_literal(e->ival);
        case E_FLOAT: return float_literal(e->fval);
        case E_NULL: return mk_i32(0);
        case E_STAR: return mk_i32(1);
        case E_ADD: case E_SUB: case E_MUL: case E_DIV:
        case E_EQ: case E_NE: case E_LT: case E_LE: case E_GT: case E_GE:
            return safe_binop(e->k, vexpr(rc, e->a), vexpr(rc, e->b));
        case E_ISNULL: return mk_i32(!vvalid(rc, e->a));
        case E_ISNOTNULL: return mk_i32(vvalid(rc, e->a));
        case E_NOT: return mk_i32(!truthy(vexpr(rc, e->a)));
        case E_AND: return mk_i32(truthy(vexpr(rc, e->a)) && truthy(vexpr(rc, e->b)));
        case E_OR: return mk_i32(truthy(vexpr(rc, e->a)) || truthy(vexpr(rc, e->b)));
        case E_CASE: {
            int take = vvalid(rc, e->a) && truthy(vexpr(rc, e->a));
            return cond_pick(take, vexpr(rc, e->b), vexpr(rc, e->c));
        }
        default: return mk_i32(0);
    }
}

static int vvalid(const rowctx* rc, const expr* e) {
    switch (e->k) {
        case E_NULL: return 0;
        case E_STAR: case E_INT: case E_FLOAT: case E_ISNULL: case E_ISNOTNULL: return 1;
        case E_COL: return col_valid(rc, e);
        case E_ADD: case E_SUB: case E_MUL: case E_DIV:
        case E_EQ: case E_NE: case E_LT: case E_LE: case E_GT: case E_GE:
            return vvalid(rc, e->a) && vvalid(rc, e->b);
        case E_NOT: return vvalid(rc, e->a);
        case E_AND: {
            int lv = vvalid(rc, e->a), rv = vvalid(rc, e->b);
            int l = truthy(vexpr(rc, e->a)), r = truthy(vexpr(rc, e->b));
            return (lv && !l) || (rv && !r) || (lv && rv);
        }
        case E_OR: {
            int lv = vvalid(rc, e->a), rv = vvalid(rc, e->b);
            int l = truthy(vexpr(rc, e->a)), r = truthy(vexpr(rc, e->b));
            return (lv && l) || (rv && r) || (lv && rv);
        }
        case E_CASE: {
            int sel = vvalid(rc, e->a) && truthy(vexpr(rc, e->a));
            return (sel && vvalid(rc, e->b)) || (!sel && vvalid(rc, e->c));
        }
        default: return 1;
    }
}

/* translateMathExprPush: the leftmost leaf is cast to unsigned long long
 * when it is a column, *, IS [NOT] NULL or a predicate; literals keep their C
 * type (visitor.cpp:1961-1968 print them uncast), right operands keep theirs,
 * operators are plain C operators. */
static cval mexpr(const rowctx* rc, const expr* e, int cast) {
    cval r;
    switch (e->k) {
        case E_COL: r = mk_i32(col_value(rc, e)); break;
        case E_INT: return int_literal(e->ival);
        case E_FLOAT: return float_literal(e->fval);
        case E_STAR: r = mk_i32(1); break;
        case E_ADD: case E_SUB: case E_MUL: case E_DIV:
            return arith(e->k, mexpr(rc, e->a, cast), mexpr(rc, e->b, 0));
        case E_ISNULL: r = mk_i32(!vvalid(rc, e->a)); break;
        case E_ISNOTNULL: r = mk_i32(vvalid(rc, e->a)); break;
        case E_EQ: case E_NE: case E_LT: case E_LE: case E_GT: case E_GE:
        case E_AND: case E_OR: case E_NOT: case E_CASE:
            r = vexpr(rc, e);
            break;
        default: return mk_i32(0); /* NULL literal: "0", uncast */
    }
    return cast ? conv(r, T_U64) : r;
}

/* BlockPred* on int items with an int compare value */
static int int_pred(int op, int32_t v, int32_t lit) {
    switch (op) {
        case E_EQ: return v == lit;
        case E_NE: return v != lit;
        case E_LT: return v < lit;
        case E_LE: return v <= lit;
        case E_GT: return v > lit;
        default: return v >= lit;
    }
}

static int fast_filter(const rowctx* rc, const expr* e) {
    if (e->k == E_AND) return fast_filter(rc, e->a) && fast_filter(rc, e->b);
    return int_pred(e->k, col_value(rc, e->a), (int32_t)(uint32_t)(uint64_t)e->b->ival);
}

static int swap_op(int op) {
    switch (op) {
        case E_LT: return E_GT;
        case E_LE: return E_GE;
        case E_GT: return E_LT;
        case E_GE: return E_LE;
        default: return op;
    }
}

static int32_t lit_int32(const expr* e) {
    if (e->k == E_INT) return (int32_t)(uint32_t)(uint64_t)e->ival;
    return (int32_t)float_literal(e->fval).f;
}

/* emitBuildFilter: AND / OR of (col op lit) or (lit op col), raw values */
static int build_filter(const otable* dim, int64_t row, const expr* e, orc_db* db) {
    if (e->k == E_AND) return build_filter(dim, row, e->a, db) && build_filter(dim, row, e->b, db);
    if (e->k == E_OR) return build_filter(dim, row, e->a, db) || build_filter(dim, row, e->b, db);
    const expr* col = e->a->k == E_COL ? e->a : e->b;
    const expr* lit = e->a->k == E_COL ? e->b : e->a;
    int op = e->a->k == E_COL ? (int)e->k : swap_op(e->k);
    ocol* c = find_col(db, col->name, NULL);
    return int_pred(op, c->data[row], lit_int32(lit));
}

/* the build rows an MHT stores: build filter, key validity, key >= minimum */
static int mht_slot(const joininfo* ji, int64_t r, orc_db* db, int64_t* slot) {
    int flag = ji->build_pred ? build_filter(ji->dim, r, ji->build_pred, db) : 1;
    if (ji->pk->validity) flag = flag && ((ji->pk->validity[r >> 6] >> (r & 63)) & 1ULL);
    if (!flag) return 0;
    int32_t x = (int32_t)((uint32_t)ji->pk->data[r] - (uint32_t)ji->kmin);
    if (x < 0) return 0;
    *slot = x % ji->len;
    return 1;
}

/* two passes as BlockBuildMHT_Count / prefix sum / BlockBuildMHT_Write */
static int build_mht(query* q, joininfo* ji) {
    ji->mkey = (int32_t*)calloc((size_t)ji->len, sizeof(int32_t));
    ji->moff = (int32_t*)calloc((size_t)ji->len, sizeof(int32_t));
    ji->mcnt = (int32_t*)calloc((size_t)ji->len, sizeof(int32_t));
    ji->mrow = (int32_t*)calloc((size_t)(ji->dim->rows + 1), sizeof(int32_t));
    int32_t* pos = (int32_t*)calloc((size_t)ji->len, sizeof(int32_t));
    if (!ji->mkey || !ji->moff || !ji->mcnt || !ji->mrow || !pos) { free(pos); FAIL("out of memory for hash table"); }
    int64_t s;
    for (int64_t r = 0; r < ji->dim->rows; ++r)
        if (mht_slot(ji, r, q->db, &s)) ++ji->mcnt[s];
    int64_t run = 0;
    for (int64_t k = 0; k < ji->len; ++k) { ji->moff[k] = (int32_t)run; run += ji->mcnt[k]; }
    for (int64_t r = 0; r < ji->dim->rows; ++r) {
        if (!mht_slot(ji, r, q->db, &s)) continue;
        ji->mrow[ji->moff[s] + pos[s]++] = (int32_t)r;
        ji->mkey[s] = ji->pk->data[r];
    }
    free(pos);
    return 0;
}

static int build_tables(query* q) {
    for (int j = 0; j < q->njoins; ++j) {
        joininfo* ji = &q->joins[j];
        if (ji->mht) {
            if (build_mht(q, ji)) return -1;
            continue;
        }
        ji->ht = (int32_t*)calloc((size_t)(2 * ji->len), sizeof(int32_t));
        if (!ji->ht) FAIL("out of memory for hash table");
        for (int64_t r = 0; r < ji->dim->rows; ++r) {
            int flag = ji->build_pred ? build_filter(ji->dim, r, ji->build_pred, q->db) : 1;
            if (ji->pk->validity) flag = flag && ((ji->pk->validity[r >> 6] >> (r & 63)) & 1ULL);
            if (!flag) continue;
            int32_t k = ji->pk->data[r];
            int32_t x = (int32_t)((uint32_t)k - (uint32_t)ji->kmin);
            if (x < 0) continue;
            int64_t s = x % ji->len;
            if (ji->payload || ji->rowid) {
                if (ji->ht[2 * s] == 0) ji->ht[2 * s] = k;
                ji->ht[2 * s + 1] = ji->rowid ? (int32_t)r : ji->payload->data[r];
            } else {
                if (ji->ht[s] == 0) ji->ht[s] = k;
            }
        }
    }
    return 0;
}

static int key_valid(const ocol* c, int64_t r) { return !c->validity || ((c->validity[r >> 6] >> (r & 63)) & 1ULL); }

/* probe with fact key alternative k */
static int probe(rowctx* rc, const joininfo* ji, int jidx, int k) {
    const int64_t r = rc->row;
    const ocol* fk = ji->fks[k];
    if (!key_valid(fk, r)) return 0;
    int32_t v = fk->data[r];
    int32_t x = (int32_t)((uint32_t)v - (uint32_t)ji->kmin);
    if (x < 0) return 0;
    int64_t s = x % ji->len;
    if (ji->payload || ji->rowid) {
        if (ji->ht[2 * s] != v) return 0;
        rc->payload[jidx] = ji->ht[2 * s + 1];
        return 1;
    }
    return ji->ht[s] == v;
}

/* ------------------------------------------------------------------ */
/* aggregation state                                                    */
/* ------------------------------------------------------------------ */

typedef struct {
    uint64_t* res;
    uint8_t* bit;      /* validity bit per slot */
    uint8_t* written;  /* group column slot written */
    uint8_t* seen;     /* scalar path: any valid input per aggregate */
    uint64_t* pv;      /* projection: output rows (values, row-major) */
    uint8_t* pb;       /* projection: validity per cell */
    int64_t prows, pcap;
} aggstate;

static int state_alloc(aggstate* s, int64_t slots) {
    s->res = (uint64_t*)calloc((size_t)slots, sizeof(uint64_t));
    s->bit = (uint8_t*)calloc((size_t)slots, 1);
    s->written = (uint8_t*)calloc((size_t)slots, 1);
    s->seen = (uint8_t*)calloc((size_t)slots, 1);
    return (s->res && s->bit && s->written && s->seen) ? 0 : -1;
}

static void state_free(aggstate* s) {
    free(s->res);
    free(s->bit);
    free(s->written);
    free(s->seen);
    free(s->pv);
    free(s->pb);
}

static void emit_row(const rowctx* rc, aggstate* s) {
    const query* q = rc->q;
    const int n = q->nexprs;
    if (s->prows == s->pcap) {
        s->pcap = s->pcap ? 2 * s->pcap : 1024;
        s->pv = (uint64_t*)realloc(s->pv, (size_t)(s->pcap * n) * sizeof(uint64_t));
        s->pb = (uint8_t*)realloc(s->pb, (size_t)(s->pcap * n));
    }
    uint64_t* v = s->pv + s->prows * n;
    uint8_t* b = s->pb + s->prows * n;
    for (int c = 0; c < n; ++c) {
        const cval x = vexpr(rc, q->exprs[c]);
        b[c] = (uint8_t)(vvalid(rc, q->exprs[c]) != 0);
        if (q->out_type[c] == ORC_T_FLOAT64) {
            const double d = conv(x, T_F64).f;
            memcpy(&v[c], &d, 8);
        } else if (q->out_type[c] == ORC_T_INT64) {
            v[c] = (uint64_t)conv(x, T_I64).i;
        } else {
            v[c] = conv(x, T_U64).u;
        }
    }
    ++s->prows;
}

static void init_min_slots(const query* q, aggstate* s) {
    for (int64_t g = 0; g < q->G; ++g)
        for (int i = 0; i < q->naggs; ++i)
            if (!strcmp(q->agg->aggfn[i], "MIN")) s->res[g * q->storage_ts + q->ngroups + i] = ~0ULL;
}

static int64_t group_hash(const rowctx* rc) {
    const query* q = rc->q;
    const plan* a = q->agg;
    int64_t h = 0;
    for (int g = 0; g < a->ngroups; ++g) {
        const expr* e = a->groups[g];
        int64_t v = conv(vexpr(rc, e), T_I64).i;
        int64_t term;
        if (q->gnullable[g]) term = vvalid(rc, e) ? v - q->gmin[g] + 1 : 0;
        else term = v - q->gmin[g];
        h = g == 0 ? term : h * (int64_t)q->gcard[g] + term;
    }
    h %= q->G;
    return (int64_t)(int32_t)h;
}

static void aggregate_row(const rowctx* rc, aggstate* s) {
    const query* q = rc->q;
    const plan* a = q->agg;
    const int ts = q->storage_ts;
    if (a->ngroups == 0) {
        for (int i = 0; i < a->naggs; ++i) {
            const char* fn = a->aggfn[i];
            int valid = a->aggs[i]->k == E_STAR ? 1 : vvalid(rc, a->aggs[i]);
            if (!valid) continue;
            uint64_t val = !strcmp(fn, "COUNT") ? 1ULL : conv(mexpr(rc, a->aggs[i], 1), T_U64).u;
            if (!strcmp(fn, "MIN")) { if (val < s->res[i]) s->res[i] = val; }
            else if (!strcmp(fn, "MAX")) { if (val > s->res[i]) s->res[i] = val; }
            else s->res[i] += val;
            if (q->hidden_slot[i] >= 0) s->res[q->hidden_slot[i]] += 1;
            s->seen[i] = 1;
        }
        return;
    }
    int64_t h = group_hash(rc);
    if (h < 0 || h >= q->G) return; /* out-of-domain codes write outside d_result in the reference */
    int64_t out = h * ts;
    for (int g = 0; g < a->ngroups; ++g) {
        const expr* e = a->groups[g];
        s->res[out + g] = conv(vexpr(rc, e), T_U64).u;
        s->written[out + g] = 1;
        if (!q->fast_sparse && vvalid(rc, e)) s->bit[out + g] = 1;
    }
    for (int i = 0; i < a->naggs; ++i) {
        const char* fn = a->aggfn[i];
        const int slot = a->ngroups + i;
        int valid = a->aggs[i]->k == E_STAR ? 1 : vvalid(rc, a->aggs[i]);
        if (valid) {
            uint64_t val = !strcmp(fn, "COUNT") ? 1ULL : conv(mexpr(rc, a->aggs[i], 1), T_U64).u;
            uint64_t* ref = &s->res[out + slot];
            if (!strcmp(fn, "MIN")) { if (val < *ref) *ref = val; }
            else if (!strcmp(fn, "MAX")) { if (val > *ref) *ref = val; }
            else *ref += val;
            if (q->hidden_slot[i] >= 0) s->res[out + q->hidden_slot[i]] += 1;
            else if (strcmp(fn, "COUNT") && !q->fast_sparse) s->bit[out + slot] = 1;
        }
        if (!strcmp(fn, "COUNT") && !q->fast_sparse) s->bit[out + slot] = 1;
    }
}

static void merge_state(const query* q, aggstate* dst, const aggstate* src) {
    const int64_t slots = q->G * q->storage_ts;
    for (int64_t k = 0; k < slots; ++k) {
        const int64_t g = k / q->storage_ts;
        const int c = (int)(k % q->storage_ts);
        (void)g;
        dst->bit[k] |= src->bit[k];
        if (c < q->ngroups) {
            if (src->written[k]) { dst->res[k] = src->res[k]; dst->written[k] = 1; }
            continue;
        }
        if (c < q->visible_ts) {
            const char* fn = q->agg->aggfn[c - q->ngroups];
            if (!strcmp(fn, "MIN")) { if (src->res[k] < dst->res[k]) dst->res[k] = src->res[k]; }
            else if (!strcmp(fn, "MAX")) { if (src->res[k] > dst->res[k]) dst->res[k] = src->res[k]; }
            else dst->res[k] += src->res[k];
            dst->seen[k] |= src->seen[k];
        } else {
            dst->res[k] += src->res[k];
        }
    }
}

/* ------------------------------------------------------------------ */
/* HAVING over result slots                                             */
/* ------------------------------------------------------------------ */

typedef struct {
    const query* q;
    const uint64_t* res;
    const uint8_t* bit;
    int64_t base;
} tuplectx;

static cval slot_value(const tuplectx* t, int slot) {
    uint64_t raw = t->res[t->base + slot];
    int ty = t->q->out_type[slot];
    if (ty == ORC_T_FLOAT64) { double d; memcpy(&d, &raw, 8); return mk_f64(d); }
    if (ty == ORC_T_INT64) return mk_i64((int64_t)raw);
    return mk_u64(raw);
}

static cval hvalue(const tuplectx* t, const expr* e);
static int hvalid(const tuplectx* t, const expr* e);

static cval hvalue(const tuplectx* t, const expr* e) {
    switch (e->k) {
        case E_SLOT: return slot_value(t, (int)e->ival);
        case E_INT: return int_literal(e->ival);
        case E_FLOAT: return float_literal(e->fval);
        case E_NULL: return mk_i32(0);
        case E_STAR: return mk_i32(1);
        case E_ADD: case E_SUB: case E_MUL: case E_DIV:
        case E_EQ: case E_NE: case E_LT: case E_LE: case E_GT: case E_GE:
            return safe_binop(e->k, hvalue(t, e->a), hvalue(t, e->b));
        case E_ISNULL: return mk_i32(!hvalid(t, e->a));
        case E_ISNOTNULL: return mk_i32(hvalid(t, e->a));
        case E_NOT: return mk_i32(!truthy(hvalue(t, e->a)));
        case E_AND: return mk_i32(truthy(hvalue(t, e->a)) && truthy(hvalue(t, e->b)));
        case E_OR: return mk_i32(truthy(hvalue(t, e->a)) || truthy(hvalue(t, e->b)));
        case E_CASE: {
            int sel = hvalid(t, e->a) && truthy(hvalue(t, e->a));
            return cond_pick(sel, hvalue(t, e->b), hvalue(t, e->c));
        }
        default: return mk_i32(0);
    }
}

static int hvalid(const tuplectx* t, const expr* e) {
    switch (e->k) {
        case E_SLOT: return t->bit[t->base + e->ival];
        case E_NULL: return 0;
        case E_STAR: case E_INT: case E_FLOAT: case E_ISNULL: case E_ISNOTNULL: return 1;
        case E_ADD: case E_SUB: case E_MUL: case E_DIV:
        case E_EQ: case E_NE: case E_LT: case E_LE: case E_GT: case E_GE:
            return hvalid(t, e->a) && hvalid(t, e->b);
        case E_NOT: return hvalid(t, e->a);
        case E_AND: {
            int lv = hvalid(t, e->a), rv = hvalid(t, e->b);
            int l = truthy(hvalue(t, e->a)), r = truthy(hvalue(t, e->b));
            return (lv && !l) || (rv && !r) || (lv && rv);
        }
        case E_OR: {
            int lv = hvalid(t, e->a), rv = hvalid(t, e->b);
            int l = truthy(hvalue(t, e->a)), r = truthy(hvalue(t, e->b));
            return (lv && l) || (rv && r) || (lv && rv);
        }
        case E_CASE: {
            int sel = hvalid(t, e->a) && truthy(hvalue(t, e->a));
            return (sel && hvalid(t, e->b)) || (!sel && hvalid(t, e->c));
        }
        default: return 1;
    }
}

/* ------------------------------------------------------------------ */
/* result shaping                                                       */
/* ------------------------------------------------------------------ */

typedef struct {
    const query* q;
    const uint64_t* val;   /* visible row-major values */
    const uint8_t* vld;
} sortctx;

static sortctx g_sort;

/* "right before left" comparator of emitRightBeforeLeftComparison, as a
 * total order with the dense row index as the final tie-break. */
static int row_cmp(int64_t l, int64_t r) {
    const sortctx* s = &g_sort;
    const plan* sp = s->q->sort;
    const int nc = s->q->visible_ts;
    for (int k = 0; k < sp->nkeys; ++k) {
        const int c = sp->keycol[k];
        if (s->q->out_nullable[c]) {
            int lv = s->vld[l * nc + c], rv = s->vld[r * nc + c];
            if (lv != rv) return lv ? -1 : 1; /* valid rows first */
            if (!lv) continue;
        }
        uint64_t a = s->val[l * nc + c], b = s->val[r * nc + c];
        int lt, gt;
        switch (s->q->out_type[c]) {
            case ORC_T_INT64: lt = (int64_t)a < (int64_t)b; gt = (int64_t)a > (int64_t)b; break;
            case ORC_T_FLOAT64: {
                double x, y;
                memcpy(&x, &a, 8);
                memcpy(&y, &b, 8);
                lt = x < y;
                gt = y < x;
                break;
            }
            default: lt = a < b; gt = a > b; break;
        }
        if (gt) return sp->keydesc[k] ? -1 : 1;
        if (lt) return sp->keydesc[k] ? 1 : -1;
    }
    return l < r ? -1 : (l > r ? 1 : 0);
}

static int row_cmp_q(const void* a, const void* b) { return row_cmp(*(const int64_t*)a, *(const int64_t*)b); }

static int shape_result(const query* q, aggstate* s, orc_result* out) {
    const plan* a = q->agg;
    const int vts = q->visible_ts;
    const int sts = q->storage_ts;
    memset(out, 0, sizeof *out);
    out->ncols = vts;
    for (int c = 0; c < vts; ++c) { out->types[c] = q->out_type[c]; out->nullable[c] = q->out_nullable[c]; }
    out->ordered = 1;

    if (q->proj) {
        /* rows in scan order (the device's within-tile order is not fixed:
         * compare unordered projections as multisets), then ORDER BY / LIMIT
         * with the row position as the final tie-break */
        const int64_t rows = s->prows;
        uint64_t* dval = (uint64_t*)calloc((size_t)(rows * vts + 1), sizeof(uint64_t));
        uint8_t* dvld = (uint8_t*)calloc((size_t)(rows * vts + 1), 1);
        for (int64_t r = 0; r < rows; ++r)
            for (int c = 0; c < vts; ++c) {
                dval[r * vts + c] = s->pv[r * vts + c];
                dvld[r * vts + c] = q->out_nullable[c] ? s->pb[r * vts + c] : 1;
            }
        int64_t out_rows = rows;
        int64_t* order = (int64_t*)malloc((size_t)(rows + 1) * sizeof(int64_t));
        for (int64_t r = 0; r < rows; ++r) order[r] = r;
        out->ordered = q->sort != NULL;
        if (q->sort) {
            if (q->sort->nkeys > 0 && rows > 1) {
                g_sort.q = q;
                g_sort.val = dval;
                g_sort.vld = dvld;
                qsort(order, (size_t)rows, sizeof(int64_t), row_cmp_q);
            }
            if (q->sort->limit >= 0 && q->sort->limit < out_rows) out_rows = q->sort->limit;
        }
        out->nrows = out_rows;
        out->values = (uint64_t*)calloc((size_t)(out_rows * vts + 1), sizeof(uint64_t));
        out->valid = (uint8_t*)calloc((size_t)(out_rows * vts + 1), 1);
        for (int64_t r = 0; r < out_rows; ++r)
            for (int c = 0; c < vts; ++c) {
                out->values[r * vts + c] = dval[order[r] * vts + c];
                out->valid[r * vts + c] = dvld[order[r] * vts + c];
            }
        free(order);
        free(dval);
        free(dvld);
        return 0;
    }

    /* scalar aggregates: validity bits as the per-block reductions set them */
    if (a->ngroups == 0) {
        for (int i = 0; i < a->naggs; ++i) {
            const char* fn = a->aggfn[i];
            if (!strcmp(fn, "COUNT")) s->bit[i] = q->fact->rows > 0;
            else if (q->hidden_slot[i] < 0) s->bit[i] = s->seen[i];
        }
    }

    int any_avg = 0;
    for (int i = 0; i < a->naggs; ++i) any_avg |= !strcmp(a->aggfn[i], "AVG");
    /* finalize_aggregate_slots: storage tuple -> visible tuple */
    uint64_t* vres = (uint64_t*)calloc((size_t)(q->G * vts + 1), sizeof(uint64_t));
    uint8_t* vbit = (uint8_t*)calloc((size_t)(q->G * vts + 1), 1);
    if (!vres || !vbit) { free(vres); free(vbit); FAIL("out of memory"); }
    for (int64_t g = 0; g < q->G; ++g) {
        const int64_t src = g * sts, dst = g * vts;
        for (int c = 0; c < vts; ++c) {
            vres[dst + c] = s->res[src + c];
            vbit[dst + c] = s->bit[src + c];
        }
        if (!any_avg) continue;
        for (int i = 0; i < a->naggs; ++i) {
            const int c = a->ngroups + i;
            const char* fn = a->aggfn[i];
            vbit[dst + c] = 0;
            vres[dst + c] = 0;
            if (!strcmp(fn, "COUNT")) {
                vres[dst + c] = s->res[src + c];
                vbit[dst + c] = a->ngroups == 0 ? 1 : s->res[src + c] != 0;
            } else if (q->hidden_slot[i] >= 0) {
                uint64_t cnt = s->res[src + q->hidden_slot[i]];
                if (cnt != 0) {
                    double avg = (double)s->res[src + c] / (double)cnt;
                    memcpy(&vres[dst + c], &avg, 8);
                    vbit[dst + c] = 1;
                }
            } else {
                vres[dst + c] = s->res[src + c];
                if (!strcmp(fn, "MIN") && vres[dst + c] == ~0ULL) vres[dst + c] = 0;
                vbit[dst + c] = s->bit[src + c];
            }
        }
    }

    if (q->fast_sparse) {
        /* sparse output: no validity bits; a tuple is present when any cell is non-zero */
        out->ordered = vts <= 1;
        int64_t rows = 0;
        if (vts <= 1) {
            rows = 1;
        } else {
            for (int64_t g = 0; g < q->G; ++g) {
                int present = 0;
                for (int c = 0; c < vts && !present; ++c) present = vres[g * vts + c] != 0;
                rows += present;
            }
        }
        out->nrows = rows;
        out->values = (uint64_t*)calloc((size_t)(rows * vts + 1), sizeof(uint64_t));
        out->valid = (uint8_t*)calloc((size_t)(rows * vts + 1), 1);
        int64_t w = 0;
        for (int64_t g = 0; g < q->G && w < rows; ++g) {
            int present = vts <= 1;
            for (int c = 0; c < vts && !present; ++c) present = vres[g * vts + c] != 0;
            if (!present) continue;
            for (int c = 0; c < vts; ++c) { out->values[w * vts + c] = vres[g * vts + c]; out->valid[w * vts + c] = 1; }
            ++w;
        }
        free(vres);
        free(vbit);
        return 0;
    }

    /* dense: present && HAVING, in group-index order */
    int64_t* idx = (int64_t*)malloc((size_t)(q->G + 1) * sizeof(int64_t));
    int64_t rows = 0;
    for (int64_t g = 0; g < q->G; ++g) {
        int present = a->ngroups == 0;
        for (int c = 0; c < vts && !present; ++c) present = vbit[g * vts + c] || vres[g * vts + c] != 0;
        if (!present) continue;
        if (a->having) {
            tuplectx t = {q, vres, vbit, g * vts};
            if (!(hvalid(&t, a->having) && truthy(hvalue(&t, a->having)))) continue;
        }
        idx[rows++] = g;
    }
    /* dense columnar rows */
    uint64_t* dval = (uint64_t*)calloc((size_t)(rows * vts + 1), sizeof(uint64_t));
    uint8_t* dvld = (uint8_t*)calloc((size_t)(rows * vts + 1), 1);
    for (int64_t r = 0; r < rows; ++r)
        for (int c = 0; c < vts; ++c) {
            dval[r * vts + c] = vres[idx[r] * vts + c];
            dvld[r * vts + c] = q->out_nullable[c] ? vbit[idx[r] * vts + c] : 1;
        }
    int64_t out_rows = rows;
    int64_t* order = (int64_t*)malloc((size_t)(rows + 1) * sizeof(int64_t));
    for (int64_t r = 0; r < rows; ++r) order[r] = r;
    if (q->sort) {
        if (q->sort->nkeys > 0 && rows > 1) {
            g_sort.q = q;
            g_sort.val = dval;
            g_sort.vld = dvld;
            qsort(order, (size_t)rows, sizeof(int64_t), row_cmp_q);
        }
        if (q->sort->limit >= 0 && q->sort->limit < out_rows) out_rows = q->sort->limit;
    }
    out->nrows = out_rows;
    out->values = (uint64_t*)calloc((size_t)(out_rows * vts + 1), sizeof(uint64_t));
    out->valid = (uint8_t*)calloc((size_t)(out_rows * vts + 1), 1);
    for (int64_t r = 0; r < out_rows; ++r)
        for (int c = 0; c < vts; ++c) {
            out->values[r * vts + c] = dval[order[r] * vts + c];
            out->valid[r * vts + c] = dvld[order[r] * vts + c];
        }
    free(order);
    free(dval);
    free(dvld);
    free(idx);
    free(vres);
    free(vbit);
    return 0;
}

/* ------------------------------------------------------------------ */
/* driver                                                               */
/* ------------------------------------------------------------------ */

/* The pipeline from op i on for one fact row, in plan order.  A join with
 * several fact key alternatives (lo_orderdate = d_datekey OR lo_commitdate =
 * d_datekey) runs the rest of the pipeline once per alternative whose key is
 * valid, differs from every valid earlier alternative's key, and matches
 * (consumeHashJoinVector, visitor.cpp:4581-4638). */
static void run_ops(const query* q, rowctx* rc, int i, aggstate* st) {
    for (; i < q->nops; ++i) {
        const pipeop* op = &q->ops[i];
        if (op->kind == 0) {
            const int flag = op->fast ? fast_filter(rc, op->pred) : (vvalid(rc, op->pred) && truthy(vexpr(rc, op->pred)));
            if (!flag) return;
            continue;
        }
        const joininfo* ji = &q->joins[op->join];
        if (ji->mht) {
            /* ProbeMultiHT, then the rest of the pipeline once per build row
             * of the slot (visitor.cpp:4671-4730) */
            const ocol* fk = ji->fks[0];
            if (!key_valid(fk, rc->row)) return;
            const int32_t v = fk->data[rc->row];
            const int32_t x = (int32_t)((uint32_t)v - (uint32_t)ji->kmin);
            if (x < 0) return;
            const int64_t s = x % ji->len;
            if (ji->mkey[s] != v || ji->mcnt[s] <= 0) return;
            for (int32_t m = 0; m < ji->mcnt[s]; ++m) {
                rc->payload[op->join] = ji->mrow[ji->moff[s] + m];
                run_ops(q, rc, i + 1, st);
            }
            return;
        }
        if (ji->nfk == 1) {
            if (!probe(rc, ji, op->join, 0)) return;
            continue;
        }
        for (int k = 0; k < ji->nfk; ++k) {
            int dup = 0;
            for (int p = 0; p < k && !dup; ++p)
                dup = key_valid(ji->fks[p], rc->row) && ji->fks[p]->data[rc->row] == ji->fks[k]->data[rc->row];
            if (dup || !probe(rc, ji, op->join, k)) continue;
            run_ops(q, rc, i + 1, st);
        }
        return;
    }
    if (q->proj) emit_row(rc, st);
    else aggregate_row(rc, st);
}

static int run_scan(query* q, aggstate* total, int threads) {
    const int64_t N = q->fact->rows;
    const int64_t slots = q->G * q->storage_ts;
    int nthr = 1;
#ifdef _OPENMP
    nthr = threads > 0 ? threads : omp_get_max_threads();
#else
    (void)threads;
#endif
    aggstate* parts = (aggstate*)calloc((size_t)nthr, sizeof(aggstate));
    for (int t = 0; t < nthr; ++t) {
        if (state_alloc(&parts[t], slots)) FAIL("out of memory for aggregate state");
        init_min_slots(q, &parts[t]);
        if (q->agg->ngroups == 0)
            for (int i = 0; i < q->naggs; ++i)
                if (!strcmp(q->agg->aggfn[i], "MIN")) parts[t].res[i] = ~0ULL;
    }
#ifdef _OPENMP
#pragma omp parallel for num_threads(nthr) schedule(static)
#endif
    for (int64_t r = 0; r < N; ++r) {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        rowctx rc;
        rc.q = q;
        rc.row = r;
        run_ops(q, &rc, 0, &parts[tid]);
    }
    *total = parts[0];
    for (int t = 1; t < nthr; ++t) {
        if (q->proj) {
            /* static schedule: thread t scanned the t-th contiguous row range */
            aggstate* a = total;
            const aggstate* b = &parts[t];
            const int n = q->nexprs;
            if (b->prows) {
                a->pv = (uint64_t*)realloc(a->pv, (size_t)((a->prows + b->prows) * n) * sizeof(uint64_t));
                a->pb = (uint8_t*)realloc(a->pb, (size_t)((a->prows + b->prows) * n));
                memcpy(a->pv + a->prows * n, b->pv, (size_t)(b->prows * n) * sizeof(uint64_t));
                memcpy(a->pb + a->prows * n, b->pb, (size_t)(b->prows * n));
                a->prows += b->prows;
                a->pcap = a->prows;
            }
        } else {
            merge_state(q, total, &parts[t]);
        }
        state_free(&parts[t]);
    }
    free(parts);
    return 0;
}

int orc_execute(orc_db* db, const char* plan_sexpr, int threads, orc_result* out, char* err, size_t errlen) {
    parser p;
    memset(&p, 0, sizeof p);
    p.s = plan_sexpr;
    g_err[0] = 0;
    memset(out, 0, sizeof *out);
    query* q = (query*)calloc(1, sizeof(query));
    int rc = -1;
    aggstate st;
    memset(&st, 0, sizeof st);
    plan* root = parse_plan(&p);
    if (!root) goto done;
    q->db = db;
    if (analyze(q, root)) goto done;
    if (build_tables(q)) goto done;
    if (run_scan(q, &st, threads)) goto done;
    if (shape_result(q, &st, out)) goto done;
    rc = 0;
done:
    if (rc && err && errlen) snprintf(err, errlen, "%s", g_err[0] ? g_err : "oracle failure");
    for (int j = 0; j < q->njoins; ++j) {
        free(q->joins[j].ht);
        free(q->joins[j].mkey);
        free(q->joins[j].moff);
        free(q->joins[j].mcnt);
        free(q->joins[j].mrow);
    }
    state_free(&st);
    free(q);
    parser_free(&p);
    return rc;
}

void orc_result_free(orc_result* r) {
    free(r->values);
    free(r->valid);
    r->values = NULL;
    r->valid = NULL;
}
