"""Pins the CPU oracle before it is trusted as the parity checker.

Each check is independent of the oracle's own formulas:
  - the reference's golden fixture test/nullable_data (expected.tsv);
  - known-answer vectors from the reference's own pred.h / load.h compiled on
    the host (tests/golden/ref_vectors.json, regenerated by make_ref_vectors.py);
  - the SQL three-valued-logic truth table of test/test_sql_3vl.cpp:24-31;
  - numpy closed forms of the star-join queries on small generated instances.
"""
import json

import numpy as np
import pytest

from oracle import oracle as orc
from tests import queries as Q
from tests.fixtures import GOLDEN, bitmap_from_bools, build_pair, nullable_expected, nullable_queries, nullable_numbers_tables


# --------------------------------------------------------------- golden fixture
def test_nullable_fixture_matches_reference_expected():
    db, odb = build_pair(nullable_numbers_tables())
    qs, exp = nullable_queries(), nullable_expected()
    got = {k: odb.execute(db.explain(sql)).tuples() for k, sql in qs.items()}
    q1 = got["q1"][0]
    assert q1[:3] == (12, 7, 145)
    assert repr(q1[3]) == exp["q1"][1][3]  # 20.714285714285715, bit-exact double
    assert [tuple(map(int, r)) for r in exp["q2"][1:]] == got["q2"]
    assert got["q3"] == [(int(exp["q3"][1][0]),)]
    assert [tuple(map(int, r)) for r in exp["q5"][1:]] == got["q5"]
    # expected.tsv lists 4 rows with flag = 1, but nullable_numbers.csv holds
    # five (ids 1, 2, 6, 8, 11); the data file is authoritative here.
    assert exp["q4"][1] == ["4"]
    flags = nullable_numbers_tables()["NULLABLE_NUMBERS"][2]["flag"]
    assert int(((flags[0] == 1) & (np.unpackbits(flags[1].view(np.uint8), bitorder="little")[:12] == 1)).sum()) == 5
    assert got["q4"] == [(5,)]


# --------------------------------------------------------------- reference primitives
@pytest.fixture(scope="module")
def ref_vectors():
    return json.loads((GOLDEN / "ref_vectors.json").read_text())


def _vec_tables(disc, qty):
    n = len(disc)
    return {"VEC": (n, True, {"vec_disc": (np.array(disc, np.int32), None),
                              "vec_qty": (np.array(qty, np.int32), None)})}


@pytest.mark.parametrize("n", [512, 300])
def test_fast_path_and_chain_matches_blockpred(ref_vectors, n):
    disc, qty = ref_vectors["discount"][:n], ref_vectors["quantity"][:n]
    db, odb = build_pair(_vec_tables(disc, qty))
    sql = "select count(*) from vec where vec_disc >= 1 and vec_disc <= 3"
    assert odb.execute(db.explain(sql)).tuples() == [(sum(ref_vectors[f"disc_ge1_le3_n{n}"][:n]),)]


def test_fast_path_strict_and_not_equal_matches_blockpred(ref_vectors):
    db, odb = build_pair(_vec_tables(ref_vectors["discount"], ref_vectors["quantity"]))
    sql = "select count(*) from vec where vec_qty <> 25 and vec_qty < 40 and vec_qty > 10"
    assert odb.execute(db.explain(sql)).tuples() == [(sum(ref_vectors["qty_ne25_lt40_gt10"]),)]


def test_build_filter_or_mask_matches_blockpred_or(ref_vectors):
    n = 300
    disc = np.array(ref_vectors["discount"][:n], np.int32)
    keys = np.arange(1, n + 1, dtype=np.int32)
    tables = {"FACTV": (n, True, {"fv_key": (keys, None)}),
              "DIMV": (n, False, {"dv_key": (keys, None), "dv_disc": (disc, None)})}
    db, odb = build_pair(tables, pks=[("DIMV", "dv_key")])
    sql = "select count(*) from factv, dimv where fv_key = dv_key and (dv_disc = 2 or dv_disc = 7 or dv_disc > 9)"
    assert "(filter (or (or" in db.explain(sql)
    assert odb.execute(db.explain(sql)).tuples() == [(sum(ref_vectors["disc_eq2_or_eq7_or_gt9_n300"]),)]


def test_validity_bitmap_matches_blockloadvalidity(ref_vectors):
    n = 300
    words = np.array([int(w) for w in ref_vectors["bitmap"]], np.uint64)
    vals = np.arange(n, dtype=np.int32)
    db, odb = build_pair({"VV": (n, True, {"vv_val": (vals, words[: (n + 63) // 64].copy())})})
    sql = "select count(*) from vv where vv_val is not null"
    assert odb.execute(db.explain(sql)).tuples() == [(sum(ref_vectors["validity_and_n300"]),)]


# --------------------------------------------------------------- three-valued logic
T, F, U = "T", "F", "U"


def _sql_and(a, b):  # test_sql_3vl.cpp:14-18
    if F in (a, b):
        return F
    return T if (a, b) == (T, T) else U


def _sql_or(a, b):
    if T in (a, b):
        return T
    return F if (a, b) == (F, F) else U


def _sql_not(a):
    return {T: F, F: T, U: U}[a]


def test_three_valued_logic_truth_table():
    # reference known answers first (test_sql_3vl.cpp:24-31)
    assert _sql_and(T, U) == U and _sql_and(F, U) == F and _sql_and(U, U) == U
    assert _sql_or(T, U) == T and _sql_or(F, U) == U and _sql_or(U, U) == U and _sql_not(U) == U
    combos = [(a, b) for a in (T, F, U) for b in (T, F, U)]
    enc = {T: 1, F: 0, U: 0}
    a = np.array([enc[x] for x, _ in combos], np.int32)
    b = np.array([enc[y] for _, y in combos], np.int32)
    va = bitmap_from_bools([x != U for x, _ in combos])
    vb = bitmap_from_bools([y != U for _, y in combos])
    db, odb = build_pair({"TV": (len(combos), True, {"tv_a": (a, va), "tv_b": (b, vb)})})
    cases = {
        "not (tv_a = 1 and tv_b = 1)": lambda x, y: _sql_not(_sql_and(x, y)),
        "tv_a = 1 or tv_b = 1": lambda x, y: _sql_or(x, y),
        "not (tv_a = 1 or tv_b = 1)": lambda x, y: _sql_not(_sql_or(x, y)),
        "not (tv_a = 1)": lambda x, y: _sql_not(x),
        "(tv_a = 1 and tv_b = 1) or tv_a is null": lambda x, y: _sql_or(_sql_and(x, y), T if x == U else F),
        "not (tv_a = 1) or not (tv_b = 1)": lambda x, y: _sql_or(_sql_not(x), _sql_not(y)),
    }
    for where, f in cases.items():
        want = sum(1 for x, y in combos if f(x, y) == T)
        got = odb.execute(db.explain(f"select count(*) from tv where {where}")).tuples()[0][0]
        assert got == want, where


# --------------------------------------------------------------- closed forms of the star joins
@pytest.fixture(scope="module")
def small_ssb():
    spec = orc.GenSpec(sf=1, fact_rows=300_000, seed=3, nullable=Q.C3_NULLABLE)
    data = orc.gen_tables(spec)
    odb = orc.load_ssb(data, spec)
    import practice_amd as pa
    db = pa.Database(-1)
    for t, cols in data.items():
        db.create_table(t, spec.rows(t), t == "LINEORDER")
        for c, (v, m) in cols.items():
            db.load_column(t, c, v, m)
        if t in orc.SSB_PK:
            db.set_primary_key(t, orc.SSB_PK[t])
    db.finalize()
    return data, odb, db


def _bits(words, n):
    return np.unpackbits(words.view(np.uint8), bitorder="little")[:n].astype(bool)


def _lookup(keys, vals, probe):
    order = np.argsort(keys)
    idx = np.searchsorted(keys[order], probe)
    idx = np.clip(idx, 0, len(keys) - 1)
    hit = keys[order][idx] == probe
    return hit, vals[order][idx]


def test_c1_closed_form(small_ssb):
    data, odb, db = small_ssb
    lo, d = data["LINEORDER"], data["DDATE"]
    hit, year = _lookup(d["d_datekey"][0], d["d_year"][0], lo["lo_orderdate"][0])
    disc, qty, ext = lo["lo_discount"][0], lo["lo_quantity"][0], lo["lo_extendedprice"][0]
    m = hit & (year == 1993) & (disc >= 1) & (disc <= 3) & (qty < 25)
    want = int((ext[m].astype(np.uint64) * disc[m].astype(np.uint64)).sum(dtype=np.uint64))
    assert odb.execute(db.explain(Q.C1)).tuples() == [(want,)]


def test_aggregate_literal_stays_int_cast_column_goes_u64(small_ssb):
    # translateMathExprPush casts only a leftmost COLUMN (or *, predicate) to
    # unsigned long long; a leftmost literal keeps C int arithmetic
    # (visitor.cpp:1954-1968), which changes what division does to negatives
    data, odb, db = small_ssb
    d = data["LINEORDER"]["lo_discount"][0].astype(np.int64)
    m64 = (1 << 64) - 1
    lit_first = sum(int(-x / 3) if x else 0 for x in d.tolist()) & m64   # int: truncates toward zero
    col_first = sum((((x - 20) & m64) // 3) for x in d.tolist()) & m64    # u64 wrap, then divide
    assert odb.execute(db.explain("select sum((0 - lo_discount) / 3) from lineorder")).tuples() == [(lit_first,)]
    assert odb.execute(db.explain("select sum((lo_discount - 20) / 3) from lineorder")).tuples() == [(col_first,)]


def test_c2_closed_form(small_ssb):
    data, odb, db = small_ssb
    lo, p, s, d = data["LINEORDER"], data["PART"], data["SUPPLIER"], data["DDATE"]
    hp, brand = _lookup(p["p_partkey"][0], p["p_brand1"][0], lo["lo_partkey"][0])
    _, cat = _lookup(p["p_partkey"][0], p["p_category"][0], lo["lo_partkey"][0])
    hs, reg = _lookup(s["s_suppkey"][0], s["s_region"][0], lo["lo_suppkey"][0])
    hd, year = _lookup(d["d_datekey"][0], d["d_year"][0], lo["lo_orderdate"][0])
    m = hp & hs & hd & (cat == 12) & (reg == 1)
    rev = lo["lo_revenue"][0]
    rv = _bits(lo["lo_revenue"][1], len(rev))
    groups = {}
    for y, b, r, v in zip(year[m], brand[m], rev[m], rv[m]):
        g = groups.setdefault((int(y), int(b)), [0, False])
        if v:
            g[0] += int(r)
            g[1] = True
    want = [(y, b, (v[0] if v[1] else None)) for (y, b), v in sorted(groups.items())]
    assert odb.execute(db.explain(Q.C2)).tuples() == want


def test_c3_closed_form_with_null_payload_quirk(small_ssb):
    """NULL s_nation loses validity through the payload (raw 0 groups with nation 0);
    NULL revenue passes the IS NULL arm of the filter but is skipped by SUM."""
    data, odb, db = small_ssb
    lo, c, s, d = data["LINEORDER"], data["CUSTOMER"], data["SUPPLIER"], data["DDATE"]
    _, cn = _lookup(c["c_custkey"][0], c["c_nation"][0], lo["lo_custkey"][0])
    _, sn = _lookup(s["s_suppkey"][0], s["s_nation"][0], lo["lo_suppkey"][0])
    _, year = _lookup(d["d_datekey"][0], d["d_year"][0], lo["lo_orderdate"][0])
    rev = lo["lo_revenue"][0].astype(np.int64)
    rv = _bits(lo["lo_revenue"][1], len(rev))
    keep = (rv & (rev > 1_000_000)) | ~rv
    key = (cn.astype(np.int64) * 25 + sn) * 7 + (year - 1992)
    sums = np.bincount(key[keep & rv], weights=None, minlength=25 * 25 * 7)
    tot = np.zeros(25 * 25 * 7, np.int64)
    np.add.at(tot, key[keep & rv], rev[keep & rv])
    present = np.bincount(key[keep], minlength=25 * 25 * 7) > 0
    anyvalid = sums > 0
    want = sorted((int(k // 175), int(k // 7 % 25), int(1992 + k % 7), int(tot[k]) if anyvalid[k] else None)
                  for k in np.nonzero(present)[0])
    got = odb.execute(db.explain(Q.C3))
    assert sorted(got.tuples()) == want
    assert got.nullable == [False, True, False, True]


# --------------------------------------------------------------- generator
def test_generator_domains_and_calendar():
    spec = orc.GenSpec(sf=20, fact_rows=100_000, seed=9, nullable="LINEORDER.lo_revenue:200")
    assert [spec.rows(t) for t in orc.SSB_TABLES] == [100_000, 1_000_000, 40_000, 600_000, 2556]
    dk, _ = orc.gen_column(spec, "DDATE", "d_datekey")
    assert dk[0] == 19920101 and dk[-1] == 19981230 and len(np.unique(dk)) == 2556
    assert 19920229 in dk and 19960229 in dk and 19930229 not in dk
    yr, _ = orc.gen_column(spec, "DDATE", "d_year")
    assert (yr == dk // 10000).all()
    cat, _ = orc.gen_column(spec, "PART", "p_category")
    brand, _ = orc.gen_column(spec, "PART", "p_brand1")
    assert (brand // 40 == cat).all() and cat.min() == 0 and cat.max() == 24
    nat, _ = orc.gen_column(spec, "SUPPLIER", "s_nation")
    reg, _ = orc.gen_column(spec, "SUPPLIER", "s_region")
    assert (reg == nat // 5).all()
    pk, _ = orc.gen_column(spec, "CUSTOMER", "c_custkey")
    assert (pk == np.arange(1, 600_001)).all()
    fk, _ = orc.gen_column(spec, "LINEORDER", "lo_partkey")
    assert fk.min() >= 1 and fk.max() <= 1_000_000
    of, _ = orc.gen_column(spec, "LINEORDER", "lo_orderdate")
    assert np.isin(of, dk).all()
    rev, valid = orc.gen_column(spec, "LINEORDER", "lo_revenue")
    v = _bits(valid, len(rev))
    assert 0.015 < 1 - v.mean() < 0.025 and (rev[~v] == 0).all()
    assert rev[v].min() >= 100_000 and rev[v].max() <= 10_000_000
    # any row range regenerates identically (counter-based streams)
    part, _ = orc.gen_column(spec, "LINEORDER", "lo_partkey", 4096, 1000)
    assert (part == fk[4096:5096]).all()


# --------------------------------------------------------------- OR-key join
def or_key_tables(seed=11, n=20000):
    """DDATE of 400 dates in 1992-1993 and a fact table whose commit date
    equals its order date for about a third of the rows, is NULL for a tenth,
    and whose keys include dates outside DDATE and key 0."""
    rng = np.random.default_rng(seed)
    days = np.array([y * 10000 + m * 100 + d for y in (1992, 1993) for m in range(1, 13) for d in range(1, 29)], np.int32)
    dates = np.sort(rng.choice(days, 400, replace=False)).astype(np.int32)
    years = (dates // 10000).astype(np.int32)
    probe = np.concatenate([dates, np.array([19910101, 0, 19940101], np.int32), days[:50]])
    od = rng.choice(probe, n).astype(np.int32)
    cd = np.where(rng.random(n) < 0.35, od, rng.choice(probe, n)).astype(np.int32)
    cd_valid = rng.random(n) > 0.1
    cd = np.where(cd_valid, cd, 0).astype(np.int32)
    rev = rng.integers(0, 100000, n).astype(np.int32)
    disc = rng.integers(0, 11, n).astype(np.int32)
    valid_words = np.packbits(np.concatenate([cd_valid, np.zeros((-n) % 64, bool)]).astype(np.uint8),
                              bitorder="little").view("<u8").copy()
    return {"DDATE": {"d_datekey": (dates, None), "d_year": (years, None)},
            "LINEORDER": {"lo_orderdate": (od, None), "lo_commitdate": (cd, valid_words), "lo_revenue": (rev, None),
                          "lo_discount": (disc, None)}}


OR_KEY_QUERIES = [
    "select d_year, count(*), sum(lo_revenue) from lineorder, ddate where lo_orderdate = d_datekey "
    "or lo_commitdate = d_datekey group by d_year",
    "select count(*), sum(lo_revenue) from lineorder, ddate where (lo_orderdate = d_datekey or lo_commitdate = d_datekey) "
    "and d_year = 1993",
    "select d_year, count(*) from lineorder, ddate where (lo_orderdate = d_datekey or lo_commitdate = d_datekey) "
    "and d_year - 1990 > lo_discount group by d_year",
]


def or_key_dbs(tables):
    import practice_amd as pa
    odb = orc.OracleDB()
    db = pa.Database(-1)
    for t in ("LINEORDER", "DDATE"):
        n = len(next(iter(tables[t].values()))[0])
        odb.add_table(t, n, t == "LINEORDER")
        db.create_table(t, n, t == "LINEORDER")
        for c, (v, m) in tables[t].items():
            odb.add_column(t, c, v, m)
            db.load_column(t, c, v, m)
    odb.set_primary_key("DDATE", "d_datekey")
    db.set_primary_key("DDATE", "d_datekey")
    odb.finalize()
    db.finalize()
    return odb, db


def or_key_closed_form(tables, q):
    """Per fact row: the order-date key, then the commit-date key when it is
    valid and differs; each that names a DDATE row (passing the query's
    filters) contributes that row's year once."""
    d = tables["DDATE"]
    year_of = dict(zip(d["d_datekey"][0].tolist(), d["d_year"][0].tolist()))
    lo = tables["LINEORDER"]
    od, cd, rev, disc = (lo[c][0].tolist() for c in ("lo_orderdate", "lo_commitdate", "lo_revenue", "lo_discount"))
    n = len(od)
    cv = np.unpackbits(lo["lo_commitdate"][1].view(np.uint8), bitorder="little")[:n].astype(bool).tolist()
    groups = {}
    for r in range(n):
        keys = [od[r]] + ([cd[r]] if cv[r] and cd[r] != od[r] else [])
        for k in keys:
            y = year_of.get(k)
            if y is None:
                continue
            if q == 1 and y != 1993:
                continue
            if q == 2 and not (y - 1990 > disc[r]):
                continue
            g = groups.setdefault(y if q != 1 else 0, [0, 0])
            g[0] += 1
            g[1] += rev[r]
    if q == 0:
        return sorted((y, c, s) for y, (c, s) in groups.items())
    if q == 1:
        c, s = groups.get(0, [0, 0])
        return [(c, s)]
    return sorted((y, c) for y, (c, _) in groups.items())


def test_or_key_join_closed_form():
    tables = or_key_tables()
    odb, db = or_key_dbs(tables)
    for q, sql in enumerate(OR_KEY_QUERIES):
        got = odb.execute(db.explain(sql)).tuples()
        assert sorted(got) == or_key_closed_form(tables, q), sql


# --------------------------------------------------------------- projection
PROJECTION_QUERIES = [
    "select lo_revenue, lo_discount * 2, lo_discount < 3 from lineorder where lo_discount < 4",
    "select lo_revenue, d_year from lineorder, ddate where lo_orderdate = d_datekey and lo_discount = 1",
    "select lo_commitdate, lo_discount from lineorder where lo_discount = 2",
    "select lo_revenue, lo_discount from lineorder where lo_discount = 3 order by lo_revenue desc limit 10",
    "select * from lineorder where lo_discount = 9",
    # every column of every scanned table, in plan order (DDATE, then LINEORDER)
    "select * from lineorder, ddate where lo_orderdate = d_datekey and lo_discount = 8",
]


def projection_closed_form(tables, q):
    lo = tables["LINEORDER"]
    od, cd, rev, disc = (lo[c][0].astype(np.int64) for c in ("lo_orderdate", "lo_commitdate", "lo_revenue", "lo_discount"))
    n = len(od)
    cv = np.unpackbits(lo["lo_commitdate"][1].view(np.uint8), bitorder="little")[:n].astype(bool)
    d = tables["DDATE"]
    year_of = dict(zip(d["d_datekey"][0].tolist(), d["d_year"][0].tolist()))
    rows = range(n)
    if q == 0:
        return [(int(rev[r]), int(disc[r]) * 2, int(disc[r] < 3)) for r in rows if disc[r] < 4]
    if q == 1:
        return [(int(rev[r]), year_of[int(od[r])]) for r in rows if disc[r] == 1 and int(od[r]) in year_of]
    if q == 2:
        return [(int(cd[r]) if cv[r] else None, 2) for r in rows if disc[r] == 2]
    if q == 3:
        sel = sorted((r for r in rows if disc[r] == 3), key=lambda r: (-rev[r], r))[:10]
        return [(int(rev[r]), 3) for r in sel]
    if q == 4:
        return [(int(od[r]), int(cd[r]) if cv[r] else None, int(rev[r]), 9) for r in rows if disc[r] == 9]
    return [(int(od[r]), year_of[int(od[r])], int(od[r]), int(cd[r]) if cv[r] else None, int(rev[r]), 8)
            for r in rows if disc[r] == 8 and int(od[r]) in year_of]


def test_projection_closed_form():
    tables = or_key_tables(seed=12, n=5000)
    odb, db = or_key_dbs(tables)
    for q, sql in enumerate(PROJECTION_QUERIES):
        got = odb.execute(db.explain(sql))
        assert got.tuples() == projection_closed_form(tables, q), sql
    r = odb.execute(db.explain(PROJECTION_QUERIES[0]))
    assert r.types == [0, 0, 1] and r.nullable == [False, False, False]
    assert odb.execute(db.explain(PROJECTION_QUERIES[2])).nullable == [True, False]


# --------------------------------------------------------------- aggregate shaping: HAVING, MIN / MAX, ORDER BY
def _year_groups(data):
    lo, d = data["LINEORDER"], data["DDATE"]
    hit, year = _lookup(d["d_datekey"][0], d["d_year"][0], lo["lo_orderdate"][0])
    rev = lo["lo_revenue"][0].astype(np.int64)
    rv = _bits(lo["lo_revenue"][1], len(rev))
    cnt, tot = {}, {}
    for y in range(1992, 1999):
        m = hit & (year == y)
        cnt[y] = int(m.sum())
        tot[y] = int(rev[m & rv].sum())
    return cnt, tot


def test_having_closed_form(small_ssb):
    """HAVING keeps the groups whose visible aggregates satisfy the predicate
    (visitor.cpp:2441-2450), thresholds chosen between the actual values so
    some years are cut by each arm."""
    data, odb, db = small_ssb
    cnt, tot = _year_groups(data)
    c_thr = sorted(cnt.values())[2]
    s_thr = sorted(tot.values())[4]
    sql = (f"select d_year, sum(lo_revenue), count(*) from lineorder, ddate where lo_orderdate = d_datekey "
           f"group by d_year having count(*) >= {c_thr} and sum(lo_revenue) > {s_thr} order by d_year")
    want = [(y, tot[y], cnt[y]) for y in range(1992, 1999) if cnt[y] >= c_thr and tot[y] > s_thr]
    assert 0 < len(want) < 7
    assert odb.execute(db.explain(sql)).tuples() == want
    sql = (f"select d_year, count(*), sum(lo_revenue) from lineorder, ddate where lo_orderdate = d_datekey "
           f"group by d_year having not (count(*) < {c_thr}) or sum(lo_revenue) <= {s_thr} order by d_year")
    want = [(y, cnt[y], tot[y]) for y in range(1992, 1999) if not (cnt[y] < c_thr) or tot[y] <= s_thr]
    assert odb.execute(db.explain(sql)).tuples() == want


def test_min_max_closed_form(small_ssb):
    """MIN / MAX over u64 values of the leftmost-cast expression (inline_math.h
    atomic min / max), NULL inputs skipped, per group and scalar."""
    data, odb, db = small_ssb
    lo = data["LINEORDER"]
    tax, qty, ext, disc = (lo[c][0].astype(np.int64) for c in ("lo_tax", "lo_quantity", "lo_extendedprice", "lo_discount"))
    rev = lo["lo_revenue"][0].astype(np.int64)
    rv = _bits(lo["lo_revenue"][1], len(rev))
    m64 = (1 << 64) - 1
    sql = ("select lo_tax, min(lo_quantity), max(lo_extendedprice), min(lo_revenue), max(lo_revenue), "
           "count(lo_revenue), min(lo_discount - 5), max(lo_discount - 5) from lineorder where lo_discount = 5 "
           "or lo_discount = 2 group by lo_tax order by lo_tax")
    want = []
    for t in sorted(set(tax.tolist())):
        m = (tax == t) & ((disc == 5) | (disc == 2))
        if not m.any():
            continue
        mv = m & rv
        wrapped = [(x - 5) & m64 for x in disc[m].tolist()]
        want.append((t, int(qty[m].min()), int(ext[m].max()), int(rev[mv].min()), int(rev[mv].max()), int(mv.sum()),
                     min(wrapped), max(wrapped)))
    assert odb.execute(db.explain(sql)).tuples() == want
    m = qty == 1
    sql = "select min(lo_revenue), max(lo_revenue), min(lo_extendedprice) from lineorder where lo_quantity = 1"
    assert odb.execute(db.explain(sql)).tuples() == [(int(rev[m & rv].min()), int(rev[m & rv].max()), int(ext[m].min()))]


def ties_tables():
    """A fact table of ten groups whose counts tie (5, 3 and 2 rows) and whose
    values are NULL for every row of groups 3 and 6."""
    counts = [5, 3, 5, 2, 5, 3, 1, 5, 4, 2]
    g = np.repeat(np.arange(10, dtype=np.int32), counts)
    rng = np.random.default_rng(4)
    g = g[rng.permutation(len(g))]
    valid = ~np.isin(g, [3, 6])
    v = np.where(valid, 1, 0).astype(np.int32)
    return {"TIES": (len(g), True, {"t_g": (g, None), "t_v": (v, bitmap_from_bools(valid))})}


# ORDER BY over the dense result: valid before NULL in either direction, then
# the value, then the group index (visitor.cpp:320-358); group index = t_g here
TIES_QUERIES = [
    ("select t_g, count(*) from ties group by t_g order by 2 desc limit 6",
     [(0, 5), (2, 5), (4, 5), (7, 5), (8, 4), (1, 3)]),
    ("select t_g, sum(t_v) from ties group by t_g order by 2 desc, t_g desc",
     [(7, 5), (4, 5), (2, 5), (0, 5), (8, 4), (5, 3), (1, 3), (9, 2), (6, None), (3, None)]),
    ("select t_g, sum(t_v) from ties group by t_g order by 2 limit 3", [(9, 2), (1, 3), (5, 3)]),
    ("select t_g, sum(t_v) from ties group by t_g order by 2", [(9, 2), (1, 3), (5, 3), (8, 4), (0, 5), (2, 5), (4, 5),
                                                                 (7, 5), (3, None), (6, None)]),
]


def test_order_by_desc_limit_tie_break_known_answers():
    tables = ties_tables()
    g = tables["TIES"][2]["t_g"][0]
    assert np.bincount(g).tolist() == [5, 3, 5, 2, 5, 3, 1, 5, 4, 2]
    db, odb = build_pair(tables)
    for sql, want in TIES_QUERIES:
        got = odb.execute(db.explain(sql))
        assert got.ordered and got.tuples() == want, sql
