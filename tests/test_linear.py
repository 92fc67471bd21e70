"""Linear form of interpreted plans (cq_lin_program, k_lin_scan): the fast-path
filters, one key-indexed probe, dense group keys and left-deep u64 aggregate
chains of a plan evaluated without per-instruction dispatch.

The crafted fact table has 100,003 rows (the last 4-row group of a thread is
partial), negative values, values near the int32 limits (so u64 products
wrap), nullable measure, group and foreign-key columns, and foreign keys
outside the dimension.  CPU tests pin which plans take the linear form and
the oracle by numpy closed forms; GPU tests compare the linear kernel, the
bytecode kernels and the compiled path with the oracle, including literal
rebinding through the prepared-statement cache and the served ad-hoc
shapes on the SSB data with nullable lo_revenue.
"""
import random

import numpy as np
import pytest

from tests import queries as Q
from tests.compare import assert_same
from tests.fixtures import bitmap_from_bools, build_pair

pa = pytest.importorskip("practice_amd")

N = 100_003
NP = 400


def lin_tables(seed=21):
    rng = np.random.default_rng(seed)
    b_ok = rng.random(N) > 0.1
    g_ok = rng.random(N) > 0.05
    k_ok = rng.random(N) > 0.03
    lo = {
        "lo_a": (rng.integers(-50, 51, N).astype(np.int32), None),
        "lo_b": (np.where(b_ok, rng.integers(0, 1_000_001, N), 0).astype(np.int32), bitmap_from_bools(b_ok)),
        "lo_c": (rng.integers(1, 51, N).astype(np.int32), None),
        "lo_big": (rng.integers(-2**31 + 1, 2**31 - 1, N).astype(np.int32), None),
        "lo_g": (np.where(g_ok, rng.integers(0, 7, N), 0).astype(np.int32), bitmap_from_bools(g_ok)),
        "lo_partkey": (np.where(k_ok, rng.integers(0, NP + 21, N), 0).astype(np.int32), bitmap_from_bools(k_ok)),
    }
    part = {
        "p_partkey": (np.arange(1, NP + 1, dtype=np.int32), None),
        "p_grp": (rng.integers(3, 10, NP).astype(np.int32), None),
        "p_val": (rng.integers(-1000, 1001, NP).astype(np.int32), None),
    }
    return {"LINEORDER": (N, True, lo), "PART": (NP, False, part)}


J = "from lineorder, part where lo_partkey = p_partkey"
LINEAR = {
    "scalar": "select count(*), sum(lo_b), min(lo_a), max(lo_a) from lineorder where lo_c < 30 and lo_a <> 0",
    "avg": "select avg(lo_c), count(lo_b), avg(lo_b) from lineorder where lo_c < 30 and lo_a <> 0",
    "chains": ("select sum(lo_big * lo_big), sum(lo_big * 3 - lo_a), sum(lo_b + lo_a - 7), sum(lo_a + 5000000000) "
               "from lineorder where lo_c >= 10"),
    "long_chain": "select sum(lo_a * lo_c - lo_big + 11), max(lo_big * 2 - lo_a) from lineorder where lo_c >= 10 and lo_c <= 45",
    "count_col": "select count(lo_b), count(*), min(lo_b + lo_a) from lineorder where lo_a > -20",
    "group_fact": "select lo_g, count(*), sum(lo_b), min(lo_big) from lineorder where lo_c <= 40 group by lo_g",
    "payload_chain": f"select sum(lo_a + p_val), count(*) {J} and lo_c > 5",
    "group_payload": f"select p_grp, count(*), sum(lo_a * lo_c), max(lo_b) {J} and lo_a < 25 group by p_grp",
    "two_keys": f"select p_grp, lo_g, count(*), max(lo_b), avg(lo_a) {J} and lo_a < 25 group by p_grp, lo_g",
    "build_filter": f"select lo_g, sum(lo_c) {J} and p_val > 0 group by lo_g",
    "having": "select lo_g, sum(lo_c) from lineorder where lo_a > 0 group by lo_g having sum(lo_c) > 30000",
    "empty": "select count(*), sum(lo_a), min(lo_b) from lineorder where lo_c > 1000",
    "empty_groups": "select lo_g, count(*), sum(lo_a) from lineorder where lo_c > 1000 group by lo_g",
}
BYTECODE = {
    "or_filter": "select count(*), sum(lo_a) from lineorder where lo_c < 5 or lo_a > 40",
    "right_product": "select sum(lo_a - lo_big * 3) from lineorder where lo_c >= 10",
    "literal_first": "select sum(2 * lo_a) from lineorder where lo_c >= 10",
    "division": "select sum(lo_b / 7) from lineorder where lo_c >= 10",
    "five_aggs": "select count(*), sum(lo_b), min(lo_a), max(lo_a), sum(lo_c) from lineorder where lo_c < 30",
    "case": "select sum(case when lo_a > 0 then lo_a else 0 end) from lineorder where lo_c > 3",
}


def _bits(words, n):
    return np.unpackbits(words.view(np.uint8), bitorder="little")[:n].astype(bool)


def _u64(x):
    return int(x) % (1 << 64)


def closed_form(tables, name):
    lo = {k: v[0].astype(np.int64) for k, v in tables["LINEORDER"][2].items()}
    bv = _bits(tables["LINEORDER"][2]["lo_b"][1], N)
    gv = _bits(tables["LINEORDER"][2]["lo_g"][1], N)
    a, b, c, big, g = lo["lo_a"], lo["lo_b"], lo["lo_c"], lo["lo_big"], lo["lo_g"]
    if name == "scalar":
        m = (c < 30) & (a != 0)
        # MIN / MAX compare the u64 values (negative values sort above)
        ua = [_u64(x) for x in a[m]]
        return [(int(m.sum()), int(b[m & bv].sum()), min(ua), max(ua))]
    if name == "avg":
        m = (c < 30) & (a != 0)
        return [(int(c[m].sum()) / int(m.sum()), int((m & bv).sum()), int(b[m & bv].sum()) / int((m & bv).sum()))]
    if name == "chains":
        m = c >= 10
        s1 = sum(_u64(x) * _u64(x) for x in big[m]) % (1 << 64)
        s2 = sum(_u64(x) * 3 - _u64(y) for x, y in zip(big[m], a[m])) % (1 << 64)
        mb = m & bv
        s3 = _u64(int((b[mb] + a[mb] - 7).sum()))
        s4 = _u64(int((a[m] + 5_000_000_000).sum()))
        return [(s1, s2, s3, s4)]
    if name == "count_col":
        m = a > -20
        return [(int((m & bv).sum()), int(m.sum()), min(_u64(x) for x in (b + a)[m & bv]))]
    if name == "group_fact":
        m = c <= 40
        out = []
        for key, ok in [(None, False)] + [(v, True) for v in range(7)]:
            sel = m & (gv if ok else ~gv)
            if ok:
                sel &= g == key
            if sel.any():
                out.append((key, int(sel.sum()), int(b[sel & bv].sum()), min(_u64(x) for x in big[sel])))
        return sorted(out, key=lambda t: (t[0] is not None, t[0] or 0))
    raise KeyError(name)


@pytest.fixture(scope="module")
def host_pair():
    tables = lin_tables()
    db, odb = build_pair(tables, pks=[("PART", "p_partkey")])
    db.set_option("interpret", 1)
    return tables, db, odb


@pytest.mark.parametrize("name", list(LINEAR))
def test_linear_form_taken(host_pair, name):
    tables, db, odb = host_pair
    d = db.describe(LINEAR[name])
    assert d["interpreted"] == 1 and d["linear"] == 1
    assert "linear form" in db.generate_code(LINEAR[name])


@pytest.mark.parametrize("name", list(BYTECODE))
def test_bytecode_only_shapes(host_pair, name):
    tables, db, odb = host_pair
    d = db.describe(BYTECODE[name])
    assert d["interpreted"] == 1 and d["linear"] == 0


def test_served_adhoc_shapes_are_linear(host_pair):
    """Every ad-hoc shape of the served mix (adhoc_query) has a linear form."""
    from oracle import oracle as orc
    spec = orc.GenSpec(sf=1, fact_rows=20_000, seed=5)
    data = orc.gen_tables(spec)
    db = pa.Database(-1)
    for t, cols in data.items():
        db.create_table(t, spec.rows(t), t == "LINEORDER")
        for c, (v, m) in cols.items():
            db.load_column(t, c, v, m)
        if t in orc.SSB_PK:
            db.set_primary_key(t, orc.SSB_PK[t])
    db.finalize()
    db.set_option("interpret", 1)
    rng = random.Random(3)
    for sql in [Q.ADHOC_1, Q.ADHOC_2] + [Q.adhoc_query(rng) for _ in range(40)]:
        assert db.describe(sql)["linear"] == 1, sql


@pytest.mark.parametrize("name", ["scalar", "avg", "chains", "count_col", "group_fact"])
def test_oracle_closed_form(host_pair, name):
    tables, db, odb = host_pair
    got = odb.execute(db.explain(LINEAR[name])).tuples()
    want = closed_form(tables, name)
    if name == "group_fact":
        got = sorted(got, key=lambda t: (t[0] is not None, t[0] or 0))
    assert got == want


def test_literal_slots_follow_emission_order(host_pair):
    """Literals keep their P.lit positions: two queries of one shape differ
    only in the literal values the plan carries."""
    tables, db, odb = host_pair
    a = db.describe("select count(*), sum(lo_a + 7) from lineorder where lo_c < 10 and lo_a > 3")
    b = db.describe("select count(*), sum(lo_a + 9) from lineorder where lo_c < 33 and lo_a > -2")
    assert a["signature"] == b["signature"] and a["linear"] == b["linear"] == 1
    assert a["lits"] != b["lits"] and len(a["lits"]) == len(b["lits"])


# ---------------------------------------------------------------- GPU

@pytest.fixture(scope="module")
def gpu_pair():
    tables = lin_tables()
    db, odb = build_pair(tables, pks=[("PART", "p_partkey")], device=0)
    yield tables, db, odb
    db.close()


MODES = {"linear": {"interpret": 1, "vm_linear": 1}, "bytecode": {"interpret": 1, "vm_linear": 0},
         "compiled": {"interpret": 0, "vm_linear": 1}}


def _set(db, opts):
    for k, v in opts.items():
        db.set_option(k, v)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", list(MODES))
def test_gpu_matches_oracle(gpu_pair, mode):
    tables, db, odb = gpu_pair
    _set(db, MODES[mode])
    try:
        for name, sql in list(LINEAR.items()) + list(BYTECODE.items()):
            print(f"[linear] mode={mode} {name}: {sql}", flush=True)
            got = db.execute(sql)
            assert_same(got, odb.execute(db.explain(sql)))
            if name in ("scalar", "avg", "chains", "count_col"):
                assert got.tuples() == closed_form(tables, name)
    finally:
        _set(db, MODES["compiled"])


@pytest.mark.gpu
def test_gpu_literal_rebinding(gpu_pair):
    """One prepared linear program serves queries that differ in constants."""
    tables, db, odb = gpu_pair
    _set(db, MODES["linear"])
    try:
        for lim, add, lo in ((10, 7, 3), (33, 9, -2), (50, -4_000_000_000, -51), (1, 0, 0)):
            sql = f"select count(*), sum(lo_a + {add}), max(lo_b) from lineorder where lo_c < {lim} and lo_a > {lo}"
            print(f"[linear] rebinding: {sql}", flush=True)
            assert_same(db.execute(sql), odb.execute(db.explain(sql)))
    finally:
        _set(db, MODES["compiled"])


@pytest.mark.gpu
def test_gpu_served_adhoc_shapes():
    """The served ad-hoc shapes on SSB data with nullable lo_revenue: the
    linear kernel against the oracle."""
    from tests.test_gpu_parity import dataset
    spec, data, odb, gdb = dataset("sf1c")
    _set(gdb, MODES["linear"])
    try:
        rng = random.Random(11)
        for sql in [Q.ADHOC_1, Q.ADHOC_2] + [Q.adhoc_query(rng) for _ in range(16)]:
            print(f"[linear] adhoc: {sql}", flush=True)
            assert gdb.describe(sql)["linear"] == 1
            assert_same(gdb.execute(sql), odb.execute(gdb.explain(sql)))
    finally:
        _set(gdb, MODES["compiled"])


@pytest.mark.gpu
def test_gpu_linear_group(gpu_pair):
    """Statements run together: the linear forms share scans (16 plans and 8
    columns at most per scan, so 26 statements take two), the bytecode-only
    shapes run alone; every result equals the statement run alone and the
    oracle."""
    tables, db, odb = gpu_pair
    _set(db, MODES["linear"])
    try:
        sqls = list(LINEAR.values()) + list(BYTECODE.values()) + [
            f"select count(*), sum(lo_a + {k}), max(lo_b) from lineorder where lo_c < {10 + 3 * k} and lo_a > {k - 20}"
            for k in range(13)]
        print(f"[linear] group of {len(sqls)} statements", flush=True)
        res = db.execute_group(sqls)
        for sql, r in zip(sqls, res):
            print(f"[linear] group member: {sql}", flush=True)
            o = odb.execute(db.explain(sql))
            assert_same(r, o)
            assert_same(db.execute(sql), o)
        for name in ("scalar", "avg", "chains", "count_col"):
            assert res[list(LINEAR).index(name)].tuples() == closed_form(tables, name)
    finally:
        _set(db, MODES["compiled"])


@pytest.mark.gpu
def test_gpu_linear_group_served_adhoc():
    """The served ad-hoc shapes on SSB data with nullable lo_revenue, run as
    one group (two shared scans), against the oracle."""
    from tests.test_gpu_parity import dataset
    spec, data, odb, gdb = dataset("sf1c")
    _set(gdb, MODES["linear"])
    try:
        rng = random.Random(12)
        sqls = [Q.ADHOC_1, Q.ADHOC_2] + [Q.adhoc_query(rng) for _ in range(22)]
        print(f"[linear] adhoc group of {len(sqls)}", flush=True)
        res = gdb.execute_group(sqls)
        for sql, r in zip(sqls, res):
            print(f"[linear] adhoc group member: {sql}", flush=True)
            assert_same(r, odb.execute(gdb.explain(sql)))
    finally:
        _set(gdb, MODES["compiled"])
