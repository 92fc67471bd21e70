"""Randomized parity over crafted star schemas: dimension keys that start
below, at or above zero, run consecutively or with gaps; fact foreign keys
that include 0, keys outside the dimension and NULLs; nullable dimension
attributes.  These are the corners of the hash-table layouts (key-indexed
tables with and without a present bit, the reference's empty-slot match of
probe key 0, calendar-free dense indexes): every statement runs through the
plain, interpreted, batched and grouped paths and is compared with the
oracle, which restates the reference's `int[2*len]` table."""
import random

import numpy as np
import pytest

from tests.compare import assert_same
from tests.fixtures import bitmap_from_bools, build_pair

pa = pytest.importorskip("practice_amd")


def random_tables(seed):
    rng = np.random.default_rng(seed)
    r = random.Random(seed)

    def dim_keys(n):
        # n keys with distinct slots k % n of the reference's table (len = n
        # rows, key base 0): consecutive, or with gaps of whole multiples of
        # n.  Colliding keys are left out: the reference keeps whichever
        # insert wins its slot, so its result is not defined there.  Keys
        # below 0 never match (on both sides), so their slots stay empty
        # and probe key 0 can meet the reference's empty-slot match.
        kmin = r.choice([-3, 0, 1, 5])
        base = np.arange(kmin, kmin + n, dtype=np.int64)
        if r.random() < 0.5:
            return base.astype(np.int32)
        return np.sort(base + n * rng.integers(0, 3, n) * (base >= 0)).astype(np.int32)

    def dense(n, k):
        # every value of [0, min(n, k)) present: a group column whose catalog
        # domain is dense.  On a sparse domain the reference's group hash
        # folds several keys into one slot and reports whichever row wrote
        # the key last, which it does not define.
        return rng.permutation(np.arange(n) % k).astype(np.int32)

    ns, nc = r.randint(8, 200), r.randint(8, 300)
    skeys, ckeys = dim_keys(ns), dim_keys(nc)
    snat_ok = rng.random(ns) > (0.1 if r.random() < 0.5 else 0.0)
    snat = dense(ns, 25)
    snat_ok[[int(np.flatnonzero(snat == v)[0]) for v in np.unique(snat)]] = True  # each value once valid
    sup = {"s_suppkey": (skeys, None),
           "s_region": (dense(ns, 5), None),
           "s_nation": (np.where(snat_ok, snat, 0).astype(np.int32), None if snat_ok.all() else bitmap_from_bools(snat_ok))}
    cus = {"c_custkey": (ckeys, None),
           "c_region": (dense(nc, 5), None),
           "c_nation": (dense(nc, 25), None)}
    n = r.randint(3000, 20000)

    def fk(keys):
        lo, hi = int(keys.min()), int(keys.max())
        v = rng.integers(min(lo, 0) - 2, hi + 3, n).astype(np.int32)
        ok = rng.random(n) > (0.05 if r.random() < 0.5 else 0.0)
        return (np.where(ok, v, 0).astype(np.int32), None if ok.all() else bitmap_from_bools(ok))

    lo = {"lo_suppkey": fk(skeys), "lo_custkey": fk(ckeys),
          "lo_revenue": (rng.integers(1, 100_000, n).astype(np.int32), None),
          "lo_quantity": (rng.integers(1, 51, n).astype(np.int32), None)}
    return {"LINEORDER": (n, True, lo), "SUPPLIER": (ns, False, sup), "CUSTOMER": (nc, False, cus)}


def random_statements(seed, n=12):
    r = random.Random(seed * 7 + 1)
    out = []
    for _ in range(n):
        dims = r.sample(["supplier", "customer"], r.randint(1, 2))
        where = [{"supplier": "lo_suppkey = s_suppkey", "customer": "lo_custkey = c_custkey"}[d] for d in dims]
        keys = []
        for d in dims:
            p = d[0]
            if r.random() < 0.6:
                a = r.choice([f"{p}_region", f"{p}_nation"])
                hi = 4 if a.endswith("region") else 24
                v = r.randint(0, hi)
                where.append(r.choice([f"{a} = {v}", f"{a} <> {v}", f"{a} < {v}", f"({a} = {v} or {a} = {r.randint(0, hi)})"]))
            if r.random() < 0.5:
                keys.append(r.choice([f"{p}_region", f"{p}_nation"]))
        if r.random() < 0.4:
            where.append(f"lo_quantity < {r.randint(1, 50)}")
        aggs = r.sample(["count(*)", "sum(lo_revenue)", "min(lo_quantity)", "max(lo_revenue)", "avg(lo_quantity)"], r.randint(1, 3))
        keys = list(dict.fromkeys(keys))
        sql = "select " + ", ".join(keys + aggs) + " from lineorder, " + ", ".join(dims) + " where " + " and ".join(where)
        if keys:
            sql += " group by " + ", ".join(keys)
        out.append(sql)
    return out


SEEDS = [1, 2, 3, 4, 5, 6]


@pytest.mark.parametrize("seed", SEEDS)
def test_random_tables_plan(seed):
    db, _ = build_pair(random_tables(seed), pks=[("SUPPLIER", "s_suppkey"), ("CUSTOMER", "c_custkey")])
    for sql in random_statements(seed):
        assert db.describe(sql)["signature"], sql
    db.close()


@pytest.mark.gpu
@pytest.mark.parametrize("seed", SEEDS)
def test_random_tables_every_path(seed):
    db, odb = build_pair(random_tables(seed), pks=[("SUPPLIER", "s_suppkey"), ("CUSTOMER", "c_custkey")], device=0)
    try:
        sqls = random_statements(seed)
        want = [odb.execute(db.explain(s)) for s in sqls]

        def check(results, path):
            for s, r, o in zip(sqls, results, want):
                print(f"[random-tables] seed={seed} {path}: {s}", flush=True)
                assert_same(r, o)

        check([db.execute(s) for s in sqls], "plain")
        check(db.batch_execute(sqls, {"graph": 0}), "batched")
        check(db.batch_execute(sqls, {"graph": 0, "max_templates": 1}), "batched per template")
        check(db.batch_execute(sqls, {"graph": 0, "masks": 0}), "batched per query")
        db.set_option("interpret", 1)
        check([db.execute(s) for s in sqls], "interpreted")
        check(db.execute_group(sqls), "interpreted group")
    finally:
        db.close()


def random_mht_tables(seed):
    """CUSTOMER without a primary key and with repeated keys (a multi-value
    join: every matching dimension row joins)."""
    rng = np.random.default_rng(seed + 100)
    r = random.Random(seed + 100)
    nc = r.randint(20, 400)
    # distinct keys with distinct slots k % nc (the reference's table has
    # one directory key per slot; colliding distinct keys are undefined
    # there), each repeated over several rows
    kmin = r.choice([0, 1, 3])
    distinct = rng.choice(np.arange(nc), r.randint(max(2, nc // 8), max(3, nc // 2)), replace=False)
    distinct = distinct + nc * rng.integers(0, 2, len(distinct))
    distinct = distinct[distinct >= kmin] if (distinct >= kmin).sum() >= 2 else distinct + kmin * nc
    span = int(distinct.max()) - kmin + 1
    cus = {"c_custkey": (rng.choice(distinct, nc).astype(np.int32), None),
           "c_region": (rng.permutation(np.arange(nc) % 5).astype(np.int32), None),
           "c_nation": (rng.permutation(np.arange(nc) % 25).astype(np.int32), None)}
    n = r.randint(2000, 12000)
    lo = {"lo_custkey": (rng.integers(min(kmin, 0) - 1, kmin + span + 2, n).astype(np.int32), None),
          "lo_revenue": (rng.integers(1, 100_000, n).astype(np.int32), None),
          "lo_quantity": (rng.integers(1, 51, n).astype(np.int32), None)}
    return {"LINEORDER": (n, True, lo), "CUSTOMER": (nc, False, cus)}


def random_mht_statements(seed, n=8):
    r = random.Random(seed + 200)
    out = []
    for _ in range(n):
        where = ["lo_custkey = c_custkey"]
        if r.random() < 0.5:
            where.append(f"c_region {r.choice(['=', '<>', '<'])} {r.randint(0, 4)}")
        if r.random() < 0.5:
            where.append(f"lo_quantity < {r.randint(1, 50)}")
        if r.random() < 0.25:
            where.append(f"lo_quantity = {r.randint(1, 50)}")
            out.append("select lo_revenue, c_nation from lineorder, customer where " + " and ".join(where))
            continue
        key = r.choice(["c_region", "c_nation", None])
        aggs = r.sample(["count(*)", "sum(lo_revenue)", "max(lo_quantity)", "min(lo_revenue)"], r.randint(1, 3))
        sql = "select " + ", ".join(([key] if key else []) + aggs) + " from lineorder, customer where " + " and ".join(where)
        if key:
            sql += f" group by {key}"
        out.append(sql)
    return out


MHT_SEEDS = [1, 2, 3]


@pytest.mark.parametrize("seed", MHT_SEEDS)
def test_random_mht_plan(seed):
    db, _ = build_pair(random_mht_tables(seed), pks=[])
    for sql in random_mht_statements(seed):
        d = db.describe(sql)
        assert d["joins"][0]["layout"] == "multi-value", sql
    db.close()


@pytest.mark.gpu
@pytest.mark.parametrize("seed", MHT_SEEDS)
def test_random_mht_every_path(seed):
    """Multi-value joins run compiled on every path: alone, with the
    interpreted tier asked for (they fall back to compiled modules) and
    through the server (cold lanes)."""
    db, odb = build_pair(random_mht_tables(seed), pks=[], device=0)
    try:
        sqls = random_mht_statements(seed)
        want = [odb.execute(db.explain(s)) for s in sqls]

        def check(results, path):
            for s, r, o in zip(sqls, results, want):
                print(f"[random-mht] seed={seed} {path}: {s}", flush=True)
                assert_same(r, o)

        check([db.execute(s) for s in sqls], "plain")
        db.set_option("interpret", 1)
        check([db.execute(s) for s in sqls], "interpreted option")
        db.set_option("interpret", 0)
        srv = pa.Server(db, {"lanes": 1, "cold_lanes": 1})
        try:
            done = [srv.wait(t, 120_000) for t in [srv.submit(s) for s in sqls]]
        finally:
            srv.stop()
        check(done, "served")
    finally:
        db.close()
