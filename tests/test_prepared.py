"""Prepared-program cache keyed by the statement with its integer literals
taken out (QueryEngine::prepareCached): for every statement, the cached
program re-bound to the statement's constants must equal a fresh plan of
that statement (signature, layouts, constants, LIMIT), and statements whose
constants shape more than the constants (LIMIT, ORDER BY position) must not
be served from the cache."""
import random

import pytest

from tests import queries as Q

pa = pytest.importorskip("practice_amd")


def _statements():
    rng = random.Random(21)
    qs = [Q.c2(c, r) for c, r in [(0, 0), (3, 4), (12, 2), (24, 1)]]
    for kind in (1, 2, 3, 4):
        qs += [Q.template_query(kind, rng) for _ in range(4)]
    qs += [Q.adhoc_query(rng) for _ in range(12)]
    qs += [
        "select c_nation, sum(lo_revenue) from lineorder, customer where lo_custkey = c_custkey "
        "group by c_nation order by 2 desc, c_nation limit 5",
        "select c_nation, sum(lo_revenue) from lineorder, customer where lo_custkey = c_custkey "
        "group by c_nation order by 2 desc, c_nation limit 7",
        "select count(*) from lineorder where lo_quantity < -3",
        "select count(*) from lineorder where lo_quantity < -4",
        "select sum(lo_revenue) from lineorder where lo_discount between 2 and 4",
    ]
    return qs


def _describe(db, q, cached):
    db.set_option("prepared_cache", 1 if cached else 0)
    try:
        return db.describe(q)
    finally:
        db.set_option("prepared_cache", 0)


def test_cached_programs_equal_fresh_plans():
    from tests.test_batch import _host_db
    db = _host_db()
    hits = 0
    for q in _statements():
        try:
            fresh = _describe(db, q, False)
        except pa.CrystalError:
            continue  # the dialect rejects it either way
        for _ in range(2):
            got = _describe(db, q, True)
            hits += got["plan"] == ""
            got.pop("plan")
            want = dict(fresh)
            want.pop("plan")
            assert got == want, q
    assert hits > 20  # the template statements are served from the cache


def test_limit_and_positions_are_not_rebound():
    from tests.test_batch import _host_db
    db = _host_db()
    a = ("select c_nation, sum(lo_revenue) from lineorder, customer where lo_custkey = c_custkey "
         "group by c_nation order by 2 desc, c_nation limit 5")
    b = a.replace("limit 5", "limit 9")
    _describe(db, a, True)
    got = _describe(db, b, True)
    assert got["limit"] == 9 and got["plan"] != ""
