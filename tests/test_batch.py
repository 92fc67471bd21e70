"""Batched shared scan: one persistent launch per group of plan signatures.

CPU tests check the generated modules (the union of fact columns is loaded
once per tile, every template module compiles for gfx950).  GPU tests check
that every query of a batch returns exactly what it returns alone and what
the oracle returns, with and without graph replay and autotuning.
"""
import pytest

from tests import queries as Q
from tests.compare import assert_results_equal, assert_same

pa = pytest.importorskip("practice_amd")


def mixed_queries():
    from tests.test_gpu_parity import EXTRA
    qs = [Q.C1, Q.C2, Q.C3]
    qs += [Q.c2(c, r) for c, r in [(0, 0), (3, 4), (12, 2), (24, 1), (7, 3)]]
    qs += EXTRA
    return qs


# ---------------------------------------------------------------- CPU

def _host_db():
    from oracle import oracle as orc
    spec = orc.GenSpec(sf=1, fact_rows=50_000, seed=1, nullable=Q.C3_NULLABLE)
    data = orc.gen_tables(spec)
    db = pa.Database(-1)
    for t, cols in data.items():
        db.create_table(t, spec.rows(t), t == "LINEORDER")
        for c, (v, m) in cols.items():
            db.load_column(t, c, v, m)
        if t in orc.SSB_PK:
            db.set_primary_key(t, orc.SSB_PK[t])
    db.finalize()
    return db


@pytest.fixture(scope="module")
def hdb():
    return _host_db()


def test_shared_scan_loads_each_column_once(hdb):
    qs = [Q.C1, Q.C2, Q.c2(3, 4), Q.C3]
    src = hdb.generate_batch_code(qs)
    # union of fact columns: C1 {orderdate, discount, quantity, extendedprice},
    # C2 {orderdate, partkey, suppkey, revenue}, C3 {custkey, suppkey, orderdate, revenue}:
    # each of the 8 is streamed from HBM once per tile (nontemporal load by the
    # last template that reads it); the 4 earlier reads of shared columns are
    # cached loads of the same tile, and so is the key column read by each
    # launch-level probe (a probe several templates make alike: DDATE on
    # lo_orderdate, made by all three, and SUPPLIER on lo_suppkey, made by
    # the C2 and C3 templates: both pack s_region and s_nation)
    assert src.count("__builtin_nontemporal_load") == 8
    launch_probes = src.count("        // launch-level probe ")
    assert launch_probes == 2 and "// launch-level probe DDATE on lo_orderdate" in src
    assert "// launch-level probe SUPPLIER on lo_suppkey" in src
    assert src.count("*((const v4i*)") == 4 + launch_probes
    # three templates (the two C2 queries share one), one block each per tile
    assert src.count("        // template ") == 3
    pa.compile_source(src)


def test_templates_with_identical_builds_stay_distinct(hdb):
    # no joins, no HAVING: both template modules hold the same kernels, but
    # their per-tile pipelines differ and must not be merged
    qs = ["select count(*) from lineorder where lo_revenue < 5000000",
          "select count(*) from lineorder where lo_revenue is null",
          "select count(*) from lineorder where lo_revenue < 7000000"]
    assert hdb.generate_batch_code(qs, part=0) == hdb.generate_batch_code(qs, part=1)
    src = hdb.generate_batch_code(qs)
    assert src.count("        // template ") == 2


def test_too_many_templates_for_one_scan(hdb):
    with pytest.raises(pa.CrystalError, match="max 16"):
        hdb.generate_batch_code(mixed_queries())


def test_template_modules_compile(hdb):
    qs = mixed_queries()
    for k in range(len(qs)):
        part = hdb.generate_batch_code(qs, part=k)
        assert "cq_dense_count" in part and "cq_scan(" not in part
        pa.compile_source(part)


def test_mask_form_for_template_queries(hdb):
    """The four serving templates reduce to query-mask tables (one pass for
    all queries of a template); masks=0 keeps the per-query loop."""
    import random
    rng = random.Random(4)
    qs = [Q.template_query(k, rng) for k in (1, 2, 3, 4) for _ in range(3)]
    src = hdb.generate_batch_code(qs)
    assert src.count("// query masks") == 4
    assert "for (int q = B.tq_begin[" not in src.split("for (long long tile")[1].split("__syncthreads")[0]
    # CUSTOMER and PART (too large for LDS): in a launch of several
    # templates their probes are issued early (their L2 round trips overlap
    # the masks); a launch of one template probes them after the other
    # joins' masks, only for items some query still accepts
    assert "if (qm[i] == 0ull) continue;" not in src
    for k, dim, key in ((3, "PART", "lo_partkey"), (4, "CUSTOMER", "lo_custkey")):
        assert f"// shared probe {dim}" in src
        one = hdb.generate_batch_code([Q.template_query(k, rng) for _ in range(3)])
        assert f"// {dim} on {key}: probed for the items some query still accepts" in one
        assert f"// shared probe {dim}" not in one and one.count("if (qm[i] == 0ull) continue;") == 1
        pa.compile_source(one)
        hdb.set_option("defer_probes", 0)
        try:
            assert "if (qm[i] == 0ull)" not in hdb.generate_batch_code([Q.template_query(k, rng) for _ in range(3)])
        finally:
            hdb.set_option("defer_probes", 1)
    plain = hdb.generate_batch_code(qs, {"masks": 0})
    assert "// query masks" not in plain
    pa.compile_source(src)
    # a filter on a column with a large domain keeps the per-query form
    wide = ["select sum(lo_revenue) from lineorder where lo_extendedprice < 5000000",
            "select sum(lo_revenue) from lineorder where lo_extendedprice < 6000000"]
    assert "// query masks" not in hdb.generate_batch_code(wide)


def test_table_placement_mask(hdb):
    """table_mask places each candidate shared table in LDS (bit set) or
    HBM: every mask compiles, 0 stages no dimension table, the default
    stages them all (the placements the autotuner times)."""
    qs = [Q.C1, Q.C2, Q.C3]
    full = hdb.generate_batch_code(qs)
    none = hdb.generate_batch_code(qs, {"table_mask": 0})
    one = hdb.generate_batch_code(qs, {"table_mask": 1})
    n_full = full.count("__shared__ u32 cq_tab")
    assert n_full >= 2 and none.count("__shared__ u32 cq_tab") == 0 and one.count("__shared__ u32 cq_tab") == 1
    for src in (full, none, one):
        pa.compile_source(src)


def test_table_lds_budget(hdb):
    """table_kb raises the static LDS the staged tables may take: a shared
    table between 32 and 64 KB is staged only with the larger budget."""
    from tests.fixtures import build_pair
    import numpy as np
    n = 40_000  # 40,000 suppliers x 8-bit entries (region + nation) = 40 KB
    sup = {"s_suppkey": (np.arange(1, n + 1, dtype=np.int32), None),
           "s_region": (np.arange(n, dtype=np.int32) % 5, None), "s_nation": (np.arange(n, dtype=np.int32) % 25, None)}
    rng = np.random.default_rng(3)
    lo = {"lo_suppkey": (rng.integers(1, n + 1, 5000).astype(np.int32), None),
          "lo_revenue": (rng.integers(1, 1000, 5000).astype(np.int32), None)}
    db, _ = build_pair({"LINEORDER": (5000, True, lo), "SUPPLIER": (n, False, sup)}, pks=[("SUPPLIER", "s_suppkey")])
    qs = ["select s_nation, sum(lo_revenue) from lineorder, supplier where lo_suppkey = s_suppkey and s_region = 1 group by s_nation",
          "select count(*) from lineorder, supplier where lo_suppkey = s_suppkey and s_nation = 3"]
    small = db.generate_batch_code(qs)
    large = db.generate_batch_code(qs, {"table_kb": 64})
    assert "__shared__ u32 cq_tab" not in small and "__shared__ u32 cq_tab0[" in large
    pa.compile_source(large)
    with pytest.raises(pa.CrystalError, match="table_kb"):
        db.generate_batch_code(qs, {"table_kb": 200})
    db.close()


@pytest.mark.parametrize("opts", [{"bt": 1024, "vpt": 4}, {"bt": 64, "vpt": 1}])
def test_launch_shapes_compile(hdb, opts):
    pa.compile_source(hdb.generate_batch_code([Q.C3, Q.C1], opts))


@pytest.mark.parametrize("opts,msg", [({"bt": 100}, "multiple of 64"), ({"vpt": 8}, "vpt"), ({"bogus": 1}, "unknown"),
                                      ({"lds_kb": 200}, "lds_kb"), ({"grid_percent": 0}, "grid_percent")])
def test_bad_options_rejected(hdb, opts, msg):
    with pytest.raises(pa.CrystalError, match=msg):
        hdb.generate_batch_code([Q.C1], opts)


# ---------------------------------------------------------------- GPU

def _check_batch(gdb, odb, qs, results):
    for sql, r in zip(qs, results):
        print(f"[batch] sql={sql}", flush=True)
        assert_results_equal(r, gdb.execute(sql), sql)
        assert_same(r, odb.execute(gdb.explain(sql)))


@pytest.mark.gpu
@pytest.mark.parametrize("graph,masks", [(0, 1), (1, 1), (0, 0)])
def test_batch_equals_alone_and_oracle(graph, masks):
    """masks 1: templates whose filters allow it run in the query-mask form
    (all their queries at once); masks 0: every template per query."""
    from tests.test_gpu_parity import dataset
    spec, data, odb, gdb = dataset("sf1")
    qs = mixed_queries()
    print(f"[batch] {len(qs)} queries graph={graph} masks={masks}", flush=True)
    res = gdb.batch_execute(qs, {"graph": graph, "masks": masks})
    rep = gdb.batch_report()
    assert rep["queries"] == len(qs) and rep["shared_scans"] >= 2
    _check_batch(gdb, odb, qs, res)


@pytest.mark.gpu
@pytest.mark.parametrize("masks", [0, 1])
def test_templates_with_many_queries(masks):
    """80 queries of each of the four templates: mask-form templates take at
    most 64 queries per launch, the rest go to another launch."""
    import random
    from tests.test_gpu_parity import dataset
    spec, data, odb, gdb = dataset("sf1")
    rng = random.Random(31)
    qs = [Q.template_query(k, rng) for k in (1, 2, 3, 4) for _ in range(80)]
    print(f"[batch] {len(qs)} template queries masks={masks}", flush=True)
    res = gdb.batch_execute(qs, {"graph": 0, "masks": masks})
    rep = gdb.batch_report()
    assert rep["queries"] == len(qs)
    for sql, r in zip(qs, res):
        assert_same(r, odb.execute(gdb.explain(sql)))


@pytest.mark.gpu
def test_deferred_probes_with_few_queries():
    """Three queries of each template, one launch per template: most items
    are rejected before the CUSTOMER and PART probes, which then skip them;
    results equal the oracle and each query alone."""
    import random
    from tests.test_gpu_parity import dataset
    spec, data, odb, gdb = dataset("sf1")
    rng = random.Random(37)
    qs = [Q.template_query(k, rng) for k in (1, 2, 3, 4) for _ in range(3)]
    res = gdb.batch_execute(qs, {"graph": 0, "masks": 1, "max_templates": 1})
    assert gdb.batch_report()["shared_scans"] == 4
    for sql, r in zip(qs, res):
        print(f"[batch] deferred: {sql}", flush=True)
        o = odb.execute(gdb.explain(sql))
        assert_same(r, o)
        assert_same(gdb.execute(sql), o)


@pytest.mark.gpu
@pytest.mark.parametrize("lds_kb", [16, 160])
def test_lds_region_sizes(lds_kb):
    """16 KiB: hash tables stay at 128 slots and C2-shaped queries (up to 280
    groups) overflow to global atomics; 160 KiB: tables grow to 1024 slots."""
    import random
    from tests.test_gpu_parity import dataset
    spec, data, odb, gdb = dataset("sf1")
    rng = random.Random(11)
    qs = [Q.template_query(k, rng) for k in (3, 3, 3, 3, 3, 3, 4, 4, 4, 1, 1, 2, 2)]
    print(f"[batch] lds_kb={lds_kb}", flush=True)
    res = gdb.batch_execute(qs, {"graph": 0, "lds_kb": lds_kb})
    _check_batch(gdb, odb, qs, res)


@pytest.mark.gpu
def test_graph_replay_with_new_constants():
    from tests.test_gpu_parity import dataset
    spec, data, odb, gdb = dataset("sf20")
    first = [Q.c2(c, r) for c, r in [(1, 0), (2, 1), (3, 2), (4, 3)]] + [Q.C1]
    second = [Q.c2(c, r) for c, r in [(20, 4), (21, 3), (22, 2), (23, 1)]] + [Q.C1.replace("1993", "1995")]
    # a batch shape runs eagerly the first time, is captured the second time
    gdb.batch_execute(first, {"graph": 1})
    assert gdb.batch_report()["scans"][0]["graph"] == "eager"
    gdb.batch_execute(first, {"graph": 1})
    assert gdb.batch_report()["scans"][0]["graph"] == "captured"
    for qs in (second, first, second):
        res = gdb.batch_execute(qs, {"graph": 1})
        assert gdb.batch_report()["scans"][0]["graph_replayed"] == 1
        _check_batch(gdb, odb, qs, res)


@pytest.mark.gpu
def test_autotuned_batch_matches():
    from tests.test_gpu_parity import dataset
    spec, data, odb, gdb = dataset("sf100")
    qs = [Q.C3, Q.C3.replace("1000000", "5000000"), Q.C1]
    res = gdb.batch_execute(qs, {"autotune": 1, "graph": 0})
    scan = gdb.batch_report()["scans"][0]
    assert scan["block_threads"] in (64, 128, 256, 512, 1024) and scan["vectors_per_thread"] in (1, 2, 4)
    assert scan["table_mask"] | 7 == 0xFFFFFFFF  # the autotuner places the first three candidates
    _check_batch(gdb, odb, qs, res)
    # every placement and table budget gives the same results
    for opts in ({"table_mask": 0}, {"table_mask": 1}, {"table_mask": 2}, {"table_kb": 64}, {"table_kb": 96}):
        print(f"[batch] {opts}", flush=True)
        _check_batch(gdb, odb, qs, gdb.batch_execute(qs, dict(opts, graph=0)))


def _key_zero_tables(first_key=1):
    """SUPPLIER keyed first_key..first_key + 15 and a fact table whose
    supplier keys include 0 and keys beyond the dimension.  Keys 1..16:
    reference slot (k - 0) % 16, key 16 shares slot 0 with a probe key 0.
    Keys 0..15: key 0 is a supplier itself."""
    import numpy as np
    rng = np.random.default_rng(7)
    n = 4096
    sup = {"s_suppkey": (np.arange(first_key, first_key + 16, dtype=np.int32), None),
           "s_region": (np.array([i % 3 for i in range(16)], np.int32), None),
           "s_nation": (np.array([i % 5 for i in range(16)], np.int32), None)}
    lo = {"lo_suppkey": (rng.integers(0, 19, n).astype(np.int32), None),
          "lo_revenue": (rng.integers(1, 1000, n).astype(np.int32), None)}
    return {"LINEORDER": (n, True, lo), "SUPPLIER": (16, False, sup)}


KEY_ZERO_QUERIES = [
    # supplier 16 has region 0: filtered out for region 1 (key 0 then matches
    # the empty slot) and kept for region 0 (key 0 does not match)
    "select count(*), sum(lo_revenue) from lineorder, supplier where lo_suppkey = s_suppkey and s_region = 1",
    "select count(*), sum(lo_revenue) from lineorder, supplier where lo_suppkey = s_suppkey and s_region = 0",
    "select s_nation, count(*) from lineorder, supplier where lo_suppkey = s_suppkey and s_region = 2 group by s_nation",
    "select s_nation, sum(lo_revenue) from lineorder, supplier where lo_suppkey = s_suppkey group by s_nation",
]


@pytest.mark.parametrize("first_key", [1, 0])
def test_launch_probe_marks_key_zero(first_key):
    """Templates that filter SUPPLIER mark probe key 0 (it matches the
    reference's empty slot when the key sharing that slot is filtered out);
    the unfiltered template reads it by the plain lookup.  The launch-level
    probe they share marks key 0 whatever the template order, and where key
    0 is a supplier itself the filtered templates probe apart."""
    from tests.fixtures import build_pair
    db, _ = build_pair(_key_zero_tables(first_key), pks=[("SUPPLIER", "s_suppkey")], device=-1)
    try:
        for qs in (KEY_ZERO_QUERIES, KEY_ZERO_QUERIES[::-1]):
            src = db.generate_batch_code(qs, {"graph": 0})
            probes = [ln for ln in src.splitlines() if "// launch-level probe SUPPLIER" in ln]
            assert len(probes) == 1, probes
            assert "if (fk == 0) { gz0[i] = true;" in src
            users = probes[0].split("(templates")[1].rstrip(")").split()
            assert len(users) == (3 if first_key == 1 else 2), probes
            pa.compile_source(src)
    finally:
        db.close()


@pytest.mark.gpu
@pytest.mark.parametrize("first_key", [1, 0])
def test_probe_key_zero_against_reference_empty_slot(first_key):
    from tests.fixtures import build_pair
    db, odb = build_pair(_key_zero_tables(first_key), pks=[("SUPPLIER", "s_suppkey")], device=0)
    try:
        res = db.batch_execute(KEY_ZERO_QUERIES, {"graph": 0})
        for sql, r in zip(KEY_ZERO_QUERIES, res):
            print(f"[key0] first_key={first_key} sql={sql}", flush=True)
            o = odb.execute(db.explain(sql))
            assert_same(db.execute(sql), o)
            assert_same(r, o)
    finally:
        db.close()


def test_templates_over_different_columns_stay_distinct(hdb):
    """Plans whose generated text is identical but which bind different fact
    columns (avg of lo_quantity and of lo_tax) are different templates: a
    launch binds one column set per template."""
    qs = ["select avg(lo_quantity) from lineorder", "select avg(lo_tax) from lineorder"]
    assert hdb.generate_batch_code(qs).count("        // template ") == 2


@pytest.mark.gpu
def test_batch_with_same_text_different_columns():
    from tests.test_gpu_parity import dataset
    spec, data, odb, gdb = dataset("sf1")
    qs = ["select avg(lo_quantity) from lineorder", "select avg(lo_tax) from lineorder",
          "select sum(lo_quantity) from lineorder where lo_discount < 3", "select sum(lo_tax) from lineorder where lo_discount < 3"]
    for opts in ({"graph": 0}, {"graph": 0, "max_templates": 1}):
        _check_batch(gdb, odb, qs, gdb.batch_execute(qs, opts))
