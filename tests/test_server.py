"""Continuous-batching server: every result is the same alone and in the mix.

Requests are submitted while earlier ones are on the GPU (several lanes,
shared scans, background compilation of new plan signatures); each result is
compared with the same query run alone through the plain path and with the
oracle.
"""
import threading

import pytest

from tests import queries as Q
from tests.compare import assert_results_equal, assert_same

pa = pytest.importorskip("practice_amd")


def test_bad_server_options_rejected_without_gpu():
    db = pa.Database(-1)
    with pytest.raises(pa.CrystalError):
        pa.Server(db, {"lanes": 2})  # host-only database: no server


@pytest.mark.gpu
def test_bad_per_template_mode_rejected():
    from tests.test_gpu_parity import dataset
    spec, data, odb, gdb = dataset("sf1")
    with pytest.raises(pa.CrystalError, match="per_template"):
        pa.Server(gdb, {"per_template": 3})


@pytest.mark.gpu
def test_mixed_requests_match_alone_and_oracle():
    from tests.test_batch import mixed_queries
    from tests.test_gpu_parity import dataset
    spec, data, odb, gdb = dataset("sf1")
    qs = mixed_queries() * 2
    srv = pa.Server(gdb, {"lanes": 2, "max_batch": 16})
    try:
        tickets = []

        def client(part):
            for q in part:
                tickets.append((q, srv.submit(q)))

        threads = [threading.Thread(target=client, args=(qs[i::3],)) for i in range(3)]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        results = [(q, srv.wait(t, 120_000)) for q, t in tickets]
        st = srv.stats()
    finally:
        srv.stop()
    assert st["completed"] == len(qs) and st["rejected"] == 0
    assert st["batches"] < len(qs)  # requests shared scans
    for q, r in results:
        print(f"[server] sql={q}", flush=True)
        assert_results_equal(r, gdb.execute(q), q)
        assert_same(r, odb.execute(gdb.explain(q)))


@pytest.mark.gpu
def test_admission_waits_for_memory_and_rejects_oversized():
    from tests.test_gpu_parity import dataset
    spec, data, odb, gdb = dataset("sf20")
    # each C2-shaped request reserves a few MB (PART / SUPPLIER tables):
    # with a 6 MB budget only one or two fit at a time
    fp = gdb.describe(Q.C2)
    srv = pa.Server(gdb, {"lanes": 2, "budget_mb": 6})
    try:
        qs = [Q.c2(c, r) for c in range(0, 25, 3) for r in range(5)][:24]
        tickets = [srv.submit(q) for q in qs]
        res = [srv.wait(t, 120_000) for t in tickets]
        # ~1M groups: over 16 MB of aggregate and output buffers
        big = srv.submit("select lo_partkey, sum(lo_revenue) from lineorder group by lo_partkey")
        with pytest.raises(pa.CrystalError, match="budget"):
            srv.wait(big, 120_000)
        st = srv.stats()
    finally:
        srv.stop()
    assert fp["groups"] > 0
    assert st["admission_waits"] > 0 and st["in_flight_bytes"] == 0
    for q, r in zip(qs, res):
        assert_results_equal(r, gdb.execute(q), q)


@pytest.mark.gpu
def test_unsupported_request_fails_alone():
    from tests.test_gpu_parity import dataset
    spec, data, odb, gdb = dataset("sf1")
    srv = pa.Server(gdb, {"lanes": 1})
    try:
        bad = srv.submit("select lo_revenue from lineorder, ddate where lo_orderdate = d_datekey "
                         "or lo_commitdate = d_datekey")
        good = srv.submit(Q.C1)
        with pytest.raises(pa.CrystalError, match="OR-key joins in a projection"):
            srv.wait(bad, 60_000)
        assert_results_equal(srv.wait(good, 60_000), gdb.execute(Q.C1), Q.C1)
    finally:
        srv.stop()


@pytest.mark.gpu
def test_or_key_requests_run_on_cold_lanes():
    """Plans the shared scan does not take (OR-key joins) run compiled on the
    cold lanes, beside batched requests, with the same results."""
    from tests.test_gpu_parity import OR_KEY_SSB, dataset
    spec, data, odb, gdb = dataset("sf1c")
    srv = pa.Server(gdb, {"lanes": 2, "max_batch": 16})
    try:
        qs = [OR_KEY_SSB[0], Q.C1, OR_KEY_SSB[2], Q.C1, OR_KEY_SSB[0]]
        tickets = [(q, srv.submit(q)) for q in qs]
        results = [(q, srv.wait(t, 120_000)) for q, t in tickets]
        st = srv.stats()
    finally:
        srv.stop()
    assert st["completed"] == len(qs) and st["rejected"] == 0
    for q, r in results:
        print(f"[server] sql={q}", flush=True)
        assert_results_equal(r, gdb.execute(q), q)
        assert_same(r, odb.execute(gdb.explain(q)))


@pytest.mark.gpu
def test_projection_requests_run_on_cold_lanes():
    from tests.test_gpu_parity import dataset
    spec, data, odb, gdb = dataset("sf1")
    q = "select lo_revenue, lo_discount from lineorder where lo_quantity = 7 and lo_discount > 8"
    srv = pa.Server(gdb, {"lanes": 1})
    try:
        r = srv.wait(srv.submit(q), 120_000)
    finally:
        srv.stop()
    assert_same(r, odb.execute(gdb.explain(q)))


def _graph_shapes(rng):
    """Two batch shapes of fresh constants: 1- and 2-dimension templates."""
    return [[Q.template_query(1, rng), Q.template_query(1, rng)],
            [Q.template_query(2, rng), Q.template_query(2, rng), Q.template_query(1, rng)]]


@pytest.mark.gpu
def test_batch_graph_replay_survives_module_eviction():
    """A captured batch keeps the code objects its kernel nodes run: with a
    one-module cache every batch evicts the other shape's modules, and the
    replays (fresh constants each round) still match the oracle."""
    import random
    from tests.test_gpu_parity import dataset
    spec, data, odb, gdb = dataset("sf1")
    rng = random.Random(21)
    modes = []
    gdb.set_option("module_cache_capacity", 1)
    try:
        for rnd in range(5):
            for qs in _graph_shapes(rng):
                print(f"[graph] round={rnd} queries={len(qs)}", flush=True)
                res = gdb.batch_execute(qs, {"graph": 1})
                modes += [s["graph"] for s in gdb.batch_report()["scans"]]
                for q, r in zip(qs, res):
                    assert_same(r, odb.execute(gdb.explain(q)))
        evictions = gdb.cache_stat("evictions")
    finally:
        gdb.set_option("module_cache_capacity", 256)
    assert evictions > 0
    assert "replayed" in modes and "capture-failed" not in modes, modes


@pytest.mark.gpu
def test_server_graph_launches_with_changing_batches():
    """graph=1 on the server: batches of changing composition on two lanes,
    cold requests (interpreted, allocating their scratch) on the cold lanes
    meanwhile; every result matches the oracle and no capture fails."""
    import random
    from tests.test_gpu_parity import dataset
    spec, data, odb, gdb = dataset("sf1")
    rng = random.Random(33)
    srv = pa.Server(gdb, {"lanes": 2, "max_batch": 8, "graph": 1, "cold_lanes": 2})
    try:
        done = []
        for wave in range(6):
            qs = [Q.template_query(rng.randint(1, 4), rng) for _ in range(rng.randint(3, 12))]
            qs += [Q.adhoc_query(rng) for _ in range(2)]
            tickets = [(q, srv.submit(q)) for q in qs]
            done += [(q, srv.wait(t, 120_000)) for q, t in tickets]
        st = srv.stats()
    finally:
        srv.stop()
    assert st["rejected"] == 0 and st["graph_failures"] == 0, st
    for q, r in done:
        print(f"[server-graph] sql={q}", flush=True)
        assert_same(r, odb.execute(gdb.explain(q)))


@pytest.mark.gpu
def test_cold_interpreted_requests_run_together():
    """Ad-hoc requests queued together on a cold lane run as one group
    (their linear forms share scans); every result matches the oracle and
    the same statement run alone."""
    import random
    from tests.test_gpu_parity import dataset
    spec, data, odb, gdb = dataset("sf1c")
    rng = random.Random(41)
    qs = [Q.adhoc_query(rng) for _ in range(40)] + [Q.template_query(k, rng) for k in (1, 2, 3, 4)]
    srv = pa.Server(gdb, {"lanes": 1, "cold_lanes": 1, "cold_group": 32})
    try:
        tickets = [(q, srv.submit(q)) for q in qs]
        done = [(q, srv.wait(t, 120_000)) for q, t in tickets]
        st = srv.stats()
    finally:
        srv.stop()
    assert st["rejected"] == 0 and st["cold_groups"] >= 1 and st["cold_grouped"] >= 2, st
    for q, r in done:
        print(f"[server-cold] sql={q}", flush=True)
        assert_same(r, odb.execute(gdb.explain(q)))


@pytest.mark.gpu
def test_adaptive_launch_form_matches_oracle():
    """per_template=2: a batch runs as one launch of the hot-set module while
    fewer than cold_backlog cold requests are queued, one launch per
    signature otherwise; waves with and without ad-hoc requests exercise
    both forms, and every result matches the oracle."""
    import random
    import time
    from tests.test_gpu_parity import dataset
    spec, data, odb, gdb = dataset("sf1")
    rng = random.Random(57)
    srv = pa.Server(gdb, {"lanes": 1, "max_batch": 32, "cold_lanes": 1, "per_template": 2})
    try:
        done = []
        # promote the four templates (each seen twice), then let the
        # hot-set module compile
        warm = [Q.template_query(k, rng) for k in (1, 2, 3, 4) for _ in range(2)]
        done += [(q, srv.wait(t, 120_000)) for q, t in [(q, srv.submit(q)) for q in warm]]
        for wave in range(12):
            if wave < 6:
                time.sleep(0.5)
            qs = [Q.template_query(rng.randint(1, 4), rng) for _ in range(rng.randint(4, 16))]
            if wave % 2:
                qs += [Q.adhoc_query(rng) for _ in range(2)]
            tickets = [(q, srv.submit(q)) for q in qs]
            done += [(q, srv.wait(t, 120_000)) for q, t in tickets]
        st = srv.stats()
    finally:
        srv.stop()
    assert st["rejected"] == 0 and st["combined_batches"] >= 1, st
    assert st["combined_batches"] < st["batches"], st
    for q, r in done:
        print(f"[server-adaptive] sql={q}", flush=True)
        assert_same(r, odb.execute(gdb.explain(q)))
