"""The tiled CPU restatement timed as the CPU baseline (oracle/ssb_tiled.c)
returns what the plan interpreter returns for the three benchmark queries."""
import pytest

from oracle import oracle as orc
from oracle import tiled
from tests import queries as Q


@pytest.fixture(scope="module")
def data_and_db():
    spec = orc.GenSpec(sf=1, fact_rows=400_003, seed=11, nullable=Q.C3_NULLABLE)
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
    yield data, odb, db
    db.close()


@pytest.mark.parametrize("threads", [1, 4])
def test_c1_matches_oracle(data_and_db, threads):
    data, odb, db = data_and_db
    want = odb.execute(db.explain(Q.C1)).tuples()
    assert [(tiled.run("C1", data, odb.column_stats, threads),)] == want


@pytest.mark.parametrize("threads", [1, 4])
def test_c2_matches_oracle(data_and_db, threads):
    data, odb, db = data_and_db
    want = odb.execute(db.explain(Q.C2)).tuples()
    got = tiled.run("C2", data, odb.column_stats, threads)
    assert len(got) > 50
    assert got == want


@pytest.mark.parametrize("threads", [1, 4])
def test_c3_matches_oracle(data_and_db, threads):
    """Nullable revenue (3VL filter, SUM skips NULL) and the catalog-nullable
    s_nation payload (+1 term, card + 1, NULL grouped under raw 0)."""
    data, odb, db = data_and_db
    assert odb.column_stats("s_nation")["nullable"]
    want = sorted(odb.execute(db.explain(Q.C3)).tuples(), key=lambda t: tuple((x is None, x) for x in t))
    got = tiled.run("C3", data, odb.column_stats, threads)
    assert len(got) == 25 * 25 * 7
    assert got == want
