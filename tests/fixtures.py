"""Helpers that load the same host columns into a host-only product catalog
(planning / code generation) and into the CPU oracle."""
import csv
import pathlib

import numpy as np

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"


def bitmap_from_bools(valid):
    valid = np.asarray(valid, bool)
    words = np.zeros((len(valid) + 63) // 64, np.uint64)
    for i in np.nonzero(valid)[0]:
        words[i >> 6] |= np.uint64(1) << np.uint64(i & 63)
    return words


def build_pair(tables, pks=(), device=-1):
    """tables: {name: (rows, is_fact, {col: (int32 values, uint64 validity words or None)})}.
    Returns (practice_amd.Database, oracle.OracleDB) holding the same data."""
    import practice_amd as pa
    from oracle import oracle as orc

    db = pa.Database(device)
    odb = orc.OracleDB()
    for name, (rows, is_fact, cols) in tables.items():
        db.create_table(name, rows, is_fact)
        odb.add_table(name, rows, is_fact)
        for c, (vals, valid) in cols.items():
            db.load_column(name, c, vals, valid)
            odb.add_column(name, c, vals, valid)
    for t, c in pks:
        db.set_primary_key(t, c)
        odb.set_primary_key(t, c)
    db.finalize()
    odb.finalize()
    return db, odb


def nullable_numbers_tables():
    """The reference's nullable mini dataset (test/nullable_data) as columns;
    empty CSV fields are NULL with raw value 0."""
    rows = list(csv.DictReader(open(GOLDEN / "nullable_data" / "nullable_numbers.csv")))
    cols = {}
    for name in ("id", "grp", "val", "flag"):
        raw = [r[name] for r in rows]
        vals = np.array([int(x) if x != "" else 0 for x in raw], np.int32)
        valid = [x != "" for x in raw]
        cols[name] = (vals, None if all(valid) else bitmap_from_bools(valid))
    return {"NULLABLE_NUMBERS": (len(rows), True, cols)}


def nullable_expected():
    """{query: [header, rows...]} from the reference's expected.tsv."""
    out = {}
    for line in open(GOLDEN / "nullable_data" / "expected.tsv"):
        parts = line.rstrip("\n").split("\t")
        out.setdefault(parts[0], []).append(parts[1:])
    return out


def nullable_queries():
    """{q1..q5: sql} from the reference's queries.sql."""
    text = open(GOLDEN / "nullable_data" / "queries.sql").read()
    out, name, buf = {}, None, []
    for line in text.splitlines():
        if line.startswith("-- q"):
            if name:
                out[name] = " ".join(buf)
            name, buf = line[3:].strip(), []
        elif line.strip():
            buf.append(line.strip())
    if name:
        out[name] = " ".join(buf)
    return out
