"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product never imports it.
"""
from __future__ import annotations

import ctypes
import pathlib
import subprocess
from typing import Dict, List, Optional

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"


def build() -> None:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


class _GenSpec(ctypes.Structure):
    _fields_ = [("sf", ctypes.c_int32), ("fact_rows", ctypes.c_int64), ("seed", ctypes.c_uint64),
                ("nullable", ctypes.c_char_p)]


class _Result(ctypes.Structure):
    _fields_ = [("ncols", ctypes.c_int), ("nrows", ctypes.c_int64), ("ordered", ctypes.c_int),
                ("types", ctypes.c_int * 64), ("nullable", ctypes.c_int * 64),
                ("values", ctypes.POINTER(ctypes.c_uint64)), ("valid", ctypes.POINTER(ctypes.c_uint8))]


def _load():
    if not LIB.exists():
        build()
    lib = ctypes.CDLL(str(LIB))
    P = ctypes.POINTER
    lib.orc_table_rows.restype = ctypes.c_int64
    lib.orc_table_rows.argtypes = [P(_GenSpec), ctypes.c_char_p]
    lib.orc_table_columns.restype = ctypes.c_int
    lib.orc_table_columns.argtypes = [ctypes.c_char_p, P(ctypes.c_char_p), ctypes.c_int]
    lib.orc_column_null_bp.restype = ctypes.c_int
    lib.orc_column_null_bp.argtypes = [P(_GenSpec), ctypes.c_char_p, ctypes.c_char_p]
    lib.orc_gen_column.restype = ctypes.c_int
    lib.orc_gen_column.argtypes = [P(_GenSpec), ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int64, ctypes.c_int64,
                                   ctypes.c_void_p, ctypes.c_void_p]
    lib.orc_datekey.restype = ctypes.c_int32
    lib.orc_datekey.argtypes = [ctypes.c_int64]
    lib.orc_db_create.restype = ctypes.c_void_p
    lib.orc_db_destroy.argtypes = [ctypes.c_void_p]
    lib.orc_db_add_table.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int64, ctypes.c_int]
    lib.orc_db_add_column.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_void_p, ctypes.c_void_p]
    lib.orc_db_set_primary_key.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_char_p]
    lib.orc_db_finalize.argtypes = [ctypes.c_void_p]
    lib.orc_db_column_stats.argtypes = [ctypes.c_void_p, ctypes.c_char_p, P(ctypes.c_int64), P(ctypes.c_int64),
                                        P(ctypes.c_uint64), P(ctypes.c_int), P(ctypes.c_int)]
    lib.orc_execute.restype = ctypes.c_int
    lib.orc_execute.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int, P(_Result), ctypes.c_char_p,
                                ctypes.c_size_t]
    lib.orc_result_free.argtypes = [P(_Result)]
    return lib


_lib = _load()

SSB_TABLES = ["LINEORDER", "PART", "SUPPLIER", "CUSTOMER", "DDATE"]
SSB_PK = {"PART": "p_partkey", "SUPPLIER": "s_suppkey", "CUSTOMER": "c_custkey", "DDATE": "d_datekey"}


class GenSpec:
    def __init__(self, sf: int = 1, fact_rows: int = 0, seed: int = 42, nullable: str = ""):
        self.sf, self.fact_rows, self.seed, self.nullable = sf, fact_rows, seed, nullable
        self._c = _GenSpec(sf, fact_rows, seed, nullable.encode())

    def rows(self, table: str) -> int:
        return _lib.orc_table_rows(ctypes.byref(self._c), table.encode())

    def null_bp(self, table: str, column: str) -> int:
        return _lib.orc_column_null_bp(ctypes.byref(self._c), table.encode(), column.encode())


def table_columns(table: str) -> List[str]:
    names = (ctypes.c_char_p * 64)()
    n = _lib.orc_table_columns(table.encode(), names, 64)
    return [names[i].decode() for i in range(n)]


def datekey(day: int) -> int:
    return _lib.orc_datekey(day)


def gen_column(spec: GenSpec, table: str, column: str, start: int = 0, count: Optional[int] = None):
    """(values int32[count], validity uint64 words or None)."""
    if count is None:
        count = spec.rows(table) - start
    out = np.empty(count, np.int32)
    valid = None
    if spec.null_bp(table, column) > 0:
        valid = np.empty((count + 63) // 64, np.uint64)
    rc = _lib.orc_gen_column(ctypes.byref(spec._c), table.encode(), column.encode(), start, count, out.ctypes.data,
                             valid.ctypes.data if valid is not None else None)
    if rc != 0:
        raise ValueError(f"unknown column {table}.{column}")
    return out, valid


def gen_tables(spec: GenSpec, tables=None, columns: Optional[Dict[str, List[str]]] = None):
    """{table: {column: (values, validity)}} for the synthetic schema."""
    out = {}
    for t in tables or SSB_TABLES:
        cols = (columns or {}).get(t) or table_columns(t)
        out[t] = {c: gen_column(spec, t, c) for c in cols}
    return out


class OracleResult:
    def __init__(self, types, nullable, ordered, values, valid):
        self.types, self.nullable, self.ordered = types, nullable, ordered
        self.values, self.valid = values, valid  # [rows, cols] uint64 / bool

    @property
    def rows(self) -> int:
        return self.values.shape[0]

    def tuples(self, null=None):
        dt = {0: np.int64, 1: np.uint64, 2: np.float64}
        cols = [self.values[:, c].view(dt[t]) for c, t in enumerate(self.types)]
        return [tuple(cols[c][r].item() if self.valid[r, c] else null for c in range(len(cols)))
                for r in range(self.rows)]


class OracleDB:
    """Oracle catalog over borrowed numpy columns."""

    def __init__(self):
        self._h = _lib.orc_db_create()
        self._keep = []

    def __del__(self):
        if getattr(self, "_h", None):
            _lib.orc_db_destroy(self._h)
            self._h = None

    def add_table(self, name: str, rows: int, is_fact: bool) -> None:
        assert _lib.orc_db_add_table(self._h, name.encode(), rows, int(is_fact)) == 0

    def add_column(self, table: str, column: str, data: np.ndarray, validity: Optional[np.ndarray] = None) -> None:
        data = np.ascontiguousarray(data, np.int32)
        if validity is not None:
            validity = np.ascontiguousarray(validity, np.uint64)
        self._keep.append((data, validity))
        assert _lib.orc_db_add_column(self._h, table.encode(), column.encode(), data.ctypes.data,
                                      validity.ctypes.data if validity is not None else None) == 0

    def set_primary_key(self, table: str, column: str) -> None:
        assert _lib.orc_db_set_primary_key(self._h, table.encode(), column.encode()) == 0

    def finalize(self) -> None:
        assert _lib.orc_db_finalize(self._h) == 0

    def column_stats(self, column: str) -> dict:
        mn, mx, card = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_uint64()
        nul, uq = ctypes.c_int(), ctypes.c_int()
        assert _lib.orc_db_column_stats(self._h, column.encode(), ctypes.byref(mn), ctypes.byref(mx),
                                        ctypes.byref(card), ctypes.byref(nul), ctypes.byref(uq)) == 0
        return {"min": mn.value, "max": mx.value, "card": card.value, "nullable": bool(nul.value),
                "unique": bool(uq.value)}

    def execute(self, plan: str, threads: int = 0) -> OracleResult:
        r = _Result()
        err = ctypes.create_string_buffer(1024)
        rc = _lib.orc_execute(self._h, plan.encode(), threads, ctypes.byref(r), err, 1024)
        if rc != 0:
            raise RuntimeError(err.value.decode())
        try:
            n, c = r.nrows, r.ncols
            if n:
                vals = np.ctypeslib.as_array(r.values, shape=(n * c,)).copy().reshape(n, c)
                vld = np.ctypeslib.as_array(r.valid, shape=(n * c,)).copy().reshape(n, c).astype(bool)
            else:
                vals = np.zeros((0, c), np.uint64)
                vld = np.zeros((0, c), bool)
            return OracleResult([r.types[i] for i in range(c)], [bool(r.nullable[i]) for i in range(c)],
                                bool(r.ordered), vals, vld)
        finally:
            _lib.orc_result_free(ctypes.byref(r))


def load_ssb(tables_data, spec: GenSpec) -> OracleDB:
    db = OracleDB()
    for t, cols in tables_data.items():
        db.add_table(t, spec.rows(t), t == "LINEORDER")
        for c, (vals, valid) in cols.items():
            db.add_column(t, c, vals, valid)
        if t in SSB_PK:
            db.set_primary_key(t, SSB_PK[t])
    db.finalize()
    return db
