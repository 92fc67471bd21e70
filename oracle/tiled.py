"""ctypes wrapper of the tiled CPU restatement (oracle/libtiled.so) of the
kernels the reference generates for the three plain benchmark queries.

TEST INFRASTRUCTURE: imported only by tests/ and the cpu_baseline leg of
bench.py.  The product never imports it.
"""
from __future__ import annotations

import ctypes
import pathlib
import subprocess
from typing import Dict, List, Tuple

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIB = HERE / "libtiled.so"

if not LIB.exists():
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
_lib = ctypes.CDLL(str(LIB))
_lib.tl_max_threads.restype = ctypes.c_int
_P = ctypes.c_void_p
_I32, _I64 = ctypes.c_int32, ctypes.c_int64
_lib.tl_c1.argtypes = [_P] * 4 + [_I64, _P, _P, _I64] + [_I32] * 4 + [ctypes.c_int, _P]
_lib.tl_c2.argtypes = ([_P] * 4 + [_I64] + [_P] * 3 + [_I64] + [_P] * 2 + [_I64] + [_P] * 2 + [_I64]
                       + [_I32, _I32, _I32, _I64, _I32, _I64, ctypes.c_int, _P, _P])
_lib.tl_c3.argtypes = ([_P] * 5 + [_I64] + [_P] * 2 + [_I64] + [_P] * 2 + [_I64] + [_P] * 2 + [_I64]
                       + [_I32, _I32, _I64, _I32, _I64, ctypes.c_int, _I32, _I64, ctypes.c_int, _P, _P])

# the constants of the three benchmark queries (tests/queries.py C1, C2, C3)
C1_CONST = {"year": 1993, "dlo": 1, "dhi": 3, "qlt": 25}
C2_CONST = {"cat": 12, "reg": 1}
C3_CONST = {"rev_gt": 1_000_000}


def max_threads() -> int:
    return _lib.tl_max_threads()


def _p(a):
    return a.ctypes.data if a is not None else None


def _col(data, table, column):
    return data[table][column][0]


def _bits(words: np.ndarray, n: int) -> np.ndarray:
    return np.unpackbits(words.view(np.uint8), bitorder="little")[:n].astype(bool)


def _rows_of(state: np.ndarray, valid: np.ndarray, tuple_size: int, ngroups: int, sum_valid=None) -> List[Tuple]:
    hit = _bits(valid, ngroups)
    st = state.reshape(ngroups, tuple_size)[hit]
    sv = _bits(sum_valid, ngroups)[hit] if sum_valid is not None else np.ones(st.shape[0], bool)
    keys = st[:, :-1].view(np.int64)
    return [tuple(int(k) for k in keys[r]) + (int(st[r, -1]) if sv[r] else None,) for r in range(st.shape[0])]


def run(config: str, data: Dict, stats, threads: int = 0):
    """Run one configuration over generated tables `data` ({table: {column:
    (values, validity)}}).  `stats(column)` gives the catalog statistics
    ({min, card, nullable}).  Returns the scalar (C1) or the group rows sorted
    (C2, C3) as tuples of Python ints."""
    threads = threads or max_threads()
    lo = data["LINEORDER"]
    n = len(lo["lo_orderdate"][0])
    dd = data["DDATE"]
    if config == "C1":
        out = np.zeros(1, np.uint64)
        c = C1_CONST
        rc = _lib.tl_c1(_p(lo["lo_orderdate"][0]), _p(lo["lo_discount"][0]), _p(lo["lo_quantity"][0]),
                        _p(lo["lo_extendedprice"][0]), n, _p(dd["d_datekey"][0]), _p(dd["d_year"][0]),
                        len(dd["d_datekey"][0]), c["year"], c["dlo"], c["dhi"], c["qlt"], threads, _p(out))
        assert rc == 0
        return int(out[0])
    if config == "C2":
        pt, su = data["PART"], data["SUPPLIER"]
        yr, br = stats("d_year"), stats("p_brand1")
        g = yr["card"] * br["card"]
        state = np.zeros(g * 3, np.uint64)
        valid = np.zeros((g + 63) // 64, np.uint64)
        c = C2_CONST
        rc = _lib.tl_c2(_p(lo["lo_orderdate"][0]), _p(lo["lo_partkey"][0]), _p(lo["lo_suppkey"][0]),
                        _p(lo["lo_revenue"][0]), n, _p(pt["p_partkey"][0]), _p(pt["p_category"][0]),
                        _p(pt["p_brand1"][0]), len(pt["p_partkey"][0]), _p(su["s_suppkey"][0]),
                        _p(su["s_region"][0]), len(su["s_suppkey"][0]), _p(dd["d_datekey"][0]), _p(dd["d_year"][0]),
                        len(dd["d_datekey"][0]), c["cat"], c["reg"], yr["min"], yr["card"], br["min"], br["card"],
                        threads, _p(state), _p(valid))
        assert rc == 0
        return sorted(_rows_of(state, valid, 3, g))
    if config == "C3":
        cu, su = data["CUSTOMER"], data["SUPPLIER"]
        cn, sn, yr = stats("c_nation"), stats("s_nation"), stats("d_year")
        g = cn["card"] * (sn["card"] + (1 if sn["nullable"] else 0)) * yr["card"]
        words = (g + 63) // 64
        state = np.zeros(g * 4, np.uint64)
        valid = np.zeros(2 * words, np.uint64)
        rc = _lib.tl_c3(_p(lo["lo_custkey"][0]), _p(lo["lo_suppkey"][0]), _p(lo["lo_orderdate"][0]),
                        _p(lo["lo_revenue"][0]), _p(lo["lo_revenue"][1]), n, _p(cu["c_custkey"][0]),
                        _p(cu["c_nation"][0]), len(cu["c_custkey"][0]), _p(su["s_suppkey"][0]), _p(su["s_nation"][0]),
                        len(su["s_suppkey"][0]), _p(dd["d_datekey"][0]), _p(dd["d_year"][0]), len(dd["d_datekey"][0]),
                        C3_CONST["rev_gt"], cn["min"], cn["card"], sn["min"], sn["card"], int(sn["nullable"]),
                        yr["min"], yr["card"], threads, _p(state), _p(valid))
        assert rc == 0
        return sorted(_rows_of(state, valid[:words], 4, g, valid[words:]))
    raise ValueError(config)


# fact / dimension columns each configuration reads
COLUMNS = {
    "C1": {"LINEORDER": ["lo_orderdate", "lo_discount", "lo_quantity", "lo_extendedprice"],
           "DDATE": ["d_datekey", "d_year"]},
    "C2": {"LINEORDER": ["lo_orderdate", "lo_partkey", "lo_suppkey", "lo_revenue"],
           "PART": ["p_partkey", "p_category", "p_brand1"], "SUPPLIER": ["s_suppkey", "s_region"],
           "DDATE": ["d_datekey", "d_year"]},
    "C3": {"LINEORDER": ["lo_custkey", "lo_suppkey", "lo_orderdate", "lo_revenue"],
           "CUSTOMER": ["c_custkey", "c_nation"], "SUPPLIER": ["s_suppkey", "s_nation"],
           "DDATE": ["d_datekey", "d_year"]},
}
