"""Result comparison helpers shared by the parity tests."""
import numpy as np


def normalize(tuples, ordered):
    return list(tuples) if ordered else sorted(tuples, key=lambda t: tuple((x is None, x) for x in t))


def assert_same(gpu, ora, ordered=None):
    """Bit-exact comparison of a GPU Result and an OracleResult (NULL cells compared as NULL)."""
    assert gpu.types == ora.types, (gpu.types, ora.types)
    assert gpu.nullable == ora.nullable, (gpu.nullable, ora.nullable)
    if ordered is None:
        ordered = ora.ordered and gpu.ordered
    g = normalize(gpu.tuples(), ordered)
    o = normalize(ora.tuples(), ordered)
    assert len(g) == len(o), (len(g), len(o))
    for i, (a, b) in enumerate(zip(g, o)):
        assert a == b, f"row {i}: gpu {a} != oracle {b}"


def assert_results_equal(a, b, what=""):
    """Two product Results (e.g. batched vs. alone) hold the same rows, bit for bit."""
    assert a.types == b.types, (what, a.types, b.types)
    assert a.nullable == b.nullable, (what, a.nullable, b.nullable)
    ordered = a.ordered and b.ordered
    x = normalize(a.tuples(), ordered)
    y = normalize(b.tuples(), ordered)
    if x != y:
        diff = [(i, u, v) for i, (u, v) in enumerate(zip(x, y)) if u != v][:5]
        raise AssertionError(f"{what}: {len(x)} rows vs {len(y)} rows; first differences (row, got, want): {diff}")
