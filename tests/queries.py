"""Benchmark configurations and their SQL (reference dialect: no BETWEEN)."""

C1 = ("select sum(lo_extendedprice*lo_discount) as revenue from lineorder, ddate "
      "where lo_orderdate = d_datekey and d_year = 1993 and lo_discount >= 1 and lo_discount <= 3 "
      "and lo_quantity < 25;")


def c2(category: int = 12, region: int = 1) -> str:
    return ("select d_year, p_brand1, sum(lo_revenue) from lineorder, part, supplier, ddate "
            "where lo_orderdate = d_datekey and lo_partkey = p_partkey and lo_suppkey = s_suppkey "
            f"and p_category = {category} and s_region = {region} "
            "group by d_year, p_brand1 order by d_year, p_brand1;")


C2 = c2()

C3 = ("select c_nation, s_nation, d_year, sum(lo_revenue) from lineorder, customer, supplier, ddate "
      "where lo_custkey = c_custkey and lo_suppkey = s_suppkey and lo_orderdate = d_datekey "
      "and (lo_revenue > 1000000 or lo_revenue is null) "
      "group by c_nation, s_nation, d_year;")

C3_NULLABLE = "LINEORDER.lo_revenue:200,SUPPLIER.s_nation:200"

# C3's pieces over the same data, for profiling what bounds it: the CUSTOMER
# join alone (one random gather per row into the 3M-key table), and the same
# four fact columns, filter, SUPPLIER / DDATE probes and grouping with no
# CUSTOMER gather (lo_custkey read by a filter that keeps every row)
C3_CUSTOMER_ONLY = ("select c_nation, sum(lo_revenue) from lineorder, customer where lo_custkey = c_custkey "
                    "and (lo_revenue > 1000000 or lo_revenue is null) group by c_nation;")
C3_NO_GATHER = ("select s_nation, d_year, sum(lo_revenue) from lineorder, supplier, ddate "
                "where lo_suppkey = s_suppkey and lo_orderdate = d_datekey and lo_custkey > 0 "
                "and (lo_revenue > 1000000 or lo_revenue is null) group by s_nation, d_year;")

# (sf, fact_rows) of each configuration at full size
FULL = {"C1": (1, 6_000_000), "C2": (20, 120_000_000), "C3": (100, 600_000_000)}


# ---------------------------------------------------------------- mixes
# Query templates of the batched and served workloads: 1 to 4 dimensions
# joined to the fact table, predicate constants drawn uniformly from each
# column's domain (year 1992-1998, discount 0-10, quantity 1-50, region 0-4,
# category 0-24).
def template_query(kind: int, rng) -> str:
    if kind == 1:
        d = rng.randint(0, 8)
        return ("select sum(lo_extendedprice*lo_discount) from lineorder, ddate "
                f"where lo_orderdate = d_datekey and d_year = {rng.randint(1992, 1998)} "
                f"and lo_discount >= {d} and lo_discount <= {d + 2} and lo_quantity < {rng.randint(2, 50)}")
    if kind == 2:
        return ("select s_nation, sum(lo_revenue) from lineorder, supplier, ddate "
                "where lo_suppkey = s_suppkey and lo_orderdate = d_datekey "
                f"and s_region = {rng.randint(0, 4)} and d_year = {rng.randint(1992, 1998)} group by s_nation")
    if kind == 3:
        return c2(rng.randint(0, 24), rng.randint(0, 4))
    y = rng.randint(1992, 1997)
    return ("select c_nation, s_nation, d_year, sum(lo_revenue) from lineorder, customer, supplier, ddate "
            "where lo_custkey = c_custkey and lo_suppkey = s_suppkey and lo_orderdate = d_datekey "
            f"and c_region = {rng.randint(0, 4)} and s_region = {rng.randint(0, 4)} "
            f"and d_year >= {y} and d_year <= {y + 1} group by c_nation, s_nation, d_year")


def template_mix(n: int, seed: int):
    """n queries, template drawn uniformly from the four, seeded."""
    import random
    rng = random.Random(seed)
    return [template_query(rng.randint(1, 4), rng) for _ in range(n)]


_ADHOC_COLS = ["lo_quantity", "lo_discount", "lo_tax", "lo_extendedprice", "lo_revenue", "lo_supplycost"]
_ADHOC_OPS = ["<", ">", "<=", ">=", "<>"]


def adhoc_query(rng) -> str:
    """A query whose plan shape (columns, operators, aggregates) is drawn at
    random, so it almost never shares a compiled module with another."""
    k = rng.randint(2, 4)
    cols = rng.sample(_ADHOC_COLS, k)
    preds = []
    for c in cols:
        op = rng.choice(_ADHOC_OPS)
        hi = 50 if c == "lo_quantity" else (10 if c in ("lo_discount", "lo_tax") else 1_000_000)
        preds.append(f"{c} {op} {rng.randint(0, hi)}")
    a, b = rng.sample(_ADHOC_COLS, 2)
    agg = rng.choice([f"sum({a})", f"sum({a} + {b})", f"count(*), sum({a})", f"max({a}), min({b})", f"sum({a} * 2 - {b})"])
    if rng.random() < 0.5:
        return f"select {agg} from lineorder where " + " and ".join(preds)
    return (f"select d_year, {agg} from lineorder, ddate where lo_orderdate = d_datekey and "
            + " and ".join(preds) + " group by d_year")


# two ad-hoc shapes of the served mix (adhoc_query), for timing the
# interpreted tier against compiled modules
ADHOC_1 = ("select count(*), sum(lo_extendedprice + lo_revenue) from lineorder where lo_quantity < 20 "
           "and lo_discount > 3 and lo_tax <> 5")
ADHOC_2 = ("select d_year, max(lo_revenue), min(lo_supplycost) from lineorder, ddate where lo_orderdate = d_datekey "
           "and lo_revenue >= 500000 and lo_extendedprice < 800000 group by d_year")

# OR-key join: a fact row counts once per distinct date key (order / commit)
# that names a DDATE row
ORKEY = ("select d_year, count(*), sum(lo_revenue) from lineorder, ddate where lo_orderdate = d_datekey "
         "or lo_commitdate = d_datekey group by d_year")

# projection: about 0.4 % of the fact rows, two columns each
PROJ = "select lo_revenue, lo_discount from lineorder where lo_quantity = 7 and lo_discount > 8"
