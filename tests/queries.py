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

# (sf, fact_rows) of each configuration at full size
FULL = {"C1": (1, 6_000_000), "C2": (20, 120_000_000), "C3": (100, 600_000_000)}
