"""Benchmark: fact-scan GB/s of the fused JIT query path on one MI355X.

Default workload: the 4-way join + GROUP BY over a 600M-row fact table with
nullable columns (scale 100, 25x25x7 groups), columns resident in HBM.  One
step = one full query execution through the C ABI (plan, cached module,
dimension builds, fused scan, materialization, result copy to the host).

value = algorithmic bytes of the query / ms per step, where
  algorithmic bytes = 4 B x rows x touched fact columns
                    + 8 B x ceil(rows/64) x nullable touched fact columns
                    + 4 B x rows x touched dimension columns.
roofline: the fused scan kernel alone (fact bytes / its HIP-event duration on
the query stream) against the 8 TB/s HBM peak.
cpu_baseline: a tiled C/OpenMP restatement of the kernels the reference
generates for the query (oracle/ssb_tiled.c), on a bounded sample of the same
workload, on all the process's cores and on one.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    # name: (sf, fact rows, nullable spec, sql key)
    "c3": (100, 600_000_000, "LINEORDER.lo_revenue:200,SUPPLIER.s_nation:200", "C3"),
    "c2": (20, 120_000_000, "", "C2"),
    "c1": (1, 6_000_000, "", "C1"),
}


def parse_args():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--workload", default="c3", choices=sorted(WORKLOADS))
    p.add_argument("--fact-rows", type=int, default=0, help="override the fact-table size")
    p.add_argument("--query", default="", help="time this query of tests/queries.py (by name) on the workload's data")
    p.add_argument("--option", action="append", default=[], metavar="NAME=VALUE",
                   help="database option (cq_db_set_option) for the timed query, repeatable")
    p.add_argument("--seed", type=int, default=42)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-sample-rows", type=int, default=120_000_000,
                   help="fact rows of the all-core CPU baseline sample (capped at the configuration's size)")
    p.add_argument("--cpu-one-core-rows", type=int, default=20_000_000,
                   help="fact rows of the one-core CPU baseline sample")
    p.add_argument("--extra", action="store_true", help="also time the other plain configurations")
    p.add_argument("--no-batched", action="store_true", help="skip the batched shared-scan measurement")
    p.add_argument("--no-served", action="store_true", help="skip the serving measurement")
    p.add_argument("--batch-queries", type=int, default=64)
    p.add_argument("--serve-rates", default="50,100,200,500,1000,2000,2700",
                   help="Poisson arrival rates (req/s), comma separated")
    p.add_argument("--serve-seconds", type=float, default=2.0, help="duration of each arrival rate")
    p.add_argument("--serve-burst", type=int, default=1024, help="requests of the closed-loop burst")
    p.add_argument("--serve-options", default="lanes=1,max_batch=128,cold_lanes=1",
                   help="server options, key=value,... (cq_server_start)")
    p.add_argument("--serve-adhoc", type=float, default=0.05, help="fraction of requests with a new plan shape")
    p.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                   help="only run a plain device-to-device copy of BYTES (HBM counter calibration)")
    return p.parse_args()


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(workload, sample_rows, one_core_rows, seed, runs=5):
    """The tiled CPU restatement of the reference's generated kernels for this
    configuration (oracle/ssb_tiled.c: 512-row tiles, int[2*len] hash tables,
    Horner group hash, u64 atomics), OpenMP over the host's cores, on a
    bounded fact sample of the same generator: one warm-up, then the median
    of `runs` timed runs (dimension builds + fact scan), on every core the
    process may use and on one core."""
    from oracle import oracle as orc
    from oracle import tiled

    sf, full_rows, nullable, key = WORKLOADS[workload]
    cols = tiled.COLUMNS[key]
    nproc = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or nproc
    out = {}
    for label, rows, th in (("all", sample_rows, threads), ("one", one_core_rows, 1)):
        spec = orc.GenSpec(sf=sf, fact_rows=rows, seed=seed, nullable=nullable)
        data = orc.gen_tables(spec, tables=list(cols), columns=cols)

        def stats(column, data=data):
            for t, cs in data.items():
                if column in cs:
                    v, m = cs[column]
                    if m is not None:
                        ok = np.unpackbits(m.view(np.uint8), bitorder="little")[:len(v)].astype(bool)
                        v = v[ok]
                    return {"min": int(v.min()), "card": int(len(np.unique(v))), "nullable": m is not None}
            raise KeyError(column)

        fact = data["LINEORDER"]
        nbytes = 4 * rows * len(fact) + 8 * ((rows + 63) // 64) * sum(m is not None for _, m in fact.values())
        nbytes += sum(4 * len(v) for t, cs in data.items() if t != "LINEORDER" for v, _ in cs.values())
        tiled.run(key, data, stats, th)  # warm-up
        times = []
        for _ in range(runs):
            t0 = time.perf_counter()
            tiled.run(key, data, stats, th)
            times.append(time.perf_counter() - t0)
        t = sorted(times)[len(times) // 2]
        out[label] = {"rows": rows, "threads": th, "ms": t * 1000, "GBps": nbytes / t / 1e9,
                      "rows_per_s": rows / t}
        del data
    a, o = out["all"], out["one"]
    return {"value": round(a["GBps"], 3), "unit": "GB/s", "cores": a["threads"], "kind": "port",
            "sample": (f"{key} over a {a['rows']:,}-row fact sample of the {full_rows:,}-row configuration (same "
                       f"generator, scale-{sf} dimensions): tiled C/OpenMP restatement of the reference's generated "
                       f"kernels (not AdaptiveCpp), dimension builds + fact scan, median of {runs} runs after a "
                       f"warm-up: {a['ms']:.0f} ms on {a['threads']} threads ({a['rows_per_s'] / 1e6:.0f} M fact "
                       f"rows/s)"),
            "one_core": {"value": round(o["GBps"], 3), "rows": o["rows"], "ms": round(o["ms"], 1),
                         "rows_per_s": round(o["rows_per_s"])},
            "host": {"cpu_model": _cpu_model(), "nproc": nproc, "os_cpu_count": os.cpu_count(),
                     "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}}


def same_rows(a, b):
    key = lambda t: tuple((x is None, x) for x in t)
    x, y = a.tuples(), b.tuples()
    if not (a.ordered and b.ordered):
        x, y = sorted(x, key=key), sorted(y, key=key)
    return a.types == b.types and x == y


def bench_batched(pa, Q, args):
    """64 template queries (1-4 dimensions, random constants) over resident
    scale-20 columns: one shared-scan batch vs. the same queries run one at a
    time through the plain path."""
    db = pa.Database(0)
    db.generate_ssb(sf=20, fact_rows=120_000_000, seed=args.seed)
    db.finalize()
    n = args.batch_queries
    mixes = [Q.template_mix(n, args.seed + k) for k in range(args.steps + 1)]
    # warm-up: template modules, shared-scan module (autotuned once), plain modules
    db.batch_execute(mixes[0], {"autotune": 1})
    tuned = db.batch_report()["scans"][0]
    for q in mixes[0]:
        db.execute(q)
    # serial plain
    t0 = time.perf_counter()
    serial = [db.execute(q) for q in mixes[1]]
    serial_s = time.perf_counter() - t0
    # batched, fresh constants every step
    t0 = time.perf_counter()
    out = None
    for k in range(1, args.steps + 1):
        res = db.batch_execute(mixes[k])
        if k == 1:
            out = res
    batched_s = (time.perf_counter() - t0) / args.steps
    rep = db.batch_report()
    exact = all(same_rows(a, b) for a, b in zip(out, serial))
    # bytes of one shared pass: every union column of every launch, read once
    union_bytes = sum(4 * 120_000_000 * sc["union_columns"] for sc in rep["scans"])
    db.close()
    return {"workload": f"{n} queries, 4 templates (2-, 3-, 4- and 4-way joins: 1, 2, 3 and 3 dimensions), random constants, "
                        "scale 20 (120,000,000 fact rows)",
            "qps": round(n / batched_s, 1), "serial_qps": round(n / serial_s, 1),
            "speedup": round(serial_s / batched_s, 2), "batch_ms": round(batched_s * 1000, 3),
            "serial_ms": round(serial_s * 1000, 3), "shared_scans": rep["shared_scans"],
            "block_threads": tuned["block_threads"], "vectors_per_thread": tuned["vectors_per_thread"],
            "table_mask": tuned.get("table_mask"), "lds_table_bytes": tuned.get("lds_table_bytes"), "table_kb": tuned.get("table_kb"),
            "graph": rep["scans"][0]["graph"], "scan_ms_last": rep["scans"][0]["scan_ms"],
            "device_ms_last": rep["scans"][0]["device_ms"], "host_ms_last": rep["scans"][0]["host_ms"],
            "plan_ms_last": rep["plan_ms"], "union_GBps": round(union_bytes / batched_s / 1e9, 1),
            "bit_exact_vs_serial": exact}


def bench_served(pa, Q, args):
    """Open-loop Poisson arrivals against the continuous-batching server over
    resident scale-100 columns (600,000,000 fact rows) at each rate of
    --serve-rates; a fraction of the requests have a plan shape never seen
    before (run at once on the interpreted tier, compiled in the background).
    Then a closed-loop burst (every request submitted at once) gives the
    saturated throughput.  Serial: the same templates one at a time through
    the plain path."""
    import random
    import threading
    db = pa.Database(0)
    db.generate_ssb(sf=100, fact_rows=600_000_000, seed=args.seed, nullable=WORKLOADS["c3"][2])
    db.finalize()
    rng = random.Random(args.seed)
    warm = [Q.template_query(k, rng) for k in (1, 2, 3, 4)]
    for q in warm:
        db.execute(q)
    ser = Q.template_mix(32, args.seed + 7)
    t0 = time.perf_counter()
    for q in ser:
        db.execute(q)
    serial_qps = len(ser) / (time.perf_counter() - t0)
    srv = pa.Server(db, dict(kv.split("=") for kv in args.serve_options.split(",") if kv))
    # warm the kernel cache: every template seen twice is compiled into the
    # hot set in the background (the ad-hoc shapes below stay cold)
    for _ in range(2):
        for t in [srv.submit(q) for q in warm]:
            srv.wait(t)
    t_end = time.perf_counter() + 60
    while srv.stats()["scan_set"] < len(warm) and time.perf_counter() < t_end:
        time.sleep(0.05)

    def draw(n):
        return [Q.adhoc_query(rng) if rng.random() < args.serve_adhoc else Q.template_query(rng.randint(1, 4), rng)
                for _ in range(n)]

    first_error = []

    def drain(tickets):
        errors = 0
        for t in tickets:
            try:
                srv.wait(t, 600_000)
            except pa.CrystalError as e:
                errors += 1
                if not first_error:
                    first_error.append(str(e)[:300])
                    print(f"[served] first error: {first_error[0]}", file=sys.stderr, flush=True)
        return errors

    rates = []
    for rate in [float(x) for x in args.serve_rates.split(",") if x]:
        srv.reset_stats()
        reqs = draw(max(1, int(rate * args.serve_seconds)))
        tickets = []

        def producer():
            t = time.perf_counter()
            for q in reqs:
                t += rng.expovariate(rate)
                d = t - time.perf_counter()
                if d > 0:
                    time.sleep(d)
                tickets.append(srv.submit(q))

        # admitted-footprint trace: bytes reserved by requests in flight,
        # sampled every 10 ms while the rate runs
        samples = []
        sampling = [True]

        def sampler(t_start):
            while sampling[0]:
                samples.append((time.perf_counter() - t_start, srv.stats()["in_flight_bytes"]))
                time.sleep(0.01)

        th = threading.Thread(target=producer)
        t0 = time.perf_counter()
        sm = threading.Thread(target=sampler, args=(t0,))
        sm.start()
        th.start()
        th.join()
        errors = drain(tickets)
        wall = time.perf_counter() - t0
        sampling[0] = False
        sm.join()
        st = srv.stats()
        rates.append({"rate": rate, "requests": len(reqs), "qps": round(st["completed"] / wall, 1),
                      "p50_ms": round(st["p50_ms"], 2), "p99_ms": round(st["p99_ms"], 2), "max_ms": round(st["max_ms"], 2),
                      "mean_batch": round(st["mean_batch"], 1), "batch_wall_ms": round(st["batch_wall_ms"], 2),
                      "cold_p50_ms": round(st["cold_p50_ms"], 2), "cold_p99_ms": round(st["cold_p99_ms"], 2),
                      "interpreted": st["interpreted_queries"],
                      "promotions": st["promotions"], "errors": errors,
                      "footprint_mb": {"max": round(max((b for _, b in samples), default=0) / 2**20, 1),
                                       "mean": round(sum(b for _, b in samples) / max(1, len(samples)) / 2**20, 1),
                                       "every_100ms": [round(b / 2**20, 1) for t, b in samples[::10]]}})
        if errors:
            rates[-1]["last_error"] = st.get("last_error", "")[:200]
        print(f"[served] rate {rate:g}: {rates[-1]}", file=sys.stderr, flush=True)
    # closed loop: every request queued at once
    srv.reset_stats()
    reqs = draw(args.serve_burst)
    t0 = time.perf_counter()
    tickets = [srv.submit(q) for q in reqs]
    errors = drain(tickets)
    wall = time.perf_counter() - t0
    st = srv.stats()
    burst = {"requests": len(reqs), "qps": round(st["completed"] / wall, 1), "p50_ms": round(st["p50_ms"], 2),
             "p99_ms": round(st["p99_ms"], 2), "mean_batch": round(st["mean_batch"], 1), "errors": errors}
    if errors:
        burst["last_error"] = st.get("last_error", "")[:200]
    print(f"[served] burst: {burst}", file=sys.stderr, flush=True)
    srv_budget = st["budget_bytes"]
    srv.stop()
    db.close()
    return {"workload": f"Poisson arrivals at {args.serve_rates} req/s for {args.serve_seconds:g} s each, then a burst of "
                        f"{args.serve_burst}; 4 templates + {args.serve_adhoc:.0%} new plan shapes, scale 100 "
                        f"(600,000,000 fact rows)",
            "server_options": args.serve_options, "budget_mb": round(srv_budget / 2**20, 1),
            "serial_qps": round(serial_qps, 1), "rates": rates, "burst": burst,
            "saturated_speedup": round(burst["qps"] / serial_qps, 2),
            "within_50ms": within_slo(rates, serial_qps, 50.0)}


def within_slo(rates, serial_qps, p99_ms):
    """The highest achieved rate whose p99 latency stays within `p99_ms`,
    and its ratio to serial execution (None when no rate does)."""
    ok = [r for r in rates if r["p99_ms"] <= p99_ms and not r.get("errors")]
    if not ok:
        return None
    best = max(ok, key=lambda r: r["qps"])
    return {"rate": best["rate"], "qps": best["qps"], "p99_ms": best["p99_ms"], "speedup": round(best["qps"] / serial_qps, 2)}


def copy_calibration(nbytes, steps):
    """A plain copy of known size: its FETCH_SIZE / WRITE_SIZE counters give
    the factor between counter readings and bytes moved."""
    import torch
    n = nbytes // 4
    a = torch.ones(n, dtype=torch.int32, device="cuda:0")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        b.copy_(a)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1000 / steps
    print(json.dumps({"copy_bytes": n * 4, "steps": steps, "ms_per_copy": round(ms, 4),
                      "GBps_read_plus_write": round(2 * n * 4 / ms / 1e6, 1)}), flush=True)


def main():
    args = parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU scaling is not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration(args.copy_calibration, max(1, args.steps))
        return
    import practice_amd as pa
    from tests import queries as Q

    sf, rows, nullable, key = WORKLOADS[args.workload]
    if args.fact_rows:
        rows = args.fact_rows
    sql = getattr(Q, args.query or key)
    db = pa.Database(0)
    t0 = time.perf_counter()
    db.generate_ssb(sf=sf, fact_rows=rows, seed=args.seed, nullable=nullable)
    db.finalize()
    gen_s = time.perf_counter() - t0
    db.set_option("kernel_timing", 1)
    for kv in args.option:
        k, v = kv.split("=")
        db.set_option(k, int(v))
    desc = db.describe(sql)
    for _ in range(args.warmup):
        db.execute(sql)
    scan_ms = []
    parts = {k: 0.0 for k in ("parse_ms", "codegen_ms", "gpu_execute_ms", "host_fetch_ms", "engine_ms")}
    t0 = time.perf_counter()
    res = None
    for _ in range(args.steps):
        res = db.execute(sql)
        scan_ms.append(res.timing["scan_kernel_ms"])
        for k in parts:
            parts[k] += res.timing[k]
    elapsed = time.perf_counter() - t0
    ms_per_step = elapsed * 1000.0 / args.steps
    gbs = desc["algorithmic_bytes"] / (ms_per_step / 1000.0) / 1e9
    scan_avg = sum(scan_ms) / len(scan_ms)
    fact_bytes = 0
    for c in desc["fact_cols"]:
        fact_bytes += 4 * rows
    nullable_cols = [c for c in desc["fact_cols"] if db.column_stats(c)["nullable"]]
    fact_bytes += 8 * ((rows + 63) // 64) * len(nullable_cols)
    achieved = fact_bytes / (scan_avg / 1000.0) / 1e9
    traffic = None
    tf = os.path.join(ROOT, "profiles", f"{args.workload}_scan_traffic.json")
    if os.path.exists(tf) and not args.fact_rows and not args.query:
        with open(tf) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
    line = {
        "metric": "fact-scan GB/s vs HBM peak; served queries/sec + p99 latency vs serial plain",
        "value": round(gbs, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32 columns, u64 aggregates",
        "data": "synthetic (seeded star schema generated in HBM)",
        "config": {"workload": f"{args.workload.upper()}: scale {sf}, {rows:,} fact rows, "
                               f"{len(desc['joins'])}-way join, {desc['groups']} group slots"
                               + (f", query {args.query}" if args.query else ""),
                   "fact_rows": rows, "groups": desc["groups"], "algorithmic_bytes": desc["algorithmic_bytes"],
                   "result_rows": res.rows, "block_threads": desc["block_threads"],
                   "lds_bytes": desc["scan_lds_bytes"], "generate_s": round(gen_s, 2)},
        "step_ms_breakdown": {k: round(v / args.steps, 4) for k, v in parts.items()},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "cq_scan", "kernel_ms": round(scan_avg, 4), "kernel_bytes": fact_bytes},
    }
    if args.extra:
        extra = {}
        for wl in ("c1", "c2"):
            if wl == args.workload:
                continue
            s2, r2, n2, k2 = WORKLOADS[wl]
            d2 = pa.Database(0)
            d2.generate_ssb(sf=s2, fact_rows=r2, seed=args.seed, nullable=n2)
            d2.finalize()
            d2.set_option("kernel_timing", 1)
            q2 = getattr(Q, k2)
            for _ in range(args.warmup):
                d2.execute(q2)
            t1 = time.perf_counter()
            sm = []
            for _ in range(args.steps):
                sm.append(d2.execute(q2).timing["scan_kernel_ms"])
            ms2 = (time.perf_counter() - t1) * 1000 / args.steps
            dd = d2.describe(q2)
            extra[wl] = {"ms_per_step": round(ms2, 4), "GBps": round(dd["algorithmic_bytes"] / ms2 / 1e6, 2),
                         "scan_kernel_ms": round(sum(sm) / len(sm), 4)}
            d2.close()
        # OR-key join over the same 600M-row fact table (two date keys probed per row)
        for _ in range(args.warmup):
            db.execute(Q.ORKEY)
        t1 = time.perf_counter()
        sm = []
        for _ in range(args.steps):
            sm.append(db.execute(Q.ORKEY).timing["scan_kernel_ms"])
        ms2 = (time.perf_counter() - t1) * 1000 / args.steps
        dd = db.describe(Q.ORKEY)
        extra["orkey_sf100"] = {"ms_per_step": round(ms2, 4), "scan_kernel_ms": round(sum(sm) / len(sm), 4),
                                "scan_GBps": round(dd["algorithmic_bytes"] / (sum(sm) / len(sm)) / 1e6, 2)}
        # projection over the same fact table (count pass + write pass)
        for _ in range(args.warmup):
            db.execute(Q.PROJ)
        t1 = time.perf_counter()
        sm = []
        for _ in range(args.steps):
            r = db.execute(Q.PROJ)
            sm.append(r.timing["scan_kernel_ms"])
        ms2 = (time.perf_counter() - t1) * 1000 / args.steps
        dd = db.describe(Q.PROJ)
        extra["projection_sf100"] = {"ms_per_step": round(ms2, 4), "kernels_ms": round(sum(sm) / len(sm), 4), "rows": r.rows,
                                     "scan_GBps": round(dd["algorithmic_bytes"] / (sum(sm) / len(sm)) / 1e6, 2)}
        line["other_configs"] = extra
    db.close()
    if not args.no_batched:
        line["batched"] = bench_batched(pa, Q, args)
    if not args.no_served:
        line["served"] = bench_served(pa, Q, args)
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(args.workload, min(args.cpu_sample_rows, rows),
                                            min(args.cpu_one_core_rows, rows), args.seed)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
