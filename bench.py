"""Benchmark of the counting path on one MI355X.

Default workload: triangle count of RMAT scale 22 (edge factor 16,
a/b/c/d = 0.57/0.19/0.19/0.05, seed 2), int32 CSR, degree-oriented, the
largest single-GPU configuration of the headline metric.  A "step" is one
triangle-count pass over the resident graph (all counting kernels, count
copied back).  Prints ONE JSON line.

  python bench.py [--steps K] [--warmup W] [--gpus 1] [--workload NAME]

The timed region holds K steps bracketed by device synchronisation; inputs
are resident in HBM before it starts.  `roofline.achieved` is the
algorithmic bytes of one pass, B_alg = 8(n+1) + 20 m+ + 4 S+, divided by the
pass's average device time from HIP events recorded on the launch stream.
`cpu_baseline` is the CPU oracle (sorted-merge counting over the same DAG,
OpenMP on the host cores) timed once on the same graph; its count is also
compared with the GPU count.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s

WORKLOADS = {
    # name: (kind, scale, edge_factor, seed)
    "tc_rmat22": ("rmat", 22, 16, 2),
    "tc_er22": ("er", 22, 16, 3),
    "tc_rmat20": ("rmat", 20, 16, 4),
    "tc_rmat16": ("rmat", 16, 16, 1),
    "tc_rmat24": ("rmat", 24, 16, 2),
}


def log(msg: str) -> None:
    """A progress line on stderr (the JSON result goes to stdout)."""
    print(f"# {msg}", file=sys.stderr, flush=True)


def make_graph(ma, kind, scale, ef, seed):
    if kind == "rmat":
        return ma.Graph.rmat(scale, ef, seed=seed)
    n = 1 << scale
    return ma.Graph.er(n, n * ef, seed)


def load_traffic(workload: str):
    """HBM bytes per pass from the committed rocprofv3 PMC summary."""
    f = ROOT / "profiles" / f"{workload}_traffic.json"
    if not f.exists():
        return None
    try:
        return json.loads(f.read_text()).get("hbm_bytes_per_pass")
    except (ValueError, OSError):
        return None


def upload_compare(ma, g, count):
    """Preparation time of a fresh device graph (CSR upload + DAG + triangle
    plan), host-built plan against device-built plan, each the second of two
    uploads (so one-time code loading is not counted), and whether the
    device-built plan gives the same count."""
    def timed(device_plan):
        for _ in range(2):
            t0 = time.perf_counter()
            dg = ma.DeviceGraph(g, 0, device_plan=device_plan)
            dt = time.perf_counter() - t0
            c = dg.tc_count()[0]
            ks = sorted(dg.tc_count()[1].kernel_ms for _ in range(5))
            dg.close()
        return dt, c, ks[2]
    host_s, _, host_ms = timed(False)
    dev_s, c, dev_ms = timed(True)
    return {"prep_host_plan_s": host_s, "prep_device_plan_s": dev_s,
            "device_plan_count_equal": c == count,
            "kernel_ms_host_plan": host_ms, "kernel_ms_device_plan": dev_ms}


def cpu_baseline(g, count_gpu, m):
    import numpy as np  # noqa: F401
    import oracle
    dag = g.orientation(128)
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()
    t0 = time.perf_counter()
    want = oracle.tc_dag(dag["dag_rowptr"], dag["dag_colidx"], threads=threads)
    dt = time.perf_counter() - t0
    cpu_model = ""
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                cpu_model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {
        "value": m / dt, "unit": "edges/s", "cores": threads, "kind": "port",
        "sample": ("full count of the same graph: sorted-merge intersection "
                   "over the same degree-oriented DAG (oracle/oracle.c "
                   "orc_tc_dag), OpenMP dynamic over vertices, "
                   f"{dt:.2f} s, {cpu_model}"),
        "seconds": dt, "count": want, "count_matches_gpu": want == count_gpu,
        "nproc": os.cpu_count(), "cpu_model": cpu_model,
    }


def cpu_small_baselines(ma) -> dict:
    """The other CPU baselines of the counting path, on the host cores of
    this run: the triangle count of RMAT-16 (ef 16, seed 1) on one thread,
    and the pattern oracle (the reference's nested set algebra, OpenMP
    over vertices) for P1..P8 on RMAT-10 (ef 16, seed 4), the largest RMAT
    scale where every pattern finishes in seconds, with its set-op work W
    and the GPU census time of the same count."""
    import oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()
    g = ma.Graph.rmat(16, 16, seed=1)
    dag = g.orientation(128)
    t0 = time.perf_counter()
    want = oracle.tc_dag(dag["dag_rowptr"], dag["dag_colidx"], threads=1)
    dt = time.perf_counter() - t0
    dg = ma.DeviceGraph(g, 0)
    gpu, st = dg.tc_count()
    dg.close()
    m = int(g.nnz) // 2
    out = {"tc_rmat16_1thread": {
        "seconds": dt, "edges_per_s": m / dt, "cores": 1, "count": want,
        "gpu_kernel_ms": st.kernel_ms, "count_matches_gpu": gpu == want}}
    g = ma.Graph.rmat(10, 16, seed=4)
    rp, ci = g.csr()
    dg = ma.DeviceGraph(g, 0)
    rows = {}
    for p in range(1, 9):
        t0 = time.perf_counter()
        c, w = oracle.pattern_work(rp, ci, p, threads=threads)
        dt = time.perf_counter() - t0
        dg.pattern_count(p)
        gc, st = dg.pattern_count(p)
        rows[f"P{p}"] = {"oracle_s": dt, "set_op_work": w,
                         "gpu_census_ms": st.kernel_ms, "count": c,
                         "count_matches_gpu": gc == c}
    dg.close()
    log("cpu baselines done")
    out["patterns_rmat10"] = {"cores": threads, "graph": "RMAT scale 10, "
                              "ef 16, seed 4", **rows}
    return out


def calibrate_copy(mib: int) -> int:
    import torch  # device memory plumbing only
    n = mib * (1 << 20) // 4
    src = torch.arange(n, dtype=torch.int32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"calibrate_copy_bytes": n * 4, "reps": 5,
                      "ok": bool(torch.equal(src, dst))}), flush=True)
    return 0


def batch_graphs(ma, count, lo, hi, seed=5):
    """`count` RMAT graphs (a/b/c/d 0.57/0.19/0.19/0.05): scale log-uniform
    in [lo, hi] (2^scale vertices, i.e. scale drawn uniformly), edge factor
    uniform in {4, 8, 16, 32}; graph i uses seed*100003 + i."""
    import random
    rng = random.Random(seed)
    specs = [(rng.randint(lo, hi), rng.choice((4, 8, 16, 32)))
             for _ in range(count)]
    return specs, [ma.Graph.rmat(s, ef, seed=seed * 100003 + i)
                   for i, (s, ef) in enumerate(specs)]


def run_batch(ma, args) -> dict:
    """The batched workload: every graph's queries (default TC, P2, P4 on
    512 RMAT graphs, scale uniform in [10, 18], edge factor in
    {4, 8, 16, 32}, seed 5) in one batch call against one call per query,
    each synchronising, with the same methods (triangle kernels; P1..P8 by
    the census).  Counts must agree between the two, and with the oracle on
    the graphs of scale <= 10."""
    import oracle
    names = [p.strip() for p in args.batch_patterns.split(",")]
    pats = [ma.PATTERN_NAMES[p] for p in names]
    lo, hi = args.batch_scales
    t0 = time.perf_counter()
    specs, graphs = batch_graphs(ma, args.batch_graphs, lo, hi)
    dgs = [ma.DeviceGraph(g, 0) for g in graphs]
    prep_s = time.perf_counter() - t0
    log(f"batch: {len(graphs)} graphs uploaded ({prep_s:.1f} s)")
    queries = [(dg, p) for dg in dgs for p in pats]
    sync = lambda: ma.native.check(ma.native.lib.mgpm_device_synchronize(0),
                                   "sync")
    med = lambda x: sorted(x)[len(x) // 2]
    # one call (and one synchronisation) per query
    one = lambda dg, p: dg.pattern_count(p)[0]

    def all_single():
        res = []
        for k, (dg, p) in enumerate(queries):
            res.append(one(dg, p))
            if (k + 1) % 256 == 0:
                log(f"batch: {k + 1}/{len(queries)} single queries done")
        return res
    if args.batch_warmup > 0:
        for _ in range(args.batch_warmup):
            single = all_single()
    else:
        # light warm-up: one census query per graph (its census DAG is built
        # and kept with the graph) and one query per pattern (kernels loaded)
        for dg in dgs:
            dg.pattern_count(1)
        for p in pats:
            one(dgs[0], p)
    sync()
    t_single = []
    for _ in range(args.batch_steps):
        t0 = time.perf_counter()
        single = all_single()
        sync()
        t_single.append(time.perf_counter() - t0)
        log(f"batch: one call per query {t_single[-1]:.2f} s")
    # the same queries in one batch call: one census per graph serves its
    # pattern queries, lanes overlap the graphs
    b = ma.Batch(0, lanes=args.batch_lanes, use_graph=args.batch_use_graph)
    if args.batch_warmup > 0:
        for _ in range(args.batch_warmup):
            counts, st = b.run(queries)
    # (without warm-up batches the lanes' workspaces grow inside the timed
    # batch: one allocation per lane, milliseconds against its seconds)
    sync()
    log("batch: warm-up batch done")
    t_batch, kms = [], []
    for _ in range(args.batch_steps):
        t0 = time.perf_counter()
        counts, st = b.run(queries)
        sync()
        t_batch.append(time.perf_counter() - t0)
        kms.append(st.kernel_ms)
        log(f"batch: one batch {t_batch[-1]:.2f} s")
    b.close()
    for dg in dgs:
        dg.close()
    # the oracle on the small graphs (its pattern algebra is cubic)
    checked, oracle_ok = 0, True
    for k, ((s_, ef), g) in enumerate(zip(specs, graphs)):
        if s_ > 10 or ef > args.batch_oracle_ef:
            continue
        rp, ci = g.csr()
        for j, p in enumerate(pats):
            want = oracle.tc(rp, ci) if p == 0 else oracle.pattern(rp, ci, p)[0]
            oracle_ok &= counts[k * len(pats) + j] == want
            checked += 1
    nq = len(queries)
    return {
        "workload": f"{args.batch_graphs} RMAT graphs, scale uniform in "
                    f"[{lo},{hi}] (2^scale vertices), edge factor uniform in "
                    "{4,8,16,32}, a/b/c/d 0.57/0.19/0.19/0.05, seed 5; "
                    f"queries {'+'.join(names)} per graph",
        "queries": nq, "lanes": args.batch_lanes,
        "hip_graph": bool(args.batch_use_graph),
        "hw_queues": os.environ.get("GPU_MAX_HW_QUEUES", "4"),
        "qps": nq / med(t_batch),
        "batch_s": med(t_batch),
        "per_graph_launch_qps": nq / med(t_single),
        "per_graph_launch_s": med(t_single),
        "speedup_vs_per_graph_launch": med(t_single) / med(t_batch),
        "counts_equal": counts == single,
        "oracle_checked_queries": checked, "oracle_equal": oracle_ok,
        "batch_kernel_ms": med(kms),
        "graph_prep_s": prep_s,
        "edges_total": sum(int(g.nnz) // 2 for g in graphs),
    }


def bench_batch(ma, args) -> int:
    r = run_batch(ma, args)
    out = {
        "metric": "batched queries/s", "value": r["qps"],
        "unit": "queries/s", "n_gpus": 1, "steps": args.batch_steps,
        "warmup": args.batch_warmup, "ms_per_step": r["batch_s"] * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "int32", "data": "synthetic",
        "config": {"workload": "batch: " + r["workload"]}, **r,
    }
    print(json.dumps(out), flush=True)
    return 0 if r["counts_equal"] and r["oracle_equal"] else 1


def bench_dag(ma, args) -> int:
    """Degree-oriented DAG + small/large division of the TC workload's graph
    (graph.cc:196-321), built on the device vs. on the host, arrays compared."""
    import numpy as np
    kind, scale, ef, seed = WORKLOADS[args.workload]
    g = make_graph(ma, kind, scale, ef, seed)
    dg = ma.DeviceGraph(g, 0)
    for _ in range(max(1, args.warmup)):
        got, st = dg.orientation(128)
    kms = []
    for _ in range(args.steps):
        got, st = dg.orientation(128)
        kms.append(st.kernel_ms)
    t0 = time.perf_counter()
    want = g.orientation(128)
    host_s = time.perf_counter() - t0
    same = all(np.array_equal(got[k], want[k]) for k in want)
    dev_ms = sorted(kms)[len(kms) // 2]
    m = int(st.m)
    out = {
        "metric": "DAG + division arcs/s", "value": 2 * m / (dev_ms / 1e3),
        "unit": "arcs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": dev_ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "int32", "data": "synthetic",
        "config": {"workload": f"{args.workload}: degree-oriented DAG and "
                               "small/large division (threshold 128) of the "
                               "undirected CSR",
                   "n": int(st.n), "m": m},
        "device_ms": dev_ms, "host_s": host_s,
        "host_threads": int(os.environ.get("OMP_NUM_THREADS", "0") or 0),
        "speedup_vs_host": host_s * 1e3 / dev_ms, "arrays_equal": same,
    }
    print(json.dumps(out), flush=True)
    return 0 if same else 1


def bench_patterns(ma, args) -> int:
    """P1..P8 counts on one RMAT graph (default RMAT-20, ef 16, seed 4, the
    pattern configuration): device time per count for each requested
    pattern and method ("census": closed forms from the subgraph census;
    "kernels": the reference's per-edge set algebra), and whether the
    methods agree."""
    scale, ef, seed = args.pattern_graph
    names = [p.strip() for p in args.patterns.split(",")]
    methods = [m.strip() for m in args.pattern_methods.split(",")]
    t0 = time.perf_counter()
    g = ma.Graph.rmat(scale, ef, seed=seed)
    dg = ma.DeviceGraph(g, 0)
    prep_s = time.perf_counter() - t0
    rows, agree = {}, True
    for name in names:
        p = ma.PATTERN_NAMES[name]
        counts = {}
        for meth in methods:
            tune = ma.make_tune(pattern_census=-1) if meth == "kernels" else None
            dg.pattern_count(p, tune)  # builds plans outside the timing
            ks = []
            for _ in range(args.steps):
                c, st = dg.pattern_count(p, tune)
                ks.append(st.kernel_ms)
            print(f"# {name} {meth}: {c} in {sorted(ks)[len(ks) // 2]:.1f} ms",
                  file=sys.stderr, flush=True)
            counts[meth] = c
            rows[f"{name}_{meth}_ms"] = sorted(ks)[len(ks) // 2]
            rows[f"{name}_count"] = c
        agree = agree and len(set(counts.values())) <= 1
    first = f"{names[0]}_{methods[0]}_ms"
    out = {
        "metric": "pattern count time", "value": rows.get(first),
        "unit": "ms", "n_gpus": 1, "steps": args.steps, "warmup": 1,
        "ms_per_step": rows.get(first), "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": f"patterns {','.join(names)} on RMAT scale "
                               f"{scale}, edge factor {ef}, seed {seed}",
                   "n": int(g.n), "m": int(g.nnz) // 2,
                   "methods": methods},
        "methods_agree": agree, "prep_s": prep_s, **rows,
    }
    print(json.dumps(out), flush=True)
    return 0 if agree else 1


def run_serve(ma, args, rate: float, n: int, single: bool = True,
              graphs_specs=None) -> dict:
    """The served stream: Poisson arrivals at `rate` req/s over resident
    RMAT graphs (default 64, scale uniform in [12, 20], ef 16, seed 6),
    pattern uniform over TC, P1..P8, through the continuous-batching server
    and through one launch per request (slots = lanes = max_batch = 1),
    same requests; latency = completion - arrival."""
    import numpy as np
    names = [p.strip() for p in args.serve_patterns.split(",")]
    pats = [ma.PATTERN_NAMES[p] for p in names]
    lo, hi = args.serve_scales
    rng = np.random.default_rng(6)
    specs = [int(rng.integers(lo, hi + 1)) for _ in range(args.serve_graphs)]
    t0 = time.perf_counter()
    if graphs_specs is None:
        graphs = [ma.Graph.rmat(s, 16, seed=6 * 100003 + i)
                  for i, s in enumerate(specs)]
    else:
        graphs = graphs_specs
    gen_s = time.perf_counter() - t0
    log(f"serve: {len(graphs)} graphs generated ({gen_s:.1f} s)")
    arrivals = np.cumsum(rng.exponential(1e6 / rate, size=n))
    req = [(int(rng.integers(len(graphs))), pats[int(rng.integers(len(pats)))])
           for _ in range(n)]

    def run(slots, max_batch, lanes, label):
        import threading
        srv = ma.Server(0, slots=slots, max_batch=max_batch, lanes=lanes)
        t0 = time.perf_counter()
        ids = [srv.add_graph(g) for g in graphs]
        add_s = time.perf_counter() - t0
        # warm-up: every pattern once on the smallest graph
        small = int(np.argmin(specs))
        srv.replay([(ids[small], p, 0) for p in pats])
        log(f"serve {label}: {len(ids)} graphs resident ({add_s:.1f} s)")
        # Open loop: a submitter thread follows the arrival schedule while
        # this thread collects the results in order.
        tickets = [None] * n
        t0 = time.perf_counter()

        def submit_all():
            for i, (g, p) in enumerate(req):
                due = t0 + arrivals[i] / 1e6
                while True:
                    left = due - time.perf_counter()
                    if left <= 0:
                        break
                    time.sleep(min(left, 1e-3))
                tickets[i] = srv.submit(ids[g], p)

        th = threading.Thread(target=submit_all)
        th.start()
        out = []
        for i in range(n):
            while tickets[i] is None:
                time.sleep(1e-4)
            out.append(srv.wait(tickets[i]))
            if (i + 1) % 50 == 0 or i + 1 == n:
                log(f"serve {label}: {i + 1}/{n} requests done "
                    f"({time.perf_counter() - t0:.1f} s)")
        th.join()
        wall = time.perf_counter() - t0
        st = srv.stats()
        srv.close()
        lat = np.array([r.latency_ms for r in out])
        ok = all(r.status == 0 for r in out)
        return {"qps": n / wall, "wall_s": wall,
                "p50_ms": float(np.percentile(lat, 50)),
                "p99_ms": float(np.percentile(lat, 99)),
                "mean_batch": float(np.mean([r.batch_size for r in out])),
                "launches": int(st.launches), "all_ok": ok,
                "busy_ms": st.busy_ms, "deferred": int(st.deferred),
                "add_graphs_s": add_s}, [r.count for r in out]

    served, counts = run(args.serve_slots, 1024, args.serve_lanes,
                         f"batched at {rate:g} req/s")
    if not single:
        return {"rate": rate, "requests": n, **served,
                "counts": counts, "graphs": graphs}
    single, counts1 = run(1, 1, 1, "one launch per request")
    # each result is the same alone: one single-query call per distinct
    # (graph, pattern) among the first requests
    dgs, alone_ok, seen = {}, True, set()
    for (g, p), c in zip(req, counts):
        if len(seen) >= args.serve_alone_checks:
            break
        if (g, p) in seen:
            continue
        seen.add((g, p))
        if g not in dgs:
            dgs[g] = ma.DeviceGraph(graphs[g], 0)
        alone_ok &= dgs[g].pattern_count(p)[0] == c
    for dg in dgs.values():
        dg.close()
    return {
        "workload": f"{args.serve_graphs} resident RMAT graphs (scale "
                    f"uniform in [{lo},{hi}], ef 16, seed 6), Poisson "
                    f"arrivals at {rate:g} req/s, {n} requests, pattern "
                    f"uniform in {{{','.join(names)}}}",
        "rate": rate, "requests": n,
        "slots": args.serve_slots, "lanes": args.serve_lanes,
        "hw_queues": os.environ.get("GPU_MAX_HW_QUEUES", "4"),
        "qps": served["qps"], "p50_ms": served["p50_ms"],
        "p99_ms": served["p99_ms"],
        "served": served, "one_launch_per_request": single,
        "qps_speedup": served["qps"] / single["qps"],
        "counts_equal_modes": counts == counts1,
        "alone_checked": len(seen), "counts_equal_alone": alone_ok,
        "graph_gen_s": gen_s, "graphs": graphs,
    }


def _guarded(what, fn, *a):
    """fn(*a), or {"error": ...} if it raised: a failing extra measurement
    must not take the triangle-count line with it."""
    try:
        return fn(*a)
    except Exception as e:  # noqa: BLE001
        log(f"{what} failed: {e!r}")
        return {"error": repr(e)}


def run_extras(ma, args) -> dict:
    """The other measurements of the default line, each a sub-object: the
    small CPU baselines, the batch workload (512 graphs, TC+P2+P4) and the
    served stream (at the offered rate, with its one-launch-per-request
    comparison, and at a rate below saturation)."""
    import copy
    extras = [e.strip() for e in args.extras.split(",") if e.strip()]
    out = {}
    if "cpu" in extras:
        out["cpu_small_baselines"] = _guarded("cpu baselines",
                                              cpu_small_baselines, ma)
    if "batch" in extras:
        # config 4 as one launch: the batch recorded into a HIP graph (the
        # timed batch plans, records and launches it; no warm-up batch)
        a = copy.copy(args)
        a.batch_warmup = 0
        a.batch_use_graph = 1
        a.batch_oracle_ef = min(args.batch_oracle_ef, 8)
        out["batch"] = _guarded("batch", run_batch, ma, a)
    if "serve" in extras:
        def serve():
            a = copy.copy(args)
            a.serve_alone_checks = args.extras_alone_checks
            r = run_serve(ma, a, args.serve_rate, args.extras_serve_requests)
            graphs = r.pop("graphs")
            rate = args.serve_low_rate or round(0.4 * r["served"]["qps"], 2)
            low = run_serve(ma, a, rate, args.serve_low_requests,
                            single=False, graphs_specs=graphs)
            low.pop("graphs")
            low.pop("counts")
            r["below_saturation"] = low
            return r
        out["serve"] = _guarded("serve", serve)
    return out


def bench_serve(ma, args) -> int:
    r = run_serve(ma, args, args.serve_rate, args.serve_requests)
    r.pop("graphs")
    out = {
        "metric": "served queries/s", "value": r["qps"], "unit": "queries/s",
        "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": r["served"]["wall_s"] * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "int32", "data": "synthetic",
        "config": {"workload": "serve: " + r["workload"]}, **r,
    }
    print(json.dumps(out), flush=True)
    return 0 if (r["counts_equal_modes"] and r["counts_equal_alone"]) else 1


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="tc_rmat22", choices=WORKLOADS)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--pattern-graph", type=int, nargs=3, default=(20, 16, 4),
                    metavar=("SCALE", "EF", "SEED"),
                    help="patterns mode: the RMAT graph")
    ap.add_argument("--patterns", default="P1,P4,P7,P8")
    ap.add_argument("--pattern-methods", default="census",
                    help="patterns mode: census and/or kernels")
    ap.add_argument("--tc-layout", type=int, default=0,
                    help="tc mode: 1 = stream laid-out partner lists, -1 = "
                         "gather them from the DAG, 0 = the library's "
                         "default (lay out when the copy fits)")
    ap.add_argument("--tc-merge-max", type=int, default=0,
                    help="tc mode: owners with d+ up to this (at most 64) "
                         "merge partner lists instead of hashing")
    ap.add_argument("--tc-half-waves", type=int, default=0,
                    help="tc mode: 1 = small-class tasks two per wave, -1 = "
                         "one per wave, 0 = the library's default")
    ap.add_argument("--tc-items", type=int, default=0,
                    help="tc mode: list items per lane per step (4 or 8; 0 = "
                         "the library's default)")
    ap.add_argument("--no-device-plan", action="store_true",
                    help="skip the host-plan vs device-plan upload timing")
    ap.add_argument("--mode", default="tc",
                    choices=("tc", "batch", "serve", "dag", "patterns"),
                    help="tc: one graph's triangle count (default); batch: "
                         "many graphs' queries in one launch vs one launch "
                         "per query; serve: Poisson request stream through "
                         "the continuous-batching server vs one launch per "
                         "request")
    ap.add_argument("--serve-graphs", type=int, default=64)
    ap.add_argument("--serve-scales", type=int, nargs=2, default=(12, 20),
                    metavar=("LO", "HI"))
    ap.add_argument("--serve-patterns", default="TC,P1,P2,P3,P4,P5,P6,P7,P8")
    ap.add_argument("--serve-rate", type=float, default=2000.0)
    ap.add_argument("--serve-requests", type=int, default=2000)
    ap.add_argument("--serve-alone-checks", type=int, default=60)
    ap.add_argument("--serve-slots", type=int, default=3)
    ap.add_argument("--serve-lanes", type=int, default=4)
    ap.add_argument("--batch-lanes", type=int, default=8)
    ap.add_argument("--hw-queues", type=int, default=16,
                    help="batch/serve modes: GPU_MAX_HW_QUEUES for this "
                         "process (0 = leave the environment's value)")
    ap.add_argument("--batch-use-graph", type=int, default=0)
    ap.add_argument("--batch-graphs", type=int, default=512)
    ap.add_argument("--batch-scales", type=int, nargs=2, default=(10, 18),
                    metavar=("LO", "HI"))
    ap.add_argument("--batch-steps", type=int, default=1)
    ap.add_argument("--batch-warmup", type=int, default=1)
    ap.add_argument("--batch-patterns", default="TC,P2,P4")
    ap.add_argument("--batch-oracle-ef", type=int, default=16,
                    help="batch: the oracle checks the graphs of scale <= 10 "
                         "and edge factor <= this")
    ap.add_argument("--extras", default="cpu,batch,serve",
                    help="tc mode: sub-objects added to the line (cpu: the "
                         "small CPU baselines; batch: the batch workload; "
                         "serve: the served stream); '' for none")
    ap.add_argument("--extras-serve-requests", type=int, default=60)
    ap.add_argument("--extras-alone-checks", type=int, default=20)
    ap.add_argument("--serve-low-rate", type=float, default=0.0,
                    help="below-saturation rate (0: 0.4 x the batched "
                         "throughput at the offered rate)")
    ap.add_argument("--serve-low-requests", type=int, default=12)
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="run only a plain device-to-device copy of MIB MiB "
                         "(5 times) to calibrate rocprofv3 FETCH_SIZE / "
                         "WRITE_SIZE against known bytes")
    args = ap.parse_args()
    if args.calibrate_copy:
        return calibrate_copy(args.calibrate_copy)
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the counting path here runs on a "
              "single MI355X and has no multi-GPU mode; run with --gpus 1",
              file=sys.stderr)
        return 2

    if (args.mode in ("batch", "serve") or
            (args.mode == "tc" and args.extras.strip())) and args.hw_queues > 0:
        # Read by the HIP runtime when it starts (before any HIP call): how
        # many hardware queues this process's streams may spread over, so
        # how many lanes of a batch can run on the device at once.
        os.environ["GPU_MAX_HW_QUEUES"] = str(args.hw_queues)
    import mercury_amd as ma
    if ma.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        return 1
    if args.mode == "batch":
        return bench_batch(ma, args)
    if args.mode == "serve":
        return bench_serve(ma, args)
    if args.mode == "dag":
        return bench_dag(ma, args)
    if args.mode == "patterns":
        return bench_patterns(ma, args)
    kind, scale, ef, seed = WORKLOADS[args.workload]
    t0 = time.perf_counter()
    g = make_graph(ma, kind, scale, ef, seed)
    gen_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    dg = ma.DeviceGraph(g, 0)
    upload_s = time.perf_counter() - t0
    kw = {k: v for k, v in (("tc_layout", args.tc_layout),
                            ("tc_merge_max", args.tc_merge_max),
                            ("tc_items", args.tc_items),
                            ("tc_half_waves", args.tc_half_waves)) if v}
    tune = ma.make_tune(**kw) if kw else None
    # plan preparation outside the warm-up and timed steps: the first count
    # lays the partner lists out (when it does), reported as prep
    prep_ms = dg.tc_count(tune)[1].prep_ms
    for _ in range(args.warmup):
        dg.tc_count(tune)
    ma.native.check(ma.native.lib.mgpm_device_synchronize(0), "sync")
    kernel_ms = []
    counts = set()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        c, st = dg.tc_count(tune)
        kernel_ms.append(st.kernel_ms)
        counts.add(c)
    ma.native.check(ma.native.lib.mgpm_device_synchronize(0), "sync")
    wall = time.perf_counter() - t0
    count = counts.pop() if len(counts) == 1 else None
    ms_per_step = wall * 1e3 / args.steps
    avg_kernel_ms = sum(kernel_ms) / len(kernel_ms)
    achieved = st.bytes_alg / (avg_kernel_ms / 1e3)
    out = {
        "metric": "TC edges/s",
        "value": st.m / (ms_per_step / 1e3),
        "unit": "edges/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {
            "workload": f"{args.workload}: triangle count, {kind.upper()} "
                        f"scale {scale}, edge factor {ef}, seed {seed}, "
                        "symmetrized, self-loops/duplicates removed, "
                        "degree-oriented DAG, int32 CSR",
            "n": st.n, "m": st.m, "m_plus": st.m_plus, "s_plus": st.s_plus,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved / 1e9,
            "peak": HBM_PEAK / 1e9,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK,
            "traffic": load_traffic(args.workload),
            "bytes_alg": st.bytes_alg,
            "kernel_ms_avg": avg_kernel_ms,
        },
        "count": count,
        "probes": st.probes,
        "laid_out_lists": st.path == 2,
        "layout_build_ms": prep_ms,
        "kernel_edges_per_s": st.m / (avg_kernel_ms / 1e3),
        "gen_s": gen_s,
        "upload_prep_s": upload_s,
    }
    if not args.no_device_plan:
        out.update(upload_compare(ma, g, count))
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(g, count, st.m)
    dg.close()
    if args.extras.strip():
        out.update(run_extras(ma, args))
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
