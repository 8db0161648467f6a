"""Benchmark: GTEPS of the fill-path symbolic factorization on one MI355X.

A step is one full symbolic factorization of the n=659,033 banded matrix
(traversal + sorted L/U structure + supernodes) with the matrix already
resident in HBM.  value = TE_alg / step time, where TE_alg = sum over sources
s of [deg(s) + sum_{v in R_<(s)} deg(v)] (deterministic, counted on device).

Prints one JSON line with roofline (algorithmic bytes of the traversal kernel
over its event-timed duration, against 8 TB/s) and a CPU baseline (the OpenMP
oracle timed on this host on a bounded sample of the same matrix).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "GTEPS + HBM GB/s at n=659K; batched matrices/s & p99 vs one-by-one"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def workload(name, n_override=0):
    import gsofa_amd as g
    import gsofa_amd.configs as cf
    if name == "band":
        n = n_override or 659033
        return cf.band(n), "band n=%d |i-j|<=256 Poisson(8) seed 3" % n
    if name == "grid3d_nd":
        return cf.grid3d_nd(), "3D 7-point 48^3 nested dissection + 5% extra, seed 2"
    if name == "grid64":
        return cf.grid64_natural(), "2D 5-point 64x64 natural order"
    if name == "bcsstk18":
        import numpy as np
        d = np.load(os.path.join(ROOT, "tests", "golden", "bcsstk18.npz"))
        return g.CsrMatrix(int(d["n"]), d["colbeg"], d["colend"], d["csr"]), "BCSSTK18"
    raise SystemExit("unknown workload " + name)


def algorithmic_bytes(n, reach, te, nnz_lu):
    # 8 B row begin/end per expanded vertex and source, 8 B per edge check
    # (CSR word + label/marker probe), 16 B per frontier pair written and read,
    # 4 B per output entry.
    return 8 * (n + reach) + 8 * te + 16 * reach + 4 * nnz_lu


def host_info():
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "cpu_model": model}


def cpu_threads():
    """All the cores this process may use: the affinity set, capped by
    OMP_NUM_THREADS where it is set (the GPU hosts give a GPU 16)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    try:
        n = min(n, int(os.environ.get("OMP_NUM_THREADS", n)))
    except ValueError:
        pass
    return max(1, n)


def cpu_full(threads):
    """The oracle's timed traversal (counters only, best of 3) over whole
    workloads: C1, BCSSTK18, C2 and the C4 batch."""
    import numpy as np

    import gsofa_amd as g
    import gsofa_amd.configs as cf
    from oracle import oracle
    d = np.load(os.path.join(ROOT, "tests", "golden", "bcsstk18.npz"))
    work = [("grid64 natural (C1)", [cf.grid64_natural()]),
            ("BCSSTK18", [g.CsrMatrix(int(d["n"]), d["colbeg"], d["colend"], d["csr"])]),
            ("3D 48^3 ND (C2)", [cf.grid3d_nd()]),
            ("256 ND 2D grids (C4)", cf.grid2d_nd_batch())]
    out = []
    for name, mats in work:
        best, te = float("inf"), 0
        for _ in range(3 if sum(M.n for M in mats) < 200000 else 1):
            sec, te = 0.0, 0
            for M in mats:
                s_, f_, t_, r_ = oracle.count_rows(M.colbeg, M.colend, M.csr, M.n, 0, M.n, threads)
                sec += s_
                te += t_
            best = min(best, sec)
        out.append({"workload": name, "seconds": round(best, 3), "te_alg": te,
                    "gteps": round(te / best / 1e9, 3)})
        log("cpu full:", json.dumps(out[-1]))
    return out


def cpu_baseline(A, target_s, threads):
    """Oracle on the top rows of the matrix (the heaviest), sized to ~target_s."""
    from oracle import oracle
    n = A.n
    rows = 8
    while True:
        sec, f, te, r = oracle.count_rows(A.colbeg, A.colend, A.csr, n, n - rows, n, threads)
        if sec >= target_s / 4 or rows >= n:
            break
        rows = min(n, max(rows * 2, int(rows * target_s / 4 / max(sec, 1e-3))))
    if sec < target_s / 2 and rows < n:
        rows = min(n, int(rows * target_s / max(sec, 1e-3)))
        sec, f, te, r = oracle.count_rows(A.colbeg, A.colend, A.csr, n, n - rows, n, threads)
    rec = {"value": te / sec / 1e9, "unit": "GTEPS", "cores": threads, "kind": "port",
           "sample": "oracle traversal (OpenMP, schedule dynamic) of rows [%d, %d) of the same "
                     "matrix: %d rows, TE=%d edge checks in %.2f s" % (n - rows, n, rows, te, sec)}
    rec.update(host_info())
    return rec


def copy_probe(nbytes=1 << 30, reps=3):
    """A plain device-to-device copy of known size: under rocprofv3 --pmc its
    FETCH_SIZE / WRITE_SIZE per dispatch calibrate the traffic counters."""
    import torch
    a = torch.ones(nbytes // 4, dtype=torch.int32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes, "reps": reps}),
          flush=True)


def batch_bench(steps, warmup):
    """Batch: 256 ND-ordered 2D grids (side log-uniform in [16,256]).

    batched:    one plan over all 256 matrices, one persistent launch per step
    one_by_one: the plain path, one plan and one launch per matrix; each plan is
                created untimed, then its one execute is timed
    Both run with the matrices resident in HBM; results stay on the device.
    """
    import numpy as np

    import gsofa_amd as g
    import gsofa_amd.configs as cf
    mats = cf.grid2d_nd_batch()
    with g.Plan(mats) as p:
        for _ in range(warmup):
            p.execute()
        times = []
        for _ in range(steps):
            t0 = time.perf_counter()
            p.execute()
            times.append(time.perf_counter() - t0)
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from digests import lu_digest
        recs = []
        for b in range(len(mats)):
            r = p.fetch_matrix(b)
            recs.append((r.fill_count, lu_digest(r.L_rowptr, r.L_col, r.U_rowptr, r.U_col)))
        fills = [f for f, _ in recs]
    batched_s = sum(times) / len(times)
    with g.Plan(mats, graph=True) as p:
        p.execute()  # runs, then captures the graph
        gtimes = []
        for _ in range(max(steps, 1)):
            t0 = time.perf_counter()
            p.execute()
            gtimes.append(time.perf_counter() - t0)
        gfills = [p.fetch_matrix(b).fill_count for b in range(len(mats))]
        replayed = p.fetch_matrix(0).graph_replay
    graph_s = sum(gtimes) / len(gtimes)
    lat = []
    for A in mats:
        with g.Plan(A) as q:
            t0 = time.perf_counter()
            q.execute()
            lat.append(time.perf_counter() - t0)
    obo_s = sum(lat)
    gold_ok = None
    gpath = os.path.join(ROOT, "tests", "golden", "goldens.json")
    if os.path.exists(gpath):
        gold = json.load(open(gpath)).get("grid2d_nd_batch")
        if gold:
            gold_ok = ([x["fill_count"] for x in gold] == fills == gfills and
                       [x["lu"] for x in gold] == [d for _, d in recs])
    lat_ms = np.array(lat) * 1e3
    return {"workload": "256 ND 2D grids, side log-uniform [16,256], seed 4",
            "matrices": len(mats), "vertices": int(sum(A.n for A in mats)),
            "batched_ms": round(batched_s * 1e3, 3),
            "batched_matrices_per_s": round(len(mats) / batched_s, 1),
            "batched_graph_ms": round(graph_s * 1e3, 3), "graph_replayed": bool(replayed),
            "batched_graph_matrices_per_s": round(len(mats) / graph_s, 1),
            "one_by_one_ms": round(obo_s * 1e3, 3),
            "one_by_one_matrices_per_s": round(len(mats) / obo_s, 1),
            "speedup_vs_one_by_one": round(obo_s / batched_s, 2),
            "one_by_one_latency_ms_p50": round(float(np.percentile(lat_ms, 50)), 3),
            "one_by_one_latency_ms_p99": round(float(np.percentile(lat_ms, 99)), 3),
            "matches_golden": gold_ok}


def bcsstk18_bench(steps, warmup):
    """BCSSTK18 (the reference's shipped input, n=11,948): one plan, `steps`
    timed executes (traversal + L/U + supernodes, matrix resident), beside the
    reference's logged times for the same input (35.63-42.12 ms per run,
    with_space_optimization/space_optimization_log_1_264_GPU_warp_centric.dat:4,16,...)."""
    import numpy as np

    import gsofa_amd as g
    d = np.load(os.path.join(ROOT, "tests", "golden", "bcsstk18.npz"))
    A = g.CsrMatrix(int(d["n"]), d["colbeg"], d["colend"], d["csr"])
    with g.Plan(A) as p:
        for _ in range(max(warmup, 1)):
            p.execute()
        times = []
        for _ in range(max(steps, 1)):
            t0 = time.perf_counter()
            p.execute()
            times.append(time.perf_counter() - t0)
        r = p.fetch()
    ms = 1e3 * sum(times) / len(times)
    return {"workload": "BCSSTK18 n=11948", "ms_per_execute": round(ms, 3),
            "traverse_ms": round(r.traverse_ms, 3), "device_ms": round(r.ms, 3),
            "gteps": round(r.te_alg / (ms / 1e3) / 1e9, 2), "fill_count": r.fill_count,
            "matches_reference_log": r.fill_count == 980136, "mode": r.mode,
            "reference_logged_ms": [35.63, 42.12]}


def serve_bench(count, rates, inflight=4, max_batch=64):
    """Serving: seeded Poisson streams (one per rate, req/s) of the batch
    family's 2D grids and 3D 7-point ND grids (side 16..64), submitted at their
    arrival times to the request server (continuous batching: `inflight`
    workers, each with a device-resident arena) and, as the one-by-one
    baseline, run through the plain one-matrix path (gsofa_symbfact: plan,
    execute, fetch, release) in arrival order.  Latency = arrival -> result on
    the host, as numpy arrays (both paths include the same conversion)."""
    import numpy as np

    import gsofa_amd as g
    import gsofa_amd.configs as cf

    def served(srv, reqs):
        # clients submit at their arrival times from a few threads (a submit
        # packs its matrix in the caller's thread), as independent callers would
        from concurrent.futures import ThreadPoolExecutor
        t0 = time.perf_counter()
        futs = []
        with ThreadPoolExecutor(max_workers=8) as pool:
            for ta, _, _, A in reqs:
                now = time.perf_counter() - t0
                if ta > now:
                    time.sleep(ta - now)
                futs.append(pool.submit(srv.submit, A))
            tickets = [f.result() for f in futs]
        out = [srv.wait(t) for t in tickets]
        span = time.perf_counter() - t0
        lat = [i["latency_ms"] + 0.0 for _, i in out]
        return [r for r, _ in out], lat, span, len({i["batch_id"] for _, i in out})

    def one_by_one(reqs):
        t0 = time.perf_counter()
        res, lat = [], []
        for ta, _, _, A in reqs:
            now = time.perf_counter() - t0
            if ta > now:
                time.sleep(ta - now)
            res.append(g.symbfact(A))
            lat.append((time.perf_counter() - t0 - ta) * 1e3)
        return res, lat, time.perf_counter() - t0

    out = {"workload": "Poisson streams, %d requests each, seed 5: 2D ND grids (side 16..256) and 3D 7-point "
                       "ND grids (side 16..64)" % count,
           "requests": count, "max_inflight": inflight, "max_batch": max_batch, "rates": []}
    g.symbfact(cf.batch_matrix(16, 0))  # warm-up of the one-shot path
    with g.Server(max_inflight=inflight, max_batch=max_batch) as srv:
        srv.wait(srv.submit(cf.batch_matrix(16, 0)))  # warm-up
        for rate in rates:
            reqs = cf.stream(count, rate)
            sres, slat, sspan, nb = served(srv, reqs)
            ores, olat, ospan = one_by_one(reqs)
            same = all(a.fill_count == b.fill_count and a.te_alg == b.te_alg and
                       np.array_equal(a.U_col, b.U_col) and np.array_equal(a.L_col, b.L_col) and
                       np.array_equal(a.super_start, b.super_start) for a, b in zip(sres, ores))
            te = sum(r.te_alg for r in sres)
            rec = {"rate": rate, "kinds_3d": sum(k for _, k, _, _ in reqs), "batches": nb,
                   "served_s": round(sspan, 3), "served_matrices_per_s": round(count / sspan, 2),
                   "served_gteps": round(te / sspan / 1e9, 2),
                   "latency_ms_p50": round(float(np.percentile(slat, 50)), 2),
                   "latency_ms_p99": round(float(np.percentile(slat, 99)), 2),
                   "one_by_one_s": round(ospan, 3), "one_by_one_matrices_per_s": round(count / ospan, 2),
                   "one_by_one_latency_ms_p50": round(float(np.percentile(olat, 50)), 2),
                   "one_by_one_latency_ms_p99": round(float(np.percentile(olat, 99)), 2),
                   "speedup_vs_one_by_one": round(ospan / sspan, 2), "results_equal_one_by_one": bool(same)}
            log("serve rate %g:" % rate, json.dumps(rec))
            out["rates"].append(rec)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="band", choices=["band", "grid3d_nd", "grid64", "bcsstk18"])
    ap.add_argument("--n", type=int, default=0, help="override n for the band workload")
    ap.add_argument("--chunk", type=int, default=0)
    ap.add_argument("--grid-blocks", type=int, default=0)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the 256-matrix batch leg")
    ap.add_argument("--no-serve", action="store_true", help="skip the request-stream leg")
    ap.add_argument("--serve-count", type=int, default=32)
    ap.add_argument("--serve-rates", default="50,200,500,1000,2000",
                    help="arrival rates (req/s) of the request-stream leg")
    ap.add_argument("--copy-probe", action="store_true",
                    help="run only a device copy of 1 GiB (HBM counter calibration)")
    ap.add_argument("--cpu-full", action="store_true",
                    help="run only the CPU oracle over whole workloads (C1, BCSSTK18, C2, C4)")
    ap.add_argument("--only", choices=["batch", "serve", "bcsstk18"],
                    help="run only this leg and print its record (profiling runs)")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe()
        return
    if args.cpu_full:
        th = cpu_threads()
        print(json.dumps({"cpu_full": cpu_full(th), "cores": th, **host_info()}), flush=True)
        return
    if args.only:
        leg = {"batch": lambda: batch_bench(args.steps, args.warmup),
               "serve": lambda: serve_bench(args.serve_count, [float(x) for x in args.serve_rates.split(",")]),
               "bcsstk18": lambda: bcsstk18_bench(args.steps, args.warmup)}[args.only]()
        print(json.dumps({args.only: leg}), flush=True)
        return
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; --gpus %d is not supported "
              "(sources shard with no exchange; see the sharding notes)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    import gsofa_amd as g
    A, desc = workload(args.workload, args.n)
    log("workload:", desc, "n=%d nnz=%d" % (A.n, A.nnz))
    plan = g.Plan(A, chunk=args.chunk, grid_blocks=args.grid_blocks)
    for i in range(args.warmup):
        t0 = time.time()
        plan.execute()
        log("warmup %d: %.1f ms" % (i, (time.time() - t0) * 1e3))
    res = plan.fetch() if args.warmup else None
    times, trav = [], []
    t_all = time.perf_counter()
    for i in range(args.steps):
        t0 = time.perf_counter()
        plan.execute()
        dt = time.perf_counter() - t0
        r = plan.fetch() if i == args.steps - 1 else None
        times.append(dt)
        log("step %d: %.1f ms" % (i, dt * 1e3))
    wall = time.perf_counter() - t_all
    res = r
    plan.close()
    ms_per_step = 1e3 * sum(times) / len(times)
    te = res.te_alg
    gteps = te / (ms_per_step / 1e3) / 1e9
    nnz_lu = int(res.L_rowptr[-1] + res.U_rowptr[-1]) if res.L_rowptr is not None else 0
    balg = algorithmic_bytes(A.n, res.reach, te, nnz_lu)
    achieved = balg / (res.traverse_ms / 1e3) / 1e9
    # HBM bytes per traversal launch, from the committed rocprofv3 FETCH_SIZE +
    # WRITE_SIZE passes over this same command, scaled by a copy kernel of
    # known size (profiles/traverse_<workload>_pmc.json)
    traffic = None
    pmc = os.path.join(ROOT, "profiles", "traverse_%s_pmc.json" % args.workload)
    if os.path.exists(pmc) and not args.n:
        try:
            traffic = json.load(open(pmc)).get("hbm_bytes_per_launch")
        except (OSError, ValueError):
            traffic = None
    # the timed run's structure against the oracle's golden: totals, the
    # digest of the sorted L/U rows, per-row fills and supernodes
    gold_ok = None
    gpath = os.path.join(ROOT, "tests", "golden", "goldens.json")
    if os.path.exists(gpath) and not args.n:
        gold = json.load(open(gpath)).get(args.workload)
        if gold and res.L_rowptr is not None:
            sys.path.insert(0, os.path.join(ROOT, "tests"))
            from digests import array_digest, lu_digest
            gold_ok = (gold["fill_count"] == res.fill_count and gold["te_alg"] == te and
                       gold["reach"] == res.reach and
                       gold["lu"] == lu_digest(res.L_rowptr, res.L_col, res.U_rowptr, res.U_col) and
                       gold["row_fill"] == array_digest(res.row_fill, "<i8") and
                       gold["super0"] == array_digest(res.super_start))
    # the autotuner's entry for this workload (applied by the library)
    from gsofa_amd import tune
    tuned = {"signature": tune.signature(A), "entry": tune.best_options(A) or None,
             "ran": {"mode": res.mode, "slots": res.chunk // max(res.grid_blocks, 1),
                     "grid_blocks": res.grid_blocks}}
    # measured HBM traffic of the traversal launch over its duration
    hbm_frac = None
    if traffic:
        hbm_frac = round(traffic / (res.traverse_ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4)
    out = {
        "metric": METRIC,
        "value": round(gteps, 3), "unit": "GTEPS", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "int32", "data": "synthetic",
        "config": {"workload": desc, "n": A.n, "nnz": A.nnz, "te_alg": te, "reach": res.reach,
                   "fill_count": res.fill_count, "nnz_LU": nnz_lu, "n_super": res.n_super,
                   "chunk": res.chunk, "grid_blocks": res.grid_blocks, "levels": res.levels,
                   "edge_checks": res.edge_checks, "traverse_ms": round(res.traverse_ms, 3),
                   "device_ms": round(res.ms, 3), "matches_golden": gold_ok,
                   "hbm_gbs_algorithmic": round(balg / (ms_per_step / 1e3) / 1e9, 1)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "frac_hbm_counters": hbm_frac},
        "tuned": tuned,
    }
    out["bcsstk18"] = bcsstk18_bench(args.steps, args.warmup)
    log("bcsstk18:", json.dumps(out["bcsstk18"]))
    if not args.no_batch:
        out["batch"] = batch_bench(args.steps, args.warmup)
        log("batch:", json.dumps(out["batch"]))
    if not args.no_serve:
        out["serve"] = serve_bench(args.serve_count, [float(x) for x in args.serve_rates.split(",")])
        log("serve:", json.dumps(out["serve"]))
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(A, args.cpu_seconds, cpu_threads())
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
