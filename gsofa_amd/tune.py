"""Autotuner for the runtime knobs of the traversal.

Every knob here leaves the results unchanged (the parity tests check this);
only time depends on them, so the tuner compares traversal times alone, with
counts only (no L/U structure).  The knobs are the ones the reference fixed
at compile time (with_space_optimization/symbfact.cuh:42,49,57,77,102):

* reach mode: sources per workgroup (64, 32, 128 or 256), threads per workgroup (1024,
  one per CU with the largest target table, or 512, two per CU) and, for
  batches, the concurrent-source chunk (64 .. 4096 sources, i.e. workgroups x
  slots);
* block mode: sources per workgroup (32, 16, 8) and the hybrid split (top
  sources run as whole-device groups);
* grid mode: concurrent-source chunk, the wave/workgroup expansion threshold
  and the frontier queue capacity.

It times the whole workload; the remaining candidates of a mode are skipped
once that mode's first candidate took over 4x the best so far.  The choice is written, per workload signature,
to a text table that the library itself reads (one line per signature,
``<signature> key=value ...``): gsofa_plan_create, gsofa_batch_create and the
request server's batches apply the entry's knobs wherever the caller left
them at their defaults.  The library reads the table named by the
GSOFA_TUNE_CACHE environment variable ("off": none), else tune_table.txt
beside libgsofa.so.  The signature buckets the largest matrix's vertex count
and the mean row length to powers of two, and a batch's matrix count to the
next power of two (gsofa_api.hip, tune_signature).
"""
import math
import os
import time

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_TABLE = os.path.join(HERE, "tune_table.txt")
MODES = {"auto": 0, "block": 1, "grid": 2, "hybrid": 3, "reach": 4}
KEYS = ("mode", "block_slots", "chunk", "grid_blocks", "top_sources", "wave_block_threshold",
        "queue_capacity", "workgroup_threads")

REACH_CANDIDATES = [dict({"mode": "reach", "block_slots": b}, **({"workgroup_threads": 512} if t == 512 else {}))
                    for t in (1024, 512) for b in (64, 32, 128, 256)]
BLOCK_CANDIDATES = ([{"mode": "block", "block_slots": s} for s in (32, 16, 8)] +
                    [{"mode": "block", "top_sources": k} for k in (2048, 8192)])
GRID_CANDIDATES = [
    {"mode": "grid", "chunk": c, "wave_block_threshold": th, "queue_capacity": q}
    for c in (2048, 8192) for th in (256, 4096) for q in (0, 1 << 22)
]
# batches: the concurrent-source chunk, largest first (the sweep stops once a
# candidate takes 4x the best so far: fewer sources in flight only get slower)
BATCH_CHUNKS = (4096, 1024, 256, 64)


def signature(A):
    """Bucketed workload signature of a matrix or a list of matrices (the
    library's tune_signature)."""
    mats = A if isinstance(A, (list, tuple)) else [A]
    n = max(M.n for M in mats)
    nnz = sum(M.nnz for M in mats)
    tot = sum(M.n for M in mats)
    deg = nnz / max(tot, 1)
    b = 1 if len(mats) <= 1 else 2 ** int(math.ceil(math.log2(len(mats))))
    return "n2^%d_deg2^%d_b%d" % (int(math.ceil(math.log2(max(n, 2)))),
                                   int(round(math.log2(max(deg, 1.0)))), b)


def read_table(path=DEFAULT_TABLE):
    """{signature: options} of a tuned table ({} if absent)."""
    out = {}
    try:
        with open(path) as f:
            for line in f:
                parts = line.split()
                if not parts or parts[0].startswith("#"):
                    continue
                opts = {}
                for kv in parts[1:]:
                    k, _, v = kv.partition("=")
                    if k in KEYS:
                        opts[k] = int(v)
                names = {v: k for k, v in MODES.items()}
                if "mode" in opts:
                    opts["mode"] = names.get(opts["mode"], opts["mode"])
                out[parts[0]] = opts
    except OSError:
        pass
    return out


def write_table(table, path=DEFAULT_TABLE):
    """Write {signature: options} in the library's format (atomically)."""
    d = os.path.dirname(path)
    if d:
        os.makedirs(d, exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        f.write("# gsofa_amd tuned knobs: <signature> key=value ... (written by gsofa_amd/tune.py)\n")
        for sig in sorted(table):
            opts = table[sig]
            kv = []
            for k in KEYS:
                if k in opts:
                    v = MODES[opts[k]] if k == "mode" else int(opts[k])
                    kv.append("%s=%d" % (k, v))
            f.write("%s %s\n" % (sig, " ".join(kv)))
    os.replace(tmp, path)


def _time_candidate(A, opts, reps):
    import gsofa_amd as g
    best = float("inf")
    with g.Plan(list(A) if isinstance(A, (list, tuple)) else A, counts_only=True, **opts) as p:
        for _ in range(reps):
            p.execute()
            best = min(best, p.fetch().traverse_ms)
    return best


def candidates(A, grid_mode=True):
    batch = isinstance(A, (list, tuple))
    if batch:
        return list(REACH_CANDIDATES) + [{"mode": "block", "block_slots": s} for s in (16, 8)]
    return list(REACH_CANDIDATES) + BLOCK_CANDIDATES + (GRID_CANDIDATES if grid_mode else [])


def tune(A, reps=1, grid_mode=True, table=DEFAULT_TABLE, log=None):
    """Time every candidate on A (a matrix, or a list of matrices as one
    batch) and return (best_options, timings).  The choice is recorded in the
    library's table at `table` (None: not recorded)."""
    batch = isinstance(A, (list, tuple))
    timings = []

    def run(c):
        t0 = time.perf_counter()
        ms = _time_candidate(A, c, reps)
        timings.append({"options": c, "traverse_ms": ms})
        if log:
            log("tune %s: %.2f ms (%.1f s)" % (c, ms, time.perf_counter() - t0))
        return ms

    slow = set()  # modes whose first candidate took over 4x the best so far
    for c in candidates(A, grid_mode):
        if c["mode"] in slow:
            timings.append({"options": c, "skipped": True})
            continue
        first = not any(t["options"]["mode"] == c["mode"] for t in timings)
        ms = run(c)
        if first and ms > 4 * min(t["traverse_ms"] for t in timings if "traverse_ms" in t):
            slow.add(c["mode"])
    if batch:  # the chunk sweep of the best reach shape
        shape = dict(min((t for t in timings if t["options"]["mode"] == "reach"),
                         key=lambda t: t["traverse_ms"])["options"])
        best_ms = min(t["traverse_ms"] for t in timings if "traverse_ms" in t)
        for chunk in BATCH_CHUNKS:
            ms = run(dict(shape, chunk=chunk))
            best_ms = min(best_ms, ms)
            if ms > 4 * best_ms:
                break
    best = min((t for t in timings if "traverse_ms" in t), key=lambda r: r["traverse_ms"])
    if table:
        t = read_table(table)
        t[signature(A)] = dict(best["options"])
        write_table(t, table)
    return dict(best["options"]), timings


def best_options(A, table=DEFAULT_TABLE):
    """Tuned options for A's signature, or {} (library defaults) if untuned."""
    return dict(read_table(table).get(signature(A), {}))


def main(argv=None):
    """python -m gsofa_amd.tune [--out FILE] [--table FILE]: tune the standard
    workloads, record the choices in the library's table and print one JSON
    record per workload."""
    import argparse
    import json
    import sys

    import numpy as np

    import gsofa_amd as g
    import gsofa_amd.configs as cf
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=None, help="also write the records to this JSON file")
    ap.add_argument("--table", default=DEFAULT_TABLE)
    ap.add_argument("--only", default=None, help="comma-separated workload names")
    args = ap.parse_args(argv)
    d = np.load(os.path.join(os.path.dirname(HERE), "tests", "golden", "bcsstk18.npz"))
    work = [("grid64", cf.grid64_natural()),
            ("bcsstk18", g.CsrMatrix(int(d["n"]), d["colbeg"], d["colend"], d["csr"])),
            ("grid3d_nd", cf.grid3d_nd()),
            ("band", cf.band()),
            ("grid2d_nd_batch", cf.grid2d_nd_batch())]
    if args.only:
        keep = set(args.only.split(","))
        work = [w for w in work if w[0] in keep]
    out = []
    for name, A in work:
        best, timings = tune(A, reps=1, table=args.table,
                             log=lambda *a: print(*a, file=sys.stderr, flush=True))
        rec = {"workload": name, "signature": signature(A), "best": best, "timings": timings}
        print(json.dumps(rec), flush=True)
        out.append(rec)
    if args.out:
        with open(args.out, "w") as f:
            json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
