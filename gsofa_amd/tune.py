"""Autotuner for the runtime knobs of the traversal.

The knobs that remain at run time are the execution mode, the persistent grid
size, the number of top sources run as whole-device groups and the sources
per workgroup (block mode),
and the concurrent-source chunk and wave/block expansion threshold (grid
mode).  Results do not depend on any knob (the parity tests
check this), only time does, so the tuner compares traversal times alone,
with counts only (no L/U structure).

It times the whole workload by default.  A sample of the sources (one
virtual-GPU share, with_space_optimization/symbfact.cuh:1084-1099) is
available with shares > 1 but misleads across modes: block mode's time on a
small sample is bounded by the latency of its deepest 32-source chunk rather
than by the work (a 1/16 share of the 48^3 ND grid takes as long as the
whole matrix in block mode), while grid mode scales with the work.

Choices are cached in a JSON file keyed by a workload signature: the vertex
count and mean row length rounded to powers of two, and for batches the
matrix count.
"""
import json
import math
import os
import time

DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "gsofa_amd", "tune.json")

# runtime knobs of gsofa_options that never change results
BLOCK_CANDIDATES = [
    {"mode": "block"},
    {"mode": "block", "grid_blocks": 384},
] + [{"mode": "block", "top_sources": k} for k in (2048, 8192, 32768)] + [
    {"mode": "block", "block_slots": s} for s in (16, 8)]
GRID_CANDIDATES = [
    {"mode": "grid", "chunk": c, "wave_block_threshold": th}
    for c in (2048, 8192) for th in (256, 4096)
]


def signature(A):
    """Bucketed workload signature of a matrix or a list of matrices."""
    mats = A if isinstance(A, (list, tuple)) else [A]
    n = max(M.n for M in mats)
    nnz = sum(M.nnz for M in mats)
    tot = sum(M.n for M in mats)
    deg = nnz / max(tot, 1)
    return "n2^%d_deg2^%d_b%d" % (int(math.ceil(math.log2(max(n, 2)))),
                                   int(round(math.log2(max(deg, 1.0)))),
                                   len(mats) if len(mats) == 1 else 2 ** int(math.ceil(math.log2(len(mats)))))


def _load(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(path, table):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def _time_candidate(A, opts, shares, reps):
    import gsofa_amd as g
    kw = dict(opts)
    best = float("inf")
    if isinstance(A, (list, tuple)):
        with g.Plan(list(A), counts_only=True, **kw) as p:
            for _ in range(reps):
                p.execute()
                best = min(best, p.fetch().traverse_ms)
        return best
    with g.Plan(A, counts_only=True, n_virtual_gpu=shares, virtual_gpu_id=0, **kw) as p:
        for _ in range(reps):
            p.execute()
            best = min(best, p.fetch().traverse_ms)
    return best


def tune(A, shares=1, reps=1, grid_mode=True, cache=DEFAULT_CACHE, log=None):
    """Time every candidate on A (a matrix, or a list of matrices as one
    batch; shares > 1: on one virtual-GPU share only) and return
    (best_options, timings).  The choice is recorded in `cache` (None: not
    recorded)."""
    batch = isinstance(A, (list, tuple))
    cands = list(BLOCK_CANDIDATES)
    if grid_mode and not batch:
        cands += GRID_CANDIDATES
    shares = max(1, min(shares, (A.n if not batch else 1)))
    timings = []
    for c in cands:
        t0 = time.perf_counter()
        ms = _time_candidate(A, c, shares, reps)
        timings.append({"options": c, "traverse_ms": ms})
        if log:
            log("tune %s: %.2f ms (%.1f s)" % (json.dumps(c), ms, time.perf_counter() - t0))
    best = min(timings, key=lambda r: r["traverse_ms"])
    if cache:
        table = _load(cache)
        table[signature(A)] = {"options": best["options"], "traverse_ms": best["traverse_ms"],
                               "shares": shares if not batch else 1}
        _store(cache, table)
    return dict(best["options"]), timings


def best_options(A, cache=DEFAULT_CACHE):
    """Tuned options for A's signature, or {} (library defaults) if untuned."""
    entry = _load(cache).get(signature(A)) if cache else None
    return dict(entry["options"]) if entry else {}


def main(argv=None):
    """python -m gsofa_amd.tune [--out FILE]: tune the standard workloads and
    print one JSON record per workload."""
    import argparse
    import sys

    import gsofa_amd.configs as cf
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=None, help="also write the records to this JSON file")
    ap.add_argument("--cache", default=DEFAULT_CACHE)
    args = ap.parse_args(argv)
    work = [("grid64 natural", cf.grid64_natural(), 1),
            ("3D 48^3 ND", cf.grid3d_nd(), 1),
            ("band n=659,033", cf.band(), 1),
            ("batch of 256 ND 2D grids", cf.grid2d_nd_batch(), 1)]
    out = []
    for name, A, shares in work:
        best, timings = tune(A, shares=shares, reps=1, cache=args.cache,
                             log=lambda *a: print(*a, file=sys.stderr, flush=True))
        rec = {"workload": name, "signature": signature(A), "shares_sampled": shares,
               "best": best, "timings": timings}
        print(json.dumps(rec), flush=True)
        out.append(rec)
    if args.out:
        with open(args.out, "w") as f:
            json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
