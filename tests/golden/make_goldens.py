"""Compute oracle goldens for the large workloads and write them as JSON.

Run in this repository (CPU only):
    python tests/golden/make_goldens.py [grid3d_nd] [band] [grid2d_nd_batch] [stream]

For each workload the CPU oracle (oracle/liboracle.so) produces the full
structure; the fixture keeps its totals and SHA-256 digests of the sorted L/U
arrays and of the supernode partitions, so the GPU parity tests can check
full-size results without rerunning the oracle.  Digest layout is defined in
tests/digests.py.
"""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import gsofa_amd.configs as cf  # noqa: E402
from oracle import oracle  # noqa: E402
from digests import oracle_record  # noqa: E402


STREAM_COUNT = 32


def run(name, A, out):
    t = time.time()
    r = oracle.symbfact(A.colbeg, A.colend, A.csr, A.n)
    rec = oracle_record(A, r)
    rec["oracle_seconds"] = round(time.time() - t, 1)
    out[name] = rec
    print(name, json.dumps(rec), flush=True)


def main(which):
    path = os.path.join(HERE, "goldens.json")
    out = json.load(open(path)) if os.path.exists(path) else {}
    if "grid3d_nd" in which:
        run("grid3d_nd", cf.grid3d_nd(), out)
    if "grid2d_nd_batch" in which:
        recs = []
        for b, A in enumerate(cf.grid2d_nd_batch()):
            r = oracle.symbfact(A.colbeg, A.colend, A.csr, A.n)
            rec = oracle_record(A, r)
            recs.append(rec)
        out["grid2d_nd_batch"] = recs
        print("grid2d_nd_batch", len(recs), flush=True)
    if "band" in which:
        run("band", cf.band(), out)
    if "stream" in which:
        # the request stream's matrices (they do not depend on the rate)
        recs = []
        t = time.time()
        for i, (_, kind, side, A) in enumerate(cf.stream(STREAM_COUNT, 500.0)):
            r = oracle.symbfact(A.colbeg, A.colend, A.csr, A.n)
            rec = oracle_record(A, r)
            rec["kind"], rec["side"] = kind, side
            recs.append(rec)
            print("stream", i, kind, side, A.n, round(time.time() - t, 1), flush=True)
        out["stream"] = recs
    with open(path, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main(sys.argv[1:] or ["grid3d_nd", "grid2d_nd_batch", "band", "stream"])
