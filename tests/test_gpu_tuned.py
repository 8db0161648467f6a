"""GPU: the library applies the autotuner's table (the tuned knobs of a
workload's signature) where the caller leaves the knobs at their defaults,
in single-matrix plans, batches and the request server; results stay
bit-exact.  Runs in a subprocess whose GSOFA_TUNE_CACHE names a table written
here (the library reads its table once per process)."""
import os
import subprocess
import sys

import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu

CHILD = r"""
import json, os, sys
sys.path.insert(0, sys.argv[1]); sys.path.insert(0, os.path.join(sys.argv[1], "tests"))
import numpy as np
import gsofa_amd as g
import gsofa_amd.configs as cf
from conftest import random_matrix
from gpu_util import assert_same_structure
from oracle import oracle
d = np.load(os.path.join(sys.argv[1], "tests", "golden", "bcsstk18.npz"))
A = g.CsrMatrix(int(d["n"]), d["colbeg"], d["colend"], d["csr"])
o = oracle.symbfact(A.colbeg, A.colend, A.csr, A.n)
r = g.symbfact(A)  # tuned: reach, 32 slots, 5 workgroups
print("single", r.mode, r.grid_blocks, r.chunk, flush=True)
assert (r.mode, r.grid_blocks, r.chunk) == ("reach", 5, 160)
assert_same_structure(r, o)
r = g.symbfact(A, block_slots=16)  # the caller's knob stands
assert r.mode == "block" and r.chunk // r.grid_blocks == 16
mats = [cf.batch_matrix(s, b) for b, s in enumerate([16, 20, 24])]
with g.Plan(mats) as p:  # tuned batch: chunk 128 = 2 workgroups of 64
    p.execute()
    for b, M in enumerate(mats):
        rb = p.fetch_matrix(b)
        assert rb.grid_blocks == 2, rb.grid_blocks
        assert_same_structure(rb, oracle.symbfact(M.colbeg, M.colend, M.csr, M.n))
with g.Server(max_inflight=1, max_batch=8) as srv:  # the server's batch takes the same entry
    ts = [srv.submit(M) for M in mats]
    for t, M in zip(ts, mats):
        rs, info = srv.wait(t)
        if info["batch_size"] == 3:  # all three in one batch: the batch's entry
            assert rs.grid_blocks == 2, (info, rs.grid_blocks)
        assert_same_structure(rs, oracle.symbfact(M.colbeg, M.colend, M.csr, M.n))
G = g.gen_grid(20, 20, 20, nd_leaf=64, extra_prob=0.05, extra_dist=2, seed=9)
r = g.symbfact(G)  # tuned: reach, 64 slots, 512-thread workgroups, 7 of them
assert (r.mode, r.grid_blocks, r.chunk) == ("reach", 7, 448), (r.mode, r.grid_blocks, r.chunk)
assert_same_structure(r, oracle.symbfact(G.colbeg, G.colend, G.csr, G.n))
print("tuned ok", flush=True)
"""


def test_library_applies_tuned_table(tmp_path):
    import numpy as np

    import gsofa_amd as g
    import gsofa_amd.configs as cf
    from gsofa_amd import tune
    d = np.load(os.path.join(ROOT, "tests", "golden", "bcsstk18.npz"))
    A = g.CsrMatrix(int(d["n"]), d["colbeg"], d["colend"], d["csr"])
    mats = [cf.batch_matrix(s, b) for b, s in enumerate([16, 20, 24])]
    G = g.gen_grid(20, 20, 20, nd_leaf=64, extra_prob=0.05, extra_dist=2, seed=9)
    table = str(tmp_path / "tune_table.txt")
    tune.write_table({tune.signature(A): {"mode": "reach", "block_slots": 32, "grid_blocks": 5},
                      tune.signature(mats): {"mode": "reach", "block_slots": 64, "chunk": 128},
                      tune.signature(G): {"mode": "reach", "block_slots": 64, "grid_blocks": 7,
                                          "workgroup_threads": 512}}, table)
    assert tune.read_table(table)[tune.signature(G)]["workgroup_threads"] == 512
    assert tune.signature(G) not in (tune.signature(A), tune.signature(mats))
    env = dict(os.environ, GSOFA_TUNE_CACHE=table)
    print("case: tuned table in a subprocess", flush=True)
    p = subprocess.run([sys.executable, "-c", CHILD, ROOT], env=env, capture_output=True, text=True, timeout=180)
    print(p.stdout[-2000:], p.stderr[-2000:], flush=True)
    assert p.returncode == 0 and "tuned ok" in p.stdout
