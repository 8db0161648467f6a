"""GPU: plans created with GSOFA_FLAG_GRAPH run their first execute eagerly,
capture the device work (traversal, structure, supernodes, counters) in a HIP
graph and replay it afterwards.  Every replay must equal the oracle or the
committed goldens, also with other eager device work and synchronisations
between the replays (a graph node that went stale would show there)."""
import json
import os
import sys

import numpy as np
import pytest

from conftest import ROOT, random_matrix
from gpu_util import assert_same_structure, record_of

pytestmark = pytest.mark.gpu

GOLD = os.path.join(ROOT, "tests", "golden", "goldens.json")


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def oracle_of(A, **kw):
    from oracle import oracle
    return oracle.symbfact(A.colbeg, A.colend, A.csr, A.n, **kw)


@pytest.mark.parametrize("kw", [dict(), dict(mode="block"), dict(mode="reach", block_slots=32),
                                dict(counts_only=True)], ids=["auto", "block", "reach32", "counts"])
def test_graph_replays_bcsstk18(bcsstk18, kw):
    import gsofa_amd as g
    o = oracle_of(bcsstk18)
    rng = np.random.default_rng(3)
    _say("case: graph replays, bcsstk18", kw)
    with g.Plan(bcsstk18, graph=True, **kw) as p:
        for it in range(4):
            p.execute()
            r = p.fetch()
            assert r.graph_replay == (it > 0), it
            if kw.get("counts_only"):
                assert r.fill_count == o.fill_count and r.te_alg == o.te_alg and r.reach == o.reach
                np.testing.assert_array_equal(r.row_fill, o.row_fill)
            else:
                assert_same_structure(r, o)
            # other device work between the replays
            A, _ = random_matrix(rng, int(rng.integers(20, 200)), 0.03)
            assert_same_structure(g.symbfact(A), oracle_of(A))


def test_graph_ignored_where_it_does_not_apply(bcsstk18):
    """Grid mode, spilled runs and explicit output capacities execute eagerly."""
    import gsofa_amd as g
    o = oracle_of(bcsstk18)
    for kw in (dict(mode="grid"), dict(mode="block", spill=True, n_virtual_gpu=3),
               dict(output_capacity=1 << 22)):
        _say("case: graph flag ignored", kw)
        with g.Plan(bcsstk18, graph=True, **kw) as p:
            for it in range(2):
                p.execute()
                r = p.fetch()
                assert not r.graph_replay
                assert_same_structure(r, o)


def test_graph_batch_replays_goldens():
    """The 256-matrix batch workload replayed from one captured graph: every
    matrix of every replay equals its golden."""
    import gsofa_amd as g
    import gsofa_amd.configs as cf
    from digests import array_digest
    if not os.path.exists(GOLD):
        pytest.skip("goldens not generated")
    golds = json.load(open(GOLD)).get("grid2d_nd_batch")
    if not golds:
        pytest.skip("batch goldens not generated")
    mats = cf.grid2d_nd_batch()
    _say("case: graph replays, batch of 256")
    with g.Plan(mats, graph=True) as p:
        for it in range(3):
            p.execute()
            for b in range(0, 256, 1 if it == 2 else 17):
                r = p.fetch_matrix(b)
                assert r.graph_replay == (it > 0)
                rec = record_of(r)
                for k in ("fill_count", "te_alg", "reach", "lu", "row_fill"):
                    assert rec[k] == golds[b][k], (it, b, k)
                assert array_digest(r.super_start) == golds[b]["super0"], (it, b)
            g.symbfact(mats[5])  # eager work between replays
