"""GSOFA_FLAG_SPILL: the shares run in turn and each share's L/U rows move to
host memory before the next, so the device holds one share's fill stream and
rows.  Every output is bit-exact against the oracle (small cases), the
in-core run and the committed 3D 48^3 ND golden, including supernodes, which are
built on the device from the per-row summaries the spilled shares leave."""
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


@pytest.mark.parametrize("mode,shares", [("block", 3), ("block", 7), ("grid", 5)])
def test_spill_bcsstk18(bcsstk18, mode, shares):
    import gsofa_amd as g
    o = oracle_of(bcsstk18)
    _say("case: bcsstk18 spill, mode", mode, "shares", shares)
    with g.Plan(bcsstk18, mode=mode, n_virtual_gpu=shares, spill=True) as p:
        for it in range(2):
            p.execute()
            r = p.fetch()
            assert r.fill_count == 980136
            assert_same_structure(r, o)


def test_spill_regrows_per_share(bcsstk18):
    """A fill stream far too small for any share: each share regrows it and
    reruns from the counters it started with."""
    import gsofa_amd as g
    _say("case: bcsstk18 spill, 4 shares, output capacity 1000")
    r = g.symbfact(bcsstk18, n_virtual_gpu=4, spill=True, output_capacity=1000)
    assert_same_structure(r, oracle_of(bcsstk18))


def test_spill_random_and_single_share(bcsstk18):
    import gsofa_amd as g
    rng = np.random.default_rng(4242)
    for trial in range(6):
        n = int(rng.integers(1, 300))
        A, _ = random_matrix(rng, n, float(rng.choice([0.01, 0.05])), sym=bool(trial % 2))
        N = int(rng.integers(1, 6))
        _say("case: random spill n", n, "shares", N)
        assert_same_structure(g.symbfact(A, n_virtual_gpu=N, spill=True, max_supernode_size=3),
                              oracle_of(A), max_size=3)
    for gid in (0, 2):
        _say("case: bcsstk18 spill, share", gid, "of 3 only")
        r = g.symbfact(bcsstk18, n_virtual_gpu=3, virtual_gpu_id=gid, spill=True)
        assert_same_structure(r, oracle_of(bcsstk18, n_virtual_gpu=3, virtual_gpu_id=gid), check_super=False)
    _say("case: bcsstk18 spill with the hybrid split")
    assert_same_structure(g.symbfact(bcsstk18, spill=True, top_sources=4096), oracle_of(bcsstk18))


def test_spill_refuses_handoff(bcsstk18):
    import gsofa_amd as g
    with g.Plan(bcsstk18, n_virtual_gpu=2, spill=True) as p:
        p.execute()
        with pytest.raises(g.GsofaError) as ei:
            p.fetch_supernodal()
        assert ei.value.code == -1


def test_spill_grid3d_nd_golden():
    import gsofa_amd as g
    import gsofa_amd.configs as cf
    from digests import array_digest
    if not os.path.exists(GOLD):
        pytest.skip("goldens not generated")
    gold = json.load(open(GOLD)).get("grid3d_nd")
    if gold is None:
        pytest.skip("golden grid3d_nd not generated")
    _say("case: 3D 48^3 ND full size, spill in 4 shares")
    r = g.symbfact(cf.grid3d_nd(), n_virtual_gpu=4, spill=True)
    rec = record_of(r)
    for k in ("fill_count", "te_alg", "reach", "nnz_L", "nnz_U", "lu", "row_fill"):
        assert rec[k] == gold[k], k
    assert r.n_super == gold["n_super0"]
    assert array_digest(r.super_start) == gold["super0"]
