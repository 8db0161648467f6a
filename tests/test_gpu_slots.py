"""Block mode with 16 and 8 source slots per workgroup: every output
bit-exact against the CPU oracle (small cases) or the committed golden of the
3D 48^3 ND grid,
across the code paths the 32-slot kernel is tested on (multi-round levels,
batches, virtual-GPU shares, the hybrid split, repeated executes)."""
import json
import os
import sys

import numpy as np
import pytest

from conftest import ROOT, random_matrix
from gpu_util import assert_same_structure, record_of

pytestmark = pytest.mark.gpu

SLOTS = [16, 8]
GOLD = os.path.join(ROOT, "tests", "golden", "goldens.json")


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def oracle_of(A, **kw):
    from oracle import oracle
    return oracle.symbfact(A.colbeg, A.colend, A.csr, A.n, **kw)


@pytest.mark.parametrize("slots", SLOTS)
def test_slots_bcsstk18_repeated(bcsstk18, slots):
    import gsofa_amd as g
    o = oracle_of(bcsstk18)
    _say("case: bcsstk18 block slots", slots)
    with g.Plan(bcsstk18, block_slots=slots) as p:
        for it in range(3):
            p.execute()
            r = p.fetch()
            assert r.fill_count == 980136
            assert_same_structure(r, o)


@pytest.mark.parametrize("slots", SLOTS)
def test_slots_random(slots):
    import gsofa_amd as g
    rng = np.random.default_rng(700 + slots)
    for trial in range(12):
        n = int(rng.integers(1, 300))
        A, _ = random_matrix(rng, n, float(rng.choice([0.005, 0.02, 0.08])), sym=bool(trial % 4 == 0))
        chunk = int(rng.choice([0, 1, 16, 700]))
        _say("case: random slots", slots, "trial", trial, "n", n, "chunk", chunk)
        assert_same_structure(g.symbfact(A, block_slots=slots, chunk=chunk), oracle_of(A))


@pytest.mark.parametrize("slots", SLOTS)
def test_slots_wide_levels(slots):
    """Wide 3D frontiers need several table rounds per level; a band has long
    chains of narrow levels."""
    import gsofa_amd as g
    A = g.gen_grid(26, 26, 26, nd_leaf=64, extra_prob=0.05, extra_dist=2, seed=11)
    _say("case: 26^3 ND grid, slots", slots)
    assert_same_structure(g.symbfact(A, block_slots=slots), oracle_of(A))
    B = g.gen_band(6000, 256, 8.0, 3)
    _say("case: band n=6000, slots", slots)
    assert_same_structure(g.symbfact(B, block_slots=slots, grid_blocks=37), oracle_of(B))


@pytest.mark.parametrize("slots", SLOTS)
def test_slots_edge_cases(slots):
    import gsofa_amd as g
    for rows in ([[0]], [[0, 1], [0, 1]], [[1], [0, 1]], [[0], [1], [2]],
                 [[0, 2], [1], [0, 2]], [[], [1, 0], [2, 1]]):
        A = g.CsrMatrix.from_rows(rows)
        _say("case: edge rows", rows, "slots", slots)
        assert_same_structure(g.symbfact(A, block_slots=slots), oracle_of(A))


@pytest.mark.parametrize("slots", SLOTS)
def test_slots_batch(slots):
    import gsofa_amd as g
    rng = np.random.default_rng(88 + slots)
    mats = []
    for k in range(30):
        n = int(rng.integers(1, 200))
        mats.append(random_matrix(rng, n, float(rng.choice([0.01, 0.05, 0.15])))[0])
    mats.append(g.gen_band(3000, 256, 8.0, 5))
    _say("case: random batch of", len(mats), "slots", slots)
    with g.Plan(mats, block_slots=slots, max_supernode_size=4) as p:
        for it in range(2):
            p.execute()
            for b, A in enumerate(mats):
                assert_same_structure(p.fetch_matrix(b), oracle_of(A), max_size=4)


def test_slots_shares_and_hybrid(bcsstk18):
    import gsofa_amd as g
    full = oracle_of(bcsstk18)
    total = 0
    for gid in range(3):
        _say("case: bcsstk18 share", gid, "of 3, slots 8")
        r = g.symbfact(bcsstk18, n_virtual_gpu=3, virtual_gpu_id=gid, chunk=97, block_slots=8)
        assert_same_structure(r, oracle_of(bcsstk18, n_virtual_gpu=3, virtual_gpu_id=gid),
                              check_super=False)
        total += r.fill_count
    assert total == full.fill_count
    for top in (100, 4096):
        _say("case: bcsstk18 hybrid top", top, "slots 16")
        with g.Plan(bcsstk18, top_sources=top, block_slots=16) as p:
            for it in range(2):
                p.execute()
                r = p.fetch()
                assert r.mode == "hybrid"
                assert_same_structure(r, full)


@pytest.mark.parametrize("slots", SLOTS)
def test_slots_grid3d_nd_full_golden(slots):
    import gsofa_amd as g
    import gsofa_amd.configs as cf
    from digests import array_digest
    if not os.path.exists(GOLD):
        pytest.skip("goldens not generated")
    gold = json.load(open(GOLD)).get("grid3d_nd")
    if gold is None:
        pytest.skip("golden grid3d_nd not generated")
    _say("case: 3D 48^3 ND full size, slots", slots)
    r = g.symbfact(cf.grid3d_nd(), block_slots=slots)
    rec = record_of(r)
    for k in ("fill_count", "te_alg", "reach", "nnz_L", "nnz_U", "lu", "row_fill"):
        assert rec[k] == gold[k], k
    assert r.n_super == gold["n_super0"]
    assert array_digest(r.super_start) == gold["super0"]


def test_auto_slots(bcsstk18):
    """Block mode with block_slots left at 0: the largest slot count whose
    chunks still cover every co-resident workgroup for a single matrix, 16 for
    batches."""
    import gsofa_amd as g
    import gsofa_amd.configs as cf
    _say("case: automatic slots")
    r = g.symbfact(bcsstk18, counts_only=True, mode="block")
    assert r.chunk // r.grid_blocks == 32 and r.fill_count == 980136
    A = g.gen_band(300, 8, 3.0, 1)
    r = g.symbfact(A, counts_only=True, mode="block")
    assert r.chunk // r.grid_blocks == 8
    assert r.fill_count == oracle_of(A).fill_count
    mats = cf.grid2d_nd_batch()[:16]
    with g.Plan(mats, counts_only=True, mode="block") as p:
        p.execute()
        r = p.fetch_matrix(0)
        assert r.chunk // r.grid_blocks == 16


@pytest.mark.parametrize("slots", [32, 16, 8])
def test_counts_only_batch_per_matrix(slots):
    """A counts-only batch reports every matrix's fill count, per-row fills, TE
    and |R| (counted by the traversal kernel, no fill stream), equal to the
    oracle's; a counts-only single plan reports its per-row fills too."""
    import gsofa_amd as g
    import gsofa_amd.configs as cf
    rng = np.random.default_rng(77 + slots)
    mats = [cf.batch_matrix(s, b) for b, s in enumerate([16, 31, 48, 70])]
    mats += [random_matrix(rng, int(rng.integers(1, 300)), 0.02)[0] for _ in range(6)]
    _say("case: counts-only batch, slots", slots)
    with g.Plan(mats, counts_only=True, block_slots=slots) as p:
        p.execute()
        for b, A in enumerate(mats):
            r = p.fetch_matrix(b)
            o = oracle_of(A)
            assert (r.fill_count, r.te_alg, r.reach) == (o.fill_count, o.te_alg, o.reach), b
            np.testing.assert_array_equal(r.row_fill, o.row_fill)
            np.testing.assert_array_equal(r.row_te, o.row_te)
            np.testing.assert_array_equal(r.row_reach, o.row_reach)
            assert r.L_rowptr is None
    A = mats[2]
    for share in (-1, 1):
        r = g.symbfact(A, counts_only=True, block_slots=slots, n_virtual_gpu=3, virtual_gpu_id=share)
        o = oracle_of(A, n_virtual_gpu=3, virtual_gpu_id=share)
        assert r.fill_count == o.fill_count
        np.testing.assert_array_equal(r.row_fill, o.row_fill)
