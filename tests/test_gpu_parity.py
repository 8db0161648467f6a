"""GPU parity: the HIP path through the C ABI against the CPU oracle.

Bit-exact on every output (sorted L/U, per-row fill counts, TE_alg, |R|,
supernodes).  Small cases run the oracle here; full-size workloads compare
against the committed goldens (tests/golden/goldens.json) and size-independent
properties.
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
MODES = ["block", "grid"]


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def oracle_of(A, **kw):
    from oracle import oracle
    return oracle.symbfact(A.colbeg, A.colend, A.csr, A.n, **kw)


@pytest.mark.parametrize("mode", MODES)
def test_bcsstk18_reference_total(bcsstk18, mode):
    import gsofa_amd as g
    _say("case: bcsstk18 mode", mode)
    r = g.symbfact(bcsstk18, mode=mode)
    assert r.mode == mode
    assert r.fill_count == 980136  # reference log LOG:2
    assert_same_structure(r, oracle_of(bcsstk18))


@pytest.mark.parametrize("mode", MODES)
def test_grid64(mode):
    import gsofa_amd as g
    import gsofa_amd.configs as cf
    A = cf.grid64_natural()
    _say("case: grid64 natural, mode", mode)
    r = g.symbfact(A, mode=mode)
    assert r.fill_count == 500094
    assert r.reach == A.n * (A.n - 1) // 2
    assert_same_structure(r, oracle_of(A))


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("seed", range(3))
def test_random_small(seed, mode):
    import gsofa_amd as g
    rng = np.random.default_rng(500 + seed)
    for trial in range(12):
        n = int(rng.integers(1, 300))
        A, _ = random_matrix(rng, n, float(rng.choice([0.005, 0.02, 0.08])),
                            sym=bool(trial % 4 == 0))
        _say("case: random seed", seed, "trial", trial, "n", n, "mode", mode)
        r = g.symbfact(A, mode=mode, chunk=int(rng.choice([0, 1, 3, 16, 700])))
        assert_same_structure(r, oracle_of(A))


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("chunk", [1, 5, 64, 1000, 0])
def test_chunk_invariance(bcsstk18, chunk, mode):
    import gsofa_amd as g
    _say("case: bcsstk18 chunk", chunk, "mode", mode)
    r = g.symbfact(bcsstk18, chunk=chunk, mode=mode)
    assert_same_structure(r, oracle_of(bcsstk18))


@pytest.mark.parametrize("qcap,chunk", [(4096, 256), (20000, 1000)])
def test_forced_queue_overflow(bcsstk18, qcap, chunk):
    """Grid mode: a small frontier queue forces the drain-and-resume path; the
    result must not change (the reference's spill path keeps its count,
    LOG:157-165).  The capacity is chosen so that the run needs a few thousand
    levels (a CPU emulation of the level loop gives 3,206 levels + 1,366 refills
    at 4096/256); capacity 64 at chunk 256 would need ~5e5 grid-wide barriers."""
    import gsofa_amd as g
    _say("case: bcsstk18 grid mode, queue", qcap, "chunk", chunk)
    r = g.symbfact(bcsstk18, queue_capacity=qcap, chunk=chunk, mode="grid")
    assert r.rescans > 0
    assert_same_structure(r, oracle_of(bcsstk18))


def test_block_rounds_and_spill():
    """Block mode: wide frontiers need several relaxation rounds per level and
    overflow the LDS owner buckets into the global spill area; results must
    not change.  The deduplicated frontier queue holds up to n entries."""
    import gsofa_amd as g
    A = g.gen_grid(26, 26, 26, nd_leaf=64, extra_prob=0.05, extra_dist=2, seed=11)
    _say("case: 26^3 ND grid, block mode")
    r = g.symbfact(A, mode="block")
    assert r.queue_capacity == A.n
    assert_same_structure(r, oracle_of(A))


@pytest.mark.parametrize("mode", MODES)
def test_block_expansion_path(bcsstk18, mode):
    import gsofa_amd as g
    _say("case: bcsstk18 block threshold 8, mode", mode)
    r = g.symbfact(bcsstk18, wave_block_threshold=8, mode=mode)
    assert_same_structure(r, oracle_of(bcsstk18))


def test_output_regrowth(bcsstk18):
    import gsofa_amd as g
    _say("case: bcsstk18 output capacity 10")
    r = g.symbfact(bcsstk18, output_capacity=10)
    assert r.passes == 2  # an explicit capacity that overflows: grown and rerun once
    assert_same_structure(r, oracle_of(bcsstk18))


@pytest.mark.parametrize("mode", MODES)
def test_fresh_plan_single_pass(mode):
    """A 32^3 ND grid with 5% extra entries fills about 200 times its nnz, far
    above a fixed reserve: sized from the row bounds, a fresh plan traverses
    once, in one execute with one host synchronisation, bit-exact; reused, it
    gives the same again."""
    import gsofa_amd as g
    A = g.gen_grid(32, 32, 32, nd_leaf=64, extra_prob=0.05, extra_dist=2, seed=6)
    o = oracle_of(A)
    assert o.fill_count > 64 * A.nnz
    _say("case: 32^3 ND fresh plan, mode", mode)
    with g.Plan(A, mode=mode) as p:
        for it in range(2):
            p.execute()
            r = p.fetch()
            assert r.passes == 1, it
            assert_same_structure(r, o)


def test_small_grid_and_budget(bcsstk18):
    import gsofa_amd as g
    _say("case: bcsstk18 grid 7 blocks, 64 MiB budget")
    r = g.symbfact(bcsstk18, grid_blocks=7, mem_gb=1 / 16)
    assert r.grid_blocks == 7
    assert_same_structure(r, oracle_of(bcsstk18))


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("shares", [3, 7])
def test_virtual_gpu_shares(bcsstk18, shares, mode):
    import gsofa_amd as g
    full = oracle_of(bcsstk18)
    total = 0
    for gid in range(shares):
        _say("case: bcsstk18 share", gid, "of", shares, "mode", mode)
        r = g.symbfact(bcsstk18, n_virtual_gpu=shares, virtual_gpu_id=gid, chunk=97, mode=mode)
        o = oracle_of(bcsstk18, n_virtual_gpu=shares, virtual_gpu_id=gid)
        assert_same_structure(r, o, check_super=False)
        assert r.n_super == -1
        total += r.fill_count
    assert total == full.fill_count
    # every share in turn on one plan (the reference's virtual-GPU loop)
    r = g.symbfact(bcsstk18, n_virtual_gpu=shares, mode=mode)
    assert_same_structure(r, full)


@pytest.mark.parametrize("max_size", [1, 4, 32])
def test_supernode_cap(bcsstk18, max_size):
    import gsofa_amd as g
    _say("case: bcsstk18 max_supernode_size", max_size)
    r = g.symbfact(bcsstk18, max_supernode_size=max_size)
    assert_same_structure(r, oracle_of(bcsstk18), max_size=max_size)


@pytest.mark.parametrize("mode", MODES)
def test_plan_repeated_execute(bcsstk18, mode):
    """Stamps and offsets carry across executions without clearing."""
    import gsofa_amd as g
    o = oracle_of(bcsstk18)
    with g.Plan(bcsstk18, chunk=512, mode=mode) as p:
        for it in range(4):
            _say("case: plan execute", it)
            p.execute()
            assert_same_structure(p.fetch(), o)


def test_counts_only(bcsstk18):
    import gsofa_amd as g
    r = g.symbfact(bcsstk18, counts_only=True)
    assert r.fill_count == 980136 and r.L_rowptr is None


@pytest.mark.parametrize("mode", MODES)
def test_edge_cases(mode):
    import gsofa_amd as g
    # single vertex; two vertices; a row without its diagonal; empty rows
    for rows in ([[0]], [[0, 1], [0, 1]], [[1], [0, 1]], [[0], [1], [2]],
                 [[0, 2], [1], [0, 2]], [[], [1, 0], [2, 1]]):
        A = g.CsrMatrix.from_rows(rows)
        _say("case: edge rows", rows, "mode", mode)
        assert_same_structure(g.symbfact(A, mode=mode), oracle_of(A))


@pytest.mark.parametrize("mode", MODES)
def test_grid3d_nd(mode):
    import gsofa_amd as g
    A = g.gen_grid(20, 20, 20, nd_leaf=64, extra_prob=0.05, extra_dist=2, seed=9)
    _say("case: 20^3 ND grid, mode", mode)
    assert_same_structure(g.symbfact(A, mode=mode), oracle_of(A))


@pytest.mark.parametrize("mode", MODES)
def test_band_small(mode):
    import gsofa_amd as g
    A = g.gen_band(6000, 256, 8.0, 3)
    _say("case: band n=6000, mode", mode)
    assert_same_structure(g.symbfact(A, chunk=1024, mode=mode), oracle_of(A))


def _golden(name):
    if not os.path.exists(GOLD):
        pytest.skip("goldens not generated")
    d = json.load(open(GOLD))
    if name not in d:
        pytest.skip("golden %s not generated" % name)
    return d[name]


@pytest.mark.parametrize("mode", MODES)
def test_grid3d_nd_full_golden(mode):
    import gsofa_amd as g
    import gsofa_amd.configs as cf
    gold = _golden("grid3d_nd")
    A = cf.grid3d_nd()
    _say("case: 3D 48^3 ND full size, mode", mode)
    r = g.symbfact(A, mode=mode)
    assert r.passes == 1  # the stream is sized from the row bounds: one traversal
    rec = record_of(r)
    for k in ("fill_count", "te_alg", "reach", "nnz_L", "nnz_U", "lu", "row_fill"):
        assert rec[k] == gold[k], k
    from digests import array_digest
    assert r.n_super == gold["n_super0"]
    assert array_digest(r.super_start) == gold["super0"]


def test_band_full_golden():
    """n=659,033 banded matrix: totals, L/U digest and supernodes vs the oracle."""
    import gsofa_amd as g
    import gsofa_amd.configs as cf
    gold = _golden("band")
    A = cf.band()
    _say("case: band n=659,033 full size, block mode")
    r = g.symbfact(A, mode="block")
    assert r.passes == 1
    rec = record_of(r)
    for k in ("fill_count", "te_alg", "reach", "nnz_L", "nnz_U", "lu", "row_fill"):
        assert rec[k] == gold[k], k
    from digests import array_digest
    assert r.n_super == gold["n_super0"]
    assert array_digest(r.super_start) == gold["super0"]
    # fill stays inside the +-256 band
    assert int(np.max(np.diff(r.U_rowptr))) <= 256 and int(np.max(np.diff(r.L_rowptr))) <= 256


def _check_gold(rec, gold, tag):
    for k in ("fill_count", "te_alg", "reach", "lu", "row_fill"):
        assert rec[k] == gold[k], (tag, k)


def test_batch_one_by_one_goldens():
    """The 256 matrices of the batch workload, one plan each, vs oracle goldens."""
    import gsofa_amd as g
    import gsofa_amd.configs as cf
    golds = _golden("grid2d_nd_batch")
    mats = cf.grid2d_nd_batch()
    assert len(golds) == len(mats) == 256
    for b in range(0, 256, 8):
        _say("case: batch matrix", b, "n", mats[b].n, "alone")
        _check_gold(record_of(g.symbfact(mats[b])), golds[b], b)


def test_batch_one_launch():
    """All 256 matrices in one plan / one persistent launch: every matrix's
    result equals its golden (and so its stand-alone result)."""
    import gsofa_amd as g
    import gsofa_amd.configs as cf
    from digests import array_digest
    golds = _golden("grid2d_nd_batch")
    mats = cf.grid2d_nd_batch()
    _say("case: batch of 256 in one plan")
    with g.Plan(mats) as p:
        assert p.matrix_count == 256
        for it in range(2):
            p.execute()
            for b in range(256):
                r = p.fetch_matrix(b)
                _check_gold(record_of(r), golds[b], b)
                assert r.n_super == golds[b]["n_super0"], b
                assert array_digest(r.super_start) == golds[b]["super0"], b


def test_random_batch_vs_oracle():
    import gsofa_amd as g
    rng = np.random.default_rng(77)
    mats = []
    for k in range(40):
        n = int(rng.integers(1, 200))
        mats.append(random_matrix(rng, n, float(rng.choice([0.01, 0.05, 0.15])))[0])
    mats.append(g.gen_band(3000, 256, 8.0, 5))
    _say("case: random batch of", len(mats))
    with g.Plan(mats, max_supernode_size=4) as p:
        p.execute()
        for b, A in enumerate(mats):
            assert_same_structure(p.fetch_matrix(b), oracle_of(A), max_size=4)


def test_autotuner_choice_is_result_neutral(bcsstk18, tmp_path):
    """Whatever the tuner picks, the
    full run equals the oracle."""
    import gsofa_amd as g
    from gsofa_amd import tune
    _say("case: autotune bcsstk18")
    best, timings = tune.tune(bcsstk18, reps=1, table=str(tmp_path / "t.txt"))
    assert len(timings) == len(tune.candidates(bcsstk18))
    assert tune.best_options(bcsstk18, table=str(tmp_path / "t.txt")) == best
    assert_same_structure(g.symbfact(bcsstk18, **best), oracle_of(bcsstk18))


@pytest.mark.parametrize("top", [1, 100, 4096, 9000, 20000])
def test_hybrid_top_sources(bcsstk18, top):
    """The top sources as grid-mode groups, the rest in block mode: bit-exact,
    repeated executes included (both label states carry over)."""
    import gsofa_amd as g
    o = oracle_of(bcsstk18)
    _say("case: bcsstk18 hybrid top", top)
    with g.Plan(bcsstk18, top_sources=top) as p:
        for it in range(2):
            p.execute()
            r = p.fetch()
            assert r.mode == "hybrid" and r.row_te is None
            assert_same_structure(r, o)


def test_hybrid_random_and_grid3d_nd():
    import gsofa_amd as g
    import gsofa_amd.configs as cf
    rng = np.random.default_rng(91)
    for trial in range(6):
        n = int(rng.integers(2, 300))
        A, _ = random_matrix(rng, n, float(rng.choice([0.01, 0.05])))
        top = int(rng.integers(1, n + 1))
        _say("case: random hybrid n", n, "top", top)
        assert_same_structure(g.symbfact(A, top_sources=top), oracle_of(A))
    gold = _golden("grid3d_nd")
    _say("case: 3D 48^3 ND hybrid top 8192")
    r = g.symbfact(cf.grid3d_nd(), top_sources=8192)
    rec = record_of(r)
    for k in ("fill_count", "te_alg", "reach", "nnz_L", "nnz_U", "lu", "row_fill"):
        assert rec[k] == gold[k], k
