"""GPU parity of the structure pass (fill stream -> per-row counts -> rows ->
sorted segments) on matrices whose rows span every length class the pass
treats differently: short rows (one wave, lane network), rows up to 1,024
entries (one wave, LDS network), up to 4,096 (one workgroup, LDS network) and
beyond (one workgroup, bitmap windows), including rows whose values span more
than one 262,144-wide window.

The matrices: the diagonal plus hub vertices of low id, each joined both ways
to a random set of higher vertices.  Each hub's set becomes a clique of fills
(every pair meets through the hub, whose id is below both), and vertices in
two sets join the sets' fills.  Results are compared bit-exactly with the
oracle.
"""
import sys

import numpy as np
import pytest

from gpu_util import assert_same_structure

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def hub_matrix(n, sizes, seed):
    import gsofa_amd as g
    rng = np.random.default_rng(seed)
    r = [np.arange(n, dtype=np.int64)]
    c = [np.arange(n, dtype=np.int64)]
    for h, k in enumerate(sizes):
        k = min(k, n - h - 1)
        nb = rng.choice(np.arange(h + 1, n), size=k, replace=False).astype(np.int64)
        r += [np.full(k, h, np.int64), nb]
        c += [nb, np.full(k, h, np.int64)]
    r = np.concatenate(r)
    c = np.concatenate(c)
    key = np.unique(r * n + c)  # sorted by row, duplicates removed
    r, c = key // n, key % n
    beg = np.searchsorted(r, np.arange(n)).astype(np.int64)
    end = np.searchsorted(r, np.arange(n), side="right").astype(np.int64)
    return g.CsrMatrix(n, beg, end, c.astype(np.int32))


def oracle_of(A, **kw):
    from oracle import oracle
    return oracle.symbfact(A.colbeg, A.colend, A.csr, A.n, **kw)


def _segment_classes(rowptr):
    ln = np.diff(rowptr)
    return {"<=64": int(((ln > 1) & (ln <= 64)).sum()), "<=1024": int(((ln > 64) & (ln <= 1024)).sum()),
            "<=4096": int(((ln > 1024) & (ln <= 4096)).sum()), ">4096": int((ln > 4096).sum())}


@pytest.mark.parametrize("mode", ["reach", "block", "grid"])
def test_hub_small(mode):
    import gsofa_amd as g
    A = hub_matrix(9000, [5000, 1200, 100, 30], seed=11)
    _say("case: hub n=9000 mode", mode)
    orc = oracle_of(A)
    cls = _segment_classes(orc.L_rowptr)
    assert cls["<=64"] and cls["<=1024"] and cls["<=4096"] and cls[">4096"], cls
    r = g.symbfact(A, mode=mode)
    assert_same_structure(r, orc)


@pytest.mark.parametrize("mode", ["reach", "block"])
def test_hub_wide_values(mode):
    """Rows longer than 4,096 entries whose values span more than one bitmap
    window (262,144 values)."""
    import gsofa_amd as g
    A = hub_matrix(300000, [6000, 1500, 300, 40], seed=12)
    _say("case: hub n=300000 mode", mode)
    orc = oracle_of(A)
    lp, lc = orc.L_rowptr, orc.L_col
    rows = np.flatnonzero(np.diff(lp) > 4096)
    span = lc[lp[rows + 1] - 1] - lc[lp[rows]]
    assert (span >= 262144).any(), "no row spans two bitmap windows"
    for p in (orc.L_rowptr, orc.U_rowptr):
        cls = _segment_classes(p)
        assert cls["<=64"] and cls["<=1024"] and cls[">4096"], cls
    r = g.symbfact(A, mode=mode)
    assert_same_structure(r, orc)


def test_hub_repeated_execute():
    """A reused plan rebuilds the same structure (the stream tables are
    cleared per tile, the row cursors per execute)."""
    import gsofa_amd as g
    A = hub_matrix(20000, [4000, 900, 50], seed=13)
    _say("case: hub n=20000 repeated execute")
    orc = oracle_of(A)
    with g.Plan(A) as p:
        for _ in range(3):
            p.execute()
            assert_same_structure(p.fetch(), orc)
