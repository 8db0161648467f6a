"""Supernodal hand-off (gsofa_plan_fetch_supernodal): L and U by columns and the
per-supernode subscripts, computed on the device, against transposes and
gathers done here with numpy from the row-major result, plus the property
that defines the subscripts (every row i of a supernode has U_i = the
supernode's subscripts above i)."""
import sys

import numpy as np
import pytest

from conftest import random_matrix

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _csc(rowptr, col, n):
    """Column-major copy: (colptr, rows sorted per column)."""
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(rowptr))
    order = np.lexsort((rows, col.astype(np.int64)))
    colptr = np.zeros(n + 1, np.int64)
    np.cumsum(np.bincount(col, minlength=n), out=colptr[1:])
    return colptr, rows[order].astype(np.int32)


def _check_handoff(r, h, max_rows_checked=None):
    n = r.n
    assert h.n == n and h.n_super == r.n_super
    assert np.array_equal(h.xsup, r.super_start)
    assert np.array_equal(h.supno, r.supno)
    xl, lr = _csc(r.L_rowptr, r.L_col, n)
    assert np.array_equal(h.xlcol, xl) and np.array_equal(h.lrow, lr)
    xu, ur = _csc(r.U_rowptr, r.U_col, n)
    assert np.array_equal(h.xucol, xu) and np.array_equal(h.urow, ur)
    # subscripts: f, then row f's U columns; every row of the supernode agrees
    ns = h.n_super
    cnt = 1 + np.diff(r.U_rowptr)[h.xsup[:-1]]
    assert np.array_equal(np.diff(h.xsub), cnt)
    check = range(ns) if max_rows_checked is None else range(min(ns, max_rows_checked))
    for s in check:
        f, l = int(h.xsup[s]), int(h.xsup[s + 1]) - 1
        sub = h.sub[h.xsub[s]:h.xsub[s + 1]]
        assert sub[0] == f
        assert np.array_equal(sub[:l - f + 1], np.arange(f, l + 1))
        for i in range(f, l + 1):
            Ui = r.U_col[r.U_rowptr[i]:r.U_rowptr[i + 1]]
            assert np.array_equal(Ui, sub[sub > i]), (s, i)


def test_handoff_small_cases(bcsstk18):
    import gsofa_amd as g
    rng = np.random.default_rng(321)
    cases = [("bcsstk18", bcsstk18, {}), ("bcsstk18 cap 4", bcsstk18, {"max_supernode_size": 4}),
             ("grid 20^3 ND", g.gen_grid(20, 20, 20, nd_leaf=64, extra_prob=0.05, extra_dist=2, seed=9), {}),
             ("band 5000", g.gen_band(5000, 64, 6.0, 2), {"block_slots": 16})]
    for k in range(6):
        n = int(rng.integers(1, 400))
        cases.append(("random n=%d" % n, random_matrix(rng, n, float(rng.choice([0.01, 0.05])),
                                                      sym=bool(k % 2))[0], {}))
    for rows in ([[0]], [[0, 1], [0, 1]], [[], [1, 0], [2, 1]]):
        cases.append(("rows %s" % rows, g.CsrMatrix.from_rows(rows), {}))
    for name, A, kw in cases:
        _say("case: hand-off", name)
        with g.Plan(A, **kw) as p:
            p.execute()
            r = p.fetch()
            h = p.fetch_supernodal()
            _check_handoff(r, h)
            # the row-major result is unchanged by the hand-off
            r2 = p.fetch()
            assert np.array_equal(r2.L_col, r.L_col) and np.array_equal(r2.U_col, r.U_col)


def test_handoff_band_properties():
    """Full size (n=659,033, 2.5e8 entries): counts, per-column order and an
    order-independent checksum of the (row, column) pairs equal between the
    row-major result and the column-major hand-off."""
    import gsofa_amd as g
    import gsofa_amd.configs as cf
    A = cf.band()
    _say("case: hand-off band n=659,033 full size")
    with g.Plan(A) as p:
        p.execute()
        r = p.fetch()
        h = p.fetch_supernodal()
    n = r.n
    P = np.uint64(1000003)
    for rp, col, xc, rows in ((r.L_rowptr, r.L_col, h.xlcol, h.lrow), (r.U_rowptr, r.U_col, h.xucol, h.urow)):
        assert xc[-1] == rp[-1]
        assert np.array_equal(np.diff(xc), np.bincount(col, minlength=n))
        cols_of = np.repeat(np.arange(n, dtype=np.uint64), np.diff(xc))
        d = np.diff(rows.astype(np.int64))
        starts = np.zeros(rows.size, bool)
        starts[xc[:-1][np.diff(xc) > 0]] = True
        assert np.all(d[~starts[1:]] > 0)  # strictly increasing inside each column
        rows_of = np.repeat(np.arange(n, dtype=np.uint64), np.diff(rp))
        a = np.sum(rows_of * P + col.astype(np.uint64), dtype=np.uint64)
        b = np.sum(rows.astype(np.uint64) * P + cols_of, dtype=np.uint64)
        assert a == b
    _check_handoff_light(r, h)


def _check_handoff_light(r, h):
    assert np.array_equal(h.xsup, r.super_start)
    cnt = 1 + np.diff(r.U_rowptr)[h.xsup[:-1]]
    assert np.array_equal(np.diff(h.xsub), cnt)
    firsts = h.sub[h.xsub[:-1]]
    assert np.array_equal(firsts, h.xsup[:-1])


def test_handoff_refusals(bcsstk18):
    import gsofa_amd as g
    _say("case: hand-off refusals")
    with g.Plan(bcsstk18) as p:
        with pytest.raises(g.GsofaError) as ei:
            p.fetch_supernodal()
        assert ei.value.code == -9  # not executed
    for kw in ({"counts_only": True}, {"n_virtual_gpu": 3, "virtual_gpu_id": 1}):
        with g.Plan(bcsstk18, **kw) as p:
            p.execute()
            with pytest.raises(g.GsofaError) as ei:
                p.fetch_supernodal()
            assert ei.value.code == -1, kw
    with g.Plan([bcsstk18, bcsstk18]) as p:
        p.execute()
        with pytest.raises(g.GsofaError) as ei:
            p.fetch_supernodal()
        assert ei.value.code == -1


def test_plan_from_mapped_binary(bcsstk18, tmp_path):
    """A plan built straight from a mapped binary file (no host copy of the
    input): the reference total and the oracle's structure."""
    import gsofa_amd as g
    from oracle import oracle
    path = str(tmp_path / "bcsstk18.bin")
    g.write_binary(bcsstk18, path)
    _say("case: bcsstk18 from a mapped binary file")
    with g.MappedMatrix(path) as mm:
        r = g.symbfact(mm.matrix)
    assert r.fill_count == 980136
    o = oracle.symbfact(bcsstk18.colbeg, bcsstk18.colend, bcsstk18.csr, bcsstk18.n)
    from gpu_util import assert_same_structure
    assert_same_structure(r, o)
