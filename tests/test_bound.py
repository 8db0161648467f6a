"""CPU tests: the output bounds of gsofa_fill_bound (include/gsofa.h), which
size the fill stream and the L/U columns before any traversal.

Pins: on structurally symmetric inputs LU without pivoting is Cholesky, so
the bounds equal the oracle's exact per-row L / U lengths; on unsymmetric
inputs they equal the oracle's lengths for the symmetrized matrix A + A^T
(whose factor they describe) and are at least A's own."""
import ctypes as C

import numpy as np
import pytest

from conftest import random_matrix


def bound(A):
    from gsofa_amd import _lib
    m = A.to_c()
    rb = np.zeros(max(A.n, 1), np.int64)
    f, l, u = C.c_uint64(), C.c_uint64(), C.c_uint64()
    rc = _lib.lib().gsofa_fill_bound(C.byref(m), rb.ctypes.data, C.byref(f), C.byref(l), C.byref(u))
    return rc, rb[:A.n], f.value, l.value, u.value


def symmetrized(A):
    import gsofa_amd as g
    rows = [set() for _ in range(A.n)]
    for i, r in enumerate(A.rows()):
        for j in r:
            rows[i].add(int(j))
            rows[int(j)].add(i)
    return g.CsrMatrix.from_rows([sorted(r) for r in rows])


def _cases(rng):
    import gsofa_amd as g
    yield g.gen_grid(12, 12, 12, nd_leaf=64, extra_prob=0.05, extra_dist=2, seed=2)
    yield g.gen_band(3000, 64, 6.0, 3)
    yield g.gen_grid(30, 30, 1, nd_leaf=64, extra_prob=0.05, extra_dist=3, seed=4)
    for _ in range(40):
        n = int(rng.integers(1, 200))
        yield random_matrix(rng, n, float(rng.choice([0.005, 0.02, 0.08])), sym=False)[0]


def test_exact_on_symmetric_inputs(oracle_mod, bcsstk18):
    import gsofa_amd as g
    import gsofa_amd.configs as cf
    rng = np.random.default_rng(11)
    cases = [bcsstk18, cf.grid64_natural(), g.gen_grid(10, 10, 10, nd_leaf=64, seed=1)]
    cases += [random_matrix(rng, int(rng.integers(2, 250)), 0.03, sym=True)[0] for _ in range(8)]
    for A in cases:
        rc, rb, f, l, u = bound(A)
        assert rc == 0
        o = oracle_mod.symbfact(A.colbeg, A.colend, A.csr, A.n)
        np.testing.assert_array_equal(rb, o.row_fill)
        assert (f, l, u) == (o.fill_count, int(o.L_rowptr[-1]), int(o.U_rowptr[-1]))
    assert bound(bcsstk18)[2] == 980136


def test_bounds_unsymmetric_inputs(oracle_mod):
    rng = np.random.default_rng(12)
    for A in _cases(rng):
        rc, rb, f, l, u = bound(A)
        assert rc == 0
        o = oracle_mod.symbfact(A.colbeg, A.colend, A.csr, A.n)
        assert np.all(rb >= o.row_fill)
        assert f >= o.fill_count and l >= o.L_rowptr[-1] and u >= o.U_rowptr[-1]
        # exact for the matrix whose Cholesky factor they describe
        S = symmetrized(A)
        s = oracle_mod.symbfact(S.colbeg, S.colend, S.csr, S.n)
        assert (l, u) == (int(s.L_rowptr[-1]), int(s.U_rowptr[-1]))
        off = np.array([np.count_nonzero(r != i) for i, r in enumerate(A.rows())], np.int64)
        np.testing.assert_array_equal(rb + off, np.diff(s.L_rowptr) + np.diff(s.U_rowptr))


def test_bound_edge_cases():
    import gsofa_amd as g
    rc, rb, f, l, u = bound(g.CsrMatrix.from_rows([]))
    assert rc == 0 and (f, l, u) == (0, 0, 0)
    # no diagonal, an empty row, duplicate entries (counted once)
    A = g.CsrMatrix.from_rows([[1, 1], [], [0, 2, 0], [3]])
    rc, rb, f, l, u = bound(A)
    assert rc == 0 and np.all(rb >= 0)
    bad = g.CsrMatrix.from_rows([[0, 4], [1]])
    assert bound(bad)[0] == -1
