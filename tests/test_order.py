"""Fill-reducing ordering (gsofa_order_nd) and symmetric permutation
(gsofa_permute), on the CPU: the ordering is a permutation on every input
shape, permuting is exact and invertible, and the ordering cuts the fill the
CPU oracle counts (against natural order, and no worse than the generator's
geometric nested dissection on the same grids)."""
import numpy as np
import pytest

from conftest import random_matrix


def _fills(A):
    from oracle import oracle
    return oracle.symbfact(A.colbeg, A.colend, A.csr, A.n).fill_count


def _rows(A):
    return [sorted(A.csr[b:e].tolist()) for b, e in zip(A.colbeg, A.colend)]


def _cases():
    import gsofa_amd as g
    rng = np.random.default_rng(99)
    out = [g.gen_grid(30, 30, 1, nd_leaf=0, seed=1), g.gen_grid(9, 9, 9, nd_leaf=0, seed=1),
           g.gen_band(3000, 40, 4.0, 2), g.CsrMatrix.from_rows([[0]]),
           g.CsrMatrix.from_rows([[], [1, 0], [2, 1]]), g.CsrMatrix(0, [], [], []),
           # two components and an isolated vertex
           g.CsrMatrix.from_rows([[0, 1], [0, 1], [2], [3, 4], [3, 4]])]
    for k in range(4):
        out.append(random_matrix(rng, int(rng.integers(2, 400)), float(rng.choice([0.005, 0.03])),
                                 sym=bool(k % 2))[0])
    return out


def test_order_is_a_permutation_and_deterministic():
    import gsofa_amd as g
    for A in _cases():
        for leaf in (1, 8, 64):
            p = g.order_nd(A, leaf)
            assert sorted(p.tolist()) == list(range(A.n)), (A.n, leaf)
            assert np.array_equal(p, g.order_nd(A, leaf))


def test_permute_is_exact_and_invertible():
    import gsofa_amd as g
    for A in _cases():
        p = g.order_nd(A, 16)
        B = g.permute(A, p)
        assert B.nnz == A.nnz
        inv = np.empty_like(p)
        inv[p] = np.arange(A.n, dtype=np.int32)
        # row k of B is row p[k] of A, columns renumbered
        for k in range(min(A.n, 200)):
            want = sorted(inv[A.csr[A.colbeg[p[k]]:A.colend[p[k]]]].tolist())
            assert B.csr[B.colbeg[k]:B.colend[k]].tolist() == want
        assert _rows(g.permute(B, inv)) == _rows(A)


def test_order_cuts_fill():
    import gsofa_amd as g
    for shape in ((40, 40, 1), (12, 12, 12)):
        A = g.gen_grid(*shape, nd_leaf=0, seed=1)
        nd = _fills(g.permute(A, g.order_nd(A, 64)))
        natural = _fills(A)
        geometric = _fills(g.gen_grid(*shape, nd_leaf=64, seed=1))
        assert nd < 0.5 * natural, (shape, nd, natural)
        assert nd <= geometric, (shape, nd, geometric)


def test_order_and_permute_refusals():
    import gsofa_amd as g
    A = g.CsrMatrix.from_rows([[0, 1], [0, 1, 2], [1, 2]])
    with pytest.raises(g.GsofaError) as ei:
        g.permute(A, [0, 0, 1])
    assert ei.value.code == -1
    with pytest.raises(g.GsofaError):
        g.permute(A, [0, 1])
    bad = g.CsrMatrix.from_rows([[0, 5], [1], [2]])
    with pytest.raises(g.GsofaError) as ei:
        g.order_nd(bad)
    assert ei.value.code == -1
