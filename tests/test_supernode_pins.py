"""CPU tests: independent pins of the oracle's supernode partition.

The reference has only a stub for supernodes
(with_space_optimization/with_supernodes/symbfact.cuh:1087-1101, knob :1577).
The partition is defined by the row rule: row i joins row i-1's supernode iff
(i-1) is in L_i, |U_i| = |U_{i-1}| - 1, and the supernode has fewer than
max_size rows.  ``orc_supernodes`` computes it from the traversal's sorted
L/U.  Two derivations that share nothing with that code pin it:

* structurally symmetric inputs: for A = A^T, LU without pivoting has L equal
  to the Cholesky factor's pattern and U = L^T, so (i-1) in L_i iff the
  elimination tree has parent(i-1) = i, and |U_i| is the column count cc(i)
  of the Cholesky factor.  The rule becomes the fundamental-supernode test of
  Liu, Ng and Peyton (SIAM J. Matrix Anal. Appl. 14, 1993):
  parent(i-1) = i and cc(i) = cc(i-1) - 1.  The elimination tree (Liu's
  ancestor algorithm, ACM TOMS 12, 1986) and the column counts (row
  subtrees) are computed here in Python from A alone;
* tiny unsymmetric inputs: struct(L+U) from dense boolean Gaussian
  elimination (``orc_brute_lu``), and the set form of the rule, the row form
  of SuperLU's T2 supernode test: (i-1) in L_i, i in U_{i-1}, and
  U_i = U_{i-1} minus {i}.

Planted mistakes in a restatement of the rule show that the pins catch
plausible misreadings (the mistakes are planted here, on the CPU).
"""
import numpy as np
import pytest

from conftest import random_matrix


def _sym_rows(A):
    return [np.unique(A.csr[b:e]) for b, e in zip(A.colbeg, A.colend)]


def etree_and_colcounts(A):
    """Elimination tree and strict column counts of the Cholesky factor of a
    structurally symmetric pattern, from the pattern alone."""
    n = A.n
    rows = _sym_rows(A)
    parent = [-1] * n
    anc = [-1] * n
    for i in range(n):
        for j in rows[i]:
            j = int(j)
            if j >= i:
                continue
            r = j
            while anc[r] != -1 and anc[r] != i:
                t = anc[r]
                anc[r] = i
                r = t
            if anc[r] == -1:
                anc[r] = i
                parent[r] = i
    # row i of L is the union of the tree paths from each j < i in A's row i
    # up to i (the row subtree); each node k on them gains row i in column k
    cc = [0] * n
    mark = [-1] * n
    for i in range(n):
        mark[i] = i
        for j in rows[i]:
            k = int(j)
            if k >= i:
                continue
            while mark[k] != i:
                mark[k] = i
                cc[k] += 1
                k = parent[k]
    return np.array(parent), np.array(cc)


def starts_from_join(join, max_size):
    """Supernode starts from a per-row join flag (join[0] is ignored)."""
    n = len(join)
    starts = [0] if n else []
    size = 1
    for i in range(1, n):
        if join[i] and (max_size <= 0 or size < max_size):
            size += 1
        else:
            starts.append(i)
            size = 1
    return np.array(starts + [n], dtype=np.int32)


def etree_supernodes(parent, cc, max_size):
    n = len(parent)
    join = [False] + [parent[i - 1] == i and cc[i] == cc[i - 1] - 1 for i in range(1, n)]
    return starts_from_join(join, max_size)


def dense_supernodes(s, max_size):
    """T2 set form on a dense struct(L+U) (diagonal ignored)."""
    n = s.shape[0]
    join = [False] * n
    for i in range(1, n):
        Ui = set(np.flatnonzero(s[i, i + 1:]) + i + 1)
        Up = set(np.flatnonzero(s[i - 1, i:]) + i) - {i}
        join[i] = bool(s[i, i - 1]) and bool(s[i - 1, i]) and Ui == Up
    return starts_from_join(join, max_size)


def symmetric_cases():
    import gsofa_amd as g
    import gsofa_amd.configs as cf
    rng = np.random.default_rng(31)
    cases = [("grid64 natural", cf.grid64_natural()),
             ("12^3 ND", g.gen_grid(12, 12, 12, nd_leaf=64, seed=2)),
             ("40x40 ND", g.gen_grid(40, 40, 1, nd_leaf=64, seed=4))]
    for k in range(6):
        n = int(rng.integers(2, 300))
        cases.append(("random sym %d" % k, random_matrix(rng, n, float(rng.choice([0.01, 0.04, 0.1])),
                                                         sym=True)[0]))
    return cases


@pytest.mark.parametrize("max_size", [0, 1, 4, 32])
def test_symmetric_supernodes_from_etree(max_size, oracle_mod, bcsstk18):
    for name, A in [("BCSSTK18", bcsstk18)] + symmetric_cases():
        parent, cc = etree_and_colcounts(A)
        want = etree_supernodes(parent, cc, max_size)
        r = oracle_mod.symbfact(A.colbeg, A.colend, A.csr, A.n)
        # the derivation's column counts are the traversal's U row lengths
        np.testing.assert_array_equal(np.diff(r.U_rowptr), cc, err_msg=name)
        ss, sn = oracle_mod.supernodes(r.L_rowptr, r.L_col, r.U_rowptr, max_size)
        np.testing.assert_array_equal(ss, want, err_msg=name)
        np.testing.assert_array_equal(sn, np.repeat(np.arange(want.size - 1), np.diff(want)))


def test_bcsstk18_etree_matches_symbolic_cholesky(oracle_mod, bcsstk18):
    """The Python elimination tree and column counts agree with the oracle's
    column-merge symbolic Cholesky (parent = first row below the diagonal)."""
    A = bcsstk18
    parent, cc = etree_and_colcounts(A)
    colptr, rowind = oracle_mod.symbolic_cholesky(A.colbeg, A.colend, A.csr, A.n)
    np.testing.assert_array_equal(np.diff(colptr), cc)
    first = np.where(np.diff(colptr) > 0, rowind[np.minimum(colptr[:-1], rowind.size - 1)], -1)
    np.testing.assert_array_equal(first, parent)


def _unsym_cases(count, seed):
    rng = np.random.default_rng(seed)
    for t in range(count):
        n = int(rng.integers(1, 40))
        dens = float(rng.choice([0.03, 0.08, 0.15, 0.3, 0.5]))
        yield random_matrix(rng, n, dens, sym=False)


@pytest.mark.parametrize("max_size", [0, 1, 4, 32])
def test_unsymmetric_supernodes_from_brute_lu(max_size, oracle_mod):
    checked = joins = 0
    for A, a in _unsym_cases(250, 700 + max_size):
        s = oracle_mod.brute_lu(a.astype(np.uint8)).astype(bool)
        np.fill_diagonal(s, False)
        want = dense_supernodes(s, max_size)
        r = oracle_mod.symbfact(A.colbeg, A.colend, A.csr, A.n)
        ss, _ = oracle_mod.supernodes(r.L_rowptr, r.L_col, r.U_rowptr, max_size)
        np.testing.assert_array_equal(ss, want, err_msg="n=%d" % A.n)
        checked += 1
        joins += A.n - (ss.size - 1)
    # the cases exercise the join rule (a cap of 1 forbids every join)
    assert checked == 250 and (joins == 0 if max_size == 1 else joins > 100)


# ----------------------------------------------------- planted mistakes ----
def _rule_variant(r, max_size, mistake):
    """The row rule restated over sorted L/U, with one switchable misreading."""
    n = r.L_rowptr.size - 1
    Ulen = np.diff(r.U_rowptr)
    Llen = np.diff(r.L_rowptr)
    join = [False] * n
    for i in range(1, n):
        Li = r.L_col[r.L_rowptr[i]:r.L_rowptr[i + 1]]
        pas = Li.size > 0 and Li[-1] == i - 1
        if mistake == "count_same":
            cnt = Ulen[i] == Ulen[i - 1]
        elif mistake == "count_L":
            cnt = Llen[i] == Llen[i - 1] + 1
        else:
            cnt = Ulen[i] == Ulen[i - 1] - 1
        if mistake == "no_pass":
            pas = True
        if mistake == "pass_any_lower":
            pas = Li.size > 0
        join[i] = pas and cnt
    if mistake == "cap_le":
        return starts_from_join(join, max_size + 1 if max_size > 0 else 0)
    return starts_from_join(join, max_size)


@pytest.mark.parametrize("mistake", [None, "count_same", "count_L", "no_pass", "pass_any_lower", "cap_le"])
def test_planted_supernode_mistakes_are_caught(mistake, oracle_mod, bcsstk18):
    """The unmutated restatement passes both pins; each misreading fails one."""
    caught = False
    sym = [bcsstk18] + [A for _, A in symmetric_cases()[3:]]
    for max_size in (0, 4):
        for A in sym:
            r = oracle_mod.symbfact(A.colbeg, A.colend, A.csr, A.n)
            parent, cc = etree_and_colcounts(A)
            if not np.array_equal(_rule_variant(r, max_size, mistake), etree_supernodes(parent, cc, max_size)):
                caught = True
        for A, a in _unsym_cases(150, 900):
            s = oracle_mod.brute_lu(a.astype(np.uint8)).astype(bool)
            np.fill_diagonal(s, False)
            r = oracle_mod.symbfact(A.colbeg, A.colend, A.csr, A.n)
            if not np.array_equal(_rule_variant(r, max_size, mistake), dense_supernodes(s, max_size)):
                caught = True
    assert caught == (mistake is not None)
