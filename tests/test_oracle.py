"""CPU tests: pin the oracle, the generators and the host-side surface.

The oracle is pinned by independent checks, none of which shares code with its
traversal:
  * the reference's own logged total for BCSSTK18 (980,136 fills,
    with_space_optimization/space_optimization_log_1_264_GPU_warp_centric.dat:2);
  * textbook boolean LU without pivoting on random unsymmetric patterns;
  * symbolic Cholesky (elimination tree) on structurally symmetric inputs;
  * closed forms for the natural-order 2D grid (|R| = n(n-1)/2, L/U symmetry);
  * a literal restatement of the supernode definition (the independent
    supernode pins are in test_supernode_pins.py).
A planted-mistake test shows the brute-force pin rejects plausible variants
of the fill rule.
"""
import hashlib
import json
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT, random_matrix


def dense_lu_rows(a):
    from oracle import oracle
    return oracle.brute_lu(a.astype(np.uint8)).astype(bool)


def structure_from_result(r, n):
    s = np.zeros((n, n), bool)
    for i in range(n):
        s[i, r.L_col[r.L_rowptr[i]:r.L_rowptr[i + 1]]] = True
        s[i, r.U_col[r.U_rowptr[i]:r.U_rowptr[i + 1]]] = True
    return s


# ---------------------------------------------------------------- goldens ----
def test_bcsstk18_reference_total(bcsstk18, oracle_mod):
    r = oracle_mod.symbfact(bcsstk18.colbeg, bcsstk18.colend, bcsstk18.csr, bcsstk18.n)
    assert r.fill_count == 980136          # reference log, LOG:2 (10 runs)
    assert int(r.row_fill.sum()) == 980136
    assert r.te_alg == 64953531
    assert r.reach == 4827690
    # fill = |struct(L+U)| - |struct(A)| (diagonal excluded)
    offdiag = sum(int(np.count_nonzero(bcsstk18.csr[b:e] != i))
                  for i, (b, e) in enumerate(zip(bcsstk18.colbeg, bcsstk18.colend)))
    assert int(r.L_rowptr[-1] + r.U_rowptr[-1]) - offdiag == 980136


def test_bcsstk18_symbolic_cholesky(bcsstk18, oracle_mod):
    """Structurally symmetric input: row s of L equals row s of Cholesky's L,
    and U = L^T (elimination tree column merge, an independent method)."""
    A = bcsstk18
    r = oracle_mod.symbfact(A.colbeg, A.colend, A.csr, A.n)
    colptr, rowind = oracle_mod.symbolic_cholesky(A.colbeg, A.colend, A.csr, A.n)
    assert int(colptr[-1]) == 558639  # strict L; 2*(558,639-68,571) = 980,136
    # U row s == Cholesky column s
    np.testing.assert_array_equal(r.U_rowptr, colptr)
    np.testing.assert_array_equal(r.U_col, rowind)
    # L row s == {j : s in column j}
    rows_of = np.repeat(np.arange(A.n), np.diff(colptr))
    order = np.lexsort((rows_of, rowind))
    Lp = np.concatenate([[0], np.cumsum(np.bincount(rowind, minlength=A.n))])
    np.testing.assert_array_equal(r.L_rowptr, Lp)
    np.testing.assert_array_equal(r.L_col, rows_of[order])


def test_grid64_closed_forms(oracle_mod):
    import gsofa_amd.configs as cf
    A = cf.grid64_natural()
    n = A.n
    r = oracle_mod.symbfact(A.colbeg, A.colend, A.csr, n)
    assert r.fill_count == 500094
    assert r.reach == n * (n - 1) // 2    # natural order: every lower vertex is reached
    assert r.te_alg == 41428864
    fills_L = int(r.L_rowptr[-1]) - (A.nnz - n) // 2
    fills_U = int(r.U_rowptr[-1]) - (A.nnz - n) // 2
    assert fills_L == fills_U == 250047
    # 5-point grid with natural order: fill stays inside the bandwidth 64
    for s in range(0, n, 97):
        row = np.concatenate([r.L_col[r.L_rowptr[s]:r.L_rowptr[s + 1]],
                              r.U_col[r.U_rowptr[s]:r.U_rowptr[s + 1]]])
        assert np.all(np.abs(row - s) <= 64)
    colptr, rowind = oracle_mod.symbolic_cholesky(A.colbeg, A.colend, A.csr, n)
    np.testing.assert_array_equal(r.U_col, rowind)


@pytest.mark.parametrize("seed", range(6))
def test_brute_force_lu(seed, oracle_mod):
    rng = np.random.default_rng(1000 + seed)
    for trial in range(40):
        n = int(rng.integers(1, 48))
        dens = float(rng.choice([0.03, 0.08, 0.15, 0.3]))
        A, a = random_matrix(rng, n, dens, sym=bool(trial % 3 == 0))
        r = oracle_mod.symbfact(A.colbeg, A.colend, A.csr, n)
        want = dense_lu_rows(a)
        np.fill_diagonal(want, False)
        got = structure_from_result(r, n)
        assert np.array_equal(got, want), (seed, trial, n)
        np.fill_diagonal(a, False)
        assert r.fill_count == int(want.sum() - a.sum())


def test_brute_force_larger(oracle_mod):
    rng = np.random.default_rng(7)
    for n, d in ((300, 0.01), (500, 0.006), (800, 0.004)):
        A, a = random_matrix(rng, n, d)
        r = oracle_mod.symbfact(A.colbeg, A.colend, A.csr, n, nthreads=4)
        want = dense_lu_rows(a)
        np.fill_diagonal(want, False)
        assert np.array_equal(structure_from_result(r, n), want)


def _py_traversal(rows, n, s, mistake=None):
    """Small pure-Python fill-path traversal with switchable plausible mistakes
    (used only to show that the brute-force pin rejects them)."""
    INF = float("inf")
    lab = [INF] * n
    mark = set(rows[s])
    fills = set()
    lab[s] = -1
    q = []
    for j in rows[s]:
        if j < s:
            lab[j] = -1
            q.append(j)
    while q:
        f = q.pop(0)
        cost = lab[f] if mistake == "cost_f" else max(f, lab[f])
        for m in rows[f]:
            if m > s and mistake == "through_higher" and m < n and lab[m] > cost:
                lab[m] = cost
                q.append(m)
            if m > s:
                if m not in mark:
                    mark.add(m)
                    fills.add(m)
                continue
            if lab[m] > cost:
                lab[m] = cost
                if m < s:
                    q.append(m)
                ok = m > cost
                if mistake == "no_L":
                    ok = False
                if ok and m not in mark:
                    mark.add(m)
                    fills.add(m)
    if mistake == "self":
        fills.add(s)
    return fills


@pytest.mark.parametrize("mistake", ["cost_f", "through_higher", "no_L", "self"])
def test_planted_mistakes_are_caught(mistake):
    """Each variant of the fill rule disagrees with boolean LU somewhere."""
    rng = np.random.default_rng(99)
    caught = False
    for trial in range(200):
        n = int(rng.integers(4, 24))
        A, a = random_matrix(rng, n, 0.15)
        rows = [list(map(int, r)) for r in A.rows()]
        want = dense_lu_rows(a)
        np.fill_diagonal(want, False)
        got = np.zeros((n, n), bool)
        good = np.zeros((n, n), bool)
        for s in range(n):
            for j in rows[s]:
                if j != s:
                    got[s, j] = good[s, j] = True
            for m in _py_traversal(rows, n, s, mistake):
                got[s, m] = True
            for m in _py_traversal(rows, n, s, None):
                good[s, m] = True
        assert np.array_equal(good, want)   # the unmutated restatement agrees
        if not np.array_equal(got, want):
            caught = True
            break
    assert caught


# ------------------------------------------------------------- supernodes ----
def _literal_supernodes(L_rowptr, L_col, U_rowptr, max_size):
    n = L_rowptr.size - 1
    starts = [0] if n else []
    size = 1
    for i in range(1, n):
        Li = set(L_col[L_rowptr[i]:L_rowptr[i + 1]].tolist())
        ui = U_rowptr[i + 1] - U_rowptr[i]
        up = U_rowptr[i] - U_rowptr[i - 1]
        if (i - 1) in Li and ui == up - 1 and (max_size <= 0 or size < max_size):
            size += 1
        else:
            starts.append(i)
            size = 1
    return np.array(starts + [n], dtype=np.int32)


@pytest.mark.parametrize("max_size", [0, 1, 2, 5, 32])
def test_supernodes_definition(max_size, oracle_mod, bcsstk18):
    rng = np.random.default_rng(5)
    cases = [bcsstk18]
    for _ in range(10):
        n = int(rng.integers(1, 120))
        cases.append(random_matrix(rng, n, float(rng.choice([0.02, 0.1, 0.3])))[0])
    for A in cases:
        r = oracle_mod.symbfact(A.colbeg, A.colend, A.csr, A.n)
        ss, sn = oracle_mod.supernodes(r.L_rowptr, r.L_col, r.U_rowptr, max_size)
        want = _literal_supernodes(r.L_rowptr, r.L_col, r.U_rowptr, max_size)
        np.testing.assert_array_equal(ss, want)
        np.testing.assert_array_equal(sn, np.repeat(np.arange(want.size - 1), np.diff(want)))
        if max_size > 0:
            assert np.all(np.diff(ss) <= max_size)


def test_supernode_pass_through_implies_u_subset(oracle_mod, bcsstk18):
    """(i-1) in L_i implies U_i contains U_{i-1} minus {i}: the count test then
    means U_i = U_{i-1} \\ {i} (row form of the fundamental-supernode test)."""
    r = oracle_mod.symbfact(bcsstk18.colbeg, bcsstk18.colend, bcsstk18.csr, bcsstk18.n)
    for i in range(1, bcsstk18.n):
        Li = r.L_col[r.L_rowptr[i]:r.L_rowptr[i + 1]]
        if Li.size and Li[-1] == i - 1:
            Ui = set(r.U_col[r.U_rowptr[i]:r.U_rowptr[i + 1]].tolist())
            Up = set(r.U_col[r.U_rowptr[i - 1]:r.U_rowptr[i]].tolist()) - {i}
            assert Up <= Ui


# ----------------------------------------------------------- invariances ----
def test_share_split_is_a_partition(oracle_mod, bcsstk18):
    A = bcsstk18
    full = oracle_mod.symbfact(A.colbeg, A.colend, A.csr, A.n)
    total = 0
    for g in range(3):
        r = oracle_mod.symbfact(A.colbeg, A.colend, A.csr, A.n, n_virtual_gpu=3, virtual_gpu_id=g)
        rows = np.arange(g, A.n, 3)
        np.testing.assert_array_equal(r.row_fill[rows], full.row_fill[rows])
        total += r.fill_count
    assert total == full.fill_count


# ------------------------------------------------------------ generators ----
def test_generator_fingerprints():
    import gsofa_amd.configs as cf
    assert cf.grid64_natural().fingerprint() == "ac2231caa27d891dcf1df8dac36ec1c69f053938b131323f61a848a807a80fb1"
    A2 = cf.grid3d_nd()
    assert (A2.n, A2.nnz) == (110592, 765883)
    assert A2.fingerprint() == "975e7eccefe6191bfb2377b623a82c046deaac43f4a799c98eec6ea814aaaa95"
    assert cf.band(20000).fingerprint() == "da1b1f02c747829033ec33420d54b30111148fc3d70a228217e6c49a1866b725"
    sides = cf.batch_sides()
    assert hashlib.sha256(json.dumps(sides).encode()).hexdigest() == \
        "e4f302f6262612110c71e4a0f62f58d332d90b1da4c2fa6fcff83eb1d8295976"
    assert min(sides) >= 16 and max(sides) <= 256
    assert cf.batch_matrix(sides[0], 0).fingerprint() == \
        "2aca5a17786135fd62c3e05170ef4471e0ff98836ca4f9d23a2bc9c27aefd424"


def test_generator_properties():
    import gsofa_amd as g
    A = g.gen_band(3000, 256, 8.0, 11)
    for i, r in enumerate(A.rows()):
        assert i in r and np.all(np.abs(r - i) <= 256) and np.unique(r).size == r.size
    # mean off-diagonal count ~ 8 (minus duplicates)
    assert 7.0 < (A.nnz - A.n) / A.n < 8.1
    G = g.gen_grid(10, 9, 8, nd_leaf=64)
    rows = G.rows()
    # ND ordering is a permutation of a symmetric 7-point pattern
    pairs = {(i, int(j)) for i, r in enumerate(rows) for j in r}
    assert all((j, i) in pairs for i, j in pairs)
    assert sorted(len(r) for r in rows)[0] == 4 and max(len(r) for r in rows) == 7


def test_nd_separator_numbered_last():
    import gsofa_amd as g
    # 1D-like grid 9x1x1 with leaf 1: split at x=4, left 0..3, right 5..8, sep 4
    G = g.gen_grid(9, 1, 1, nd_leaf=1)
    rows = G.rows()
    # the separator (x=4) is numbered last: its row has two neighbours
    # x order 0,1,3,2 | 5,6,8,7 | 4 -> new ids; x=4 neighbours x=3 (2) and x=5 (4)
    assert sorted(rows[8].tolist()) == [2, 4, 8]


# ------------------------------------------------------------- C surface ----
def test_header_symbols_exported():
    from gsofa_amd import _lib
    hdr = open(os.path.join(ROOT, "include", "gsofa.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    names = sorted(set(re.findall(r"\b(gsofa_[a-z0-9_]+)\s*\(", hdr)))
    assert names, "no declarations found"
    assert sorted(_lib.EXPORTS) == names
    L = _lib.lib()
    for nm in names:
        assert hasattr(L, nm), nm


def test_ascii_roundtrip(tmp_path, bcsstk18):
    import gsofa_amd as g
    p = [str(tmp_path / x) for x in ("b.dat", "e.dat", "c.dat")]
    g.write_ascii(bcsstk18, *p)
    B = g.read_ascii(*p)
    assert B.n == bcsstk18.n and B.nnz == bcsstk18.nnz
    np.testing.assert_array_equal(B.colbeg, bcsstk18.colbeg)
    np.testing.assert_array_equal(B.colend, bcsstk18.colend)
    np.testing.assert_array_equal(B.csr, bcsstk18.csr)


def test_ascii_reader_errors(tmp_path):
    import gsofa_amd as g
    with pytest.raises(g.GsofaError):
        g.read_ascii(str(tmp_path / "nope1"), str(tmp_path / "nope2"), str(tmp_path / "nope3"))


def test_cli_usage_error():
    exe = os.path.join(ROOT, "gsofa_amd", "bin", "gsofa")
    p = subprocess.run([exe, "a", "b"], capture_output=True, text=True)
    assert p.returncode == 1 and "usage" in p.stderr


def test_invalid_matrix_rejected_before_device():
    """Malformed input is rejected with a status code (no exit())."""
    import gsofa_amd as g
    A = g.CsrMatrix(2, [0, 1], [1, 3], [0, 1])  # colend beyond nnz
    with pytest.raises(g.GsofaError) as ei:
        g.symbfact(A)
    assert ei.value.code in (-1, -6)


def test_server_argument_checks():
    """Server creation validates its options before touching a device; without
    a GPU it reports GSOFA_ERR_NO_DEVICE rather than failing later."""
    import gsofa_amd as g
    with pytest.raises(g.GsofaError) as ei:
        g.Server(max_inflight=0)
    assert ei.value.code == -1
    # spilled output is the server's own choice per request: the flag is refused
    import ctypes as C
    from gsofa_amd import _lib
    o = _lib.ServerOptions()
    _lib.lib().gsofa_server_options_default(C.byref(o))
    o.flags = g.GSOFA_FLAG_SPILL
    h = C.c_void_p()
    assert _lib.lib().gsofa_server_create(C.byref(o), C.byref(h)) == -1 and not h.value
    try:
        import torch
        has_gpu = torch.cuda.is_available()
    except ImportError:
        has_gpu = False
    if not has_gpu:
        with pytest.raises(g.GsofaError) as ei:
            g.Server()
        assert ei.value.code == -6


def test_block_slots_argument_check():
    """Slots per workgroup other than 32, 16 or 8 are refused before any
    device work (plan creation and the footprint query)."""
    import ctypes as C
    import gsofa_amd as g
    from gsofa_amd import _lib
    A = g.CsrMatrix.from_rows([[0, 1], [0, 1, 2], [1, 2]])
    for bad in (1, 12, 64, -4):
        with pytest.raises(g.GsofaError) as ei:
            g.symbfact(A, block_slots=bad)
        assert ei.value.code == -1, bad
        o = g._options(block_slots=bad)
        m = A.to_c()
        assert _lib.lib().gsofa_plan_footprint(C.byref(m), C.byref(o)) == -1, bad
    assert g._options(block_slots=8).block_slots == 8
    assert g._options().block_slots == 0


def test_spill_refused_for_batches():
    """GSOFA_FLAG_SPILL is for single matrices; a batch with it is refused
    before any device work."""
    import gsofa_amd as g
    A = g.CsrMatrix.from_rows([[0, 1], [0, 1, 2], [1, 2]])
    with pytest.raises(g.GsofaError) as ei:
        g.Plan([A, A], spill=True)
    assert ei.value.code == -1
    assert g._options(spill=True).flags & g.GSOFA_FLAG_SPILL


def test_tune_signature_and_table(tmp_path):
    """Autotuner host logic: signatures bucket sizes, the library's text table
    round-trips, every candidate is a symbfact option."""
    import gsofa_amd as g
    import gsofa_amd.configs as cf
    from gsofa_amd import tune
    A = g.CsrMatrix.from_rows([[0, 1], [0, 1, 2], [1, 2]])
    B = g.CsrMatrix.from_rows([[0], [1], [2], [3]])
    assert tune.signature(A) == tune.signature(A)
    assert tune.signature([A, B]) != tune.signature(A)
    assert tune.signature(cf.grid64_natural()) == "n2^12_deg2^2_b1"
    assert tune.signature(cf.grid2d_nd_batch()) == "n2^16_deg2^2_b256"
    path = str(tmp_path / "t.txt")
    assert tune.best_options(A, table=path) == {}
    tune.write_table({tune.signature(A): {"mode": "reach", "block_slots": 32, "grid_blocks": 3},
                      tune.signature(B): {"mode": "grid", "chunk": 2048, "queue_capacity": 1 << 22}}, path)
    assert tune.best_options(A, table=path) == {"mode": "reach", "block_slots": 32, "grid_blocks": 3}
    assert tune.best_options(B, table=path) == {"mode": "grid", "chunk": 2048, "queue_capacity": 1 << 22}
    line = [x for x in open(path) if x.startswith(tune.signature(A))][0].split()
    assert line[1:] == ["mode=4", "block_slots=32", "grid_blocks=3"]
    for c in tune.REACH_CANDIDATES + tune.BLOCK_CANDIDATES + tune.GRID_CANDIDATES:
        g._options(**c)
