import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# the tests check the library's own defaults; the tuned table it would apply
# is exercised on its own (test_gpu_tuned.py, in a subprocess)
os.environ.setdefault("GSOFA_TUNE_CACHE", "off")
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and libgsofa.so")
    # GPU runs: the oracle computes in its own CPU-only process (oracle.py)
    mark = config.getoption("-m") or ""
    if "gpu" in mark and "not gpu" not in mark:
        os.environ.setdefault("ORACLE_PROCESS", "1")


@pytest.fixture(scope="session")
def bcsstk18():
    import gsofa_amd as g
    d = np.load(os.path.join(ROOT, "tests", "golden", "bcsstk18.npz"))
    return g.CsrMatrix(int(d["n"]), d["colbeg"], d["colend"], d["csr"])


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle
    return oracle


def random_matrix(rng, n, density, sym=False, shuffle_rows=True):
    """Random pattern with the diagonal present, rows in random order and a
    non-monotone colbeg (the reference dataset's layout)."""
    import gsofa_amd as g
    a = rng.random((n, n)) < density
    if sym:
        a = a | a.T
    np.fill_diagonal(a, True)
    rows = [np.flatnonzero(a[i]).astype(np.int32) for i in range(n)]
    for r in rows:
        rng.shuffle(r)
    order = rng.permutation(n) if shuffle_rows else np.arange(n)
    lens = np.array([rows[i].size for i in range(n)])
    beg = np.zeros(n, np.int64)
    pos = 0
    parts = []
    for i in order:
        beg[i] = pos
        parts.append(rows[i])
        pos += lens[i]
    csr = np.concatenate(parts) if parts else np.zeros(0, np.int32)
    return g.CsrMatrix(n, beg, beg + lens, csr), a
