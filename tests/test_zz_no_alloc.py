"""GPU: gsofa_counts_noalloc, the allocation-free counts-only run, through the
C ABI against the existing counts-only entry point (gsofa_symbfact, block
mode) and the oracle: two calls in a row, the second on a smaller matrix, so
the scratch's leftovers from the first call must not leak into the second.
Device buffers for inputs and outputs are the caller's (torch here)."""
import sys

import numpy as np
import pytest

from conftest import random_matrix

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _packed(A):
    rows = [np.unique(A.csr[A.colbeg[i]:A.colend[i]]) for i in range(A.n)]
    rp = np.zeros(A.n + 1, np.int32)
    rp[1:] = np.cumsum([len(r) for r in rows])
    ci = np.concatenate(rows).astype(np.int32) if rows else np.zeros(0, np.int32)
    return rp, ci


def _noalloc(A):
    import torch

    from gsofa_amd import _lib
    rp, ci = _packed(A)
    d_rp = torch.from_numpy(rp).cuda()
    d_ci = torch.from_numpy(ci).cuda()
    outs = [torch.full((A.n,), -7, dtype=torch.int64, device="cuda") for _ in range(3)]
    tot = torch.zeros(4, dtype=torch.int64, device="cuda")
    torch.cuda.synchronize()
    rc = _lib.lib().gsofa_counts_noalloc(A.n, int(ci.size), d_rp.data_ptr(), d_ci.data_ptr(),
                                         outs[0].data_ptr(), outs[1].data_ptr(), outs[2].data_ptr(),
                                         tot.data_ptr(), 0, None)
    assert rc == 0, rc
    torch.cuda.synchronize()
    return [o.cpu().numpy() for o in outs], tot.cpu().numpy()


def test_noalloc_matches_counts_only_entry_point(bcsstk18):
    import gsofa_amd as g
    from oracle import oracle
    rng = np.random.default_rng(11)
    small, _ = random_matrix(rng, 300, 0.02)
    for A in (bcsstk18, small):  # two calls in a row, the second smaller
        _say("case: counts without allocation, n", A.n)
        (fill, te, reach), tot = _noalloc(A)
        r = g.symbfact(A, counts_only=True, mode="block")
        o = oracle.symbfact(A.colbeg, A.colend, A.csr, A.n)
        np.testing.assert_array_equal(fill, r.row_fill)
        np.testing.assert_array_equal(te, r.row_te)
        np.testing.assert_array_equal(reach, r.row_reach)
        np.testing.assert_array_equal(fill, o.row_fill)
        assert (int(tot[0]), int(tot[1]), int(tot[2])) == (r.fill_count, r.te_alg, r.reach)
        assert int(tot[3]) >= int(tot[1])
    assert int(_noalloc(bcsstk18)[1][0]) == 980136  # and a third, larger again


def test_noalloc_rejects_what_does_not_fit():
    """A matrix whose one workgroup does not fit the scratch is refused
    (GSOFA_ERR_BUDGET) before any work is queued."""
    import torch

    from gsofa_amd import _lib
    d = torch.zeros(16, dtype=torch.int64, device="cuda")
    p = d.data_ptr()
    n = 1_000_000
    assert _lib.lib().gsofa_counts_noalloc(n, 1, p, p, p, p, p, p, 0, None) == -7
    assert _lib.lib().gsofa_counts_noalloc(0, 0, p, p, p, p, p, p, 0, None) == -1
