"""The device path on inputs reordered by gsofa_order_nd: bit-exact against
the oracle, and the fill the ordering saves shows up in the GPU's count."""
import sys

import pytest

from gpu_util import assert_same_structure

pytestmark = pytest.mark.gpu


def test_reordered_grids_vs_oracle():
    import gsofa_amd as g
    from oracle import oracle
    for shape in ((64, 64, 1), (16, 16, 16)):
        A = g.gen_grid(*shape, nd_leaf=0, seed=1)
        B = g.permute(A, g.order_nd(A, 64))
        print("case: graph-ND reordered grid", shape, flush=True)
        sys.stdout.flush()
        r = g.symbfact(B)
        o = oracle.symbfact(B.colbeg, B.colend, B.csr, B.n)
        assert_same_structure(r, o)
        assert r.fill_count < g.symbfact(A, counts_only=True).fill_count
