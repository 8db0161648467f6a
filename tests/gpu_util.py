"""Helpers shared by the GPU parity tests."""
import numpy as np

from digests import array_digest, lu_digest


def assert_same_structure(res, orc, check_super=True, max_size=0):
    """Bit-exact comparison of a GPU result with the oracle's."""
    from oracle import oracle
    assert res.fill_count == orc.fill_count
    assert res.te_alg == orc.te_alg
    assert res.reach == orc.reach
    np.testing.assert_array_equal(res.L_rowptr, orc.L_rowptr)
    np.testing.assert_array_equal(res.U_rowptr, orc.U_rowptr)
    np.testing.assert_array_equal(res.L_col, orc.L_col)
    np.testing.assert_array_equal(res.U_col, orc.U_col)
    np.testing.assert_array_equal(res.row_fill, orc.row_fill)
    if getattr(res, "row_te", None) is not None:  # block mode reports per-row TE and |R|
        np.testing.assert_array_equal(res.row_te, orc.row_te)
        np.testing.assert_array_equal(res.row_reach, orc.row_reach)
    if check_super:
        ss, sn = oracle.supernodes(orc.L_rowptr, orc.L_col, orc.U_rowptr, max_size)
        assert res.n_super == ss.size - 1
        np.testing.assert_array_equal(res.super_start, ss)
        np.testing.assert_array_equal(res.supno, sn)


def record_of(res):
    return {
        "fill_count": res.fill_count, "te_alg": res.te_alg, "reach": res.reach,
        "nnz_L": int(res.L_rowptr[-1]), "nnz_U": int(res.U_rowptr[-1]),
        "lu": lu_digest(res.L_rowptr, res.L_col, res.U_rowptr, res.U_col),
        "row_fill": array_digest(res.row_fill, "<i8"),
    }
