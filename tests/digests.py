"""Digests of symbolic-factorization results, shared by goldens and tests.

A result is summarised by totals and SHA-256 digests of the sorted structure:
  lu      = sha256(L_rowptr int64 | L_col int32 | U_rowptr int64 | U_col int32)
  super0  = sha256(super_start int32) with no size cap
  super32 = sha256(super_start int32) with max_supernode_size = 32
All arrays little-endian, concatenated in that order.
"""
import hashlib

import numpy as np


def lu_digest(L_rowptr, L_col, U_rowptr, U_col):
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(L_rowptr, dtype="<i8").tobytes())
    h.update(np.ascontiguousarray(L_col, dtype="<i4").tobytes())
    h.update(np.ascontiguousarray(U_rowptr, dtype="<i8").tobytes())
    h.update(np.ascontiguousarray(U_col, dtype="<i4").tobytes())
    return h.hexdigest()


def array_digest(a, dtype="<i4"):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=dtype).tobytes()).hexdigest()


def oracle_record(A, r):
    from oracle import oracle
    ss0, _ = oracle.supernodes(r.L_rowptr, r.L_col, r.U_rowptr, 0)
    ss32, _ = oracle.supernodes(r.L_rowptr, r.L_col, r.U_rowptr, 32)
    return {
        "n": int(A.n), "nnz": int(A.nnz), "matrix": A.fingerprint(),
        "fill_count": int(r.fill_count), "te_alg": int(r.te_alg), "reach": int(r.reach),
        "nnz_L": int(r.L_rowptr[-1]), "nnz_U": int(r.U_rowptr[-1]),
        "lu": lu_digest(r.L_rowptr, r.L_col, r.U_rowptr, r.U_col),
        "row_fill": array_digest(r.row_fill, "<i8"),
        "n_super0": int(ss0.size - 1), "super0": array_digest(ss0),
        "n_super32": int(ss32.size - 1), "super32": array_digest(ss32),
    }
