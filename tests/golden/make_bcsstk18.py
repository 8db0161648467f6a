"""Write tests/golden/bcsstk18.npz from the reference's shipped dataset.

Input (read here only, never at test time): the three ASCII files of the
reference's BCSSTK18 dump, with_space_optimization/dataset/{colbeg,colend,
csr}_11948.dat (csr file: nnz, n, then nnz column indices; colbeg/colend: n
integers each).  BCSSTK18 is a public Harwell-Boeing matrix; the reference's
README (README.md:27-30) says it was reordered by SuperLU_DIST before the dump.
The arrays are stored unchanged (rows unsorted, colbeg non-monotone).
"""
import os
import sys

import numpy as np

REF = "/root/reference/with_space_optimization/dataset"


def read_ints(path):
    with open(path) as f:
        return np.array(f.read().split(), dtype=np.int64)


def main(out=None):
    out = out or os.path.join(os.path.dirname(os.path.abspath(__file__)), "bcsstk18.npz")
    csr_raw = read_ints(os.path.join(REF, "csr_11948.dat"))
    nnz, n = int(csr_raw[0]), int(csr_raw[1])
    csr = csr_raw[2:2 + nnz].astype(np.int32)
    colbeg = read_ints(os.path.join(REF, "colbeg_11948.dat"))[:n].astype(np.int32)
    colend = read_ints(os.path.join(REF, "colend_11948.dat"))[:n].astype(np.int32)
    assert csr.size == nnz and colbeg.size == n and colend.size == n
    np.savez_compressed(out, n=np.int64(n), nnz=np.int64(nnz), colbeg=colbeg,
                        colend=colend, csr=csr)
    print(out, n, nnz)


if __name__ == "__main__":
    main(*sys.argv[1:])
