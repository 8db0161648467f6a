"""ctypes binding of the CPU oracle (oracle/liboracle.so).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker, never as the thing
measured or shipped.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None


class _Matrix(C.Structure):
    _fields_ = [("n", C.c_int32), ("nnz", C.c_int64), ("colbeg", C.c_void_p),
                ("colend", C.c_void_p), ("csr", C.c_void_p)]


class _Result(C.Structure):
    _fields_ = [("n", C.c_int32), ("fill_count", C.c_uint64), ("te_alg", C.c_uint64),
                ("reach", C.c_uint64), ("edge_checks", C.c_uint64),
                ("row_fill", C.c_void_p), ("row_te", C.c_void_p), ("row_reach", C.c_void_p),
                ("L_rowptr", C.c_void_p), ("L_col", C.c_void_p),
                ("U_rowptr", C.c_void_p), ("U_col", C.c_void_p)]


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        L.orc_symbfact.argtypes = [C.POINTER(_Matrix), C.c_int, C.c_int, C.c_int, C.POINTER(_Result)]
        L.orc_symbfact.restype = C.c_int
        L.orc_result_free.argtypes = [C.POINTER(_Result)]
        L.orc_count_rows.argtypes = [C.POINTER(_Matrix), C.c_int32, C.c_int32, C.c_int,
                                     C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
        L.orc_count_rows.restype = C.c_double
        L.orc_supernodes.argtypes = [C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32,
                                     C.c_void_p, C.c_void_p]
        L.orc_supernodes.restype = C.c_int32
        L.orc_brute_lu.argtypes = [C.c_int32, C.c_void_p]
        L.orc_brute_lu.restype = C.c_int
        L.orc_symbolic_cholesky.argtypes = [C.POINTER(_Matrix), C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)]
        L.orc_symbolic_cholesky.restype = C.c_int64
        L.orc_free.argtypes = [C.c_void_p]
        _lib = L
    return _lib


def _mat(colbeg, colend, csr, n):
    colbeg = np.ascontiguousarray(colbeg, dtype=np.int32)
    colend = np.ascontiguousarray(colend, dtype=np.int32)
    csr = np.ascontiguousarray(csr, dtype=np.int32)
    m = _Matrix(int(n), int(csr.size), colbeg.ctypes.data, colend.ctypes.data, csr.ctypes.data)
    return m, (colbeg, colend, csr)


def _arr(ptr, count, dtype):
    if count == 0:
        return np.zeros(0, dtype=dtype)
    ct = {np.int64: C.c_int64, np.int32: C.c_int32}[dtype]
    return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(ct)), shape=(count,)).copy()


class OracleResult:
    """Per-row sorted strictly-lower (L) / strictly-upper (U) structure of L+U."""

    def __init__(self, **kw):
        self.__dict__.update(kw)


# ORACLE_PROCESS=1: symbfact runs in a separate CPU-only worker process
# (oracle/worker.py, one, started on first use), so the checker's memory is
# never shared with the process that drives the GPU.
_proc = None


def _in_worker(*args):
    global _proc
    import pickle
    import sys
    if _proc is None:
        import atexit
        _proc = subprocess.Popen([sys.executable, "-m", "oracle.worker"], stdin=subprocess.PIPE,
                                 stdout=subprocess.PIPE, cwd=os.path.dirname(_HERE))
        atexit.register(_proc.stdin.close)
    pickle.dump(args, _proc.stdin, protocol=pickle.HIGHEST_PROTOCOL)
    _proc.stdin.flush()
    ok, res = pickle.load(_proc.stdout)
    if not ok:
        raise ValueError(res)
    return res


def symbfact(colbeg, colend, csr, n, n_virtual_gpu=1, virtual_gpu_id=-1, nthreads=0):
    if os.environ.get("ORACLE_PROCESS") == "1":
        return _in_worker(np.asarray(colbeg), np.asarray(colend), np.asarray(csr), int(n),
                          n_virtual_gpu, virtual_gpu_id, nthreads)
    return _symbfact(colbeg, colend, csr, n, n_virtual_gpu, virtual_gpu_id, nthreads)


def _symbfact(colbeg, colend, csr, n, n_virtual_gpu=1, virtual_gpu_id=-1, nthreads=0):
    m, keep = _mat(colbeg, colend, csr, n)
    r = _Result()
    rc = lib().orc_symbfact(C.byref(m), n_virtual_gpu, virtual_gpu_id, nthreads, C.byref(r))
    if rc != 0:
        raise ValueError("oracle rejected the matrix (code %d)" % rc)
    try:
        n = int(r.n)
        L_rowptr = _arr(r.L_rowptr, n + 1, np.int64)
        U_rowptr = _arr(r.U_rowptr, n + 1, np.int64)
        out = OracleResult(
            n=n, fill_count=int(r.fill_count), te_alg=int(r.te_alg), reach=int(r.reach),
            edge_checks=int(r.edge_checks),
            row_fill=_arr(r.row_fill, n, np.int64), row_te=_arr(r.row_te, n, np.int64),
            row_reach=_arr(r.row_reach, n, np.int64),
            L_rowptr=L_rowptr, L_col=_arr(r.L_col, int(L_rowptr[-1]), np.int32),
            U_rowptr=U_rowptr, U_col=_arr(r.U_col, int(U_rowptr[-1]), np.int32))
    finally:
        lib().orc_result_free(C.byref(r))
    return out


def count_rows(colbeg, colend, csr, n, row_begin, row_end, nthreads=0):
    """Counters only (fills, TE_alg, |R|) for rows [row_begin, row_end); returns
    (seconds, fills, te_alg, reach)."""
    m, keep = _mat(colbeg, colend, csr, n)
    f, t, r = C.c_uint64(), C.c_uint64(), C.c_uint64()
    sec = lib().orc_count_rows(C.byref(m), row_begin, row_end, nthreads, C.byref(f), C.byref(t), C.byref(r))
    if sec < 0:
        raise ValueError("oracle rejected the matrix")
    return sec, int(f.value), int(t.value), int(r.value)


def supernodes(L_rowptr, L_col, U_rowptr, max_size=0):
    L_rowptr = np.ascontiguousarray(L_rowptr, dtype=np.int64)
    L_col = np.ascontiguousarray(L_col, dtype=np.int32)
    U_rowptr = np.ascontiguousarray(U_rowptr, dtype=np.int64)
    n = L_rowptr.size - 1
    ss = np.zeros(n + 1, dtype=np.int32)
    sn = np.zeros(max(n, 1), dtype=np.int32)
    ns = lib().orc_supernodes(n, L_rowptr.ctypes.data, L_col.ctypes.data, U_rowptr.ctypes.data,
                              int(max_size), ss.ctypes.data, sn.ctypes.data)
    return ss[:ns + 1].copy(), sn[:n].copy()


def brute_lu(dense):
    a = np.ascontiguousarray(dense, dtype=np.uint8).copy()
    n = a.shape[0]
    if lib().orc_brute_lu(n, a.ctypes.data) != 0:
        raise ValueError("brute_lu failed")
    return a


def symbolic_cholesky(colbeg, colend, csr, n):
    """Strictly-lower column structure of L for a structurally symmetric pattern."""
    m, keep = _mat(colbeg, colend, csr, n)
    cp, ri = C.c_void_p(), C.c_void_p()
    nnz = lib().orc_symbolic_cholesky(C.byref(m), C.byref(cp), C.byref(ri))
    if nnz < 0:
        raise ValueError("symbolic_cholesky failed")
    colptr = _arr(cp.value, n + 1, np.int64)
    rowind = _arr(ri.value, int(nnz), np.int32)
    lib().orc_free(cp)
    lib().orc_free(ri)
    return colptr, rowind
