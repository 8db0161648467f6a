"""gsofa_amd -- MI355X fill-path symbolic factorization (Python host mirror).

A thin layer over the C ABI in include/gsofa.h (libgsofa.so).  It mirrors the
reference's operator surface for this path:

* ``symbfact(A, n_virtual_gpu=..., mem_gb=...)`` stands for the reference's
  ``symbfact_min_id`` (with_space_optimization/symbfact.cuh:1553): same inputs
  (colbeg/colend/csr, N_GPU, MEM_GB, max_supernode_size), same fill count, plus
  the sorted L/U structure and the supernode partition;
* ``read_ascii`` / ``write_ascii`` handle the reference's three-file format
  (with_space_optimization/symbfact.cuh:1591-1640);
* ``Plan`` keeps a matrix (or a batch of matrices) resident on the device for
  repeated execution;
* ``Server`` admits requests as they arrive and runs them in continuously
  formed batches (no counterpart in the reference).

Errors surface as ``GsofaError`` carrying the library status code.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import GSOFA_FLAG_COUNTS_ONLY, GSOFA_FLAG_GRAPH, GSOFA_FLAG_SPILL

__all__ = ["CsrMatrix", "SymbfactResult", "GsofaError", "Plan", "Server", "symbfact", "gen_grid",
           "gen_band", "read_ascii", "write_ascii", "footprint", "GSOFA_FLAG_COUNTS_ONLY", "GSOFA_FLAG_SPILL",
           "GSOFA_FLAG_GRAPH"]


class GsofaError(RuntimeError):
    def __init__(self, code, what=""):
        self.code = code
        msg = _lib.lib().gsofa_strerror(code).decode()
        super().__init__("%s%s (status %d)" % (what + ": " if what else "", msg, code))


def _check(rc, what=""):
    if rc != 0:
        raise GsofaError(rc, what)


class CsrMatrix:
    """Row i's structure is csr[colbeg[i]:colend[i]] (the reference's layout)."""

    def __init__(self, n, colbeg, colend, csr):
        self.n = int(n)
        self.colbeg = np.ascontiguousarray(colbeg, dtype=np.int32)
        self.colend = np.ascontiguousarray(colend, dtype=np.int32)
        self.csr = np.ascontiguousarray(csr, dtype=np.int32)
        assert self.colbeg.size == self.n and self.colend.size == self.n

    @property
    def nnz(self):
        return int(self.csr.size)

    @classmethod
    def from_rows(cls, rows):
        lens = np.array([len(r) for r in rows], dtype=np.int64)
        end = np.cumsum(lens)
        beg = end - lens
        csr = np.concatenate([np.asarray(r, dtype=np.int32) for r in rows]) if rows else np.zeros(0, np.int32)
        return cls(len(rows), beg, end, csr)

    def rows(self):
        return [self.csr[b:e] for b, e in zip(self.colbeg, self.colend)]

    def to_c(self):
        return _lib.Matrix(self.n, self.nnz, self.colbeg.ctypes.data, self.colend.ctypes.data,
                           self.csr.ctypes.data)

    def fingerprint(self):
        """Order-independent-of-layout digest of the pattern (rows sorted)."""
        import hashlib
        h = hashlib.sha256()
        h.update(np.int64(self.n).tobytes())
        for b, e in zip(self.colbeg, self.colend):
            r = np.sort(self.csr[b:e])
            h.update(np.int64(r.size).tobytes())
            h.update(r.astype(np.int32).tobytes())
        return h.hexdigest()


def _from_csr(c):
    try:
        n, nnz = int(c.n), int(c.nnz)

        def arr(ptr, cnt):
            if cnt == 0:
                return np.zeros(0, np.int32)
            return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(C.c_int32)), shape=(cnt,)).copy()
        return CsrMatrix(n, arr(c.colbeg, n), arr(c.colend, n), arr(c.csr, nnz))
    finally:
        _lib.lib().gsofa_csr_free(C.byref(c))


def gen_grid(nx, ny, nz=1, nd_leaf=0, extra_prob=0.0, extra_dist=0, seed=0):
    c = _lib.Csr()
    _check(_lib.lib().gsofa_gen_grid(nx, ny, nz, nd_leaf, float(extra_prob), extra_dist, seed,
                                     C.byref(c)), "gen_grid")
    return _from_csr(c)


def gen_band(n, band, lam, seed):
    c = _lib.Csr()
    _check(_lib.lib().gsofa_gen_band(n, band, float(lam), seed, C.byref(c)), "gen_band")
    return _from_csr(c)


def read_ascii(colbeg_path, colend_path, csr_path):
    c = _lib.Csr()
    _check(_lib.lib().gsofa_read_ascii(colbeg_path.encode(), colend_path.encode(),
                                       csr_path.encode(), C.byref(c)), "read_ascii")
    return _from_csr(c)


def read_binary(path):
    """Binary triple (include/gsofa.h, gsofa_read_binary) as a CsrMatrix."""
    c = _lib.Csr()
    _check(_lib.lib().gsofa_read_binary(path.encode(), C.byref(c)), "read_binary")
    return _from_csr(c)


def order_nd(A, leaf=64):
    """Nested-dissection ordering of A's symmetrized graph (gsofa_order_nd):
    perm[k] = original index of the k-th vertex."""
    m = A.to_c()
    perm = np.zeros(max(A.n, 1), np.int32)
    _check(_lib.lib().gsofa_order_nd(C.byref(m), int(leaf), perm.ctypes.data), "order_nd")
    return perm[:A.n].copy()


def permute(A, perm):
    """B = P A P^T (gsofa_permute): row k of B is row perm[k] of A."""
    m = A.to_c()
    p = np.ascontiguousarray(perm, dtype=np.int32)
    if p.size != A.n:
        raise GsofaError(-1, "permute: perm has %d entries for n=%d" % (p.size, A.n))
    c = _lib.Csr()
    _check(_lib.lib().gsofa_permute(C.byref(m), p.ctypes.data if p.size else None, C.byref(c)), "permute")
    return _from_csr(c)


def write_binary(A, path):
    m = A.to_c()
    _check(_lib.lib().gsofa_write_binary(C.byref(m), path.encode()), "write_binary")


class MappedMatrix:
    """A binary triple mapped read-only (gsofa_map_binary).  ``matrix`` is a
    CsrMatrix whose arrays view the mapping: valid until close()."""

    def __init__(self, path):
        h = C.c_void_p()
        v = _lib.Matrix()
        _check(_lib.lib().gsofa_map_binary(path.encode(), C.byref(h), C.byref(v)), "map_binary")
        self._h = h
        n, nnz = int(v.n), int(v.nnz)

        def view(ptr, cnt):
            if cnt == 0:
                return np.zeros(0, np.int32)
            a = np.ctypeslib.as_array(C.cast(ptr, C.POINTER(C.c_int32)), shape=(cnt,))
            a.flags.writeable = False
            return a
        self.matrix = CsrMatrix(n, view(v.colbeg, n), view(v.colend, n), view(v.csr, nnz))

    def close(self):
        if getattr(self, "_h", None):
            self.matrix = None
            _lib.lib().gsofa_unmap_binary(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        self.close()


def write_ascii(A, colbeg_path, colend_path, csr_path):
    m = A.to_c()
    _check(_lib.lib().gsofa_write_ascii(C.byref(m), colbeg_path.encode(), colend_path.encode(),
                                        csr_path.encode()), "write_ascii")


def _options(n_virtual_gpu=1, virtual_gpu_id=-1, mem_gb=0.0, mem_budget_bytes=0,
             max_supernode_size=0, chunk=0, wave_block_threshold=0, device=0, queue_capacity=0,
             output_capacity=0, counts_only=False, grid_blocks=0, mode=0, top_sources=0,
             block_slots=0, spill=False, graph=False, workgroup_threads=0):
    o = _lib.Options()
    _lib.lib().gsofa_options_default(C.byref(o))
    o.n_virtual_gpu = n_virtual_gpu
    o.virtual_gpu_id = virtual_gpu_id
    o.mem_budget_bytes = int(mem_budget_bytes or mem_gb * (1 << 30))
    o.max_supernode_size = max_supernode_size
    o.chunk = chunk
    if wave_block_threshold:
        o.wave_block_threshold = wave_block_threshold
    o.device = device
    o.queue_capacity = queue_capacity
    o.output_capacity = output_capacity
    o.flags = ((GSOFA_FLAG_COUNTS_ONLY if counts_only else 0) | (GSOFA_FLAG_SPILL if spill else 0) |
               (GSOFA_FLAG_GRAPH if graph else 0))
    o.grid_blocks = grid_blocks
    o.mode = {"auto": 0, "block": 1, "grid": 2, "reach": 4}.get(mode, mode)
    o.top_sources = top_sources
    o.block_slots = block_slots
    o.workgroup_threads = workgroup_threads
    return o


class _Owner:
    """Holds a library-owned gsofa_result and releases it (gsofa_result_free)
    when the last array viewing its memory is gone."""

    def __init__(self, r):
        self.r = r

    def __del__(self):
        if self.r is not None:
            try:
                _lib.lib().gsofa_result_free(C.byref(self.r))
            except Exception:  # interpreter shutdown: the process releases it
                pass
            self.r = None


class _View:
    """numpy view of `cnt` elements at `ptr`, keeping `owner` alive."""

    def __init__(self, ptr, cnt, dt, owner):
        self.owner = owner
        self.__array_interface__ = {"data": (int(ptr), False), "shape": (int(cnt),),
                                    "typestr": np.dtype(dt).str, "version": 3}


class SymbfactResult:
    """Sorted strictly-lower (L) / strictly-upper (U) structure of L+U by row.

    Takes ownership of the library result `r`: the arrays are views of its
    memory (no copy), released with the last of them."""

    def __init__(self, r):
        owner = _Owner(r)
        n = int(r.n)
        self.n = n
        self.fill_count = int(r.fill_count)
        self.te_alg = int(r.te_alg)
        self.reach = int(r.reach)
        self.edge_checks = int(r.edge_checks)
        self.ms = float(r.ms)
        self.traverse_ms = float(r.traverse_ms)
        self.share_ms_min = float(r.share_ms_min)
        self.share_ms_max = float(r.share_ms_max)
        self.chunk = int(r.chunk)
        self.mode = {1: "block", 2: "grid", 3: "hybrid", 4: "reach"}.get(int(r.mode), int(r.mode))
        self.grid_blocks = int(r.grid_blocks)
        self.queue_capacity = int(r.queue_capacity)
        self.levels = int(r.levels)
        self.rescans = int(r.rescans)
        self.n_super = int(r.n_super)
        self.passes = int(r.passes)
        self.graph_replay = bool(r.graph_replay)

        def arr(ptr, cnt, ct, dt):
            if not ptr or cnt <= 0:
                return np.zeros(max(cnt, 0), dt) if ptr or cnt == 0 else None
            return np.asarray(_View(ptr, cnt, dt, owner))
        self.row_fill = arr(r.row_fill, n, C.c_int64, np.int64) if r.row_fill else None
        self.row_te = arr(r.row_te, n, C.c_int64, np.int64) if r.row_te else None
        self.row_reach = arr(r.row_reach, n, C.c_int64, np.int64) if r.row_reach else None
        self.L_rowptr = arr(r.L_rowptr, n + 1, C.c_int64, np.int64) if r.L_rowptr else None
        self.U_rowptr = arr(r.U_rowptr, n + 1, C.c_int64, np.int64) if r.U_rowptr else None
        self.L_col = (arr(r.L_col, int(self.L_rowptr[-1]), C.c_int32, np.int32)
                      if self.L_rowptr is not None and r.L_col else None)
        self.U_col = (arr(r.U_col, int(self.U_rowptr[-1]), C.c_int32, np.int32)
                      if self.U_rowptr is not None and r.U_col else None)
        if self.n_super >= 0 and r.super_start:
            self.super_start = arr(r.super_start, self.n_super + 1, C.c_int32, np.int32)
            self.supno = arr(r.supno, n, C.c_int32, np.int32) if r.supno else np.zeros(0, np.int32)
        else:
            self.super_start = None
            self.supno = None


def _carr(ptr, cnt, ct, dt):
    if cnt <= 0:
        return np.zeros(0, dt)
    return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(ct)), shape=(cnt,)).copy()


class SupernodalResult:
    """Supernodal hand-off (include/gsofa.h, gsofa_supernodal): supernode
    starts and ids, per-supernode subscripts, and L and U by columns."""

    def __init__(self, h):
        n, ns = int(h.n), int(h.n_super)
        self.n, self.n_super = n, ns
        self.xsup = _carr(h.xsup, ns + 1, C.c_int32, np.int32)
        self.supno = _carr(h.supno, n, C.c_int32, np.int32)
        self.xsub = _carr(h.xsub, ns + 1, C.c_int64, np.int64)
        self.sub = _carr(h.sub, int(self.xsub[-1]), C.c_int32, np.int32)
        self.xlcol = _carr(h.xlcol, n + 1, C.c_int64, np.int64)
        self.lrow = _carr(h.lrow, int(self.xlcol[-1]), C.c_int32, np.int32)
        self.xucol = _carr(h.xucol, n + 1, C.c_int64, np.int64)
        self.urow = _carr(h.urow, int(self.xucol[-1]), C.c_int32, np.int32)


def symbfact(A, **kw):
    """One-shot symbolic factorization of ``A`` (a CsrMatrix) on the GPU."""
    m = A.to_c()
    o = _options(**kw)
    r = _lib.Result()
    _check(_lib.lib().gsofa_symbfact(C.byref(m), C.byref(o), C.byref(r)), "symbfact")
    return SymbfactResult(r)


def footprint(A, **kw):
    m = A.to_c()
    o = _options(**kw)
    return int(_lib.lib().gsofa_plan_footprint(C.byref(m), C.byref(o)))


class Plan:
    """A matrix (or a batch of matrices) uploaded once; execute() runs the
    device path, fetch() / fetch_matrix(i) copy results to the host."""

    def __init__(self, A, **kw):
        self._A = A  # keep the borrowed arrays alive during create
        o = _options(**kw)
        h = C.c_void_p()
        if isinstance(A, (list, tuple)):
            mats = (_lib.Matrix * len(A))(*[M.to_c() for M in A])
            _check(_lib.lib().gsofa_batch_create(mats, len(A), C.byref(o), C.byref(h)), "batch_create")
        else:
            m = A.to_c()
            _check(_lib.lib().gsofa_plan_create(C.byref(m), C.byref(o), C.byref(h)), "plan_create")
        self._h = h

    @property
    def matrix_count(self):
        return int(_lib.lib().gsofa_plan_matrix_count(self._h))

    def fetch_matrix(self, index):
        r = _lib.Result()
        _check(_lib.lib().gsofa_plan_fetch_matrix(self._h, index, C.byref(r)), "plan_fetch_matrix")
        return SymbfactResult(r)

    def execute(self):
        _check(_lib.lib().gsofa_plan_execute(self._h), "plan_execute")

    def fetch(self):
        r = _lib.Result()
        _check(_lib.lib().gsofa_plan_fetch(self._h, C.byref(r)), "plan_fetch")
        return SymbfactResult(r)

    def fetch_supernodal(self):
        """SuperLU_DIST-style hand-off of the last execute (single matrix,
        every share, structure built)."""
        h = _lib.Supernodal()
        _check(_lib.lib().gsofa_plan_fetch_supernodal(self._h, C.byref(h)), "plan_fetch_supernodal")
        try:
            return SupernodalResult(h)
        finally:
            _lib.lib().gsofa_supernodal_free(C.byref(h))

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().gsofa_plan_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        self.close()


class Server:
    """Request server with continuous batching (include/gsofa.h, gsofa_server_*).

    submit(A) copies the matrix and returns a ticket; wait(ticket) / poll(ticket)
    return (SymbfactResult, info) once the request is final, where info holds
    queued_ms, latency_ms, batch_ms, batch_id, batch_size and spill_shares.
    """

    def __init__(self, device=0, mem_gb=0.0, max_inflight=2, max_batch=256, max_supernode_size=0,
                 counts_only=False):
        o = _lib.ServerOptions()
        _lib.lib().gsofa_server_options_default(C.byref(o))
        o.device = device
        o.mem_budget_bytes = int(mem_gb * (1 << 30))
        o.max_inflight = max_inflight
        o.max_batch = max_batch
        o.max_supernode_size = max_supernode_size
        o.flags = GSOFA_FLAG_COUNTS_ONLY if counts_only else 0
        h = C.c_void_p()
        _check(_lib.lib().gsofa_server_create(C.byref(o), C.byref(h)), "server_create")
        self._h = h

    def submit(self, A):
        m = A.to_c()
        t = C.c_int64()
        _check(_lib.lib().gsofa_server_submit(self._h, C.byref(m), C.byref(t)), "server_submit")
        return int(t.value)

    def _final(self, rc, r, info):
        _check(rc, "server request")
        res = SymbfactResult(r)
        return res, {"queued_ms": info.queued_ms, "latency_ms": info.latency_ms,
                     "batch_ms": info.batch_ms, "batch_id": info.batch_id,
                     "batch_size": info.batch_size, "spill_shares": info.spill_shares}

    def poll(self, ticket):
        """None while the request runs, else (result, info)."""
        r, info = _lib.Result(), _lib.RequestInfo()
        rc = _lib.lib().gsofa_server_poll(self._h, ticket, C.byref(r), C.byref(info))
        if rc == _lib.GSOFA_PENDING:
            return None
        return self._final(rc, r, info)

    def wait(self, ticket):
        r, info = _lib.Result(), _lib.RequestInfo()
        rc = _lib.lib().gsofa_server_wait(self._h, ticket, C.byref(r), C.byref(info))
        return self._final(rc, r, info)

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().gsofa_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        self.close()
