"""ctypes binding of libgsofa.so (the C ABI declared in include/gsofa.h).

The library is built in-tree by ``__graft_entry__.build()`` (or ``make -C
gsofa_amd/csrc``).  There is no fallback: if the shared library is missing,
importing the product API raises.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgsofa.so")

GSOFA_FLAG_COUNTS_ONLY = 0x1
GSOFA_FLAG_SPILL = 0x2
GSOFA_FLAG_GRAPH = 0x4
MODE_AUTO, MODE_BLOCK, MODE_GRID, MODE_HYBRID, MODE_REACH = 0, 1, 2, 3, 4

# Every entry point include/gsofa.h declares (kept in sync by a CPU test).
EXPORTS = [
    "gsofa_strerror", "gsofa_csr_free", "gsofa_options_default", "gsofa_result_free",
    "gsofa_symbfact", "gsofa_plan_create", "gsofa_plan_execute", "gsofa_plan_fetch",
    "gsofa_plan_destroy", "gsofa_plan_footprint", "gsofa_read_ascii", "gsofa_write_ascii",
    "gsofa_gen_grid", "gsofa_gen_band", "gsofa_gen_side_loguniform", "gsofa_gen_uniform_int",
    "gsofa_gen_unit", "gsofa_batch_create", "gsofa_plan_matrix_count", "gsofa_plan_fetch_matrix",
    "gsofa_server_options_default", "gsofa_server_create", "gsofa_server_submit",
    "gsofa_server_poll", "gsofa_server_wait", "gsofa_server_destroy",
    "gsofa_plan_fetch_supernodal", "gsofa_supernodal_free",
    "gsofa_write_binary", "gsofa_read_binary", "gsofa_map_binary", "gsofa_unmap_binary",
    "gsofa_order_nd", "gsofa_permute", "gsofa_fill_bound", "gsofa_counts_noalloc",
]
GSOFA_PENDING = 1


class Matrix(C.Structure):
    _fields_ = [("n", C.c_int32), ("nnz", C.c_int64), ("colbeg", C.c_void_p),
                ("colend", C.c_void_p), ("csr", C.c_void_p)]


class Csr(C.Structure):
    _fields_ = [("n", C.c_int32), ("nnz", C.c_int64), ("colbeg", C.c_void_p),
                ("colend", C.c_void_p), ("csr", C.c_void_p)]


class Options(C.Structure):
    _fields_ = [("n_virtual_gpu", C.c_int32), ("virtual_gpu_id", C.c_int32),
                ("mem_budget_bytes", C.c_int64), ("max_supernode_size", C.c_int32),
                ("chunk", C.c_int32), ("wave_block_threshold", C.c_int32), ("device", C.c_int32),
                ("queue_capacity", C.c_int64), ("output_capacity", C.c_int64),
                ("flags", C.c_int32), ("grid_blocks", C.c_int32), ("mode", C.c_int32),
                ("top_sources", C.c_int32),
                ("block_slots", C.c_int32), ("workgroup_threads", C.c_int32)]


class Result(C.Structure):
    _fields_ = [("n", C.c_int32), ("fill_count", C.c_uint64), ("te_alg", C.c_uint64),
                ("reach", C.c_uint64), ("edge_checks", C.c_uint64),
                ("row_fill", C.c_void_p), ("L_rowptr", C.c_void_p), ("L_col", C.c_void_p),
                ("U_rowptr", C.c_void_p), ("U_col", C.c_void_p), ("n_super", C.c_int32),
                ("super_start", C.c_void_p), ("supno", C.c_void_p), ("ms", C.c_double),
                ("traverse_ms", C.c_double), ("share_ms_min", C.c_double),
                ("share_ms_max", C.c_double), ("chunk", C.c_int32), ("mode", C.c_int32),
                ("grid_blocks", C.c_int32),
                ("queue_capacity", C.c_int64), ("levels", C.c_uint64), ("rescans", C.c_uint64),
                ("row_te", C.c_void_p), ("row_reach", C.c_void_p), ("passes", C.c_int32),
                ("graph_replay", C.c_int32)]


class Supernodal(C.Structure):
    _fields_ = [("n", C.c_int32), ("n_super", C.c_int32), ("xsup", C.c_void_p), ("supno", C.c_void_p),
                ("xsub", C.c_void_p), ("sub", C.c_void_p), ("xlcol", C.c_void_p), ("lrow", C.c_void_p),
                ("xucol", C.c_void_p), ("urow", C.c_void_p)]


class ServerOptions(C.Structure):
    _fields_ = [("device", C.c_int32), ("mem_budget_bytes", C.c_int64), ("max_inflight", C.c_int32),
                ("max_batch", C.c_int32), ("max_supernode_size", C.c_int32), ("flags", C.c_int32)]


class RequestInfo(C.Structure):
    _fields_ = [("queued_ms", C.c_double), ("latency_ms", C.c_double), ("batch_ms", C.c_double),
                ("batch_id", C.c_int32), ("batch_size", C.c_int32), ("spill_shares", C.c_int32)]


_lib = None


def lib():
    """Load libgsofa.so; raises OSError when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise OSError("libgsofa.so not found at %s; run __graft_entry__.build() "
                      "or make -C gsofa_amd/csrc" % LIB_PATH)
    L = C.CDLL(LIB_PATH)
    P = C.POINTER
    L.gsofa_strerror.argtypes = [C.c_int]
    L.gsofa_strerror.restype = C.c_char_p
    L.gsofa_csr_free.argtypes = [P(Csr)]
    L.gsofa_options_default.argtypes = [P(Options)]
    L.gsofa_result_free.argtypes = [P(Result)]
    L.gsofa_symbfact.argtypes = [P(Matrix), P(Options), P(Result)]
    L.gsofa_symbfact.restype = C.c_int
    L.gsofa_plan_create.argtypes = [P(Matrix), P(Options), P(C.c_void_p)]
    L.gsofa_plan_create.restype = C.c_int
    L.gsofa_plan_execute.argtypes = [C.c_void_p]
    L.gsofa_plan_execute.restype = C.c_int
    L.gsofa_plan_fetch.argtypes = [C.c_void_p, P(Result)]
    L.gsofa_plan_fetch.restype = C.c_int
    L.gsofa_plan_destroy.argtypes = [C.c_void_p]
    L.gsofa_batch_create.argtypes = [P(Matrix), C.c_int32, P(Options), P(C.c_void_p)]
    L.gsofa_batch_create.restype = C.c_int
    L.gsofa_plan_matrix_count.argtypes = [C.c_void_p]
    L.gsofa_plan_matrix_count.restype = C.c_int32
    L.gsofa_plan_fetch_matrix.argtypes = [C.c_void_p, C.c_int32, P(Result)]
    L.gsofa_plan_fetch_matrix.restype = C.c_int
    L.gsofa_plan_fetch_supernodal.argtypes = [C.c_void_p, P(Supernodal)]
    L.gsofa_plan_fetch_supernodal.restype = C.c_int
    L.gsofa_supernodal_free.argtypes = [P(Supernodal)]
    L.gsofa_plan_footprint.argtypes = [P(Matrix), P(Options)]
    L.gsofa_plan_footprint.restype = C.c_int64
    L.gsofa_order_nd.argtypes = [P(Matrix), C.c_int32, C.c_void_p]
    L.gsofa_order_nd.restype = C.c_int
    L.gsofa_permute.argtypes = [P(Matrix), C.c_void_p, P(Csr)]
    L.gsofa_permute.restype = C.c_int
    L.gsofa_fill_bound.argtypes = [P(Matrix), C.c_void_p, P(C.c_uint64), P(C.c_uint64), P(C.c_uint64)]
    L.gsofa_fill_bound.restype = C.c_int
    L.gsofa_counts_noalloc.argtypes = [C.c_int32, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                       C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p]
    L.gsofa_counts_noalloc.restype = C.c_int
    L.gsofa_write_binary.argtypes = [P(Matrix), C.c_char_p]
    L.gsofa_write_binary.restype = C.c_int
    L.gsofa_read_binary.argtypes = [C.c_char_p, P(Csr)]
    L.gsofa_read_binary.restype = C.c_int
    L.gsofa_map_binary.argtypes = [C.c_char_p, P(C.c_void_p), P(Matrix)]
    L.gsofa_map_binary.restype = C.c_int
    L.gsofa_unmap_binary.argtypes = [C.c_void_p]
    L.gsofa_read_ascii.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, P(Csr)]
    L.gsofa_read_ascii.restype = C.c_int
    L.gsofa_write_ascii.argtypes = [P(Matrix), C.c_char_p, C.c_char_p, C.c_char_p]
    L.gsofa_write_ascii.restype = C.c_int
    L.gsofa_gen_grid.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_double,
                                 C.c_int32, C.c_uint64, P(Csr)]
    L.gsofa_gen_grid.restype = C.c_int
    L.gsofa_gen_band.argtypes = [C.c_int32, C.c_int32, C.c_double, C.c_uint64, P(Csr)]
    L.gsofa_gen_band.restype = C.c_int
    L.gsofa_gen_side_loguniform.argtypes = [C.c_uint64, C.c_uint64, C.c_int32, C.c_int32]
    L.gsofa_gen_side_loguniform.restype = C.c_int32
    L.gsofa_gen_uniform_int.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, C.c_int32, C.c_int32]
    L.gsofa_gen_uniform_int.restype = C.c_int32
    L.gsofa_gen_unit.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64]
    L.gsofa_gen_unit.restype = C.c_double
    L.gsofa_server_options_default.argtypes = [P(ServerOptions)]
    L.gsofa_server_create.argtypes = [P(ServerOptions), P(C.c_void_p)]
    L.gsofa_server_create.restype = C.c_int
    L.gsofa_server_submit.argtypes = [C.c_void_p, P(Matrix), P(C.c_int64)]
    L.gsofa_server_submit.restype = C.c_int
    L.gsofa_server_poll.argtypes = [C.c_void_p, C.c_int64, P(Result), P(RequestInfo)]
    L.gsofa_server_poll.restype = C.c_int
    L.gsofa_server_wait.argtypes = [C.c_void_p, C.c_int64, P(Result), P(RequestInfo)]
    L.gsofa_server_wait.restype = C.c_int
    L.gsofa_server_destroy.argtypes = [C.c_void_p]
    _lib = L
    return L
