/*
 * gsofa.h -- C ABI of the MI355X fill-path symbolic factorization engine.
 *
 * Drop-in boundary for the reference's single entry point
 *     void symbfact_min_id(int argc, char** argv, int myrank,
 *                          ull_t& fill_count, double& time)
 * (with_space_optimization/symbfact.cuh:1553, called from
 *  with_space_optimization/main.cu:31), which reads three ASCII files
 * (with_space_optimization/symbfact.cuh:1591-1640), runs the multi-source
 * fill-path traversal for every virtual GPU share
 * (with_space_optimization/symbfact.cuh:1995-2462) and prints
 * "Total Number of fill-ins detected: N" (with_space_optimization/symbfact.cuh:2526).
 *
 * Here the same work is split into plain C entry points with status codes
 * (never exit()), borrowed read-only inputs and library-owned outputs.
 * Threading: one handle (plan) is used by one host thread at a time.
 */
#ifndef GSOFA_H
#define GSOFA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (the reference prints and calls exit(); barrier.cuh:11-20,
 *      symbfact.cuh:1559,1592-1597) ---------------------------------------- */
#define GSOFA_OK 0
#define GSOFA_ERR_INVALID -1      /* bad argument or malformed matrix          */
#define GSOFA_ERR_ALLOC -2        /* host or device allocation failed          */
#define GSOFA_ERR_HIP -3          /* a HIP runtime call failed                 */
#define GSOFA_ERR_TOO_LARGE -4    /* exceeds 32-bit index limits               */
#define GSOFA_ERR_IO -5           /* file could not be read or written         */
#define GSOFA_ERR_NO_DEVICE -6    /* no usable GPU                             */
#define GSOFA_ERR_BUDGET -7       /* does not fit the device-memory budget     */
#define GSOFA_ERR_DEVICE_TIMEOUT -8 /* a device-side wait exceeded its bound   */
#define GSOFA_ERR_STATE -9        /* call not valid in the object's state      */

const char* gsofa_strerror(int status);

/* ---- input ------------------------------------------------------------------
 * Borrowed, read-only, the reference's layout: row i's structure (an edge
 * i -> j for every j listed) is csr[colbeg[i] .. colend[i]).  Rows may be
 * unsorted and colbeg need not be monotone (with_space_optimization/dataset/).
 * Entries must be distinct within a row; the diagonal is expected present. */
typedef struct gsofa_matrix {
    int32_t n;
    int64_t nnz; /* length of csr; must be < 2^31 (block mode: < 2^30 after packing) */
    const int32_t* colbeg;
    const int32_t* colend;
    const int32_t* csr;
} gsofa_matrix;

/* Library-owned matrix (generators, file reader); free with gsofa_csr_free. */
typedef struct gsofa_csr {
    int32_t n;
    int64_t nnz;
    int32_t* colbeg;
    int32_t* colend;
    int32_t* csr;
} gsofa_csr;
void gsofa_csr_free(gsofa_csr* m);

/* ---- options ---------------------------------------------------------------
 * n_virtual_gpu / mem_budget_bytes are the reference CLI's N_GPU and MEM_GB
 * (with_space_optimization/symbfact.cuh:1565,1571); max_supernode_size is the
 * supernode variant's extra knob (with_space_optimization/with_supernodes/
 * symbfact.cuh:1577).  The remaining fields replace compile-time #defines. */
#define GSOFA_FLAG_COUNTS_ONLY 0x1  /* fill counts only: no L/U or supernodes  */
#define GSOFA_FLAG_SPILL 0x2        /* single matrix: after each of the         */
                                    /* n_virtual_gpu shares, move its L/U rows  */
                                    /* to host memory, so device output holds   */
                                    /* one share (supernodes stay on device)    */
#define GSOFA_FLAG_GRAPH 0x4        /* plans: the first execute runs and then   */
                                    /* captures its device work (traversal,     */
                                    /* structure, supernodes) in a HIP graph;   */
                                    /* later executes replay it, one launch and */
                                    /* one host wait each.  Block and reach     */
                                    /* modes with output sized from the row     */
                                    /* bounds; otherwise it is ignored.  No     */
                                    /* counterpart in the reference, whose host */
                                    /* relaunches every level (with_space_      */
                                    /* optimization/symbfact.cuh:1247-1311)     */

typedef struct gsofa_options {
    int32_t n_virtual_gpu;        /* sources split cyclically into this many shares (>=1) */
    int32_t virtual_gpu_id;       /* share to run, or -1 for every share in turn          */
    int64_t mem_budget_bytes;     /* labels+markers+queues pool; 0 = automatic             */
    int32_t max_supernode_size;   /* 0 = unbounded                                          */
    int32_t chunk;                /* concurrent sources per group; 0 = automatic            */
    int32_t wave_block_threshold; /* row length from which a whole workgroup expands a row  */
    int32_t device;               /* HIP device ordinal                                     */
    int64_t queue_capacity;       /* frontier pairs per queue; 0 = automatic                */
    int64_t output_capacity;      /* fill entries reserved; 0 = automatic (grown on demand) */
    int32_t flags;                /* GSOFA_FLAG_*                                            */
    int32_t grid_blocks;          /* persistent grid size; 0 = occupancy-bounded maximum    */
    int32_t mode;                 /* GSOFA_MODE_*                                           */
    int32_t top_sources;          /* block mode, one matrix, one share: the top sources     */
                                  /* (rounded up to whole groups) first run as grid-mode    */
                                  /* groups (GSOFA_MODE_HYBRID); 0 = none                   */
    int32_t block_slots;          /* block mode: sources per workgroup, 32, 16 or 8; 0 = 32 */
                                  /* (fewer slots, larger per-round target table).  Reach   */
                                  /* mode: 32, 64, 128 or 256 (bits of a visited word);      */
                                  /* 0 = 64, or 32 when fewer chunks than workgroups        */
    int32_t workgroup_threads;    /* reach mode: threads per workgroup, 1024 (one per CU,   */
                                  /* the largest target table) or 512 (two per CU, half    */
                                  /* the table each); 0 = 1024                              */
} gsofa_options;

/* Execution modes of the traversal (identical results):
 * BLOCK: every workgroup of a persistent grid owns a few source slots and runs
 *        their frontier levels alone (LDS queues, workgroup barriers only);
 *        chunk = total concurrent sources.
 * GRID:  one group of chunk sources at a time; all workgroups expand each
 *        frontier level together, separated by a grid-wide generation barrier
 *        (the reference's persistent kernel, without_space_optimization/
 *        symbfact.cuh:706-962).
 * HYBRID (reported, chosen with top_sources): the deepest sources, whose
 *        traversals are long chains of wide levels, run as GRID groups that
 *        spread every level over the whole device; the rest run in BLOCK mode.
 * REACH: two label-free reachability passes with a visited bit per source
 *        slot: over A (U rows, TE_alg, |R|) and over A^T (L rows, since
 *        struct(L) of A is struct(U) of A^T transposed).  Every vertex is
 *        expanded at most once per source.  Whole-matrix runs only (one share
 *        or every share in one execute, in core); per-row TE and |R| are not
 *        reported (row_te = row_reach = NULL), their totals are.
 * AUTO:  the tuned table's mode for the workload if it has one; else REACH
 *        when n_virtual_gpu is 1, nothing spills and block_slots and
 *        top_sources are 0; else BLOCK (GRID above 2^30 packed entries). */
#define GSOFA_MODE_AUTO 0
#define GSOFA_MODE_BLOCK 1
#define GSOFA_MODE_GRID 2
#define GSOFA_MODE_HYBRID 3
#define GSOFA_MODE_REACH 4

void gsofa_options_default(gsofa_options* opt);

/* ---- result ----------------------------------------------------------------
 * Library-owned; release with gsofa_result_free.  L is the strictly lower and
 * U the strictly upper part of struct(L+U), sorted by column within each row.
 * Rows outside the processed share are empty and have row_fill = -1.
 * fill_count = |struct(L+U)| - |struct(A)| over processed rows, diagonal
 * excluded (the reference's "fill-ins"). */
typedef struct gsofa_result {
    int32_t n;
    uint64_t fill_count;
    uint64_t te_alg;      /* sum_s [deg(s) + sum_{v in R_<(s)} deg(v)] over processed rows */
    uint64_t reach;       /* sum_s |R_<(s)|                                                */
    uint64_t edge_checks; /* edges actually examined (>= te_alg; re-expansions included)  */
    int64_t* row_fill;    /* n */
    int64_t* L_rowptr;    /* n+1 */
    int32_t* L_col;
    int64_t* U_rowptr;    /* n+1 */
    int32_t* U_col;
    int32_t n_super;      /* -1 when not computed (single share or counts only)      */
    int32_t* super_start; /* n_super+1 entries, last = n                              */
    int32_t* supno;       /* n entries: supernode id of each row                       */
    double ms;            /* device time of traversal + structure + supernodes         */
    double traverse_ms;   /* device time of the traversal kernel(s) alone              */
    double share_ms_min;  /* fastest / slowest share (reference log, symbfact.cuh:2530) */
    double share_ms_max;
    int32_t chunk;        /* concurrent sources actually used                          */
    int32_t mode;         /* the mode that ran: GSOFA_MODE_REACH, _BLOCK, _GRID, _HYBRID */
    int32_t grid_blocks;  /* persistent grid actually used                             */
    int64_t queue_capacity;
    uint64_t levels;      /* frontier levels summed over groups                        */
    uint64_t rescans;     /* queue-overflow drain-and-resume passes                    */
    int64_t* row_te;      /* n: TE of each processed row (block mode only; else NULL)  */
    int64_t* row_reach;   /* n: |R_<(s)| of each processed row (block mode only)       */
    int32_t passes;       /* traversal passes of the execute: 1, or more when an       */
                          /* explicit output_capacity overflowed and the run repeated  */
    int32_t graph_replay; /* 1: the execute replayed the plan's captured graph          */
                          /* (GSOFA_FLAG_GRAPH); ms then times the whole replay and     */
                          /* traverse_ms is 0                                           */
} gsofa_result;

void gsofa_result_free(gsofa_result* r);

/* Upper bounds of the output, on the host, before any traversal (no
 * counterpart in the reference, which sizes its pool from MEM_GB and spills
 * when it overflows, with_space_optimization/symbfact.cuh:2079,2170-2380).
 * A fill path of A is a path of the graph of A + A^T, so struct(L+U) lies in
 * the pattern of the Cholesky factor C of A + A^T and its transpose: row i
 * has at most rc(i) L entries and cc(i) U entries (strict row / column counts
 * of C, from its elimination tree in near-linear time).
 *   row_fill_bound  n entries (caller-owned, may be NULL): >= row_fill[i]
 *   fill_bound      >= fill_count;  L_bound >= nnz(L);  U_bound >= nnz(U)
 * All equal the exact values when A is structurally symmetric.  Plans size
 * their fill stream and output from these, so no traversal is repeated. */
int gsofa_fill_bound(const gsofa_matrix* A, int64_t* row_fill_bound, uint64_t* fill_bound,
                     uint64_t* L_bound, uint64_t* U_bound);

/* One-shot: upload, run, fetch, release.  Replaces symbfact_min_id
 * (with_space_optimization/symbfact.cuh:1553). */
int gsofa_symbfact(const gsofa_matrix* A, const gsofa_options* opt, gsofa_result* out);

/* Fill counts of one matrix with no allocation at all.  No counterpart in the
 * reference; it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * counts-only block-mode execute, with every work array taken from a static
 * device scratch of GSOFA_NOALLOC_SCRATCH_BYTES that each GPU the library
 * runs on holds until the process ends.  All pointers are device memory:
 *   d_rowptr  n+1 int32, d_col nnz int32: the rows of A packed in order, no
 *             duplicate entries (any order within a row);
 *   d_row_fill, d_row_te, d_row_reach  n int64 each: per-row fills, TE_alg
 *             and |R_<(s)|, as gsofa_result's;
 *   d_totals  4 uint64: fill_count, te_alg, reach, edge checks.
 * The work is queued on `stream` (a hipStream_t; NULL = the default stream)
 * and the call returns; wait on the stream before reading the outputs.
 * GSOFA_ERR_BUDGET when one workgroup's arrays (72 bytes per vertex) and the
 * work table do not fit the scratch (n above about 900,000).  Calls that
 * overlap on one device share the scratch and race: serialise them. */
#define GSOFA_NOALLOC_SCRATCH_BYTES (64u << 20)
int gsofa_counts_noalloc(int32_t n, int64_t nnz, const int32_t* d_rowptr, const int32_t* d_col,
                         int64_t* d_row_fill, int64_t* d_row_te, int64_t* d_row_reach, uint64_t* d_totals,
                         int32_t device, void* stream);

/* ---- device-resident plan ---------------------------------------------------
 * create: validate, upload the CSR and size the workspace (no traversal).
 * execute: traversal + L/U structure + supernodes on the device; the inputs
 *          stay resident and results stay on the device.
 * fetch:   copy the last execute's results to a gsofa_result.  */
typedef struct gsofa_plan gsofa_plan;
int gsofa_plan_create(const gsofa_matrix* A, const gsofa_options* opt, gsofa_plan** plan);
int gsofa_plan_execute(gsofa_plan* plan);
int gsofa_plan_fetch(gsofa_plan* plan, gsofa_result* out);
void gsofa_plan_destroy(gsofa_plan* plan);
/* Device bytes a plan for A would hold (CSR + workspace + output reserve). */
int64_t gsofa_plan_footprint(const gsofa_matrix* A, const gsofa_options* opt);

/* ---- batches ---------------------------------------------------------------
 * Many independent matrices in one plan: they are laid out block-diagonally
 * (matrix b's vertices follow matrix b-1's), so one persistent launch works
 * through a table of (matrix, source-chunk) items.  No traversal can leave its
 * matrix, so every matrix's result equals its stand-alone result.  No
 * counterpart in the reference, whose driver handles one matrix per run
 * (with_space_optimization/symbfact.cuh:1553). */
int gsofa_batch_create(const gsofa_matrix* mats, int32_t count, const gsofa_options* opt,
                       gsofa_plan** plan);
int32_t gsofa_plan_matrix_count(const gsofa_plan* plan);
/* Result of matrix `index` of the last execute, in that matrix's own numbering. */
int gsofa_plan_fetch_matrix(gsofa_plan* plan, int32_t index, gsofa_result* out);

/* ---- supernodal hand-off ------------------------------------------------------
 * No counterpart in the reference, which stops at the fill count; this is the
 * input a supernodal numeric factorization takes next (the reference's inputs
 * come from SuperLU_DIST, README.md "Dataset generation").  The arrays follow
 * SuperLU_DIST's symbolic output (Glu_persist_t xsup/supno, Glu_freeable_t
 * compressed subscripts and column-oriented L and U), for the row supernodes
 * of gsofa_result:
 *   xsup/supno  supernode starts (n_super+1, last = n) and per-row ids;
 *   xsub/sub    supernode s (rows f = xsup[s] .. l = xsup[s+1]-1): f, then the
 *               sorted U columns of row f.  Every row i of s has U_i = those
 *               columns above i, so the list is {f..l} followed by the block
 *               row's columns beyond l;
 *   xlcol/lrow  L by columns: rows i > j with j in L_i, sorted per column j;
 *   xucol/urow  U by columns: rows i < j with j in U_i, sorted per column j.
 * Computed on the device from the last execute of a single-matrix plan that
 * built supernodes (every share, not counts-only); otherwise
 * GSOFA_ERR_INVALID.  Release with gsofa_supernodal_free. */
typedef struct gsofa_supernodal {
    int32_t n;
    int32_t n_super;
    int32_t* xsup;    /* n_super+1 */
    int32_t* supno;   /* n */
    int64_t* xsub;    /* n_super+1 */
    int32_t* sub;
    int64_t* xlcol;   /* n+1 */
    int32_t* lrow;
    int64_t* xucol;   /* n+1 */
    int32_t* urow;
} gsofa_supernodal;
int gsofa_plan_fetch_supernodal(gsofa_plan* plan, gsofa_supernodal* out);
void gsofa_supernodal_free(gsofa_supernodal* s);

/* ---- request server (continuous batching) -----------------------------------
 * No counterpart in the reference (one matrix per process run).  Requests are
 * submitted one at a time from any host thread; the server copies the matrix
 * and returns a ticket.  max_inflight worker threads each own a HIP stream:
 * whenever a worker is free it admits the waiting requests, oldest first, up
 * to max_batch and while the device bytes of all in-flight batches stay within
 * mem_budget_bytes, runs them as one batch plan (one persistent launch) and
 * publishes every request's result.  Batches of different workers overlap on
 * the device.  Each result equals the request's stand-alone result.
 * gsofa_server_submit checks the request as gsofa_plan_create would (index
 * ranges, size limits) and returns any error to its caller alone.  A request
 * whose own footprint exceeds budget / max_inflight runs alone once the
 * device drains, with the whole budget, its output in core or in spilled
 * shares (gsofa_request_info.spill_shares); only one with room for no
 * traversal workgroup fails, with GSOFA_ERR_BUDGET.  GSOFA_FLAG_SPILL in the
 * server's flags is refused (GSOFA_ERR_INVALID): the server chooses it. */
#define GSOFA_PENDING 1 /* gsofa_server_poll: the request is not finished */

typedef struct gsofa_server gsofa_server;

typedef struct gsofa_server_options {
    int32_t device;
    int64_t mem_budget_bytes;   /* device bytes for all in-flight batches; 0 = 80% of free */
    int32_t max_inflight;       /* batches on the device at once (worker threads) */
    int32_t max_batch;          /* requests per batch */
    int32_t max_supernode_size; /* as gsofa_options */
    int32_t flags;              /* GSOFA_FLAG_* */
} gsofa_server_options;

typedef struct gsofa_request_info {
    double queued_ms;   /* submit -> its batch starts */
    double latency_ms;  /* submit -> result published */
    double batch_ms;    /* device time of its batch (traversal + structure + supernodes) */
    int32_t batch_id;   /* 1, 2, ... in admission order */
    int32_t batch_size; /* requests in its batch */
    int32_t spill_shares; /* 1: in core; N > 1: ran alone in N spilled shares */
} gsofa_request_info;

void gsofa_server_options_default(gsofa_server_options* opt);
int gsofa_server_create(const gsofa_server_options* opt, gsofa_server** server);
/* Copies A; *ticket identifies the request until its result is taken. */
int gsofa_server_submit(gsofa_server* server, const gsofa_matrix* A, int64_t* ticket);
/* GSOFA_PENDING while running; otherwise the request's final status, with its
 * result moved to *out (on GSOFA_OK) and the ticket released. */
int gsofa_server_poll(gsofa_server* server, int64_t ticket, gsofa_result* out,
                      gsofa_request_info* info);
/* Blocks until the request is final, then behaves as gsofa_server_poll. */
int gsofa_server_wait(gsofa_server* server, int64_t ticket, gsofa_result* out,
                      gsofa_request_info* info);
/* Finishes every admitted and waiting request, then releases the server;
 * results not taken are freed. */
void gsofa_server_destroy(gsofa_server* server);

/* ---- inputs: reference file format and generators ---------------------------
 * ASCII triple (with_space_optimization/symbfact.cuh:1591-1640): the csr file
 * holds nnz, n and nnz indices; colbeg/colend hold n integers each. */
int gsofa_read_ascii(const char* colbeg_path, const char* colend_path, const char* csr_path,
                     gsofa_csr* out);
int gsofa_write_ascii(const gsofa_matrix* A, const char* colbeg_path, const char* colend_path,
                      const char* csr_path);

/* Binary triple: a 64-byte little-endian header ("GSOFACS1", version 1,
 * header size 64, n, nnz, byte offsets of colbeg, colend and csr), then the
 * three int32 arrays, each on a 64-byte boundary.  No counterpart in the
 * reference, which reads only the ASCII triple; this is its fast form.
 * write: to path.tmp, then renamed over path.
 * read:  a library-owned copy (free with gsofa_csr_free).
 * map:   a read-only mapping; `view` points into it and stays valid until
 *        gsofa_unmap_binary.  Malformed or truncated files give
 *        GSOFA_ERR_IO. */
typedef struct gsofa_mapped gsofa_mapped;
int gsofa_write_binary(const gsofa_matrix* A, const char* path);
int gsofa_read_binary(const char* path, gsofa_csr* out);
int gsofa_map_binary(const char* path, gsofa_mapped** handle, gsofa_matrix* view);
void gsofa_unmap_binary(gsofa_mapped* handle);

/* Fill-reducing ordering for inputs that come without one (the reference
 * takes SuperLU_DIST's ordering as its input; no counterpart there): nested
 * dissection on breadth-first level structures of the symmetrized graph of A
 * (George and Liu, SIAM J. Numer. Anal. 15, 1978).  Regions of at most `leaf`
 * vertices are numbered by minimum degree; separators come after both sides.
 * perm[k] = original index of the k-th vertex of the new order (n entries,
 * caller-owned).  gsofa_permute builds B = P A P^T for such a perm: row k of
 * B is row perm[k] of A, columns renumbered and sorted (free with
 * gsofa_csr_free); a perm that is not a permutation gives GSOFA_ERR_INVALID. */
int gsofa_order_nd(const gsofa_matrix* A, int32_t leaf, int32_t* perm);
int gsofa_permute(const gsofa_matrix* A, const int32_t* perm, gsofa_csr* out);

/* 5-point (nz == 1) or 7-point Laplacian pattern on an nx*ny*nz grid.  nd_leaf
 * > 0 orders it by geometric nested dissection with that leaf size (0 =
 * natural row-major).  With probability extra_prob a row gains one extra
 * one-directional entry to a uniformly chosen grid point at Manhattan distance
 * 2..extra_dist.  Rows are sorted. */
int gsofa_gen_grid(int32_t nx, int32_t ny, int32_t nz, int32_t nd_leaf, double extra_prob,
                   int32_t extra_dist, uint64_t seed, gsofa_csr* out);
/* Diagonal plus Poisson(lambda) off-diagonal columns per row drawn uniformly
 * from |i-j| <= band (j != i), duplicates removed.  Rows are sorted. */
int gsofa_gen_band(int32_t n, int32_t band, double lambda, uint64_t seed, gsofa_csr* out);
/* Deterministic scalar draws used by the batch / stream workloads. */
int32_t gsofa_gen_side_loguniform(uint64_t seed, uint64_t index, int32_t lo, int32_t hi);
int32_t gsofa_gen_uniform_int(uint64_t seed, uint64_t purpose, uint64_t index, int32_t lo,
                              int32_t hi);
double gsofa_gen_unit(uint64_t seed, uint64_t purpose, uint64_t index);

#ifdef __cplusplus
}
#endif
#endif /* GSOFA_H */
