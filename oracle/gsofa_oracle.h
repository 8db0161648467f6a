/*
 * gsofa_oracle.h -- CPU oracle for the gSoFa fill-path symbolic factorization.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (gsofa_amd/, include/) may
 * link, load or call this library.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg use it, and only as the checker / CPU baseline.
 *
 * The oracle is a sequential restatement of the reference's per-source max-id
 * label-correcting traversal (OpenMP over independent sources):
 *   seeding        with_space_optimization/symbfact.cuh:639-721
 *   relaxation     with_space_optimization/symbfact.cuh:883-1082 (warp-centric)
 *                  and :723-882 (thread-centric; identical edge arithmetic)
 *   offset trick   with_space_optimization/symbfact.cuh:1928-1932, 2388-2407
 *   source map     with_space_optimization/symbfact.cuh:1084-1099
 *
 * Independent pins (no shared code with the traversal):
 *   orc_brute_lu              boolean Gaussian elimination without pivoting
 *   orc_symbolic_cholesky     elimination-tree column merge (symmetric inputs)
 *   orc_supernodes            the written supernode definition, restated; pinned
 *                             by tests/test_supernode_pins.py (elimination tree
 *                             + column counts on symmetric inputs, the T2 set
 *                             form over orc_brute_lu on unsymmetric ones)
 */
#ifndef GSOFA_ORACLE_H
#define GSOFA_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Row i's structure is csr[colbeg[i] : colend[i]) (reference layout, rows may be
 * unsorted and colbeg non-monotone). */
typedef struct {
    int32_t n;
    int64_t nnz;
    const int32_t* colbeg;
    const int32_t* colend;
    const int32_t* csr;
} orc_matrix;

typedef struct {
    int32_t n;
    uint64_t fill_count;     /* sum over processed rows of |fills_s|            */
    uint64_t te_alg;         /* sum_s [deg(s) + sum_{v in R_<(s)} deg(v)]        */
    uint64_t reach;          /* sum_s |R_<(s)|                                   */
    uint64_t edge_checks;    /* edge checks actually performed (FIFO order)      */
    /* per row (n entries); rows outside the processed share are empty / -1 */
    int64_t* row_fill;       /* fills in row s, -1 if row not processed          */
    int64_t* row_te;
    int64_t* row_reach;
    /* sorted strictly-lower (L) and strictly-upper (U) structure of L+U, by row */
    int64_t* L_rowptr;       /* n+1 */
    int32_t* L_col;
    int64_t* U_rowptr;       /* n+1 */
    int32_t* U_col;
} orc_result;

/* Process the rows s with s % n_virtual_gpu == virtual_gpu_id (all rows when
 * virtual_gpu_id < 0).  nthreads <= 0 means the OpenMP default.
 * Returns 0 on success, negative on bad input. */
int orc_symbfact(const orc_matrix* A, int n_virtual_gpu, int virtual_gpu_id,
                 int nthreads, orc_result* out);
void orc_result_free(orc_result* r);

/* Same traversal restricted to rows [row_begin, row_end) with no structure
 * output: only counters (fill_count, te_alg, reach).  Used as the bounded CPU
 * baseline sample. Returns wall seconds of the traversal, or < 0 on error. */
double orc_count_rows(const orc_matrix* A, int32_t row_begin, int32_t row_end,
                      int nthreads, uint64_t* fill_count, uint64_t* te_alg,
                      uint64_t* reach);

/* Supernode partition from sorted L/U (definition: row i joins row i-1's
 * supernode iff (i-1) in L_i, |U_i| == |U_{i-1}| - 1 and the supernode has
 * fewer than max_size rows; max_size <= 0 means unbounded).
 * super_start has room for n+1 entries, supno for n.  Returns n_super. */
int32_t orc_supernodes(int32_t n, const int64_t* L_rowptr, const int32_t* L_col,
                       const int64_t* U_rowptr, int32_t max_size,
                       int32_t* super_start, int32_t* supno);

/* Boolean LU without pivoting on a dense n x n pattern (row-major bytes,
 * a[i*n+j] != 0 iff edge i->j).  Overwrites a with struct(L+U). n <= 8192. */
int orc_brute_lu(int32_t n, uint8_t* a);

/* Symbolic Cholesky for a structurally symmetric pattern: returns the strictly
 * lower column structure of L (column j: rows i > j), sorted, library-owned.
 * Free with orc_free. Returns nnz of strict L or < 0 on error. */
int64_t orc_symbolic_cholesky(const orc_matrix* A, int64_t** colptr,
                              int32_t** rowind);
void orc_free(void* p);

#ifdef __cplusplus
}
#endif
#endif
