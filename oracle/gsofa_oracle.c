/*
 * gsofa_oracle.c -- CPU oracle (test infrastructure only; see gsofa_oracle.h).
 *
 * Each function cites the reference lines it restates.  Reference paths are
 * relative to the reference repository root:
 *   WS = with_space_optimization/symbfact.cuh
 *   WO = without_space_optimization/symbfact.cuh
 */
#include "gsofa_oracle.h"

#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define ORC_MAXV 0xFFFFFFFFu /* reference MAX_VAL = UINT_MAX (WS:29) */

typedef struct {
    int32_t n;
    uint32_t* lab;      /* max-id labels, one per vertex (WO:1115 unpacked form) */
    uint32_t* mark;     /* fill markers stamped with s+1 (WS:699, WS:1006)        */
    uint32_t last_stamp;
    uint32_t off;       /* max_id_offset (WS:1928)                               */
    uint32_t used;      /* sources processed since the last re-init (WS:2387)    */
    uint32_t loops;     /* UINT_MAX / n  (WS:1932)                               */
    int32_t* q;         /* FIFO ring of frontier vertices                         */
    int64_t qcap, qhead, qlen;
    int32_t* fills;     /* fills of the current source                            */
    int64_t fcap, nf;
} orc_ws;

static int ws_init(orc_ws* w, int32_t n) {
    memset(w, 0, sizeof(*w));
    w->n = n;
    w->lab = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1));
    w->mark = (uint32_t*)calloc((size_t)(n > 0 ? n : 1), sizeof(uint32_t));
    w->qcap = 1024;
    w->q = (int32_t*)malloc(sizeof(int32_t) * (size_t)w->qcap);
    w->fcap = 1024;
    w->fills = (int32_t*)malloc(sizeof(int32_t) * (size_t)w->fcap);
    if (!w->lab || !w->mark || !w->q || !w->fills) return -1;
    for (int32_t i = 0; i < n; i++) w->lab[i] = ORC_MAXV;
    /* max_id_offset = MAX_VAL - vert_count; group_loops = MAX_VAL / vert_count
     * (WS:1928-1932) */
    w->off = ORC_MAXV - (uint32_t)n;
    w->loops = n > 0 ? ORC_MAXV / (uint32_t)n : 1;
    w->used = 0;
    return 0;
}

static void ws_free(orc_ws* w) {
    free(w->lab);
    free(w->mark);
    free(w->q);
    free(w->fills);
}

static inline void q_push(orc_ws* w, int32_t v) {
    if (w->qlen == w->qcap) {
        int64_t nc = w->qcap * 2;
        int32_t* nq = (int32_t*)malloc(sizeof(int32_t) * (size_t)nc);
        for (int64_t i = 0; i < w->qlen; i++) nq[i] = w->q[(w->qhead + i) % w->qcap];
        free(w->q);
        w->q = nq;
        w->qcap = nc;
        w->qhead = 0;
    }
    w->q[(w->qhead + w->qlen) % w->qcap] = v;
    w->qlen++;
}

static inline int32_t q_pop(orc_ws* w) {
    int32_t v = w->q[w->qhead];
    w->qhead = (w->qhead + 1) % w->qcap;
    w->qlen--;
    return v;
}

static inline void fill_push(orc_ws* w, int32_t m) {
    if (w->nf == w->fcap) {
        w->fcap *= 2;
        w->fills = (int32_t*)realloc(w->fills, sizeof(int32_t) * (size_t)w->fcap);
    }
    w->fills[w->nf++] = m;
}

typedef struct {
    uint64_t fills, te, reach, checks;
} src_stats;

/*
 * One source s: the reference's seeding followed by label-correcting
 * relaxation to a fixed point.  The reference runs the relaxation one frontier
 * level per kernel launch; the fixed point (and therefore the stamped set) does
 * not depend on the processing order, so a FIFO is used here.
 *
 *   seeding    WS:639-721  label(s) = offset; stamp every A-neighbour j; if j<s
 *                          label(j) = offset and enqueue j.
 *   relax      WS:883-1082 cost = max(f+offset, label(f)); for m in row f:
 *                          m > s  -> stamp m (U fill iff previously unstamped);
 *                          label(m) > cost -> label(m) = cost; if m < s enqueue;
 *                          if m+offset > cost stamp m (L fill iff unstamped).
 */
static src_stats traverse_source(const orc_matrix* A, orc_ws* w, int32_t s, int keep_fills) {
    src_stats st = {0, 0, 0, 0};
    const int32_t* cb = A->colbeg;
    const int32_t* ce = A->colend;
    const int32_t* csr = A->csr;
    const uint32_t n = (uint32_t)A->n;
    const uint32_t off = w->off;
    const uint32_t stale = off + n; /* labels >= stale belong to earlier sources */
    uint32_t stamp = (uint32_t)s + 1u;
    if (stamp <= w->last_stamp) { /* sources must increase per workspace */
        memset(w->mark, 0, sizeof(uint32_t) * n);
    }
    w->last_stamp = stamp;
    w->qhead = 0;
    w->qlen = 0;
    w->nf = 0;

    st.te = (uint64_t)(ce[s] - cb[s]);
    st.checks = st.te;
    w->lab[s] = off; /* WS:662 cost_array_d[offset+source] = max_id_offset */
    for (int64_t k = cb[s]; k < ce[s]; k++) {
        int32_t j = csr[k];
        w->mark[j] = stamp; /* WS:699 */
        if ((uint32_t)j < (uint32_t)s) {
            if (w->lab[j] >= stale) {
                st.reach++;
                st.te += (uint64_t)(ce[j] - cb[j]);
            }
            w->lab[j] = off; /* WS:707 */
            q_push(w, j);    /* WS:710-713 */
        }
    }
    while (w->qlen > 0) {
        int32_t f = q_pop(w);
        uint32_t lf = w->lab[f];
        uint32_t cost = ((uint32_t)f + off) > lf ? ((uint32_t)f + off) : lf; /* WS:924 */
        for (int64_t k = cb[f]; k < ce[f]; k++) {
            uint32_t m = (uint32_t)csr[k];
            st.checks++;
            if (m > (uint32_t)s) { /* WS:992-1009 */
                if (w->mark[m] < stamp) {
                    w->mark[m] = stamp;
                    st.fills++;
                    if (keep_fills) fill_push(w, (int32_t)m);
                }
                continue;
            }
            if (w->lab[m] > cost) { /* WS:1018-1021 */
                uint32_t old = w->lab[m];
                w->lab[m] = cost;
                if (old >= stale) {
                    st.reach++;
                    st.te += (uint64_t)(ce[m] - cb[m]);
                }
                if (m < (uint32_t)s) q_push(w, (int32_t)m); /* WS:1023-1052 */
                if (m + off > cost) {                        /* WS:1054-1063 */
                    if (w->mark[m] < stamp) {
                        w->mark[m] = stamp;
                        st.fills++;
                        if (keep_fills) fill_push(w, (int32_t)m);
                    }
                }
            }
        }
    }
    /* advance the no-reset offset (WS:2388-2407) */
    w->off -= n;
    w->used++;
    if (w->used >= w->loops) {
        w->off = ORC_MAXV - n;
        for (uint32_t i = 0; i < n; i++) w->lab[i] = ORC_MAXV;
        w->used = 0;
    }
    return st;
}

static int cmp_i32(const void* a, const void* b) {
    int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
    return (x > y) - (x < y);
}

static int check_matrix(const orc_matrix* A) {
    if (!A || A->n < 0 || A->nnz < 0) return -1;
    if (A->n > 0 && (!A->colbeg || !A->colend || (A->nnz > 0 && !A->csr))) return -1;
    for (int32_t i = 0; i < A->n; i++) {
        if (A->colbeg[i] < 0 || A->colend[i] < A->colbeg[i] || A->colend[i] > A->nnz) return -2;
        for (int64_t k = A->colbeg[i]; k < A->colend[i]; k++)
            if (A->csr[k] < 0 || A->csr[k] >= A->n) return -3;
    }
    return 0;
}

/* Source map: the rows of virtual GPU g are g, g+N, g+2N, ... -- the union over
 * groups of source_d[k] = gpu_id + k*N_gpu + begin_group (WS:1084-1099). */
int orc_symbfact(const orc_matrix* A, int n_virtual_gpu, int virtual_gpu_id, int nthreads,
                 orc_result* out) {
    int rc = check_matrix(A);
    if (rc) return rc;
    if (n_virtual_gpu < 1) n_virtual_gpu = 1;
    const int32_t n = A->n;
    memset(out, 0, sizeof(*out));
    out->n = n;
    out->row_fill = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n + 1));
    out->row_te = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n + 1));
    out->row_reach = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n + 1));
    int32_t** rows = (int32_t**)calloc((size_t)(n + 1), sizeof(int32_t*));
    int64_t* rlen = (int64_t*)calloc((size_t)(n + 1), sizeof(int64_t));
    for (int32_t i = 0; i < n; i++) out->row_fill[i] = out->row_te[i] = out->row_reach[i] = -1;

    int64_t count;
    int32_t first;
    int32_t stride = n_virtual_gpu;
    if (virtual_gpu_id < 0) {
        first = 0;
        stride = 1;
        count = n;
    } else {
        first = virtual_gpu_id;
        count = first < n ? ((int64_t)n - first + stride - 1) / stride : 0;
    }
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    uint64_t tf = 0, tt = 0, tr = 0, tc = 0;
    int bad = 0;
#pragma omp parallel num_threads(nthreads) reduction(+ : tf, tt, tr, tc)
    {
        orc_ws w;
        if (ws_init(&w, n)) {
#pragma omp atomic write
            bad = 1;
        }
#pragma omp for schedule(dynamic, 16)
        for (int64_t k = 0; k < count; k++) {
            if (bad) continue;
            int32_t s = (int32_t)(first + k * stride);
            src_stats st = traverse_source(A, &w, s, 1);
            tf += st.fills;
            tt += st.te;
            tr += st.reach;
            tc += st.checks;
            out->row_fill[s] = (int64_t)st.fills;
            out->row_te[s] = (int64_t)st.te;
            out->row_reach[s] = (int64_t)st.reach;
            int64_t deg = A->colend[s] - A->colbeg[s];
            int32_t* row = (int32_t*)malloc(sizeof(int32_t) * (size_t)(deg + w.nf + 1));
            int64_t len = 0;
            for (int64_t e = A->colbeg[s]; e < A->colend[s]; e++)
                if (A->csr[e] != s) row[len++] = A->csr[e];
            for (int64_t e = 0; e < w.nf; e++) row[len++] = w.fills[e];
            qsort(row, (size_t)len, sizeof(int32_t), cmp_i32);
            rows[s] = row;
            rlen[s] = len;
        }
        ws_free(&w);
    }
    if (bad) {
        for (int32_t i = 0; i < n; i++) free(rows[i]);
        free(rows);
        free(rlen);
        orc_result_free(out);
        return -4;
    }
    out->fill_count = tf;
    out->te_alg = tt;
    out->reach = tr;
    out->edge_checks = tc;

    /* assemble sorted strictly-lower / strictly-upper CSR */
    out->L_rowptr = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n + 1));
    out->U_rowptr = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n + 1));
    int64_t nl = 0, nu = 0;
    out->L_rowptr[0] = out->U_rowptr[0] = 0;
    for (int32_t s = 0; s < n; s++) {
        for (int64_t e = 0; e < rlen[s]; e++) {
            if (rows[s][e] < s) nl++;
            else nu++;
        }
        out->L_rowptr[s + 1] = nl;
        out->U_rowptr[s + 1] = nu;
    }
    out->L_col = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nl + 1));
    out->U_col = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nu + 1));
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads)
    for (int32_t s = 0; s < n; s++) {
        int64_t pl = out->L_rowptr[s], pu = out->U_rowptr[s];
        for (int64_t e = 0; e < rlen[s]; e++) {
            if (rows[s][e] < s) out->L_col[pl++] = rows[s][e];
            else out->U_col[pu++] = rows[s][e];
        }
        free(rows[s]);
    }
    free(rows);
    free(rlen);
    return 0;
}

void orc_result_free(orc_result* r) {
    if (!r) return;
    free(r->row_fill);
    free(r->row_te);
    free(r->row_reach);
    free(r->L_rowptr);
    free(r->L_col);
    free(r->U_rowptr);
    free(r->U_col);
    memset(r, 0, sizeof(*r));
}

double orc_count_rows(const orc_matrix* A, int32_t row_begin, int32_t row_end, int nthreads,
                      uint64_t* fill_count, uint64_t* te_alg, uint64_t* reach) {
    if (check_matrix(A)) return -1.0;
    if (row_begin < 0) row_begin = 0;
    if (row_end > A->n) row_end = A->n;
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    uint64_t tf = 0, tt = 0, tr = 0;
    int bad = 0;
    double t0 = omp_get_wtime();
#pragma omp parallel num_threads(nthreads) reduction(+ : tf, tt, tr)
    {
        orc_ws w;
        if (ws_init(&w, A->n)) {
#pragma omp atomic write
            bad = 1;
        }
#pragma omp for schedule(dynamic, 4)
        for (int32_t s = row_begin; s < row_end; s++) {
            if (bad) continue;
            src_stats st = traverse_source(A, &w, s, 0);
            tf += st.fills;
            tt += st.te;
            tr += st.reach;
        }
        ws_free(&w);
    }
    double t1 = omp_get_wtime();
    if (bad) return -4.0;
    if (fill_count) *fill_count = tf;
    if (te_alg) *te_alg = tt;
    if (reach) *reach = tr;
    return t1 - t0;
}

/* The supernode rule as written for the build (the reference only has a stub:
 * SN/symbfact.cuh:1087-1101 names pass_through and nz_row_U per row, and
 * SN/symbfact.cuh:1577 parses max_supernode_size). */
int32_t orc_supernodes(int32_t n, const int64_t* L_rowptr, const int32_t* L_col,
                       const int64_t* U_rowptr, int32_t max_size, int32_t* super_start,
                       int32_t* supno) {
    if (n <= 0) {
        if (super_start) super_start[0] = 0;
        return 0;
    }
    int32_t ns = 1, size = 1;
    super_start[0] = 0;
    supno[0] = 0;
    for (int32_t i = 1; i < n; i++) {
        int64_t lb = L_rowptr[i], le = L_rowptr[i + 1];
        int pass = (le > lb) && (L_col[le - 1] == i - 1);
        int64_t ui = U_rowptr[i + 1] - U_rowptr[i];
        int64_t up = U_rowptr[i] - U_rowptr[i - 1];
        int joins = pass && (ui == up - 1) && (max_size <= 0 || size < max_size);
        if (joins) {
            size++;
        } else {
            super_start[ns++] = i;
            size = 1;
        }
        supno[i] = ns - 1;
    }
    super_start[ns] = n;
    return ns;
}

/* Textbook structural Gaussian elimination without pivoting:
 * for k = 0..n-1, for i > k with a(i,k): a(i,j) |= a(k,j) for all j > k. */
int orc_brute_lu(int32_t n, uint8_t* a) {
    if (n < 0 || n > 8192) return -1;
    if (n == 0) return 0;
    const int32_t W = (n + 63) / 64;
    uint64_t* b = (uint64_t*)calloc((size_t)n * (size_t)W, sizeof(uint64_t));
    if (!b) return -2;
    for (int32_t i = 0; i < n; i++)
        for (int32_t j = 0; j < n; j++)
            if (a[(size_t)i * n + j]) b[(size_t)i * W + j / 64] |= 1ull << (j % 64);
    for (int32_t k = 0; k < n; k++) {
        const uint64_t* rk = b + (size_t)k * W;
        int32_t w0 = (k + 1) / 64;
        uint64_t m0 = ((k + 1) % 64) ? ~((1ull << ((k + 1) % 64)) - 1) : ~0ull;
        for (int32_t i = k + 1; i < n; i++) {
            uint64_t* ri = b + (size_t)i * W;
            if (!((ri[k / 64] >> (k % 64)) & 1ull)) continue;
            if (w0 < W) ri[w0] |= rk[w0] & m0;
            for (int32_t w = w0 + 1; w < W; w++) ri[w] |= rk[w];
        }
    }
    for (int32_t i = 0; i < n; i++)
        for (int32_t j = 0; j < n; j++)
            a[(size_t)i * n + j] = (uint8_t)((b[(size_t)i * W + j / 64] >> (j % 64)) & 1ull);
    free(b);
    return 0;
}

/* Symbolic Cholesky by elimination-tree column merging (George & Liu):
 *   struct(L_{*j}) = {i > j : a(i,j)} U  union_{parent(c)=j} struct(L_{*c}) \ {j},
 *   parent(j) = min struct(L_{*j}).
 * Requires a structurally symmetric pattern; column j's lower part is then row
 * j's entries greater than j. */
int64_t orc_symbolic_cholesky(const orc_matrix* A, int64_t** colptr_out, int32_t** rowind_out) {
    if (check_matrix(A)) return -1;
    const int32_t n = A->n;
    int64_t* colptr = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n + 1));
    int64_t cap = A->nnz + 16, len = 0;
    int32_t* rowind = (int32_t*)malloc(sizeof(int32_t) * (size_t)cap);
    int32_t* head = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n + 1)); /* child lists */
    int32_t* next = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n + 1));
    int32_t* seen = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n + 1));
    if (!colptr || !rowind || !head || !next || !seen) return -2;
    for (int32_t i = 0; i < n; i++) {
        head[i] = -1;
        seen[i] = -1;
    }
    colptr[0] = 0;
    for (int32_t j = 0; j < n; j++) {
        int64_t start = len;
#define PUSH_ROW(r)                                                            \
    do {                                                                       \
        int32_t r_ = (r);                                                      \
        if (r_ > j && seen[r_] != j) {                                         \
            seen[r_] = j;                                                      \
            if (len == cap) {                                                  \
                cap *= 2;                                                      \
                rowind = (int32_t*)realloc(rowind, sizeof(int32_t) * (size_t)cap); \
            }                                                                  \
            rowind[len++] = r_;                                                \
        }                                                                      \
    } while (0)
        for (int64_t k = A->colbeg[j]; k < A->colend[j]; k++) PUSH_ROW(A->csr[k]);
        for (int32_t c = head[j]; c >= 0; c = next[c])
            for (int64_t k = colptr[c]; k < colptr[c + 1]; k++) PUSH_ROW(rowind[k]);
#undef PUSH_ROW
        qsort(rowind + start, (size_t)(len - start), sizeof(int32_t), cmp_i32);
        colptr[j + 1] = len;
        if (len > start) {
            int32_t p = rowind[start];
            next[j] = head[p];
            head[p] = j;
        }
    }
    free(head);
    free(next);
    free(seen);
    *colptr_out = colptr;
    *rowind_out = rowind;
    return len;
}

void orc_free(void* p) { free(p); }
