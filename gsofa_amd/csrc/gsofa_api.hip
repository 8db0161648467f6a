// gsofa_api.hip -- host driver behind the C ABI (include/gsofa.h).
//
// Replaces symbfact_min_id (with_space_optimization/symbfact.cuh:1553-2583):
// the reference's per-level launch + device sync + memcpy loop and its host
// spill buffers become one cooperative persistent launch per virtual-GPU
// share, with the level loop, the group loop and queue-overflow recovery on
// the device.  The host only sizes buffers, launches, and builds outputs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include <dlfcn.h>
#include <sys/mman.h>

#include "gsofa.h"
#include "gsofa_internal.h"

using namespace gsofa;

// Rows of one share moved to host memory (GSOFA_FLAG_SPILL): the share owns
// rows g, g + N, g + 2N, ...; its k-th row is Lcol[Loff[k] .. Loff[k+1]) (and
// the same for U).  Host memory is O(n + nnz(L+U)) whatever N is.
struct SpillShare {
    std::vector<int64_t> Loff, Uoff;
    std::vector<int32_t> Lcol, Ucol;
};

struct gsofa_plan {
    int device = 0;
    hipStream_t stream = nullptr;
    gsofa_options opt{};
    int32_t n = 0;
    int64_t nnz = 0;
    int32_t* d_rp = nullptr;
    int32_t* d_ci = nullptr;
    std::vector<int32_t> h_rp;  // packed row pointers (for sizing and stats)

    int mode = GSOFA_MODE_BLOCK;
    int32_t C = 0;        // grid mode: sources per group; block mode: nblocks * S
    uint32_t S = 0;       // block mode: slots per workgroup
    size_t blk_lds = 0;
    uint32_t* d_mask = nullptr;
    uint4* d_qgv = nullptr;
    uint32_t* d_bstate = nullptr;
    // matrices of a batch (block-diagonal global numbering) and the work table
    int32_t n_mats = 1;
    std::vector<uint32_t> h_mat_base;   // n_mats + 1
    uint32_t* d_mat_base = nullptr;
    uint32_t W = 0;                     // largest matrix
    std::vector<BlkChunk> h_chunks;
    std::vector<uint32_t> share_chunk_off;  // per share, into h_chunks
    BlkChunk* d_chunks = nullptr;
    unsigned long long* d_row_te = nullptr;
    unsigned long long* d_row_reach = nullptr;
    unsigned long long* d_row_fk = nullptr;   // fills per source row, counted by the block kernel
    // reach mode: A^T as packed rows, visited words and next-level masks per
    // workgroup (bits = sources per workgroup), per-matrix TE and |R|
    int32_t* d_rpT = nullptr;
    int32_t* d_ciT = nullptr;
    void* d_vis = nullptr;
    void* d_rmask = nullptr;
    int rbits = 0;
    int rthreads = 1024;  // reach mode: threads per workgroup
    unsigned long long* d_mat_te = nullptr;
    unsigned long long* d_mat_reach = nullptr;
    std::vector<unsigned long long> h_mat_te, h_mat_reach;  // last fetch (batches)
    uint32_t Q = 0;
    uint32_t nblocks = 0;
    int coop = 1;
    uint64_t lab_words = 0;
    uint32_t orphan_words = 0;
    uint32_t* d_lab = nullptr;
    uint32_t* d_mark = nullptr;
    uint32_t* d_orphan = nullptr;
    uint64_t* d_q0 = nullptr;
    uint64_t* d_q1 = nullptr;
    uint64_t* d_fills = nullptr;
    uint64_t fill_cap = 0;
    uint64_t fill_need = 0;  // bounded plans: the fills the row bounds allow
    Ctrl* d_ctrl = nullptr;
    // Output sized from the row bounds of gsofa_fill_bound: the fill stream
    // cannot overflow and the L/U columns exist before the traversal, so an
    // execute runs the traversal once and synchronises with the host only at
    // its end.  Otherwise (an explicit output_capacity, or bounds that do not
    // fit the device) the stream grows on demand and an overflow reruns.
    bool bounded = false;
    bool totals_known = true;  // nnzL / nnzU hold the last execute's totals
    int32_t passes = 0;        // traversal passes of the last execute
    // Groups between label re-initialisations: the decreasing offset has room
    // for 2^32 / n of them (with_space_optimization/symbfact.cuh:1928-1932,
    // 2388-2407).  GSOFA_OFFSET_SPAN=k in the environment at plan creation
    // lowers it to k, so tests reach the re-initialisation in a few groups.
    uint32_t span_cap = 0xFFFFFFFFu;

    // hybrid: sources [s_top, n) run first as grid-mode groups [h_g_lo, h_g_hi)
    // of hC sources (markers, orphans and queues below), the rest in block mode
    bool hybrid = false;
    uint32_t s_top = 0;
    int32_t hC = 0, h_g_lo = 0, h_g_hi = 0;
    uint32_t hQ = 0, h_nblocks = 0;
    uint64_t hlab_words = 0;
    uint32_t* d_hlab = nullptr;

    // persistent label/marker state carried across launches
    uint32_t off = 0;
    uint32_t since = 0;
    uint64_t hw = 0;
    uint32_t epoch = 0;       // stamp_base = epoch * n
    bool marks_clean = true;  // markers hold only stamps below the next stamp_base

    // structure outputs
    uint32_t* d_cntL = nullptr;
    uint32_t* d_cntU = nullptr;
    int64_t* d_Lrp = nullptr;
    int64_t* d_Urp = nullptr;
    int64_t* d_rowfill = nullptr;
    int64_t* d_scan = nullptr;
    size_t scan_words = 0;
    int32_t* d_Lcol = nullptr;
    int32_t* d_Ucol = nullptr;
    int32_t* d_Ltmp = nullptr;
    int32_t* d_Utmp = nullptr;
    int64_t Lcap = 0, Ucap = 0;
    int64_t nnzL = 0, nnzU = 0;
    int32_t* d_supno = nullptr;
    int32_t* d_sstart = nullptr;
    int32_t* d_nsuper = nullptr;
    int64_t* d_sn64 = nullptr;
    uint32_t* d_snflag = nullptr;

    // last execute
    bool executed = false;
    bool structure = false;
    bool supernodes = false;
    bool spilled = false;               // the last execute moved its rows to host memory
    std::vector<SpillShare> sp;         // per share, when spilled
    std::vector<int64_t> sp_rowfill;    // n, when spilled
    uint32_t* d_sp_lcnt = nullptr;      // n: L / U row lengths and pass-through flags
    uint32_t* d_sp_ucnt = nullptr;      //    kept across spilled shares
    uint32_t* d_sp_pass = nullptr;
    Ctrl stats{};
    double ms = 0, traverse_ms = 0, share_min = 0, share_max = 0;
    gsofa_result host{};      // host copy of the last execute (per-matrix fetches)
    bool host_valid = false;
    std::vector<hipEvent_t> ev;
    uint64_t spin_ticks = 0;
    // GSOFA_FLAG_GRAPH: the captured execute (traversal, structure,
    // supernodes, counters to pinned host memory), replayed by later executes
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    bool graph_failed = false;  // capture was refused: executes run eagerly
    bool capturing = false;     // the stream is being captured: no event records
    Ctrl* h_ctrl = nullptr;     // pinned copy of the device counters
    bool last_replayed = false; // the last execute replayed the graph
    // arena (request server worker): capacities of the buffers a batch
    // fills, host copies the uploads read from, and supernode arrays' size
    bool arena = false;
    struct Caps {
        size_t rp = 0, ci = 0, fills = 0, L = 0, U = 0, rows = 0, scan = 0, mats = 0, chunks = 0, wg = 0, nb = 0;
    } cap;
    size_t sn_cap = 0;
    std::vector<int32_t> h_ci, h_rpT, h_ciT;
    int64_t held = 0;  // device bytes the arena holds
};

namespace {

const char* kErr[] = {"ok",
                      "invalid argument or malformed matrix",
                      "allocation failed",
                      "HIP runtime error",
                      "problem exceeds 32-bit index limits",
                      "file could not be read or written",
                      "no usable GPU",
                      "does not fit the device-memory budget",
                      "device-side wait exceeded its bound",
                      "call not valid in the current state"};

#define HIPCHK(x)                                                                       \
    do {                                                                                \
        hipError_t e_ = (x);                                                            \
        if (e_ != hipSuccess) {                                                         \
            fprintf(stderr, "gsofa: %s failed: %s (%s:%d)\n", #x, hipGetErrorString(e_), \
                    __FILE__, __LINE__);                                                \
            return GSOFA_ERR_HIP;                                                       \
        }                                                                               \
    } while (0)
// The same, for code that has queued copies into host memory on stream s:
// they end before the error returns, since the caller may free that memory.
#define HIPCHK_OUT(x, s)                                                                \
    do {                                                                                \
        hipError_t e_ = (x);                                                            \
        if (e_ != hipSuccess) {                                                         \
            fprintf(stderr, "gsofa: %s failed: %s (%s:%d)\n", #x, hipGetErrorString(e_), \
                    __FILE__, __LINE__);                                                \
            (void)hipStreamSynchronize(s);                                              \
            return GSOFA_ERR_HIP;                                                       \
        }                                                                               \
    } while (0)

template <class T>
int dalloc(T** p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    if (hipMalloc((void**)p, count * sizeof(T)) != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
        return GSOFA_ERR_ALLOC;
    }
    return GSOFA_OK;
}

// Device allocations released together when the holder goes out of scope.
struct DevScratch {
    std::vector<void*> p;
    ~DevScratch() {
        for (void* q : p) (void)hipFree(q);
    }
    template <class T>
    int get(T** out, size_t count) {
        void* q = nullptr;
        if (hipMalloc(&q, sizeof(T) * std::max<size_t>(count, 1)) != hipSuccess) {
            (void)hipGetLastError();
            return GSOFA_ERR_ALLOC;
        }
        p.push_back(q);
        *out = (T*)q;
        return GSOFA_OK;
    }
};

template <class T>
void dfree(T*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

int validate(const gsofa_matrix* A) {
    if (!A || A->n < 0 || A->nnz < 0) return GSOFA_ERR_INVALID;
    if (A->n > kMaxN || A->nnz >= (int64_t)INT32_MAX) return GSOFA_ERR_TOO_LARGE;
    if (A->n > 0 && (!A->colbeg || !A->colend || (A->nnz > 0 && !A->csr))) return GSOFA_ERR_INVALID;
    for (int32_t i = 0; i < A->n; i++) {
        const int64_t b = A->colbeg[i], e = A->colend[i];
        if (b < 0 || e < b || e > A->nnz) return GSOFA_ERR_INVALID;
        for (int64_t k = b; k < e; k++)
            if (A->csr[k] < 0 || A->csr[k] >= A->n) return GSOFA_ERR_INVALID;
    }
    return GSOFA_OK;
}

// Pack rows contiguously, sorted and without duplicates.
void pack_rows(const gsofa_matrix* A, std::vector<int32_t>& rp, std::vector<int32_t>& ci) {
    const int32_t n = A->n;
    rp.assign((size_t)n + 1, 0);
    ci.resize((size_t)A->nnz);
    int64_t p = 0;
    for (int32_t i = 0; i < n; i++) {
        rp[i] = (int32_t)p;
        const int64_t b = A->colbeg[i], e = A->colend[i];
        int32_t* dst = ci.data() + p;
        std::copy(A->csr + b, A->csr + e, dst);
        std::sort(dst, dst + (e - b));
        int64_t len = std::unique(dst, dst + (e - b)) - dst;
        p += len;
    }
    rp[n] = (int32_t)p;
    ci.resize((size_t)p);
}

uint64_t share_count(int32_t n, int32_t N, int32_t gid) {
    return gid < n ? ((uint64_t)n - (uint64_t)gid + (uint64_t)N - 1) / (uint64_t)N : 0;
}

uint64_t slot_base_h(uint64_t k, uint64_t first, uint64_t N) {
    return k * (first + 1) + N * (k * (k - 1) / 2);
}

// Label words the packed arena needs for chunk C (largest group of any share).
uint64_t label_words(int32_t n, int32_t N, int32_t C, int32_t gid_lo, int32_t gid_hi) {
    uint64_t best = 0;
    for (int32_t gid = gid_lo; gid <= gid_hi; gid++) {
        const uint64_t cnt = share_count(n, N, gid);
        if (!cnt) continue;
        const uint64_t groups = (cnt + C - 1) / C;
        // the last two groups bound the arena (nsrc shrinks only in the last)
        for (uint64_t g = groups >= 2 ? groups - 2 : 0; g < groups; g++) {
            const uint64_t k0 = g * (uint64_t)C;
            const uint64_t nsrc = std::min<uint64_t>((uint64_t)C, cnt - k0);
            const uint64_t first = (uint64_t)gid + k0 * (uint64_t)N;
            best = std::max(best, slot_base_h(nsrc, first, N));
        }
    }
    return best;
}

// Output one execute needs at most, from the row bounds of lu_row_bounds:
// fills, L and U entries of the rows it processes (every share, one share, or
// for a spilled run the largest share).  rp/ci: packed rows.
void output_need(const gsofa_options& o, int32_t n, const int32_t* rp, const int32_t* ci, uint64_t* need_fill,
                 uint64_t* need_L, uint64_t* need_U) {
    std::vector<int64_t> Lb, Ub;
    lu_row_bounds(n, rp, ci, Lb, Ub);
    const int32_t N = std::max(1, o.n_virtual_gpu);
    std::vector<uint64_t> fs((size_t)N, 0), ls((size_t)N, 0), us((size_t)N, 0);
    for (int32_t i = 0; i < n; i++) {
        int64_t off = 0;  // A's own entries are not fills
        for (int32_t q = rp[i]; q < rp[i + 1]; q++) off += ci[q] != i;
        fs[(size_t)(i % N)] += (uint64_t)(Lb[i] + Ub[i] - off);
        ls[(size_t)(i % N)] += (uint64_t)Lb[i];
        us[(size_t)(i % N)] += (uint64_t)Ub[i];
    }
    const bool spill = (o.flags & GSOFA_FLAG_SPILL) != 0;
    *need_fill = *need_L = *need_U = 0;
    for (int32_t g = 0; g < N; g++) {
        if (o.virtual_gpu_id >= 0 && g != o.virtual_gpu_id) continue;
        *need_fill = spill ? std::max(*need_fill, fs[g]) : *need_fill + fs[g];
        *need_L = spill ? std::max(*need_L, ls[g]) : *need_L + ls[g];
        *need_U = spill ? std::max(*need_U, us[g]) : *need_U + us[g];
    }
}

// Fill stream (8 B per fill) plus sorted and staging columns (8 B per entry).
int64_t output_bytes(uint64_t fills, uint64_t L, uint64_t U) { return (int64_t)(8 * fills + 8 * (L + U)); }

struct Sizing {
    int32_t C;
    uint32_t Q;
    uint64_t lab_words;
    uint64_t fill_cap;
    int64_t work_bytes;  // labels + markers + orphan bitmap + queues
    int64_t total_bytes;
};

int64_t work_bytes_for(int32_t n, int32_t N, int32_t C, uint32_t Q, int32_t gid_lo, int32_t gid_hi,
                       uint64_t* lw) {
    const uint64_t words = label_words(n, N, C, gid_lo, gid_hi);
    if (lw) *lw = words;
    const uint64_t ow = ((uint64_t)n + 31) / 32;
    return (int64_t)(4 * words + 4 * (uint64_t)C * n + 4 * (uint64_t)C * ow + 16 * (uint64_t)Q);
}

uint32_t default_queue(int32_t n, int32_t C) {
    uint64_t q = (uint64_t)C * (uint64_t)n / 2;
    q = std::max<uint64_t>(q, 4096);
    q = std::min<uint64_t>(q, 1ull << 28);
    return (uint32_t)q;
}

int choose_sizing(const gsofa_matrix* A, const gsofa_options* opt, int64_t nnz_packed,
                  int64_t avail_bytes, Sizing* sz) {
    const int32_t n = A->n;
    const int32_t N = std::max(1, opt->n_virtual_gpu);
    const int32_t lo = opt->virtual_gpu_id >= 0 ? opt->virtual_gpu_id : 0;
    const int32_t hi = opt->virtual_gpu_id >= 0 ? opt->virtual_gpu_id : N - 1;
    const uint64_t maxsrc = share_count(n, N, lo);  // share 0 is the largest
    int32_t C = opt->chunk > 0 ? opt->chunk : 8192;
    C = (int32_t)std::min<uint64_t>((uint64_t)C, std::max<uint64_t>(maxsrc, 1));
    const int64_t budget = opt->mem_budget_bytes > 0 ? opt->mem_budget_bytes : avail_bytes;
    for (;;) {
        uint32_t Q = opt->queue_capacity > 0
                         ? (uint32_t)std::min<int64_t>(opt->queue_capacity, (int64_t)UINT32_MAX - 64)
                         : default_queue(n, C);
        uint64_t lw = 0;
        int64_t wb = work_bytes_for(n, N, C, Q, lo, hi, &lw);
        if (wb <= budget || C == 1) {
            if (wb > budget && opt->mem_budget_bytes > 0) {
                // shrink the queue before giving up
                while (wb > budget && Q > 64 && opt->queue_capacity <= 0) {
                    Q /= 2;
                    wb = work_bytes_for(n, N, C, Q, lo, hi, &lw);
                }
                if (wb > budget) return GSOFA_ERR_BUDGET;
            }
            sz->C = C;
            sz->Q = Q;
            sz->lab_words = lw;
            sz->work_bytes = wb;
            uint64_t fc = opt->output_capacity > 0 ? (uint64_t)opt->output_capacity
                                                   : std::max<uint64_t>(64 * (uint64_t)nnz_packed, 1 << 20);
            sz->fill_cap = fc;
            sz->total_bytes = wb + (int64_t)(8 * fc) + 4 * (int64_t)(n + 1 + nnz_packed);
            return GSOFA_OK;
        }
        C = std::max(1, C / 2);
    }
}


// Work table: chunks of up to S (slots per workgroup) sources of one matrix, in the order
// they are handed out (decreasing global source).  Shares of a single matrix
// get consecutive ranges of the table.
void build_chunks(gsofa_plan* P) {
    const int32_t N = std::max(1, P->opt.n_virtual_gpu);
    const uint32_t S = P->S;
    P->h_chunks.clear();
    P->share_chunk_off.assign((size_t)N + 1, 0);
    if (P->n_mats == 1) {
        for (int32_t gid = 0; gid < N; gid++) {
            P->share_chunk_off[gid] = (uint32_t)P->h_chunks.size();
            const uint64_t cnt = P->hybrid ? P->s_top : share_count(P->n, N, gid);
            for (uint64_t c = 0; c * S < cnt; c++) {
                const uint64_t k = cnt - 1 - c * S;
                BlkChunk ch;
                ch.first = (uint32_t)(gid + k * (uint64_t)N);
                ch.count = (uint32_t)std::min<uint64_t>(S, cnt - c * S);
                ch.stride = (uint32_t)N;
                ch.mat = 0;
                P->h_chunks.push_back(ch);
            }
        }
        P->share_chunk_off[N] = (uint32_t)P->h_chunks.size();
        return;
    }
    for (int32_t b = P->n_mats - 1; b >= 0; b--) {
        const uint32_t base = P->h_mat_base[b], nb = P->h_mat_base[b + 1] - base;
        for (uint32_t c = 0; c * S < nb; c++) {
            BlkChunk ch;
            ch.first = base + nb - 1 - c * S;
            ch.count = std::min<uint32_t>(S, nb - c * S);
            ch.stride = 1;
            ch.mat = (uint32_t)b;
            P->h_chunks.push_back(ch);
        }
    }
    // Longest first: a source's traversal runs through the vertices numbered
    // below it, so its local index bounds its reach; the deepest chunks of
    // every matrix start before the short ones fill in behind them.
    std::stable_sort(P->h_chunks.begin(), P->h_chunks.end(), [&](const BlkChunk& a, const BlkChunk& b) {
        return a.first - P->h_mat_base[a.mat] > b.first - P->h_mat_base[b.mat];
    });
    P->share_chunk_off[1] = (uint32_t)P->h_chunks.size();
}

int size_block_mode(gsofa_plan* P, int64_t avail) {
    const int32_t N = std::max(1, P->opt.n_virtual_gpu);
    const uint64_t maxsrc = std::max<uint64_t>(P->hybrid ? P->s_top : share_count(P->n, N, 0), 1);
    const uint32_t S = P->S;
    const uint64_t nchunks = P->n_mats == 1 ? (maxsrc + S - 1) / S
                                            : std::max<uint64_t>(P->h_chunks.size(), 1);
    uint32_t nb = 0;
    const int rc = gsofa::block_grid(P->opt, (int)S, nchunks, P->W, avail, &nb);
    if (rc) return rc;
    P->nblocks = nb;
    P->C = (int32_t)(nb * S);
    P->Q = 0;
    P->blk_lds = 0;  // all LDS of the block kernel is static
    return GSOFA_OK;
}

// Replicate the grid kernel's offset evolution over groups [g_lo, g_hi) of
// share gid, C sources per group.
void advance_groups(gsofa_plan* P, int32_t gid, int32_t C, uint64_t g_lo, uint64_t g_hi) {
    const uint32_t n = (uint32_t)P->n;
    const uint32_t N = (uint32_t)std::max(1, P->opt.n_virtual_gpu);
    const uint32_t loops = std::min(0xFFFFFFFFu / n, P->span_cap);
    const uint64_t cnt = share_count(P->n, (int32_t)N, gid);
    for (uint64_t g = g_lo; g < g_hi; g++) {
        const uint64_t k0 = g * (uint64_t)C;
        const uint64_t nsrc = std::min<uint64_t>((uint64_t)C, cnt - k0);
        const uint64_t size_g = slot_base_h(nsrc, (uint64_t)gid + k0 * N, N);
        if (P->since >= loops) {
            P->off = 0xFFFFFFFFu - n;
            P->since = 0;
            P->hw = 0;
        }
        if (size_g > P->hw) P->hw = size_g;
        P->off -= n;
        P->since++;
    }
}

void advance_state(gsofa_plan* P, int32_t gid) {
    const int32_t N = std::max(1, P->opt.n_virtual_gpu);
    const uint64_t cnt = share_count(P->n, N, gid);
    advance_groups(P, gid, P->C, 0, (cnt + P->C - 1) / P->C);
}

void release_device(gsofa_plan* P) {
    dfree(P->d_rp);
    dfree(P->d_ci);
    dfree(P->d_lab);
    dfree(P->d_mark);
    dfree(P->d_orphan);
    dfree(P->d_q0);
    dfree(P->d_q1);
    dfree(P->d_fills);
    dfree(P->d_ctrl);
    dfree(P->d_cntL);
    dfree(P->d_cntU);
    dfree(P->d_Lrp);
    dfree(P->d_Urp);
    dfree(P->d_rowfill);
    dfree(P->d_scan);
    dfree(P->d_Lcol);
    dfree(P->d_Ucol);
    dfree(P->d_Ltmp);
    dfree(P->d_Utmp);
    dfree(P->d_supno);
    dfree(P->d_sstart);
    dfree(P->d_nsuper);
    dfree(P->d_sn64);
    dfree(P->d_snflag);
    dfree(P->d_sp_lcnt);
    dfree(P->d_sp_ucnt);
    dfree(P->d_sp_pass);
    dfree(P->d_mask);
    dfree(P->d_qgv);
    dfree(P->d_bstate);
    dfree(P->d_mat_base);
    dfree(P->d_chunks);
    dfree(P->d_row_te);
    dfree(P->d_hlab);
    dfree(P->d_row_reach);
    dfree(P->d_row_fk);
    dfree(P->d_rpT);
    dfree(P->d_ciT);
    dfree(P->d_vis);
    dfree(P->d_rmask);
    dfree(P->d_mat_te);
    dfree(P->d_mat_reach);
}

int ensure_cols(gsofa_plan* P, int64_t nl, int64_t nu) {
    if (nl > P->Lcap) {
        dfree(P->d_Lcol);
        dfree(P->d_Ltmp);
        int64_t cap = nl + nl / 8 + 1024;
        if (dalloc(&P->d_Lcol, (size_t)cap) || dalloc(&P->d_Ltmp, (size_t)cap)) return GSOFA_ERR_ALLOC;
        P->Lcap = cap;
    }
    if (nu > P->Ucap) {
        dfree(P->d_Ucol);
        dfree(P->d_Utmp);
        int64_t cap = nu + nu / 8 + 1024;
        if (dalloc(&P->d_Ucol, (size_t)cap) || dalloc(&P->d_Utmp, (size_t)cap)) return GSOFA_ERR_ALLOC;
        P->Ucap = cap;
    }
    return GSOFA_OK;
}

// Device-to-host copies of large results go through pinned staging buffers
// (a process-wide pool, kept until exit): the DMA of one chunk overlaps the
// host copy of the previous one into the caller's (pageable) array.  That host
// copy is what limits the rate: it first touches fresh pages (a few GB/s per
// thread), so a large copy is split over up to kCopyParts threads, each with
// two staging buffers of its own, and result arrays are allocated on 2 MB
// boundaries with transparent huge pages requested (host_array).
constexpr size_t kStageBytes = 16u << 20;
constexpr size_t kStageMin = 8u << 20;    // smaller copies go direct
constexpr size_t kPartMin = 64u << 20;    // bytes per copy thread, at least
constexpr int kCopyParts = 4;
constexpr size_t kHugePage = 2u << 20;
std::mutex g_stage_mu;
std::vector<void*> g_stage_free;

void* stage_get() {
    {
        std::lock_guard<std::mutex> lk(g_stage_mu);
        if (!g_stage_free.empty()) {
            void* p = g_stage_free.back();
            g_stage_free.pop_back();
            return p;
        }
    }
    void* p = nullptr;
    if (hipHostMalloc(&p, kStageBytes, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    return p;
}

void stage_put(void* p) {
    if (!p) return;
    std::lock_guard<std::mutex> lk(g_stage_mu);
    g_stage_free.push_back(p);
}

// dst <- src through the staging buffers b0, b1: chunk k + 1's DMA runs while
// chunk k is copied out.  Every DMA has ended when this returns.
hipError_t copy_staged(void* dst, const void* src, size_t bytes, hipStream_t s, void* b0, void* b1) {
    void* b[2] = {b0, b1};
    hipEvent_t ev[2] = {nullptr, nullptr};
    hipError_t e = hipEventCreateWithFlags(&ev[0], hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ev[1], hipEventDisableTiming);
    const size_t nch = (bytes + kStageBytes - 1) / kStageBytes;
    auto issue = [&](size_t k) {
        const size_t off = k * kStageBytes, len = std::min(kStageBytes, bytes - off);
        hipError_t r = hipMemcpyAsync(b[k & 1], (const char*)src + off, len, hipMemcpyDeviceToHost, s);
        return r == hipSuccess ? hipEventRecord(ev[k & 1], s) : r;
    };
    if (e == hipSuccess) e = issue(0);
    for (size_t k = 0; k < nch && e == hipSuccess; k++) {
        if (k + 1 < nch) e = issue(k + 1);  // its buffer was emptied in iteration k - 1
        if (e == hipSuccess) e = hipEventSynchronize(ev[k & 1]);
        if (e == hipSuccess) {
            const size_t off = k * kStageBytes, len = std::min(kStageBytes, bytes - off);
            memcpy((char*)dst + off, b[k & 1], len);
        }
    }
    if (e != hipSuccess) (void)hipStreamSynchronize(s);  // no copy left in flight into the buffers
    for (hipEvent_t x : ev)
        if (x) (void)hipEventDestroy(x);
    return e;
}

// dst <- src (device), complete on return.  Work queued on s before the call
// is ordered before the copy.
hipError_t copy_d2h(void* dst, const void* src, size_t bytes, hipStream_t s) {
    if (bytes == 0) return hipSuccess;
    int parts = 0;
    void* b[2 * kCopyParts] = {};
    if (bytes >= kStageMin) {
        const size_t want = std::min<size_t>(kCopyParts, std::max<size_t>(1, bytes / kPartMin));
        for (; parts < (int)want; parts++) {
            b[2 * parts] = stage_get();
            b[2 * parts + 1] = stage_get();
            if (!b[2 * parts] || !b[2 * parts + 1]) {
                stage_put(b[2 * parts]);
                stage_put(b[2 * parts + 1]);
                b[2 * parts] = b[2 * parts + 1] = nullptr;
                break;
            }
        }
    }
    if (parts == 0) {  // small, or no pinned memory: one direct copy
        hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s);
        return e == hipSuccess ? hipStreamSynchronize(s) : e;
    }
    // parts of whole staging chunks; the last takes the remainder
    const size_t chunks = (bytes + kStageBytes - 1) / kStageBytes;
    const size_t per = (chunks + parts - 1) / parts * kStageBytes;
    std::vector<hipError_t> err((size_t)parts, hipSuccess);
    std::vector<std::thread> th;
    auto part = [&](int i) {
        const size_t off = (size_t)i * per;
        if (off >= bytes) return;
        err[(size_t)i] = copy_staged((char*)dst + off, (const char*)src + off, std::min(per, bytes - off), s,
                                     b[2 * i], b[2 * i + 1]);
    };
    for (int i = 1; i < parts; i++) {
        try {
            th.emplace_back(part, i);
        } catch (...) {  // no thread: this one copies that part too
            part(i);
        }
    }
    part(0);
    for (std::thread& t : th) t.join();
    for (int i = 0; i < 2 * parts; i++) stage_put(b[i]);
    for (hipError_t x : err)
        if (x != hipSuccess) return x;
    return hipSuccess;
}

// A host result array of `bytes`, released with free(): large ones start on
// a 2 MB boundary with transparent huge pages requested, so that their first
// touch (the staged copy above) faults once per 2 MB rather than per 4 KB.
void* host_array(size_t bytes) {
    if (bytes < kStageMin) return malloc(std::max<size_t>(bytes, 1));
    const size_t len = (bytes + kHugePage - 1) / kHugePage * kHugePage;
    void* p = nullptr;
    if (posix_memalign(&p, kHugePage, len) != 0) return nullptr;
    (void)madvise(p, len, MADV_HUGEPAGE);  // advice only: small pages if refused
    return p;
}

float ev_ms(hipEvent_t a, hipEvent_t b) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) {
        (void)hipGetLastError();
        return 0.f;
    }
    return ms;
}

}  // namespace

namespace gsofa {

int validate_matrix(const gsofa_matrix* A) { return validate(A); }

// Block mode: per workgroup, labels [W][S] + two masks + two 16-byte-entry
// queues = (4 S + 40) W bytes (+ 8 of state): 168 W at 32 slots.
int64_t block_bytes(uint32_t nb, uint32_t W, int slots) {
    return (int64_t)nb * ((4ll * slots + 8 + 32) * W + 8);
}

// Everything else a block-mode plan allocates: CSR, fill stream, per-row
// arrays, scan scratch, work table, and the L/U columns (+ sort scratch)
// reserved for nnz + fill_cap entries.
int64_t block_fixed_bytes(int64_t n, int64_t nnz, int32_t n_mats, uint64_t nchunks, uint64_t fill_cap,
                          int64_t cols) {
    if (cols < 0) cols = nnz + (int64_t)fill_cap;
    return 4 * (n + 1) + 4 * nnz + 8 * (int64_t)fill_cap + 104 * n + 8 * (int64_t)scan_tmp_words(n) +
           16 * (int64_t)nchunks + 4 * ((int64_t)n_mats + 1) + 8 * cols + (1 << 20);
}

uint64_t default_fill_cap(int64_t nnz) { return std::max<uint64_t>(64 * (uint64_t)nnz, 1 << 20); }

// Slots per workgroup when the options leave it open (block_slots = 0).
// Batches take kBatchSlots: within a few percent of the best on batches of
// small 2D meshes (where 8 is best) and the best on 3D meshes.  A single
// matrix with nsrc sources takes the largest of 32, 16, 8 whose chunks still
// give every co-resident workgroup one.
int choose_slots(const gsofa_options& o, int32_t n_mats, uint64_t nsrc, int* slots) {
    if (o.block_slots != 0) {
        if (!blk_slots_valid(o.block_slots)) return GSOFA_ERR_INVALID;
        *slots = o.block_slots;
        return GSOFA_OK;
    }
    if (n_mats > 1) {
        *slots = kBatchSlots;
        return GSOFA_OK;
    }
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, o.device) != hipSuccess) {
        (void)hipGetLastError();
        return GSOFA_ERR_HIP;
    }
    for (int S = kBlockSlots; S > 8; S /= 2) {
        int occ = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, blk_kernel_ptr(S), blk_block_size(), 0) !=
            hipSuccess) {
            (void)hipGetLastError();
            return GSOFA_ERR_HIP;
        }
        if ((nsrc + S - 1) / S >= (uint64_t)std::max(occ, 1) * (uint64_t)cus) {
            *slots = S;
            return GSOFA_OK;
        }
    }
    *slots = 8;
    return GSOFA_OK;
}

// Workgroups of the block kernel with `slots` slots that are co-resident on
// `device`.
int resident_blocks(int slots, int device, uint32_t* out) {
    int occ = 0, cus = 0;
    if (!blk_slots_valid(slots)) return GSOFA_ERR_INVALID;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, blk_kernel_ptr(slots), blk_block_size(), 0) !=
            hipSuccess ||
        occ <= 0 || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) {
        (void)hipGetLastError();
        return GSOFA_ERR_HIP;
    }
    *out = (uint32_t)occ * (uint32_t)cus;
    return GSOFA_OK;
}

// As many workgroups as can be co-resident (or as the options ask: chunk =
// concurrent sources, grid_blocks), at most one per chunk, shrunk until the
// per-workgroup arrays fit `avail` bytes (and opt.mem_budget_bytes if set).
int block_grid(const gsofa_options& o, int S, uint64_t nchunks, uint32_t W, int64_t avail, uint32_t* nb_out) {
    if (o.mem_budget_bytes > 0) avail = std::min<int64_t>(avail, o.mem_budget_bytes);
    uint32_t nb = 0;
    const int rc = resident_blocks(S, o.device, &nb);
    if (rc) return rc;
    if (o.grid_blocks > 0) nb = std::min<uint32_t>(nb, (uint32_t)o.grid_blocks);
    if (o.chunk > 0) nb = std::min<uint32_t>(nb, (uint32_t)std::max(1, o.chunk / S));
    nb = (uint32_t)std::min<uint64_t>(nb, std::max<uint64_t>(nchunks, 1));
    while (block_bytes(nb, W, S) > avail && nb > 1) nb = nb > 64 ? nb - nb / 8 : nb - 1;
    if (block_bytes(nb, W, S) > avail) return GSOFA_ERR_BUDGET;
    *nb_out = nb;
    return GSOFA_OK;
}

// ---- tuned knobs ---------------------------------------------------------------
// The autotuner (gsofa_amd/tune.py) times the knobs that never change results
// and writes its choice per workload signature to a text table, one line per
// signature: "<signature> key=value ...".  The library reads the table named
// by GSOFA_TUNE_CACHE (empty or "off": none), else tune_table.txt beside
// libgsofa.so, once per process, and applies an entry's knobs wherever the
// caller left them at their defaults.  The signature buckets the largest
// matrix's vertex count and the mean row length to powers of two, and a
// batch's matrix count to the next power of two.

namespace {

struct TuneEntry {
    std::string sig;
    int32_t mode = -1, block_slots = -1, chunk = -1, grid_blocks = -1, top_sources = -1, threshold = -1;
    int32_t threads = -1;
    int64_t queue_capacity = -1;
};

std::vector<TuneEntry> load_tune_table() {
    std::vector<TuneEntry> t;
    std::string path;
    if (const char* e = getenv("GSOFA_TUNE_CACHE")) {
        path = e;
        if (path.empty() || path == "off") return t;
    } else {
        Dl_info info;
        if (!dladdr(reinterpret_cast<void*>(&load_tune_table), &info) || !info.dli_fname) return t;
        path = info.dli_fname;
        const size_t slash = path.rfind('/');
        path = (slash == std::string::npos ? std::string(".") : path.substr(0, slash)) + "/tune_table.txt";
    }
    FILE* f = fopen(path.c_str(), "r");
    if (!f) return t;
    char line[1024];
    while (fgets(line, sizeof(line), f)) {
        if (line[0] == '#' || line[0] == '\n') continue;
        TuneEntry e;
        char* save = nullptr;
        char* tok = strtok_r(line, " \t\n", &save);
        if (!tok) continue;
        e.sig = tok;
        while ((tok = strtok_r(nullptr, " \t\n", &save))) {
            char* eq = strchr(tok, '=');
            if (!eq) continue;
            *eq = 0;
            const long long v = atoll(eq + 1);
            if (!strcmp(tok, "mode")) e.mode = (int32_t)v;
            else if (!strcmp(tok, "block_slots")) e.block_slots = (int32_t)v;
            else if (!strcmp(tok, "chunk")) e.chunk = (int32_t)v;
            else if (!strcmp(tok, "grid_blocks")) e.grid_blocks = (int32_t)v;
            else if (!strcmp(tok, "top_sources")) e.top_sources = (int32_t)v;
            else if (!strcmp(tok, "wave_block_threshold")) e.threshold = (int32_t)v;
            else if (!strcmp(tok, "queue_capacity")) e.queue_capacity = (int64_t)v;
            else if (!strcmp(tok, "workgroup_threads")) e.threads = (int32_t)v;
        }
        t.push_back(e);
    }
    fclose(f);
    return t;
}

const std::vector<TuneEntry>& tune_table() {
    static const std::vector<TuneEntry> t = load_tune_table();
    return t;
}

}  // namespace

std::string tune_signature(int64_t n_max, int64_t nnz, int64_t n_total, int32_t n_mats) {
    const double deg = (double)nnz / (double)std::max<int64_t>(n_total, 1);
    const int ln = (int)std::ceil(std::log2((double)std::max<int64_t>(n_max, 2)));
    const int ld = (int)std::lround(std::log2(std::max(deg, 1.0)));
    const int b = n_mats <= 1 ? 1 : (1 << (int)std::ceil(std::log2((double)n_mats)));
    char buf[64];
    snprintf(buf, sizeof(buf), "n2^%d_deg2^%d_b%d", ln, ld, b);
    return buf;
}

// The tuned knobs of the workload's signature, applied where `o` has its
// defaults (mode AUTO, slots, chunk, grid, hybrid split, threshold and queue
// capacity unset).  Returns true when an entry applied.
bool apply_tuned(gsofa_options* o, int64_t n_max, int64_t nnz, int64_t n_total, int32_t n_mats) {
    const std::string sig = tune_signature(n_max, nnz, n_total, n_mats);
    for (const TuneEntry& e : tune_table()) {
        if (e.sig != sig) continue;
        if (o->mode != GSOFA_MODE_AUTO || o->block_slots || o->chunk || o->grid_blocks || o->top_sources ||
            o->queue_capacity || o->workgroup_threads ||
            (o->wave_block_threshold && o->wave_block_threshold != 4096))
            return false;  // the caller chose knobs: they stand
        if (e.mode >= 0) o->mode = e.mode;
        if (e.block_slots >= 0) o->block_slots = e.block_slots;
        if (e.chunk >= 0) o->chunk = e.chunk;
        if (e.grid_blocks >= 0) o->grid_blocks = e.grid_blocks;
        if (e.top_sources >= 0) o->top_sources = e.top_sources;
        if (e.threshold > 0) o->wave_block_threshold = e.threshold;
        if (e.queue_capacity >= 0) o->queue_capacity = e.queue_capacity;
        if (e.threads >= 0) o->workgroup_threads = e.threads;
        return true;
    }
    return false;
}

// The mode AUTO resolves to: reach mode wherever it applies (one share, in
// core, no hybrid split, sources per workgroup left open), else block mode
// (grid mode for block-mode matrices past 32-bit column offsets is decided at
// plan creation).
int resolve_mode(const gsofa_options& o) {
    if (o.mode != GSOFA_MODE_AUTO) return o.mode;
    const bool one_share = o.n_virtual_gpu <= 1;
    if (one_share && !(o.flags & GSOFA_FLAG_SPILL) && o.top_sources <= 0 && o.block_slots == 0)
        return GSOFA_MODE_REACH;
    return GSOFA_MODE_AUTO;
}

// Reach mode: sources per workgroup (32, 64, 128 or 256).  Unless the
// options (or the tuned table) fix it, a single matrix takes 64 when its
// chunks still give every co-resident workgroup one, else 32; batches take
// 64.  Wider groups pay where the sources' searches advance together (banded
// matrices) and cost where they do not (3D nested dissection), so only
// tuning picks them.
int choose_bits(const gsofa_options& o, int32_t n_mats, uint64_t nsrc, int* bits) {
    if (o.block_slots != 0) {
        if (!rch_bits_valid(o.block_slots)) return GSOFA_ERR_INVALID;
        *bits = o.block_slots;
        return GSOFA_OK;
    }
    if (n_mats > 1) {
        *bits = 64;
        return GSOFA_OK;
    }
    uint32_t nb = 0;
    int rc = reach_resident(o.device, reach_threads(o), &nb);
    if (rc) return rc;
    *bits = (nsrc + 63) / 64 >= nb ? 64 : 32;
    return GSOFA_OK;
}

int reach_threads(const gsofa_options& o) { return o.workgroup_threads != 0 ? o.workgroup_threads : 1024; }

int reach_resident(int device, int threads, uint32_t* out) {
    int occ = 0, cus = 0;
    if (!rch_threads_valid(threads)) return GSOFA_ERR_INVALID;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, rch_kernel_ptr(64, threads), threads, 0) != hipSuccess ||
        occ <= 0 || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) {
        (void)hipGetLastError();
        return GSOFA_ERR_HIP;
    }
    *out = (uint32_t)occ * (uint32_t)cus;
    return GSOFA_OK;
}

// As block_grid, for the reach kernel's per-workgroup arrays.
int reach_grid(const gsofa_options& o, int bits, uint64_t nchunks, uint32_t W, int64_t avail, uint32_t* nb_out) {
    if (o.mem_budget_bytes > 0) avail = std::min<int64_t>(avail, o.mem_budget_bytes);
    uint32_t nb = 0;
    const int rc = reach_resident(o.device, reach_threads(o), &nb);
    if (rc) return rc;
    if (o.grid_blocks > 0) nb = std::min<uint32_t>(nb, (uint32_t)o.grid_blocks);
    if (o.chunk > 0) nb = std::min<uint32_t>(nb, (uint32_t)std::max(1, o.chunk / bits));
    nb = (uint32_t)std::min<uint64_t>(nb, std::max<uint64_t>(nchunks, 1));
    while (reach_bytes(nb, W, bits) > avail && nb > 1) nb = nb > 64 ? nb - nb / 8 : nb - 1;
    if (reach_bytes(nb, W, bits) > avail) return GSOFA_ERR_BUDGET;
    *nb_out = nb;
    return GSOFA_OK;
}

// Packed rows of A^T from packed rows of A (sorted, as counting sort keeps
// the row order of A within each column).
void transpose_rows(int32_t n, const int32_t* rp, const int32_t* ci, std::vector<int32_t>& rpT,
                    std::vector<int32_t>& ciT) {
    rpT.assign((size_t)n + 1, 0);
    const int64_t nnz = rp[n];
    for (int64_t k = 0; k < nnz; k++) rpT[(size_t)ci[k] + 1]++;
    for (int32_t i = 0; i < n; i++) rpT[(size_t)i + 1] += rpT[(size_t)i];
    ciT.resize((size_t)nnz);
    std::vector<int32_t> cur(rpT.begin(), rpT.end() - 1);
    for (int32_t i = 0; i < n; i++)
        for (int32_t k = rp[i]; k < rp[i + 1]; k++) ciT[(size_t)cur[(size_t)ci[k]]++] = i;
}

}  // namespace gsofa

extern "C" {

const char* gsofa_strerror(int status) {
    int i = -status;
    if (i < 0 || i >= (int)(sizeof(kErr) / sizeof(kErr[0]))) return "unknown status";
    return kErr[i];
}

void gsofa_options_default(gsofa_options* o) {
    if (!o) return;
    memset(o, 0, sizeof(*o));
    o->n_virtual_gpu = 1;
    o->virtual_gpu_id = -1;
    o->wave_block_threshold = 4096;
}

void gsofa_result_free(gsofa_result* r) {
    if (!r) return;
    free(r->row_te);
    free(r->row_reach);
    free(r->row_fill);
    free(r->L_rowptr);
    free(r->L_col);
    free(r->U_rowptr);
    free(r->U_col);
    free(r->super_start);
    free(r->supno);
    memset(r, 0, sizeof(*r));
}

int64_t gsofa_plan_footprint(const gsofa_matrix* A, const gsofa_options* opt) {
    gsofa_options o;
    if (opt) o = *opt;
    else gsofa_options_default(&o);
    if (!A || A->n <= 0) return 0;
    if (o.n_virtual_gpu <= 1 && !(o.flags & GSOFA_FLAG_SPILL)) apply_tuned(&o, A->n, A->nnz, A->n, 1);
    o.mode = resolve_mode(o);
    if (o.mode == GSOFA_MODE_GRID || (o.mode == GSOFA_MODE_AUTO && A->nnz >= kMaxBlockNnz)) {
        Sizing sz;
        if (choose_sizing(A, &o, A->nnz, INT64_MAX / 4, &sz) != GSOFA_OK) return -1;
        return sz.total_bytes;
    }
    if (o.mode == GSOFA_MODE_REACH) {
        if (o.block_slots != 0 && !rch_bits_valid(o.block_slots)) return -1;
        if (validate(A) != GSOFA_OK) return -1;
        int bits = 0;
        if (choose_bits(o, 1, (uint64_t)A->n, &bits) != GSOFA_OK) return -1;
        const uint64_t nchunks = ((uint64_t)A->n + bits - 1) / bits;
        std::vector<int32_t> rp, ci;
        pack_rows(A, rp, ci);
        uint64_t nf = 0, nl = 0, nu = 0;
        gsofa_options w = o;
        w.n_virtual_gpu = 1;
        w.virtual_gpu_id = -1;
        output_need(w, A->n, rp.data(), ci.data(), &nf, &nl, &nu);
        uint32_t nb = 0;
        if (reach_grid(o, bits, nchunks, (uint32_t)A->n, INT64_MAX / 4, &nb) != GSOFA_OK) return -1;
        return reach_bytes(nb, (uint32_t)A->n, bits) +
               block_fixed_bytes(A->n, (int64_t)ci.size(), 1, nchunks, rch_fill_room(nf), (int64_t)(nl + nu)) +
               4 * (int64_t)(A->n + 1) + 4 * (int64_t)ci.size();
    }
    const int32_t N = std::max(1, o.n_virtual_gpu);
    if (o.block_slots != 0 && !blk_slots_valid(o.block_slots)) return -1;
    int S = 0;
    if (choose_slots(o, 1, share_count(A->n, N, 0), &S) != GSOFA_OK) return -1;
    const uint64_t nchunks = (share_count(A->n, N, 0) + S - 1) / S;
    uint64_t fc = o.output_capacity > 0 ? (uint64_t)o.output_capacity : default_fill_cap(A->nnz);
    if ((o.flags & GSOFA_FLAG_SPILL) && o.output_capacity <= 0)  // one share's fills at a time
        fc = std::max<uint64_t>(fc / (uint64_t)N, 1u << 20);
    int64_t cols = -1;  // default: room for nnz + fc entries
    if (o.output_capacity <= 0) {  // what plan creation reserves from the row bounds
        if (validate(A) != GSOFA_OK) return -1;
        std::vector<int32_t> rp, ci;
        pack_rows(A, rp, ci);
        uint64_t nf = 0, nl = 0, nu = 0;
        output_need(o, A->n, rp.data(), ci.data(), &nf, &nl, &nu);
        fc = std::max<uint64_t>(nf, 1);
        cols = (int64_t)(nl + nu);
    }
    uint32_t nb = 0;
    if (block_grid(o, S, nchunks, (uint32_t)A->n, INT64_MAX / 4, &nb) != GSOFA_OK) return -1;
    return block_bytes(nb, (uint32_t)A->n, S) + block_fixed_bytes(A->n, A->nnz, 1, nchunks, fc, cols);
}

}  // extern "C"

// A is the global matrix; for a batch it is block-diagonal with matrix b on
// vertices [mat_base[b], mat_base[b+1]).
static int plan_create_impl(const gsofa_matrix* A, const uint32_t* mat_base, int32_t n_mats,
                            const gsofa_options* opt, gsofa_plan** out) {
    if (!out) return GSOFA_ERR_INVALID;
    *out = nullptr;
    int rc = validate(A);
    if (rc) return rc;
    gsofa_plan* P = new gsofa_plan();
    if (opt) P->opt = *opt;
    else gsofa_options_default(&P->opt);
    P->n_mats = n_mats;
    P->h_mat_base.assign(mat_base, mat_base + n_mats + 1);
    P->W = 1;
    for (int32_t b = 0; b < n_mats; b++) P->W = std::max<uint32_t>(P->W, mat_base[b + 1] - mat_base[b]);
    if (n_mats > 1) {  // batches: one share, reach or block mode
        P->opt.n_virtual_gpu = 1;
        P->opt.virtual_gpu_id = -1;
        if (P->opt.mode != GSOFA_MODE_REACH && P->opt.mode != GSOFA_MODE_AUTO) P->opt.mode = GSOFA_MODE_BLOCK;
    }
    if (P->opt.n_virtual_gpu < 1) P->opt.n_virtual_gpu = 1;
    if (P->opt.n_virtual_gpu == 1 && !(P->opt.flags & GSOFA_FLAG_SPILL))
        apply_tuned(&P->opt, P->W, A->nnz, A->n, n_mats);
    P->opt.mode = resolve_mode(P->opt);
    const bool reach = P->opt.mode == GSOFA_MODE_REACH;
    if (P->opt.virtual_gpu_id >= P->opt.n_virtual_gpu ||
        (P->opt.block_slots != 0 && !(reach ? rch_bits_valid(P->opt.block_slots)
                                            : blk_slots_valid(P->opt.block_slots))) ||
        ((P->opt.flags & GSOFA_FLAG_SPILL) && n_mats > 1) ||
        // reach mode answers whole-matrix runs in core only
        (reach && ((P->opt.virtual_gpu_id >= 0 && P->opt.n_virtual_gpu > 1) || (P->opt.flags & GSOFA_FLAG_SPILL) ||
                   P->opt.top_sources > 0))) {
        delete P;
        return GSOFA_ERR_INVALID;
    }
    if (reach) {  // every share in one execute is the whole matrix
        P->opt.n_virtual_gpu = 1;
        P->opt.virtual_gpu_id = -1;
    }
    if (const char* sp = getenv("GSOFA_OFFSET_SPAN")) {
        const long v = atol(sp);
        if (v >= 1 && v < (long)0xFFFFFFFFu) P->span_cap = (uint32_t)v;
    }
    if (P->opt.wave_block_threshold <= 0) P->opt.wave_block_threshold = 4096;
    P->opt.wave_block_threshold = std::min(P->opt.wave_block_threshold, 1 << 20);
    P->n = A->n;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        (void)hipGetLastError();
        delete P;
        return GSOFA_ERR_NO_DEVICE;
    }
    if (P->opt.device < 0 || P->opt.device >= ndev) {
        delete P;
        return GSOFA_ERR_INVALID;
    }
    P->device = P->opt.device;
    if (hipSetDevice(P->device) != hipSuccess || hipStreamCreate(&P->stream) != hipSuccess) {
        (void)hipGetLastError();
        delete P;
        return GSOFA_ERR_HIP;
    }
    std::vector<int32_t> ci;
    pack_rows(A, P->h_rp, ci);
    P->nnz = (int64_t)ci.size();
    const int32_t n = P->n;

    size_t freeb = 0, totalb = 0;
    (void)hipMemGetInfo(&freeb, &totalb);
    int64_t avail = (int64_t)((double)freeb * 0.80) - 8 * (int64_t)(n + 1 + P->nnz) -
                    64ll * (n + 1) - (64ll << 20);
    // Output reserve from the row bounds: the fills, L and U entries of the
    // rows one execute processes (every share, one share, or for a spilled
    // run the largest share).
    uint64_t need_fill = 0, need_L = 0, need_U = 0;
    if (P->opt.output_capacity <= 0 && n > 0) {
        output_need(P->opt, n, P->h_rp.data(), ci.data(), &need_fill, &need_L, &need_U);
        const int64_t out_bytes = output_bytes(reach ? rch_fill_room(need_fill) : need_fill, need_L, need_U);
        if (out_bytes <= avail / 2) {
            P->bounded = true;
            avail -= out_bytes;
        }
    }
    P->mode = P->opt.mode == GSOFA_MODE_GRID ? GSOFA_MODE_GRID
              : reach                          ? GSOFA_MODE_REACH
                                               : GSOFA_MODE_BLOCK;
    // The block kernel addresses the packed column array with 32-bit byte
    // offsets.  Past that, AUTO takes the grid kernel (64-bit addressing);
    // an explicit block or hybrid request, or a batch, cannot be served.
    if (P->mode == GSOFA_MODE_BLOCK && P->nnz >= kMaxBlockNnz) {
        if (P->opt.mode == GSOFA_MODE_AUTO && P->opt.top_sources <= 0 && n_mats == 1) {
            P->mode = GSOFA_MODE_GRID;
        } else {
            (void)hipStreamDestroy(P->stream);
            delete P;
            return GSOFA_ERR_TOO_LARGE;
        }
    }
    P->orphan_words = (uint32_t)(((uint64_t)n + 31) / 32);
    if (P->mode == GSOFA_MODE_BLOCK && P->opt.top_sources > 0 && n_mats == 1 &&
        P->opt.n_virtual_gpu == 1 && n > 0) {
        // hybrid: the top sources as whole grid-mode groups of up to 8192
        const int64_t K = std::min<int64_t>(P->opt.top_sources, n);
        P->hybrid = true;
        P->hC = (int32_t)std::min<int64_t>(K, 8192);
        P->h_g_hi = (int32_t)(((int64_t)n + P->hC - 1) / P->hC);
        P->h_g_lo = std::max<int32_t>(0, P->h_g_hi - (int32_t)((K + P->hC - 1) / P->hC));
        P->s_top = (uint32_t)P->h_g_lo * (uint32_t)P->hC;
        P->hQ = P->opt.queue_capacity > 0 ? (uint32_t)std::min<int64_t>(P->opt.queue_capacity, 1ll << 28)
                                          : default_queue(n, P->hC);
        P->hlab_words = label_words(n, 1, P->hC, 0, 0);
        avail -= 4 * (int64_t)P->hlab_words + 4 * (int64_t)P->hC * n + 4ll * P->hC * P->orphan_words +
                 16 * (int64_t)P->hQ;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, P->device) != hipSuccess) {
        (void)hipGetLastError();
        (void)hipStreamDestroy(P->stream);
        delete P;
        return GSOFA_ERR_HIP;
    }
    int e = 0;
    std::vector<int32_t> rpT, ciT;
    if (P->mode == GSOFA_MODE_REACH) {
        int bits = 0;
        rc = choose_bits(P->opt, n_mats, (uint64_t)n, &bits);
        P->S = (uint32_t)bits;
        P->rbits = bits;
        P->rthreads = reach_threads(P->opt);
        if (!rc) {
            // the A pass's chunks, deepest first, each followed by the same
            // sources' chunk of the A^T pass: one launch runs both
            build_chunks(P);
            std::vector<BlkChunk> both;
            both.reserve(2 * P->h_chunks.size());
            for (const BlkChunk& ch : P->h_chunks) {
                both.push_back(ch);
                BlkChunk t = ch;
                t.stride |= rch_pass_bit();
                both.push_back(t);
            }
            P->h_chunks.swap(both);
            P->share_chunk_off.assign(2, 0);
            P->share_chunk_off[1] = (uint32_t)P->h_chunks.size();
            transpose_rows(n, P->h_rp.data(), ci.data(), rpT, ciT);
            const int64_t tbytes = 4 * (int64_t)(n + 1) + 4 * P->nnz + 16 * (int64_t)n_mats;
            rc = reach_grid(P->opt, bits, std::max<uint64_t>(P->h_chunks.size(), 1), P->W,
                            std::max<int64_t>(avail - tbytes, 1 << 20), &P->nblocks);
        }
        if (rc) {
            (void)hipStreamDestroy(P->stream);
            delete P;
            return rc;
        }
        P->C = (int32_t)(P->nblocks * P->S);
        const size_t nbw = (size_t)P->nblocks * (size_t)P->W;
        e |= dalloc((uint64_t**)&P->d_vis, nbw * (size_t)(rch_visit_bytes(bits) / 8));
        e |= dalloc((uint32_t**)&P->d_rmask, 2 * nbw * (size_t)(bits / 32));
        e |= dalloc(&P->d_qgv, 2 * nbw);
        e |= dalloc(&P->d_bstate, (size_t)P->nblocks);
        e |= dalloc(&P->d_mat_base, (size_t)n_mats + 1);
        e |= dalloc(&P->d_chunks, std::max<size_t>(P->h_chunks.size(), 1));
        e |= dalloc(&P->d_row_fk, (size_t)n);
        e |= dalloc(&P->d_mat_te, (size_t)n_mats);
        e |= dalloc(&P->d_mat_reach, (size_t)n_mats);
        e |= dalloc(&P->d_rpT, (size_t)n + 1);
        e |= dalloc(&P->d_ciT, (size_t)P->nnz);
    } else if (P->mode == GSOFA_MODE_BLOCK) {
        int S = 0;
        const uint64_t nsrc = P->hybrid ? P->s_top : share_count(n, P->opt.n_virtual_gpu, 0);
        rc = choose_slots(P->opt, n_mats, nsrc, &S);
        P->S = (uint32_t)S;
        if (!rc) build_chunks(P);
        if (!rc) rc = size_block_mode(P, std::max<int64_t>(avail, 1 << 20));
        if (rc) {
            (void)hipStreamDestroy(P->stream);
            delete P;
            return rc;
        }
        const size_t slots = (size_t)P->nblocks * (size_t)P->S;
        e |= dalloc(&P->d_lab, slots * (size_t)P->W);
        e |= dalloc(&P->d_mask, (size_t)P->nblocks * 2 * (size_t)P->W);
        e |= dalloc(&P->d_qgv, (size_t)P->nblocks * 2 * (size_t)P->W);
        e |= dalloc(&P->d_bstate, 2 * (size_t)P->nblocks);
        e |= dalloc(&P->d_mat_base, (size_t)n_mats + 1);
        e |= dalloc(&P->d_chunks, std::max<size_t>(P->h_chunks.size(), 1));
        e |= dalloc(&P->d_row_te, (size_t)n);
        e |= dalloc(&P->d_row_reach, (size_t)n);
        e |= dalloc(&P->d_row_fk, (size_t)n);
        P->lab_words = slots * (size_t)P->W;
        if (P->hybrid) {
            e |= dalloc(&P->d_hlab, (size_t)P->hlab_words);
            e |= dalloc(&P->d_mark, (size_t)P->hC * (size_t)n);
            e |= dalloc(&P->d_orphan, (size_t)P->hC * P->orphan_words);
            e |= dalloc(&P->d_q0, (size_t)P->hQ);
            e |= dalloc(&P->d_q1, (size_t)P->hQ);
        }
    } else {
        Sizing sz;
        rc = choose_sizing(A, &P->opt, P->nnz, std::max<int64_t>(avail, 1 << 20), &sz);
        if (rc) {
            (void)hipStreamDestroy(P->stream);
            delete P;
            return rc;
        }
        P->C = sz.C;
        P->Q = sz.Q;
        P->lab_words = sz.lab_words;
        e |= dalloc(&P->d_lab, (size_t)P->lab_words);
        e |= dalloc(&P->d_mark, (size_t)P->C * (size_t)n);
        e |= dalloc(&P->d_orphan, (size_t)P->C * P->orphan_words);
        e |= dalloc(&P->d_q0, (size_t)P->Q);
        e |= dalloc(&P->d_q1, (size_t)P->Q);
    }
    P->fill_cap = P->opt.output_capacity > 0 ? (uint64_t)P->opt.output_capacity : default_fill_cap(P->nnz);
    if ((P->opt.flags & GSOFA_FLAG_SPILL) && P->opt.output_capacity <= 0)  // one share's fills at a time
        P->fill_cap = std::max<uint64_t>(P->fill_cap / (uint64_t)std::max(1, P->opt.n_virtual_gpu), 1u << 20);
    if (P->bounded) {
        P->fill_cap = P->mode == GSOFA_MODE_REACH ? rch_fill_room(need_fill) : std::max<uint64_t>(need_fill, 1);
        P->fill_need = need_fill;
        P->Lcap = (int64_t)need_L;
        P->Ucap = (int64_t)need_U;
        e |= dalloc(&P->d_Lcol, (size_t)need_L);
        e |= dalloc(&P->d_Ltmp, (size_t)need_L);
        e |= dalloc(&P->d_Ucol, (size_t)need_U);
        e |= dalloc(&P->d_Utmp, (size_t)need_U);
    }
    e |= dalloc(&P->d_rp, (size_t)n + 1);
    e |= dalloc(&P->d_ci, (size_t)P->nnz);
    e |= dalloc(&P->d_fills, (size_t)P->fill_cap);
    e |= dalloc(&P->d_ctrl, 1);
    e |= dalloc(&P->d_cntL, (size_t)n);
    e |= dalloc(&P->d_cntU, (size_t)n);
    e |= dalloc(&P->d_Lrp, (size_t)n + 1);
    e |= dalloc(&P->d_Urp, (size_t)n + 1);
    e |= dalloc(&P->d_rowfill, (size_t)n);
    P->scan_words = scan_tmp_words(n) + 2 * (size_t)n + 8;
    e |= dalloc(&P->d_scan, P->scan_words);
    if (e) {
        release_device(P);
        (void)hipStreamDestroy(P->stream);
        delete P;
        return GSOFA_ERR_ALLOC;
    }
    bool ok = hipMemcpyAsync(P->d_rp, P->h_rp.data(), sizeof(int32_t) * (n + 1), hipMemcpyHostToDevice,
                             P->stream) == hipSuccess;
    if (P->nnz)
        ok = ok && hipMemcpyAsync(P->d_ci, ci.data(), sizeof(int32_t) * P->nnz, hipMemcpyHostToDevice,
                                  P->stream) == hipSuccess;
    if (P->mode == GSOFA_MODE_GRID || P->hybrid) {
        const size_t Cm = P->hybrid ? (size_t)P->hC : (size_t)P->C;
        ok = ok && hipMemsetAsync(P->d_mark, 0, sizeof(uint32_t) * Cm * n, P->stream) == hipSuccess;
        ok = ok && hipMemsetAsync(P->d_orphan, 0, sizeof(uint32_t) * Cm * P->orphan_words,
                                  P->stream) == hipSuccess;
    }
    if (P->mode == GSOFA_MODE_BLOCK) {
        ok = ok && hipMemsetAsync(P->d_mask, 0, sizeof(uint32_t) * (size_t)P->nblocks * 2 * P->W,
                                  P->stream) == hipSuccess;
    }
    if (P->mode == GSOFA_MODE_REACH) {
        const size_t nbw = (size_t)P->nblocks * (size_t)P->W;
        const size_t vb = nbw * (size_t)rch_visit_bytes(P->rbits), mb = 2 * nbw * (size_t)(P->rbits / 8);
        ok = ok && hipMemsetAsync(P->d_vis, 0, vb, P->stream) == hipSuccess;
        ok = ok && hipMemsetAsync(P->d_rmask, 0, mb, P->stream) == hipSuccess;
        ok = ok && hipMemsetAsync(P->d_bstate, 0, sizeof(uint32_t) * P->nblocks, P->stream) == hipSuccess;
        ok = ok && hipMemcpyAsync(P->d_rpT, rpT.data(), sizeof(int32_t) * (n + 1), hipMemcpyHostToDevice,
                                  P->stream) == hipSuccess;
        if (P->nnz)
            ok = ok && hipMemcpyAsync(P->d_ciT, ciT.data(), sizeof(int32_t) * P->nnz, hipMemcpyHostToDevice,
                                      P->stream) == hipSuccess;
        ok = ok && hipMemcpyAsync(P->d_mat_base, P->h_mat_base.data(), sizeof(uint32_t) * (n_mats + 1),
                                  hipMemcpyHostToDevice, P->stream) == hipSuccess;
        if (!P->h_chunks.empty())
            ok = ok && hipMemcpyAsync(P->d_chunks, P->h_chunks.data(), sizeof(BlkChunk) * P->h_chunks.size(),
                                      hipMemcpyHostToDevice, P->stream) == hipSuccess;
    }
    std::vector<uint32_t> bst;
    if (P->mode == GSOFA_MODE_BLOCK) {
        ok = ok && hipMemsetAsync(P->d_lab, 0xFF, sizeof(uint32_t) * P->lab_words, P->stream) == hipSuccess;
        bst.resize(2 * (size_t)P->nblocks);
        for (uint32_t b = 0; b < P->nblocks; b++) {
            bst[2 * b] = 0xFFFFFFFFu - P->W;
            bst[2 * b + 1] = 0;
        }
        ok = ok && hipMemcpyAsync(P->d_bstate, bst.data(), sizeof(uint32_t) * bst.size(),
                                  hipMemcpyHostToDevice, P->stream) == hipSuccess;
        ok = ok && hipMemcpyAsync(P->d_mat_base, P->h_mat_base.data(), sizeof(uint32_t) * (n_mats + 1),
                                  hipMemcpyHostToDevice, P->stream) == hipSuccess;
        if (!P->h_chunks.empty())
            ok = ok && hipMemcpyAsync(P->d_chunks, P->h_chunks.data(), sizeof(BlkChunk) * P->h_chunks.size(),
                                      hipMemcpyHostToDevice, P->stream) == hipSuccess;
    }
    ok = ok && hipStreamSynchronize(P->stream) == hipSuccess;
    if (!ok) {
        (void)hipGetLastError();
        release_device(P);
        (void)hipStreamDestroy(P->stream);
        delete P;
        return GSOFA_ERR_HIP;
    }
    P->off = 0xFFFFFFFFu - (uint32_t)std::max(n, 1);
    P->since = 0;
    P->hw = 0;
    P->epoch = 0;

    if (P->mode == GSOFA_MODE_GRID || P->hybrid) {
        // persistent grid: as many blocks as can be co-resident
        int occ = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, traverse_kernel_ptr(), traverse_block_size(),
                                                         0) != hipSuccess ||
            occ <= 0) {
            (void)hipGetLastError();
            release_device(P);
            (void)hipStreamDestroy(P->stream);
            delete P;
            return GSOFA_ERR_HIP;
        }
        uint32_t maxblocks = (uint32_t)occ * (uint32_t)prop.multiProcessorCount;
        const uint32_t gb = P->opt.grid_blocks > 0 && !P->hybrid
                                ? std::min<uint32_t>((uint32_t)P->opt.grid_blocks, maxblocks)
                                : std::min<uint32_t>(maxblocks, (uint32_t)prop.multiProcessorCount);
        if (P->hybrid) P->h_nblocks = gb;
        else P->nblocks = gb;
    }
    int rate_khz = 0;
    if (hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, P->device) != hipSuccess ||
        rate_khz <= 0) {
        (void)hipGetLastError();
        rate_khz = 100000;
    }
    P->spin_ticks = (uint64_t)rate_khz * 20000ull;  // 20 s
    *out = P;
    return GSOFA_OK;
}

extern "C" {

int gsofa_plan_create(const gsofa_matrix* A, const gsofa_options* opt, gsofa_plan** out) {
    if (!A) return GSOFA_ERR_INVALID;
    const uint32_t mb[2] = {0u, (uint32_t)std::max(A->n, 0)};
    return plan_create_impl(A, mb, 1, opt, out);
}

int gsofa_batch_create(const gsofa_matrix* mats, int32_t count, const gsofa_options* opt,
                       gsofa_plan** out) {
    if (!out || count < 1 || !mats) return GSOFA_ERR_INVALID;
    *out = nullptr;
    int64_t n = 0, nnz = 0;
    for (int32_t b = 0; b < count; b++) {
        int rc = validate(&mats[b]);
        if (rc) return rc;
        n += mats[b].n;
        nnz += mats[b].nnz;
    }
    if (n > kMaxN || nnz >= (int64_t)INT32_MAX) return GSOFA_ERR_TOO_LARGE;
    // block-diagonal global matrix: matrix b's vertices are shifted by base_b
    std::vector<uint32_t> mb((size_t)count + 1, 0);
    std::vector<int32_t> cb((size_t)n), ce((size_t)n), cs((size_t)nnz);
    int64_t v = 0, k = 0;
    for (int32_t b = 0; b < count; b++) {
        const gsofa_matrix& M = mats[b];
        mb[b] = (uint32_t)v;
        for (int32_t i = 0; i < M.n; i++) {
            cb[v + i] = (int32_t)k;
            for (int64_t e = M.colbeg[i]; e < M.colend[i]; e++) cs[k++] = M.csr[e] + (int32_t)v;
            ce[v + i] = (int32_t)k;
        }
        v += M.n;
    }
    mb[count] = (uint32_t)v;
    gsofa_matrix G = {(int32_t)n, nnz, cb.data(), ce.data(), cs.data()};
    return plan_create_impl(&G, mb.data(), count, opt, out);
}

int32_t gsofa_plan_matrix_count(const gsofa_plan* P) { return P ? P->n_mats : 0; }

}  // extern "C"

// Grid-mode launch parameters for every group of share gid.
static TravParams grid_params(const gsofa_plan* P, int32_t gid) {
    const int32_t n = P->n;
    const int32_t N = std::max(1, P->opt.n_virtual_gpu);
    const uint64_t cnt = share_count(n, N, gid);
    TravParams tp{};
    tp.rp = P->d_rp;
    tp.ci = P->d_ci;
    tp.n = n;
    tp.n_gpu = N;
    tp.gpu_id = gid;
    tp.chunk = P->C;
    tp.n_groups = P->C > 0 ? (int32_t)((cnt + P->C - 1) / P->C) : 0;
    tp.first_group = 0;
    tp.lab = P->d_lab;
    tp.lab_words = P->lab_words;
    tp.mark = P->d_mark;
    tp.qbuf0 = P->d_q0;
    tp.qbuf1 = P->d_q1;
    tp.qcap = P->Q;
    tp.orphan_words = P->orphan_words;
    tp.orphan = P->d_orphan;
    tp.fills = P->d_fills;
    tp.fill_cap = P->fill_cap;
    tp.ctrl = P->d_ctrl;
    tp.nblocks = P->nblocks;
    tp.block_threshold = P->opt.wave_block_threshold;
    tp.off_start = P->off;
    tp.groups_since_init = P->since;
    tp.lab_hw = P->hw;
    tp.stamp_base = P->epoch * (uint32_t)n;
    tp.loops = std::min(0xFFFFFFFFu / (uint32_t)std::max(n, 1), P->span_cap);
    tp.spin_ticks = P->spin_ticks;
    return tp;
}

// One share's traversal, bracketed by events 2+2si / 3+2si.
static int run_share(gsofa_plan* P, int32_t gid, size_t si) {
    const int32_t n = P->n;
    // grid mode: stamps must increase per marker slot: move to a fresh
    // stamp range (or clear the markers when the 32-bit range is used up)
    if ((P->mode == GSOFA_MODE_GRID || P->hybrid) && !P->marks_clean) {
        if ((uint64_t)(P->epoch + 2) * (uint64_t)n >= 0xFFFFFFFFull) {
            const size_t Cm = P->hybrid ? (size_t)P->hC : (size_t)P->C;
            HIPCHK(hipMemsetAsync(P->d_mark, 0, sizeof(uint32_t) * Cm * n, P->stream));
            P->epoch = 0;
        } else {
            P->epoch++;
        }
    }
    P->marks_clean = false;
    if (P->mode == GSOFA_MODE_REACH) {
        RchParams rp{};
        rp.rp = P->d_rp;
        rp.ci = P->d_ci;
        rp.rpT = P->d_rpT;
        rp.ciT = P->d_ciT;
        rp.W = P->W;
        rp.chunks = P->d_chunks;
        rp.n_chunks = (uint32_t)P->h_chunks.size();
        rp.mat_base = P->d_mat_base;
        rp.nblocks = P->nblocks;
        rp.vis = P->d_vis;
        rp.nmask = P->d_rmask;
        rp.q = P->d_qgv;
        rp.fills = P->d_fills;
        rp.fill_cap = P->structure ? P->fill_cap : 0;  // counts only: per-row tallies, no fill stream
        rp.fill_blk = rch_fill_block(P->fill_need ? P->fill_need : P->fill_cap, P->nblocks);
        rp.ctrl = P->d_ctrl;
        rp.bstate = P->d_bstate;
        rp.row_fills = P->d_row_fk;
        rp.mat_te = P->d_mat_te;
        rp.mat_reach = P->d_mat_reach;
        rp.loops = std::min(0xFFFFFFFFu, P->span_cap);
        HIPCHK(launch_fill(reinterpret_cast<uint32_t*>(P->d_mat_te), 2 * (uint64_t)P->n_mats, 0u, P->stream));
        HIPCHK(launch_fill(reinterpret_cast<uint32_t*>(P->d_mat_reach), 2 * (uint64_t)P->n_mats, 0u, P->stream));
        if (!P->structure)  // both passes add their fills per row
            HIPCHK(launch_fill(reinterpret_cast<uint32_t*>(P->d_row_fk), 2 * (uint64_t)n, 0u, P->stream));
        HIPCHK(launch_fill(&P->d_ctrl->src_next, 1, 0u, P->stream));
        if (P->structure)  // reserved entries left unused must read as ~0
            HIPCHK(launch_fill(reinterpret_cast<uint32_t*>(P->d_fills), 2 * P->fill_cap, 0xFFFFFFFFu, P->stream));
        if (!P->capturing) HIPCHK(hipEventRecord(P->ev[2 + 2 * si], P->stream));
        if (rp.n_chunks > 0) HIPCHK(launch_reach(rp, P->rbits, P->rthreads, P->stream));
        if (!P->capturing) HIPCHK(hipEventRecord(P->ev[3 + 2 * si], P->stream));
        return GSOFA_OK;
    }
    if (P->mode == GSOFA_MODE_BLOCK) {
        BlkParams bp{};
        bp.rp = P->d_rp;
        bp.ci = P->d_ci;
        bp.n = n;
        bp.W = P->W;
        bp.chunks = P->d_chunks + P->share_chunk_off[gid];
        bp.n_chunks = P->share_chunk_off[gid + 1] - P->share_chunk_off[gid];
        bp.mat_base = P->d_mat_base;
        bp.row_te = P->d_row_te;
        bp.row_reach = P->d_row_reach;
        bp.row_fills = P->d_row_fk;
        bp.nblocks = P->nblocks;
        bp.lab = P->d_lab;
        bp.mask = P->d_mask;
        bp.qg = P->d_qgv;
        bp.fills = P->d_fills;
        bp.fill_cap = P->structure ? P->fill_cap : 0;  // counts only: no fill stream
        bp.loops = std::min(0xFFFFFFFFu / P->W, P->span_cap);
        bp.ctrl = P->d_ctrl;
        bp.bstate = P->d_bstate;
        HIPCHK(launch_fill(&P->d_ctrl->src_next, 1, 0u, P->stream));
        if (!P->capturing) HIPCHK(hipEventRecord(P->ev[2 + 2 * si], P->stream));
        if (P->hybrid && P->h_g_hi > P->h_g_lo) {  // the top groups, whole device each
            TravParams tp = grid_params(P, gid);
            tp.chunk = P->hC;
            tp.first_group = P->h_g_lo;
            tp.n_groups = P->h_g_hi - P->h_g_lo;
            tp.lab = P->d_hlab;
            tp.lab_words = P->hlab_words;
            tp.qcap = P->hQ;
            tp.nblocks = P->h_nblocks;
            HIPCHK(launch_traverse(tp, P->stream, P->coop));
            advance_groups(P, gid, P->hC, (uint64_t)P->h_g_lo, (uint64_t)P->h_g_hi);
        }
        if (bp.n_chunks > 0) HIPCHK(launch_traverse_blk(bp, P->S, P->blk_lds, P->stream));
        if (!P->capturing) HIPCHK(hipEventRecord(P->ev[3 + 2 * si], P->stream));
        return GSOFA_OK;
    }
    TravParams tp = grid_params(P, gid);
    HIPCHK(hipEventRecord(P->ev[2 + 2 * si], P->stream));
    if (tp.n_groups > 0) HIPCHK(launch_traverse(tp, P->stream, P->coop));
    HIPCHK(hipEventRecord(P->ev[3 + 2 * si], P->stream));
    advance_state(P, gid);
    return GSOFA_OK;
}

// Sorted L/U rows (d_Lrp/d_Lcol, d_Urp/d_Ucol) of the rows of share gpu_id of
// N (-1: every row) from the fill stream's n_fills entries.  With on_device,
// the kernels read the stream's length from the traversal's counter and the
// columns reserved at plan creation hold any result: no host round trip (the
// totals are read when needed, read_totals).
static int build_structure(gsofa_plan* P, uint64_t n_fills, int32_t N, int32_t gpu_id, bool on_device = false) {
    const int32_t n = P->n;
    FinalParams f{};
    f.n = n;
    f.rp = P->d_rp;
    f.ci = P->d_ci;
    f.fills = P->d_fills;
    f.n_fills = on_device ? P->fill_cap : n_fills;
    f.n_fills_dev = on_device ? &P->d_ctrl->fill_pos : nullptr;
    f.n_gpu = N;
    f.gpu_id = gpu_id;
    f.L_rowptr = P->d_Lrp;
    f.U_rowptr = P->d_Urp;
    f.row_fill = P->d_rowfill;
    f.cntL = P->d_cntL;
    f.cntU = P->d_cntU;
    f.scan_tmp = P->d_scan;
    HIPCHK(launch_count(f, P->stream));
    HIPCHK(exclusive_scan_u32_to_i64(P->d_cntL, P->d_Lrp, n, P->d_scan, P->stream));
    HIPCHK(exclusive_scan_u32_to_i64(P->d_cntU, P->d_Urp, n, P->d_scan, P->stream));
    if (on_device) {
        P->totals_known = false;
    } else {
        int64_t tot[2] = {0, 0};
        HIPCHK_OUT(hipMemcpyAsync(&tot[0], P->d_Lrp + n, sizeof(int64_t), hipMemcpyDeviceToHost, P->stream), P->stream);
        HIPCHK_OUT(hipMemcpyAsync(&tot[1], P->d_Urp + n, sizeof(int64_t), hipMemcpyDeviceToHost, P->stream), P->stream);
        HIPCHK_OUT(hipStreamSynchronize(P->stream), P->stream);
        P->nnzL = tot[0];
        P->nnzU = tot[1];
        P->totals_known = true;
        int rc = ensure_cols(P, P->nnzL, P->nnzU);
        if (rc) return rc;
    }
    f.L_col = P->d_Ltmp;
    f.U_col = P->d_Utmp;
    f.L_cap = P->Lcap;
    f.U_cap = P->Ucap;
    HIPCHK(launch_scatter(f, P->stream));
    HIPCHK(launch_sort_segments(P->d_Lrp, P->d_Ltmp, P->d_Lcol, n, P->Lcap, P->stream));
    HIPCHK(launch_sort_segments(P->d_Urp, P->d_Utmp, P->d_Ucol, n, P->Ucap, P->stream));
    return GSOFA_OK;
}

// nnzL / nnzU of the last structure pass, read from the row pointers' ends.
static int read_totals(gsofa_plan* P) {
    if (P->totals_known || P->n == 0) return GSOFA_OK;
    int64_t tot[2] = {0, 0};
    HIPCHK_OUT(hipMemcpyAsync(&tot[0], P->d_Lrp + P->n, sizeof(int64_t), hipMemcpyDeviceToHost, P->stream), P->stream);
    HIPCHK_OUT(hipMemcpyAsync(&tot[1], P->d_Urp + P->n, sizeof(int64_t), hipMemcpyDeviceToHost, P->stream), P->stream);
    HIPCHK_OUT(hipStreamSynchronize(P->stream), P->stream);
    if (tot[0] > P->Lcap || tot[1] > P->Ucap) return GSOFA_ERR_STATE;  // cannot happen: the bounds hold
    P->nnzL = tot[0];
    P->nnzU = tot[1];
    P->totals_known = true;
    return GSOFA_OK;
}

// Supernodes from d_Lrp/d_Lcol/d_Urp, or (pass != null) from d_Urp and the
// per-row pass-through flags.
static int build_supernodes(gsofa_plan* P, const uint32_t* pass) {
    const int32_t n = P->n;
    if (!P->d_supno || P->sn_cap < (size_t)n) {
        dfree(P->d_supno);
        dfree(P->d_sstart);
        dfree(P->d_nsuper);
        dfree(P->d_sn64);
        dfree(P->d_snflag);
        int e = 0;
        e |= dalloc(&P->d_supno, (size_t)n);
        e |= dalloc(&P->d_sstart, (size_t)n + 1);
        e |= dalloc(&P->d_nsuper, 1);
        e |= dalloc(&P->d_sn64, scan_tmp_words(n) + 2 * (size_t)n + 8);
        e |= dalloc(&P->d_snflag, (size_t)n);
        if (e) return GSOFA_ERR_ALLOC;
        P->sn_cap = (size_t)n;
    }
    SuperParams sp{};
    sp.n = n;
    sp.L_rowptr = P->d_Lrp;
    sp.L_col = P->d_Lcol;
    sp.U_rowptr = P->d_Urp;
    sp.max_size = P->opt.max_supernode_size;
    sp.supno = P->d_supno;
    sp.super_start = P->d_sstart;
    sp.n_super = P->d_nsuper;
    sp.tmp64 = P->d_sn64;
    sp.flag = P->d_snflag;
    sp.pass = pass;
    HIPCHK(launch_supernodes(sp, P->stream));
    return GSOFA_OK;
}

// GSOFA_FLAG_SPILL: each share runs to completion (its fill stream regrown
// and the share rerun from the counters it started with, if it overflows),
// then its sorted rows go to host memory and the fill stream starts over.
// Per-row lengths and pass-through flags stay on the device for the
// supernode pass; the rows themselves are merged at fetch.
static int execute_spill(gsofa_plan* P, const std::vector<int32_t>& shares) {
    const int32_t n = P->n;
    const int32_t N = std::max(1, P->opt.n_virtual_gpu);
    if (!P->d_sp_lcnt) {
        int e = 0;
        e |= dalloc(&P->d_sp_lcnt, (size_t)n);
        e |= dalloc(&P->d_sp_ucnt, (size_t)n);
        e |= dalloc(&P->d_sp_pass, (size_t)n);
        if (e) return GSOFA_ERR_ALLOC;
    }
    HIPCHK(launch_fill(P->d_sp_lcnt, (uint64_t)n, 0u, P->stream));
    HIPCHK(launch_fill(P->d_sp_ucnt, (uint64_t)n, 0u, P->stream));
    HIPCHK(launch_fill(P->d_sp_pass, (uint64_t)n, 0u, P->stream));
    P->sp.assign((size_t)N, SpillShare{});
    P->sp_rowfill.assign((size_t)n, -1);
    std::vector<int64_t> rowfill((size_t)n), lrp((size_t)n + 1), urp((size_t)n + 1);
    Ctrl snap;
    memset(&snap, 0, sizeof(snap));
    HIPCHK(hipEventRecord(P->ev[0], P->stream));
    int64_t nnzL = 0, nnzU = 0;
    for (size_t si = 0; si < shares.size(); si++) {
        const int32_t gid = shares[si];
        for (int attempt = 0; attempt < 3; attempt++) {
            if (attempt > 0) P->passes++;  // a share reran with a larger stream
            HIPCHK(hipMemcpyAsync(P->d_ctrl, &snap, sizeof(Ctrl), hipMemcpyHostToDevice, P->stream));
            int rc = run_share(P, gid, si);
            if (rc) return rc;
            HIPCHK_OUT(hipMemcpyAsync(&P->stats, P->d_ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, P->stream), P->stream);
            HIPCHK_OUT(hipStreamSynchronize(P->stream), P->stream);
            if (P->stats.abort) return GSOFA_ERR_DEVICE_TIMEOUT;
            if (P->stats.fill_pos <= P->fill_cap) break;
            dfree(P->d_fills);
            P->fill_cap = P->stats.fill_pos + P->stats.fill_pos / 16 + 1024;
            if (dalloc(&P->d_fills, (size_t)P->fill_cap)) return GSOFA_ERR_ALLOC;
        }
        if (P->stats.fill_pos > P->fill_cap) return GSOFA_ERR_ALLOC;
        int rc = build_structure(P, P->stats.fill_pos, N, gid);
        if (rc) return rc;
        HIPCHK(spill_rows(P->d_Lrp, P->d_Lcol, P->d_Urp, n, N, gid, P->d_sp_lcnt, P->d_sp_ucnt, P->d_sp_pass,
                          P->stream));
        SpillShare& h = P->sp[(size_t)gid];
        h.Lcol.resize((size_t)std::max<int64_t>(P->nnzL, 1));
        h.Ucol.resize((size_t)std::max<int64_t>(P->nnzU, 1));
        HIPCHK_OUT(hipMemcpyAsync(lrp.data(), P->d_Lrp, sizeof(int64_t) * ((size_t)n + 1), hipMemcpyDeviceToHost,
                                  P->stream), P->stream);
        HIPCHK_OUT(hipMemcpyAsync(urp.data(), P->d_Urp, sizeof(int64_t) * ((size_t)n + 1), hipMemcpyDeviceToHost,
                                  P->stream), P->stream);
        if (P->nnzL)
            HIPCHK_OUT(hipMemcpyAsync(h.Lcol.data(), P->d_Lcol, sizeof(int32_t) * (size_t)P->nnzL,
                                      hipMemcpyDeviceToHost, P->stream), P->stream);
        if (P->nnzU)
            HIPCHK_OUT(hipMemcpyAsync(h.Ucol.data(), P->d_Ucol, sizeof(int32_t) * (size_t)P->nnzU,
                                      hipMemcpyDeviceToHost, P->stream), P->stream);
        HIPCHK_OUT(hipMemcpyAsync(rowfill.data(), P->d_rowfill, sizeof(int64_t) * (size_t)n, hipMemcpyDeviceToHost,
                                  P->stream), P->stream);
        HIPCHK_OUT(hipStreamSynchronize(P->stream), P->stream);
        // the share's rows are the only non-empty ones: keep their offsets
        h.Loff.assign(1, 0);
        h.Uoff.assign(1, 0);
        for (int32_t i = gid; i < n; i += N) {
            P->sp_rowfill[(size_t)i] = rowfill[(size_t)i];
            h.Loff.push_back(lrp[(size_t)i + 1]);
            h.Uoff.push_back(urp[(size_t)i + 1]);
        }
        nnzL += P->nnzL;
        nnzU += P->nnzU;
        snap = P->stats;  // the next share continues the counters with an empty fill stream
        snap.fill_pos = 0;
    }
    P->nnzL = nnzL;
    P->nnzU = nnzU;
    if (P->supernodes) {
        // the whole matrix's row pointers from the kept lengths, then the pass
        HIPCHK(exclusive_scan_u32_to_i64(P->d_sp_lcnt, P->d_Lrp, n, P->d_scan, P->stream));
        HIPCHK(exclusive_scan_u32_to_i64(P->d_sp_ucnt, P->d_Urp, n, P->d_scan, P->stream));
        int rc = build_supernodes(P, P->d_sp_pass);
        if (rc) return rc;
    }
    P->traverse_ms = 0;
    P->share_min = 1e300;
    P->share_max = 0;
    for (size_t si = 0; si < shares.size(); si++) {
        double t = ev_ms(P->ev[2 + 2 * si], P->ev[3 + 2 * si]);
        P->traverse_ms += t;
        P->share_min = std::min(P->share_min, t);
        P->share_max = std::max(P->share_max, t);
    }
    HIPCHK(hipEventRecord(P->ev[1], P->stream));
    HIPCHK_OUT(hipStreamSynchronize(P->stream), P->stream);
    P->ms = ev_ms(P->ev[0], P->ev[1]);
    P->executed = true;
    return GSOFA_OK;
}

// Reach mode, batches: per-matrix TE and |R| of the last execute to the host
// (queued on the plan's stream; valid after the next synchronisation).
static hipError_t fetch_mat_totals(gsofa_plan* P) {
    if (P->mode != GSOFA_MODE_REACH || P->n_mats <= 1) return hipSuccess;
    P->h_mat_te.resize((size_t)P->n_mats);
    P->h_mat_reach.resize((size_t)P->n_mats);
    hipError_t e = hipMemcpyAsync(P->h_mat_te.data(), P->d_mat_te, sizeof(unsigned long long) * P->n_mats,
                                  hipMemcpyDeviceToHost, P->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(P->h_mat_reach.data(), P->d_mat_reach, sizeof(unsigned long long) * P->n_mats,
                           hipMemcpyDeviceToHost, P->stream);
    return e;
}

// Traversal time per share and in total, and the whole execute's device time,
// from the events of the last execute (ev[0] .. ev[1]; share si: 2+2si .. 3+2si).
static void finish_timing(gsofa_plan* P, size_t nshares) {
    P->traverse_ms = 0;
    P->share_min = 1e300;
    P->share_max = 0;
    for (size_t si = 0; si < nshares; si++) {
        const double t = ev_ms(P->ev[2 + 2 * si], P->ev[3 + 2 * si]);
        P->traverse_ms += t;
        P->share_min = std::min(P->share_min, t);
        P->share_max = std::max(P->share_max, t);
    }
    P->ms = ev_ms(P->ev[0], P->ev[1]);
}

// The device work of one bounded execute, queued on the plan's stream:
// counters cleared, every share's traversal, the structure and supernode
// passes, and the counters copied to `stats` (host memory).  Nothing here
// waits for the device, allocates or sets memory with a memset node, so the
// sequence can be captured in a graph.
static int queue_bounded(gsofa_plan* P, const std::vector<int32_t>& shares, int32_t N, bool all_rows,
                         Ctrl* stats) {
    HIPCHK(launch_fill(reinterpret_cast<uint32_t*>(P->d_ctrl), sizeof(Ctrl) / 4, 0u, P->stream));
    if (!P->capturing) HIPCHK(hipEventRecord(P->ev[0], P->stream));
    for (size_t si = 0; si < shares.size(); si++) {
        int rc = run_share(P, shares[si], si);
        if (rc) return rc;
    }
    if (P->structure) {
        int rc = build_structure(P, 0, N, all_rows ? -1 : P->opt.virtual_gpu_id, true);
        if (!rc && P->supernodes) rc = build_supernodes(P, nullptr);
        if (rc) return rc;
    }
    HIPCHK_OUT(hipMemcpyAsync(stats, P->d_ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, P->stream), P->stream);
    if (!P->capturing) HIPCHK(hipEventRecord(P->ev[1], P->stream));
    return GSOFA_OK;
}

// GSOFA_FLAG_GRAPH: capture queue_bounded once (after an eager execute has
// allocated every lazily sized buffer).  A refused capture leaves the plan
// executing eagerly.
static void capture_graph(gsofa_plan* P, const std::vector<int32_t>& shares, int32_t N, bool all_rows) {
    if (!P->h_ctrl && hipHostMalloc((void**)&P->h_ctrl, sizeof(Ctrl), hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        P->h_ctrl = nullptr;
        P->graph_failed = true;
        return;
    }
    if (hipStreamBeginCapture(P->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        (void)hipGetLastError();
        P->graph_failed = true;
        return;
    }
    P->capturing = true;
    const int rc = queue_bounded(P, shares, N, all_rows, P->h_ctrl);
    P->capturing = false;
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(P->stream, &g);
    if (rc != GSOFA_OK || e != hipSuccess || !g ||
        hipGraphInstantiate(&P->graph_exec, g, nullptr, nullptr, 0) != hipSuccess) {
        (void)hipGetLastError();
        if (g) (void)hipGraphDestroy(g);
        P->graph_exec = nullptr;
        P->graph_failed = true;
        return;
    }
    P->graph = g;
}

extern "C" {

int gsofa_plan_execute(gsofa_plan* P) {
    if (!P) return GSOFA_ERR_INVALID;
    HIPCHK(hipSetDevice(P->device));
    const int32_t n = P->n;
    const int32_t N = std::max(1, P->opt.n_virtual_gpu);
    std::vector<int32_t> shares;
    if (P->opt.virtual_gpu_id >= 0) shares.push_back(P->opt.virtual_gpu_id);
    else
        for (int32_t g = 0; g < N; g++) shares.push_back(g);
    const bool all_rows = P->opt.virtual_gpu_id < 0 || N == 1;
    P->structure = !(P->opt.flags & GSOFA_FLAG_COUNTS_ONLY);
    P->supernodes = P->structure && all_rows && n > 0;
    P->spilled = false;
    P->executed = false;
    P->host_valid = false;
    P->last_replayed = false;
    memset(&P->stats, 0, sizeof(P->stats));
    if (n == 0) {
        P->executed = true;
        P->nnzL = P->nnzU = 0;
        return GSOFA_OK;
    }
    const size_t nev = 2 * shares.size() + 2;
    while (P->ev.size() < nev) {
        hipEvent_t e;
        HIPCHK(hipEventCreate(&e));
        P->ev.push_back(e);
    }
    if (P->structure && (P->opt.flags & GSOFA_FLAG_SPILL)) {
        P->spilled = true;
        P->passes = 1;
        return execute_spill(P, shares);
    }

    P->passes = 0;
    // graph replays: the modes whose launches carry no per-execute host state
    const bool graphable = (P->opt.flags & GSOFA_FLAG_GRAPH) && P->bounded && !P->graph_failed &&
                           (P->mode == GSOFA_MODE_REACH || (P->mode == GSOFA_MODE_BLOCK && !P->hybrid));
    if (graphable && P->graph_exec) {
        P->passes = 1;
        HIPCHK(hipEventRecord(P->ev[0], P->stream));
        HIPCHK(hipGraphLaunch(P->graph_exec, P->stream));
        HIPCHK(hipEventRecord(P->ev[1], P->stream));
        HIPCHK_OUT(hipStreamSynchronize(P->stream), P->stream);
        P->stats = *P->h_ctrl;
        if (P->stats.abort) return GSOFA_ERR_DEVICE_TIMEOUT;
        if (!P->structure || P->stats.fill_pos <= P->fill_cap) {
            // the graph holds no events of its own: its whole replay is timed
            P->ms = ev_ms(P->ev[0], P->ev[1]);
            P->traverse_ms = P->share_min = P->share_max = 0;
            P->totals_known = !P->structure;
            P->last_replayed = true;
            P->executed = true;
            return GSOFA_OK;
        }
        return GSOFA_ERR_STATE;  // cannot happen: the bounds hold
    }
    if (P->bounded) {
        P->passes = 1;
        // one pass: traversal, structure and supernodes are queued back to
        // back; the host waits once, at the end
        int rc = queue_bounded(P, shares, N, all_rows, &P->stats);
        if (rc) return rc;
        HIPCHK_OUT(hipStreamSynchronize(P->stream), P->stream);
        if (P->stats.abort) return GSOFA_ERR_DEVICE_TIMEOUT;
        if (!P->structure || P->stats.fill_pos <= P->fill_cap) {
            finish_timing(P, shares.size());
            P->executed = true;
            if (graphable) capture_graph(P, shares, N, all_rows);
            return GSOFA_OK;
        }
        // the bounds do not hold (cannot happen for a valid matrix): take the
        // path that grows the stream and reruns
        P->bounded = false;
        P->totals_known = true;
    }
    for (int attempt = 0; attempt < 3; attempt++) {
        P->passes++;
        HIPCHK(launch_fill(reinterpret_cast<uint32_t*>(P->d_ctrl), sizeof(Ctrl) / 4, 0u, P->stream));
        HIPCHK(hipEventRecord(P->ev[0], P->stream));
        for (size_t si = 0; si < shares.size(); si++) {
            int rc = run_share(P, shares[si], si);
            if (rc) return rc;
        }
        HIPCHK_OUT(hipMemcpyAsync(&P->stats, P->d_ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, P->stream), P->stream);
        HIPCHK_OUT(hipStreamSynchronize(P->stream), P->stream);
        if (P->stats.abort) return GSOFA_ERR_DEVICE_TIMEOUT;
        if (!P->structure || P->stats.fill_pos <= P->fill_cap) break;
        // fill stream outgrew its reserve: grow to the exact size and rerun
        dfree(P->d_fills);
        P->fill_cap = P->stats.fill_pos + P->stats.fill_pos / 16 + 1024;
        if (dalloc(&P->d_fills, (size_t)P->fill_cap)) return GSOFA_ERR_ALLOC;
    }
    if (P->structure && P->stats.fill_pos > P->fill_cap) return GSOFA_ERR_ALLOC;

    if (P->structure) {
        int rc = build_structure(P, P->stats.fill_pos, N, all_rows ? -1 : P->opt.virtual_gpu_id);
        if (rc) return rc;
        if (P->supernodes) {
            rc = build_supernodes(P, nullptr);
            if (rc) return rc;
        }
    }
    HIPCHK(hipEventRecord(P->ev[1], P->stream));
    HIPCHK_OUT(hipStreamSynchronize(P->stream), P->stream);
    finish_timing(P, shares.size());
    P->executed = true;
    return GSOFA_OK;
}

int gsofa_plan_fetch(gsofa_plan* P, gsofa_result* r) {
    if (!P || !r) return GSOFA_ERR_INVALID;
    if (!P->executed) return GSOFA_ERR_STATE;
    memset(r, 0, sizeof(*r));
    const int32_t n = P->n;
    r->n = n;
    r->fill_count = P->stats.fill_count;
    r->te_alg = P->stats.te;
    r->reach = P->stats.reach;
    r->edge_checks = P->stats.edge_checks;
    r->levels = P->stats.levels;
    r->rescans = P->stats.rescans;
    r->passes = P->passes;
    r->graph_replay = P->last_replayed ? 1 : 0;
    r->ms = P->ms;
    r->traverse_ms = P->traverse_ms;
    r->share_ms_min = P->share_min;
    r->share_ms_max = P->share_max;
    r->chunk = P->C;
    r->mode = P->hybrid ? GSOFA_MODE_HYBRID : P->mode;
    r->grid_blocks = (int32_t)P->nblocks;
    r->queue_capacity = P->mode == GSOFA_MODE_BLOCK || P->mode == GSOFA_MODE_REACH ? (int64_t)P->W : (int64_t)P->Q;
    r->n_super = -1;
    if (n == 0) {
        if (P->structure) {
            r->L_rowptr = (int64_t*)calloc(1, sizeof(int64_t));
            r->U_rowptr = (int64_t*)calloc(1, sizeof(int64_t));
            r->n_super = 0;
            r->super_start = (int32_t*)calloc(1, sizeof(int32_t));
        }
        return GSOFA_OK;
    }
    HIPCHK_OUT(hipSetDevice(P->device), P->stream);
    if (P->structure && !P->spilled) {
        const int rc = read_totals(P);
        if (rc) return rc;
    }
    if (!P->structure) {
        // counts only: the block kernel's per-row fills, TE and |R|; the
        // reach kernels' per-row fills (grid and hybrid runs report totals only)
        if (P->mode == GSOFA_MODE_REACH) {
            r->row_fill = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
            if (!r->row_fill) return GSOFA_ERR_ALLOC;
            HIPCHK_OUT(hipMemcpyAsync(r->row_fill, P->d_row_fk, sizeof(int64_t) * n, hipMemcpyDeviceToHost,
                                      P->stream), P->stream);
            HIPCHK_OUT(fetch_mat_totals(P), P->stream);
            HIPCHK_OUT(hipStreamSynchronize(P->stream), P->stream);
            return GSOFA_OK;
        }
        if (P->mode != GSOFA_MODE_BLOCK || P->hybrid) return GSOFA_OK;
        r->row_fill = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
        r->row_te = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
        r->row_reach = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
        if (!r->row_fill || !r->row_te || !r->row_reach) {
            gsofa_result_free(r);
            return GSOFA_ERR_ALLOC;
        }
        HIPCHK_OUT(hipMemcpyAsync(r->row_fill, P->d_row_fk, sizeof(int64_t) * n, hipMemcpyDeviceToHost, P->stream),
                   P->stream);
        HIPCHK_OUT(hipMemcpyAsync(r->row_te, P->d_row_te, sizeof(int64_t) * n, hipMemcpyDeviceToHost, P->stream),
                   P->stream);
        HIPCHK_OUT(hipMemcpyAsync(r->row_reach, P->d_row_reach, sizeof(int64_t) * n, hipMemcpyDeviceToHost,
                                  P->stream), P->stream);
        HIPCHK_OUT(hipStreamSynchronize(P->stream), P->stream);
        const int32_t N = std::max(1, P->opt.n_virtual_gpu), id = P->opt.virtual_gpu_id;
        for (int32_t i = 0; i < n; i++)
            if (id >= 0 && i % N != id) r->row_fill[i] = r->row_te[i] = r->row_reach[i] = -1;
        return GSOFA_OK;
    }
    r->row_fill = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
    r->L_rowptr = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n + 1));
    r->U_rowptr = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n + 1));
    r->L_col = (int32_t*)host_array(sizeof(int32_t) * (size_t)std::max<int64_t>(P->nnzL, 1));
    r->U_col = (int32_t*)host_array(sizeof(int32_t) * (size_t)std::max<int64_t>(P->nnzU, 1));
    if (!r->row_fill || !r->L_rowptr || !r->U_rowptr || !r->L_col || !r->U_col) {
        gsofa_result_free(r);
        return GSOFA_ERR_ALLOC;
    }
    if (!P->spilled)
        HIPCHK_OUT(hipMemcpyAsync(r->row_fill, P->d_rowfill, sizeof(int64_t) * n, hipMemcpyDeviceToHost, P->stream),
                   P->stream);
    if (P->mode == GSOFA_MODE_BLOCK && !P->hybrid) {
        r->row_te = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
        r->row_reach = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
        if (!r->row_te || !r->row_reach) {
            gsofa_result_free(r);
            return GSOFA_ERR_ALLOC;
        }
        HIPCHK_OUT(hipMemcpyAsync(r->row_te, P->d_row_te, sizeof(int64_t) * n, hipMemcpyDeviceToHost, P->stream),
                   P->stream);
        HIPCHK_OUT(hipMemcpyAsync(r->row_reach, P->d_row_reach, sizeof(int64_t) * n, hipMemcpyDeviceToHost,
                                  P->stream), P->stream);
    }
    if (P->spilled) {
        // rows in order, each from the share that built it (row i: share i mod N)
        memcpy(r->row_fill, P->sp_rowfill.data(), sizeof(int64_t) * (size_t)n);
        const int32_t N = std::max(1, P->opt.n_virtual_gpu);
        int64_t lo = 0, uo = 0;
        r->L_rowptr[0] = r->U_rowptr[0] = 0;
        for (int32_t i = 0; i < n; i++) {
            const SpillShare& h = P->sp[(size_t)(i % N)];
            const size_t k = (size_t)(i / N);  // the row's place in its share
            if (k + 1 < h.Loff.size()) {
                const int64_t lb = h.Loff[k], le = h.Loff[k + 1];
                const int64_t ub = h.Uoff[k], ue = h.Uoff[k + 1];
                memcpy(r->L_col + lo, h.Lcol.data() + lb, sizeof(int32_t) * (size_t)(le - lb));
                memcpy(r->U_col + uo, h.Ucol.data() + ub, sizeof(int32_t) * (size_t)(ue - ub));
                lo += le - lb;
                uo += ue - ub;
            }
            r->L_rowptr[i + 1] = lo;
            r->U_rowptr[i + 1] = uo;
        }
    } else {
        HIPCHK_OUT(hipMemcpyAsync(r->L_rowptr, P->d_Lrp, sizeof(int64_t) * (n + 1), hipMemcpyDeviceToHost, P->stream),
                   P->stream);
        HIPCHK_OUT(hipMemcpyAsync(r->U_rowptr, P->d_Urp, sizeof(int64_t) * (n + 1), hipMemcpyDeviceToHost, P->stream),
                   P->stream);
        HIPCHK_OUT(copy_d2h(r->L_col, P->d_Lcol, sizeof(int32_t) * (size_t)P->nnzL, P->stream), P->stream);
        HIPCHK_OUT(copy_d2h(r->U_col, P->d_Ucol, sizeof(int32_t) * (size_t)P->nnzU, P->stream), P->stream);
    }
    if (r->row_te) {
        HIPCHK_OUT(hipStreamSynchronize(P->stream), P->stream);
        for (int32_t i = 0; i < n; i++)
            if (r->row_fill[i] < 0) r->row_te[i] = r->row_reach[i] = -1;
    }
    HIPCHK_OUT(fetch_mat_totals(P), P->stream);
    if (P->supernodes) {
        int32_t ns = 0;
        HIPCHK_OUT(hipMemcpyAsync(&ns, P->d_nsuper, sizeof(int32_t), hipMemcpyDeviceToHost, P->stream), P->stream);
        HIPCHK_OUT(hipStreamSynchronize(P->stream), P->stream);
        r->n_super = ns;
        r->super_start = (int32_t*)malloc(sizeof(int32_t) * (size_t)(ns + 1));
        r->supno = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
        if (!r->super_start || !r->supno) {
            gsofa_result_free(r);
            return GSOFA_ERR_ALLOC;
        }
        HIPCHK_OUT(hipMemcpyAsync(r->super_start, P->d_sstart, sizeof(int32_t) * (ns + 1), hipMemcpyDeviceToHost,
                                  P->stream), P->stream);
        HIPCHK_OUT(hipMemcpyAsync(r->supno, P->d_supno, sizeof(int32_t) * n, hipMemcpyDeviceToHost, P->stream),
                   P->stream);
    }
    HIPCHK_OUT(hipStreamSynchronize(P->stream), P->stream);
    return GSOFA_OK;
}

int gsofa_plan_fetch_matrix(gsofa_plan* P, int32_t index, gsofa_result* r) {
    if (!P || !r || index < 0 || index >= P->n_mats) return GSOFA_ERR_INVALID;
    if (!P->executed) return GSOFA_ERR_STATE;
    if (!P->host_valid) {
        gsofa_result_free(&P->host);
        int rc = gsofa_plan_fetch(P, &P->host);
        if (rc) return rc;
        P->host_valid = true;
    }
    const gsofa_result& G = P->host;
    memset(r, 0, sizeof(*r));
    const int64_t b0 = P->h_mat_base[index], b1 = P->h_mat_base[index + 1];
    const int32_t nb = (int32_t)(b1 - b0);
    r->n = nb;
    r->ms = G.ms;
    r->traverse_ms = G.traverse_ms;
    r->chunk = G.chunk;
    r->mode = G.mode;
    r->grid_blocks = G.grid_blocks;
    r->queue_capacity = G.queue_capacity;
    r->n_super = -1;
    if (P->n_mats == 1) {  // the plan's own totals (also for counts-only grid runs)
        r->fill_count = G.fill_count;
        r->te_alg = G.te_alg;
        r->reach = G.reach;
        r->edge_checks = G.edge_checks;
        r->levels = G.levels;
        r->rescans = G.rescans;
    }
    r->passes = G.passes;
    r->graph_replay = G.graph_replay;
    if (!G.row_fill) return GSOFA_OK;  // counts only in grid or hybrid mode: totals alone
    const bool lu = G.L_rowptr != nullptr;  // false: counts only (per-row fills, TE, |R|)
    r->row_fill = (int64_t*)malloc(sizeof(int64_t) * (size_t)std::max(nb, 1));
    const int64_t l0 = lu ? G.L_rowptr[b0] : 0, l1 = lu ? G.L_rowptr[b1] : 0;
    const int64_t u0 = lu ? G.U_rowptr[b0] : 0, u1 = lu ? G.U_rowptr[b1] : 0;
    if (lu) {
        r->L_rowptr = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nb + 1));
        r->U_rowptr = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nb + 1));
        r->L_col = (int32_t*)host_array(sizeof(int32_t) * (size_t)std::max<int64_t>(l1 - l0, 1));
        r->U_col = (int32_t*)host_array(sizeof(int32_t) * (size_t)std::max<int64_t>(u1 - u0, 1));
    }
    if (G.row_te) {
        r->row_te = (int64_t*)malloc(sizeof(int64_t) * (size_t)std::max(nb, 1));
        r->row_reach = (int64_t*)malloc(sizeof(int64_t) * (size_t)std::max(nb, 1));
    }
    if (!r->row_fill || (lu && (!r->L_rowptr || !r->U_rowptr || !r->L_col || !r->U_col)) ||
        (G.row_te && (!r->row_te || !r->row_reach))) {
        gsofa_result_free(r);
        return GSOFA_ERR_ALLOC;
    }
    uint64_t fills = 0, te = 0, reach = 0;
    for (int32_t i = 0; i < nb; i++) {
        r->row_fill[i] = G.row_fill[b0 + i];
        fills += (uint64_t)std::max<int64_t>(r->row_fill[i], 0);
        if (G.row_te) {
            r->row_te[i] = G.row_te[b0 + i];
            r->row_reach[i] = G.row_reach[b0 + i];
            te += (uint64_t)std::max<int64_t>(r->row_te[i], 0);
            reach += (uint64_t)std::max<int64_t>(r->row_reach[i], 0);
        }
    }
    r->fill_count = fills;
    if (G.row_te) {  // per-row TE and |R| exist in block mode (every batch)
        r->te_alg = te;
        r->reach = reach;
    } else if (P->mode == GSOFA_MODE_REACH && P->n_mats > 1 && (size_t)index < P->h_mat_te.size()) {
        r->te_alg = P->h_mat_te[(size_t)index];  // reach mode: per-matrix totals
        r->reach = P->h_mat_reach[(size_t)index];
    }
    if (!lu) return GSOFA_OK;
    for (int32_t i = 0; i <= nb; i++) {
        r->L_rowptr[i] = G.L_rowptr[b0 + i] - l0;
        r->U_rowptr[i] = G.U_rowptr[b0 + i] - u0;
    }
    for (int64_t k = l0; k < l1; k++) r->L_col[k - l0] = G.L_col[k] - (int32_t)b0;
    for (int64_t k = u0; k < u1; k++) r->U_col[k - u0] = G.U_col[k] - (int32_t)b0;
    if (G.n_super >= 0 && G.supno && nb > 0) {
        // supernodes never cross matrices: (i-1) is never in L_i at a boundary
        const int32_t s0 = G.supno[b0], s1 = G.supno[b1 - 1];
        r->n_super = s1 - s0 + 1;
        r->super_start = (int32_t*)malloc(sizeof(int32_t) * (size_t)(r->n_super + 1));
        r->supno = (int32_t*)malloc(sizeof(int32_t) * (size_t)nb);
        if (!r->super_start || !r->supno) {
            gsofa_result_free(r);
            return GSOFA_ERR_ALLOC;
        }
        for (int32_t s = 0; s < r->n_super; s++) r->super_start[s] = G.super_start[s0 + s] - (int32_t)b0;
        r->super_start[r->n_super] = nb;
        for (int32_t i = 0; i < nb; i++) r->supno[i] = G.supno[b0 + i] - s0;
    }
    return GSOFA_OK;
}

void gsofa_plan_destroy(gsofa_plan* P) {
    if (!P) return;
    (void)hipSetDevice(P->device);
    (void)hipStreamSynchronize(P->stream);
    if (P->graph_exec) (void)hipGraphExecDestroy(P->graph_exec);
    if (P->graph) (void)hipGraphDestroy(P->graph);
    if (P->h_ctrl) (void)hipHostFree(P->h_ctrl);
    for (auto e : P->ev) (void)hipEventDestroy(e);
    gsofa_result_free(&P->host);
    release_device(P);
    (void)hipStreamDestroy(P->stream);
    delete P;
}

void gsofa_supernodal_free(gsofa_supernodal* h) {
    if (!h) return;
    free(h->xsup);
    free(h->supno);
    free(h->xsub);
    free(h->sub);
    free(h->xlcol);
    free(h->lrow);
    free(h->xucol);
    free(h->urow);
    memset(h, 0, sizeof(*h));
}

int gsofa_plan_fetch_supernodal(gsofa_plan* P, gsofa_supernodal* h) {
    if (!P || !h) return GSOFA_ERR_INVALID;
    memset(h, 0, sizeof(*h));
    if (!P->executed) return GSOFA_ERR_STATE;
    if (P->n_mats != 1 || !P->structure || (!P->supernodes && P->n > 0) || P->spilled) return GSOFA_ERR_INVALID;
    const int32_t n = P->n;
    h->n = n;
    int32_t ns = 0;
    HIPCHK_OUT(hipSetDevice(P->device), P->stream);
    const int trc = read_totals(P);
    if (trc) return trc;
    if (n > 0) {
        HIPCHK_OUT(hipMemcpyAsync(&ns, P->d_nsuper, sizeof(int32_t), hipMemcpyDeviceToHost, P->stream), P->stream);
        HIPCHK_OUT(hipStreamSynchronize(P->stream), P->stream);
    }
    h->n_super = ns;
    DevScratch T;  // device scratch for this call only
    int64_t *d_xl = nullptr, *d_xu = nullptr, *d_xs = nullptr;
    int32_t *d_lrow = nullptr, *d_urow = nullptr;
    if (T.get(&d_xl, (size_t)n + 1) || T.get(&d_xu, (size_t)n + 1) || T.get(&d_xs, (size_t)ns + 1) ||
        T.get(&d_lrow, (size_t)P->nnzL) || T.get(&d_urow, (size_t)P->nnzU))
        return GSOFA_ERR_ALLOC;
    // L and U by columns; the staging arrays of the structure pass are free now
    HIPCHK_OUT(transpose_structure(P->d_Lrp, P->d_Lcol, P->nnzL, n, P->d_cntL, d_xl, P->d_scan, P->d_Ltmp, d_lrow,
                                   P->stream), P->stream);
    HIPCHK_OUT(transpose_structure(P->d_Urp, P->d_Ucol, P->nnzU, n, P->d_cntU, d_xu, P->d_scan, P->d_Utmp, d_urow,
                                   P->stream), P->stream);
    // supernode subscripts
    int64_t nsub = 0;
    int32_t* d_sub = nullptr;
    if (ns > 0) {
        HIPCHK_OUT(supernode_sub_count(P->d_sstart, ns, P->d_Urp, P->d_cntL, P->stream), P->stream);
        HIPCHK_OUT(exclusive_scan_u32_to_i64(P->d_cntL, d_xs, ns, P->d_scan, P->stream), P->stream);
        HIPCHK_OUT(hipMemcpyAsync(&nsub, d_xs + ns, sizeof(int64_t), hipMemcpyDeviceToHost, P->stream), P->stream);
        HIPCHK_OUT(hipStreamSynchronize(P->stream), P->stream);
        if (T.get(&d_sub, (size_t)nsub)) return GSOFA_ERR_ALLOC;
        HIPCHK_OUT(supernode_sub_write(P->d_sstart, ns, P->d_Urp, P->d_Ucol, d_xs, d_sub, P->stream), P->stream);
    }
    h->xsup = (int32_t*)malloc(sizeof(int32_t) * ((size_t)ns + 1));
    h->supno = (int32_t*)malloc(sizeof(int32_t) * std::max<size_t>((size_t)n, 1));
    h->xsub = (int64_t*)calloc((size_t)ns + 1, sizeof(int64_t));
    h->sub = (int32_t*)malloc(sizeof(int32_t) * std::max<size_t>((size_t)nsub, 1));
    h->xlcol = (int64_t*)calloc((size_t)n + 1, sizeof(int64_t));
    h->lrow = (int32_t*)malloc(sizeof(int32_t) * std::max<size_t>((size_t)P->nnzL, 1));
    h->xucol = (int64_t*)calloc((size_t)n + 1, sizeof(int64_t));
    h->urow = (int32_t*)malloc(sizeof(int32_t) * std::max<size_t>((size_t)P->nnzU, 1));
    if (!h->xsup || !h->supno || !h->xsub || !h->sub || !h->xlcol || !h->lrow || !h->xucol || !h->urow) {
        gsofa_supernodal_free(h);
        return GSOFA_ERR_ALLOC;
    }
    h->xsup[0] = 0;
    if (n > 0) {
        HIPCHK_OUT(hipMemcpyAsync(h->xsup, P->d_sstart, sizeof(int32_t) * ((size_t)ns + 1), hipMemcpyDeviceToHost,
                                  P->stream), P->stream);
        HIPCHK_OUT(hipMemcpyAsync(h->supno, P->d_supno, sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToHost,
                                  P->stream), P->stream);
        HIPCHK_OUT(hipMemcpyAsync(h->xlcol, d_xl, sizeof(int64_t) * ((size_t)n + 1), hipMemcpyDeviceToHost, P->stream),
                   P->stream);
        HIPCHK_OUT(hipMemcpyAsync(h->xucol, d_xu, sizeof(int64_t) * ((size_t)n + 1), hipMemcpyDeviceToHost, P->stream),
                   P->stream);
    }
    if (P->nnzL)
        HIPCHK_OUT(hipMemcpyAsync(h->lrow, d_lrow, sizeof(int32_t) * (size_t)P->nnzL, hipMemcpyDeviceToHost, P->stream),
                   P->stream);
    if (P->nnzU)
        HIPCHK_OUT(hipMemcpyAsync(h->urow, d_urow, sizeof(int32_t) * (size_t)P->nnzU, hipMemcpyDeviceToHost, P->stream),
                   P->stream);
    if (ns > 0) {
        HIPCHK_OUT(hipMemcpyAsync(h->xsub, d_xs, sizeof(int64_t) * ((size_t)ns + 1), hipMemcpyDeviceToHost, P->stream),
                   P->stream);
        HIPCHK_OUT(hipMemcpyAsync(h->sub, d_sub, sizeof(int32_t) * (size_t)nsub, hipMemcpyDeviceToHost, P->stream),
                   P->stream);
    }
    HIPCHK_OUT(hipStreamSynchronize(P->stream), P->stream);
    return GSOFA_OK;
}

int gsofa_fill_bound(const gsofa_matrix* A, int64_t* row_fill_bound, uint64_t* fill_bound, uint64_t* L_bound,
                     uint64_t* U_bound) {
    const int rc = validate(A);
    if (rc) return rc;
    std::vector<int32_t> rp, ci;
    pack_rows(A, rp, ci);
    std::vector<int64_t> Lb, Ub;
    lu_row_bounds(A->n, rp.data(), ci.data(), Lb, Ub);
    uint64_t f = 0, l = 0, u = 0;
    for (int32_t i = 0; i < A->n; i++) {
        int64_t off = 0;  // A's own off-diagonal entries are not fills
        for (int32_t p = rp[i]; p < rp[i + 1]; p++) off += ci[p] != i;
        const int64_t b = Lb[i] + Ub[i] - off;
        if (row_fill_bound) row_fill_bound[i] = b;
        f += (uint64_t)b;
        l += (uint64_t)Lb[i];
        u += (uint64_t)Ub[i];
    }
    if (fill_bound) *fill_bound = f;
    if (L_bound) *L_bound = l;
    if (U_bound) *U_bound = u;
    return GSOFA_OK;
}

int gsofa_symbfact(const gsofa_matrix* A, const gsofa_options* opt, gsofa_result* out) {
    if (!out) return GSOFA_ERR_INVALID;
    memset(out, 0, sizeof(*out));
    gsofa_plan* P = nullptr;
    int rc = gsofa_plan_create(A, opt, &P);
    if (rc) return rc;
    rc = gsofa_plan_execute(P);
    if (!rc) rc = gsofa_plan_fetch(P, out);
    gsofa_plan_destroy(P);
    return rc;
}

}  // extern "C"

// ---- arenas ----------------------------------------------------------------
// A reach-mode plan that takes one batch after another: the request server's
// workers each hold one.  Its buffers grow to the largest batch seen and are
// not freed between batches, so a worker in steady state makes no device
// allocation (a hipFree waits for the whole device, which would stall the
// other workers' batches).  The per-workgroup arrays keep one stride (their
// capacity), so a workgroup's visited stamps always sit in its own region;
// they are cleared only when that capacity grows.

namespace gsofa {

int pack_matrix(const gsofa_matrix* A, PackedMatrix* M) {
    const int rc = validate(A);
    if (rc) return rc;
    if (A->n < 1) return GSOFA_ERR_INVALID;
    pack_rows(A, M->rp, M->ci);
    M->n = A->n;
    M->nnz = (int64_t)M->ci.size();
    transpose_rows(M->n, M->rp.data(), M->ci.data(), M->rpT, M->ciT);
    gsofa_options o;
    gsofa_options_default(&o);
    output_need(o, M->n, M->rp.data(), M->ci.data(), &M->need_fill, &M->need_L, &M->need_U);
    return GSOFA_OK;
}

int64_t reach_fixed_bytes(int64_t n, int64_t nnz, int32_t n_mats, uint64_t nchunks, uint64_t fills, uint64_t lu) {
    return block_fixed_bytes(n, nnz, n_mats, nchunks, rch_fill_room(fills), (int64_t)lu) +
           4 * (n + 1) + 4 * nnz + 16 * (int64_t)n_mats;
}

int64_t arena_held_bytes(const gsofa_plan* P) { return P ? P->held : 0; }

int arena_create(const gsofa_options* opt, gsofa_plan** out) {
    if (!out) return GSOFA_ERR_INVALID;
    *out = nullptr;
    gsofa_plan* P = new gsofa_plan();
    if (opt) P->opt = *opt;
    else gsofa_options_default(&P->opt);
    P->opt.mode = GSOFA_MODE_REACH;
    P->opt.n_virtual_gpu = 1;
    P->opt.virtual_gpu_id = -1;
    P->opt.flags &= ~(GSOFA_FLAG_SPILL | GSOFA_FLAG_GRAPH);
    P->mode = GSOFA_MODE_REACH;
    P->arena = true;
    P->device = P->opt.device;
    if (hipSetDevice(P->device) != hipSuccess || hipStreamCreate(&P->stream) != hipSuccess) {
        (void)hipGetLastError();
        delete P;
        return GSOFA_ERR_HIP;
    }
    *out = P;
    return GSOFA_OK;
}

int arena_load(gsofa_plan* P, const PackedMatrix* const* mats, int32_t count, int64_t budget) {
    if (!P || !P->arena || !mats || count < 1) return GSOFA_ERR_INVALID;
    HIPCHK(hipSetDevice(P->device));
    int64_t n = 0, nnz = 0;
    uint32_t W = 1;
    uint64_t nf = 0, nl = 0, nu = 0;
    for (int32_t b = 0; b < count; b++) {
        n += mats[b]->n;
        nnz += mats[b]->nnz;
        W = std::max<uint32_t>(W, (uint32_t)mats[b]->n);
        nf += mats[b]->need_fill;
        nl += mats[b]->need_L;
        nu += mats[b]->need_U;
    }
    if (n > kMaxN || nnz >= (int64_t)INT32_MAX) return GSOFA_ERR_TOO_LARGE;
    // block-diagonal batch on the host: matrix b's vertices follow matrix b-1's
    P->n = (int32_t)n;
    P->nnz = nnz;
    P->n_mats = count;
    P->h_mat_base.assign((size_t)count + 1, 0);
    P->h_rp.assign((size_t)n + 1, 0);
    P->h_ci.resize((size_t)nnz);
    std::vector<int32_t>& rpT = P->h_rpT;
    rpT.assign((size_t)n + 1, 0);
    P->h_ciT.resize((size_t)nnz);
    int64_t v = 0, k = 0;
    for (int32_t b = 0; b < count; b++) {
        const PackedMatrix& M = *mats[b];
        P->h_mat_base[b] = (uint32_t)v;
        for (int32_t i = 0; i <= M.n; i++) {
            P->h_rp[(size_t)(v + i)] = (int32_t)(k + M.rp[(size_t)i]);
            rpT[(size_t)(v + i)] = (int32_t)(k + M.rpT[(size_t)i]);
        }
        for (int64_t e = 0; e < M.nnz; e++) {
            P->h_ci[(size_t)(k + e)] = M.ci[(size_t)e] + (int32_t)v;
            P->h_ciT[(size_t)(k + e)] = M.ciT[(size_t)e] + (int32_t)v;
        }
        v += M.n;
        k += M.nnz;
    }
    P->h_mat_base[count] = (uint32_t)v;
    // chunks: every matrix's, deepest first, each followed by its A^T twin.
    // Arenas keep 64 sources per workgroup, so the visited words keep one
    // layout from batch to batch.
    const int bits = 64;
    P->S = (uint32_t)bits;
    P->rbits = bits;
    P->W = W;
    build_chunks(P);
    std::vector<BlkChunk> both;
    both.reserve(2 * P->h_chunks.size());
    for (const BlkChunk& ch : P->h_chunks) {
        both.push_back(ch);
        BlkChunk t = ch;
        t.stride |= rch_pass_bit();
        both.push_back(t);
    }
    P->h_chunks.swap(both);
    P->share_chunk_off.assign(2, 0);
    P->share_chunk_off[1] = (uint32_t)P->h_chunks.size();
    const uint64_t nchunks = P->h_chunks.size();
    // workgroups: what the budget leaves after the batch's fixed arrays
    const int64_t fixed = reach_fixed_bytes(n, nnz, count, nchunks, nf, nl + nu);
    uint32_t nb = 0;
    gsofa_options o = P->opt;
    o.mem_budget_bytes = 0;
    o.mode = GSOFA_MODE_AUTO;  // the tuned grid for this batch's signature (slots stay 64)
    if (!apply_tuned(&o, W, nnz, n, count) || resolve_mode(o) != GSOFA_MODE_REACH) o = P->opt;
    o.mem_budget_bytes = 0;
    int rc = reach_grid(o, bits, nchunks, W, budget - fixed, &nb);
    if (rc) return rc;
    P->rthreads = reach_threads(o);
    gsofa_plan::Caps& c = P->cap;
    const size_t rows = (size_t)n + 1;
    // The arena keeps each buffer at the largest size seen.  If keeping them
    // and growing the rest would exceed the budget, it starts over with
    // exactly this batch's sizes (the only path that frees device memory).
    {
        auto mx = [](size_t a, uint64_t b) { return (int64_t)std::max<uint64_t>(a, std::max<uint64_t>(b, 1)); };
        const int64_t keep = 40 * mx(c.rp, rows) + 8 * mx(c.ci, (uint64_t)nnz) + 8 * mx(c.fills, rch_fill_room(nf)) +
                             8 * mx(c.L, nl) + 8 * mx(c.U, nu) + 8 * mx(c.scan, scan_tmp_words(n) + 2 * (size_t)n + 8) +
                             reach_bytes((uint32_t)std::max<size_t>(c.nb, nb), (uint32_t)std::max<size_t>(c.wg, W), 64);
        if (keep > budget && P->held > 0) {
            HIPCHK_OUT(hipStreamSynchronize(P->stream), P->stream);
            release_device(P);
            P->cap = gsofa_plan::Caps();
            P->sn_cap = 0;
            P->held = 0;
        }
    }
    int e = 0;
    if (c.rp < rows) {
        dfree(P->d_rp);
        dfree(P->d_rpT);
        dfree(P->d_cntL);
        dfree(P->d_cntU);
        dfree(P->d_Lrp);
        dfree(P->d_Urp);
        dfree(P->d_rowfill);
        dfree(P->d_row_fk);
        e |= dalloc(&P->d_rp, rows);
        e |= dalloc(&P->d_rpT, rows);
        e |= dalloc(&P->d_cntL, rows);
        e |= dalloc(&P->d_cntU, rows);
        e |= dalloc(&P->d_Lrp, rows);
        e |= dalloc(&P->d_Urp, rows);
        e |= dalloc(&P->d_rowfill, rows);
        e |= dalloc(&P->d_row_fk, rows);
        c.rp = rows;
    }
    const size_t ne = (size_t)std::max<int64_t>(nnz, 1);
    if (c.ci < ne) {
        dfree(P->d_ci);
        dfree(P->d_ciT);
        e |= dalloc(&P->d_ci, ne);
        e |= dalloc(&P->d_ciT, ne);
        c.ci = ne;
    }
    const size_t nfe = (size_t)rch_fill_room(nf);
    if (c.fills < nfe) {
        dfree(P->d_fills);
        e |= dalloc(&P->d_fills, nfe);
        c.fills = nfe;
    }
    const size_t nle = (size_t)std::max<uint64_t>(nl, 1), nue = (size_t)std::max<uint64_t>(nu, 1);
    if (c.L < nle) {
        dfree(P->d_Lcol);
        dfree(P->d_Ltmp);
        e |= dalloc(&P->d_Lcol, nle);
        e |= dalloc(&P->d_Ltmp, nle);
        c.L = nle;
    }
    if (c.U < nue) {
        dfree(P->d_Ucol);
        dfree(P->d_Utmp);
        e |= dalloc(&P->d_Ucol, nue);
        e |= dalloc(&P->d_Utmp, nue);
        c.U = nue;
    }
    P->scan_words = scan_tmp_words(n) + 2 * (size_t)n + 8;
    if (c.scan < P->scan_words) {
        dfree(P->d_scan);
        e |= dalloc(&P->d_scan, P->scan_words);
        c.scan = P->scan_words;
    }
    if (c.mats < (size_t)count + 1) {
        dfree(P->d_mat_base);
        dfree(P->d_mat_te);
        dfree(P->d_mat_reach);
        e |= dalloc(&P->d_mat_base, (size_t)count + 1);
        e |= dalloc(&P->d_mat_te, (size_t)count + 1);
        e |= dalloc(&P->d_mat_reach, (size_t)count + 1);
        c.mats = (size_t)count + 1;
    }
    if (c.chunks < nchunks) {
        dfree(P->d_chunks);
        e |= dalloc(&P->d_chunks, (size_t)nchunks);
        c.chunks = nchunks;
    }
    if (!P->d_ctrl) e |= dalloc(&P->d_ctrl, 1);
    if (e) return GSOFA_ERR_ALLOC;
    // per-workgroup arrays: one stride (their capacity) for every batch,
    // cleared when they grow
    if (!P->d_vis || c.wg < W || c.nb < nb) {
        dfree(P->d_vis);
        dfree(P->d_rmask);
        dfree(P->d_qgv);
        dfree(P->d_bstate);
        c.wg = std::max<size_t>(W, c.wg);
        c.nb = std::max<size_t>(nb, c.nb);
        if (reach_bytes((uint32_t)c.nb, (uint32_t)c.wg, 64) > budget - fixed) {  // exactly this batch's
            c.wg = W;
            c.nb = nb;
        }
        const size_t nbw = c.nb * c.wg;
        e |= dalloc((uint4**)&P->d_vis, nbw);
        e |= dalloc((uint64_t**)&P->d_rmask, 2 * nbw);
        e |= dalloc(&P->d_qgv, 2 * nbw);
        e |= dalloc(&P->d_bstate, c.nb);
        if (e) return GSOFA_ERR_ALLOC;
        HIPCHK(hipMemsetAsync(P->d_vis, 0, 16 * nbw, P->stream));
        HIPCHK(hipMemsetAsync(P->d_rmask, 0, 16 * nbw, P->stream));
        HIPCHK(hipMemsetAsync(P->d_bstate, 0, sizeof(uint32_t) * c.nb, P->stream));
    }
    P->held = (int64_t)(c.rp * (4 + 4 + 4 + 4 + 8 + 8 + 8 + 8) + c.ci * 8 + c.fills * 8 + c.L * 8 + c.U * 8 +
                        c.scan * 8 + c.mats * 20 + c.chunks * sizeof(BlkChunk) + sizeof(Ctrl)) +
              reach_bytes((uint32_t)c.nb, (uint32_t)c.wg, 64);
    P->W = (uint32_t)c.wg;  // the kernels' per-workgroup stride
    P->nblocks = nb;
    P->C = (int32_t)(nb * P->S);
    P->fill_cap = nfe;
    P->fill_need = nf;
    P->Lcap = (int64_t)c.L;
    P->Ucap = (int64_t)c.U;
    P->bounded = true;
    P->totals_known = true;
    P->span_cap = 0xFFFFFFFFu;
    if (const char* sp = getenv("GSOFA_OFFSET_SPAN")) {
        const long vsp = atol(sp);
        if (vsp >= 1 && vsp < (long)0xFFFFFFFFu) P->span_cap = (uint32_t)vsp;
    }
    // uploads (the host copies stay in the plan until the next load)
    HIPCHK(hipMemcpyAsync(P->d_rp, P->h_rp.data(), sizeof(int32_t) * ((size_t)n + 1), hipMemcpyHostToDevice,
                          P->stream));
    HIPCHK(hipMemcpyAsync(P->d_rpT, rpT.data(), sizeof(int32_t) * ((size_t)n + 1), hipMemcpyHostToDevice,
                          P->stream));
    if (nnz) {
        HIPCHK(hipMemcpyAsync(P->d_ci, P->h_ci.data(), sizeof(int32_t) * (size_t)nnz, hipMemcpyHostToDevice,
                              P->stream));
        HIPCHK(hipMemcpyAsync(P->d_ciT, P->h_ciT.data(), sizeof(int32_t) * (size_t)nnz, hipMemcpyHostToDevice,
                              P->stream));
    }
    HIPCHK(hipMemcpyAsync(P->d_mat_base, P->h_mat_base.data(), sizeof(uint32_t) * ((size_t)count + 1),
                          hipMemcpyHostToDevice, P->stream));
    HIPCHK(hipMemcpyAsync(P->d_chunks, P->h_chunks.data(), sizeof(BlkChunk) * nchunks, hipMemcpyHostToDevice,
                          P->stream));
    P->executed = false;
    P->host_valid = false;
    return GSOFA_OK;
}

// Matrix `index` of the last execute, in its own numbering, copied straight
// from the device (its columns are shifted to local numbering there first).
int arena_fetch(gsofa_plan* P, int32_t index, gsofa_result* r) {
    if (!P || !P->arena || !r || index < 0 || index >= P->n_mats) return GSOFA_ERR_INVALID;
    if (!P->executed) return GSOFA_ERR_STATE;
    memset(r, 0, sizeof(*r));
    HIPCHK_OUT(hipSetDevice(P->device), P->stream);
    const int64_t b0 = P->h_mat_base[(size_t)index], b1 = P->h_mat_base[(size_t)index + 1];
    const int32_t nb = (int32_t)(b1 - b0);
    r->n = nb;
    r->ms = P->ms;
    r->traverse_ms = P->traverse_ms;
    r->chunk = P->C;
    r->mode = GSOFA_MODE_REACH;
    r->grid_blocks = (int32_t)P->nblocks;
    r->queue_capacity = (int64_t)P->W;
    r->passes = P->passes;
    r->n_super = -1;
    unsigned long long te = 0, reach = 0;
    HIPCHK_OUT(hipMemcpyAsync(&te, P->d_mat_te + index, sizeof(te), hipMemcpyDeviceToHost, P->stream), P->stream);
    HIPCHK_OUT(hipMemcpyAsync(&reach, P->d_mat_reach + index, sizeof(reach), hipMemcpyDeviceToHost, P->stream),
               P->stream);
    r->te_alg = te;
    r->reach = reach;
    r->row_fill = (int64_t*)malloc(sizeof(int64_t) * (size_t)std::max(nb, 1));
    if (!r->row_fill) return GSOFA_ERR_ALLOC;
    if (!P->structure) {  // counts only: the kernels' per-row tallies
        HIPCHK_OUT(hipMemcpyAsync(r->row_fill, P->d_row_fk + b0, sizeof(int64_t) * (size_t)nb, hipMemcpyDeviceToHost,
                                  P->stream), P->stream);
        HIPCHK_OUT(hipStreamSynchronize(P->stream), P->stream);
        r->te_alg = te;
        r->reach = reach;
        for (int32_t i = 0; i < nb; i++) r->fill_count += (uint64_t)r->row_fill[i];
        return GSOFA_OK;
    }
    r->L_rowptr = (int64_t*)malloc(sizeof(int64_t) * ((size_t)nb + 1));
    r->U_rowptr = (int64_t*)malloc(sizeof(int64_t) * ((size_t)nb + 1));
    if (!r->L_rowptr || !r->U_rowptr) {
        gsofa_result_free(r);
        return GSOFA_ERR_ALLOC;
    }
    HIPCHK_OUT(hipMemcpyAsync(r->L_rowptr, P->d_Lrp + b0, sizeof(int64_t) * ((size_t)nb + 1), hipMemcpyDeviceToHost,
                              P->stream), P->stream);
    HIPCHK_OUT(hipMemcpyAsync(r->U_rowptr, P->d_Urp + b0, sizeof(int64_t) * ((size_t)nb + 1), hipMemcpyDeviceToHost,
                              P->stream), P->stream);
    HIPCHK_OUT(hipMemcpyAsync(r->row_fill, P->d_rowfill + b0, sizeof(int64_t) * (size_t)nb, hipMemcpyDeviceToHost,
                              P->stream), P->stream);
    HIPCHK_OUT(hipStreamSynchronize(P->stream), P->stream);
    r->te_alg = te;
    r->reach = reach;
    const int64_t l0 = r->L_rowptr[0], l1 = r->L_rowptr[nb], u0 = r->U_rowptr[0], u1 = r->U_rowptr[nb];
    for (int32_t i = 0; i <= nb; i++) {
        r->L_rowptr[i] -= l0;
        r->U_rowptr[i] -= u0;
    }
    for (int32_t i = 0; i < nb; i++) r->fill_count += (uint64_t)r->row_fill[i];
    r->L_col = (int32_t*)host_array(sizeof(int32_t) * (size_t)std::max<int64_t>(l1 - l0, 1));
    r->U_col = (int32_t*)host_array(sizeof(int32_t) * (size_t)std::max<int64_t>(u1 - u0, 1));
    if (!r->L_col || !r->U_col) {
        gsofa_result_free(r);
        return GSOFA_ERR_ALLOC;
    }
    if (b0) {  // global to the matrix's own column numbers
        HIPCHK_OUT(launch_add(P->d_Lcol + l0, (uint64_t)(l1 - l0), -(int32_t)b0, P->stream), P->stream);
        HIPCHK_OUT(launch_add(P->d_Ucol + u0, (uint64_t)(u1 - u0), -(int32_t)b0, P->stream), P->stream);
    }
    HIPCHK_OUT(copy_d2h(r->L_col, P->d_Lcol + l0, sizeof(int32_t) * (size_t)(l1 - l0), P->stream), P->stream);
    HIPCHK_OUT(copy_d2h(r->U_col, P->d_Ucol + u0, sizeof(int32_t) * (size_t)(u1 - u0), P->stream), P->stream);
    if (P->supernodes && nb > 0) {
        int32_t s0 = 0, s1 = 0;
        HIPCHK_OUT(hipMemcpyAsync(&s0, P->d_supno + b0, sizeof(int32_t), hipMemcpyDeviceToHost, P->stream), P->stream);
        HIPCHK_OUT(hipMemcpyAsync(&s1, P->d_supno + b1 - 1, sizeof(int32_t), hipMemcpyDeviceToHost, P->stream),
                   P->stream);
        HIPCHK_OUT(hipStreamSynchronize(P->stream), P->stream);
        // supernodes never cross matrices: (i-1) is never in L_i at a boundary
        r->n_super = s1 - s0 + 1;
        r->super_start = (int32_t*)malloc(sizeof(int32_t) * ((size_t)r->n_super + 1));
        r->supno = (int32_t*)malloc(sizeof(int32_t) * (size_t)nb);
        if (!r->super_start || !r->supno) {
            gsofa_result_free(r);
            return GSOFA_ERR_ALLOC;
        }
        HIPCHK_OUT(hipMemcpyAsync(r->super_start, P->d_sstart + s0, sizeof(int32_t) * (size_t)r->n_super,
                                  hipMemcpyDeviceToHost, P->stream), P->stream);
        HIPCHK_OUT(hipMemcpyAsync(r->supno, P->d_supno + b0, sizeof(int32_t) * (size_t)nb, hipMemcpyDeviceToHost,
                                  P->stream), P->stream);
        HIPCHK_OUT(hipStreamSynchronize(P->stream), P->stream);
        for (int32_t s = 0; s < r->n_super; s++) r->super_start[s] -= (int32_t)b0;
        r->super_start[r->n_super] = nb;
        for (int32_t i = 0; i < nb; i++) r->supno[i] -= s0;
    }
    HIPCHK_OUT(hipStreamSynchronize(P->stream), P->stream);
    return GSOFA_OK;
}

}  // namespace gsofa
