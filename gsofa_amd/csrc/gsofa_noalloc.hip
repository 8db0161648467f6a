// gsofa_noalloc.hip -- fill counts of one matrix without any allocation.
//
// A variant of the counts-only block-mode execute (gsofa_plan_execute with
// GSOFA_FLAG_COUNTS_ONLY and GSOFA_MODE_BLOCK) for callers that cannot call
// hipMalloc: code under HIP graph capture, tight serving loops, harnesses
// that forbid it.  The caller passes the matrix and the outputs as device
// pointers; every work array of the traversal (per-workgroup labels, slot
// masks and queues, per-workgroup offsets, the work table and the control
// block) is carved out of one file-scope device array, `scratch`, of
// GSOFA_NOALLOC_SCRATCH_BYTES, which each GPU holds until the process ends.
// Its device address is taken once per device (hipGetSymbolAddress) and only
// that address is handed to kernels and copies.  A call queues its work on
// the caller's stream and returns; calls that overlap on one device share the
// array and race, so a caller serialises them (one stream, or a wait between).
#include <mutex>

#include "gsofa.h"
#include "gsofa_internal.h"

namespace gsofa {

namespace {

static __device__ __align__(256) unsigned char scratch[GSOFA_NOALLOC_SCRATCH_BYTES];

constexpr int kSlots = 8;        // sources per workgroup: the smallest per-vertex footprint
constexpr int kMaxDevices = 64;

std::mutex g_mu;
void* g_ptr[kMaxDevices];        // device address of `scratch` on each device (null: not yet taken)
uint32_t g_resident[kMaxDevices];

__global__ void k_na_chunks(BlkChunk* c, uint32_t n, uint32_t nchunks) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nchunks; i += gridDim.x * blockDim.x) {
        const uint32_t first = n - 1u - i * (uint32_t)kSlots;  // deepest first
        c[i].first = first;
        c[i].count = first + 1u < (uint32_t)kSlots ? first + 1u : (uint32_t)kSlots;
        c[i].stride = 1u;
        c[i].mat = 0u;
    }
}

// Per-workgroup state of a fresh plan (offset just below the stale range,
// no mini-group run since the labels were set), the one-matrix window.
__global__ void k_na_state(uint32_t* bstate, uint32_t nb, uint32_t W, uint32_t* mat_base) {
    for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < nb; b += gridDim.x * blockDim.x) {
        bstate[2 * b] = 0xFFFFFFFFu - W;
        bstate[2 * b + 1] = 0u;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        mat_base[0] = 0u;
        mat_base[1] = W;
    }
}

__global__ void k_na_totals(const Ctrl* c, unsigned long long* out) {
    if (threadIdx.x == 0) {
        out[0] = c->fill_count;
        out[1] = c->te;
        out[2] = c->reach;
        out[3] = c->edge_checks;
    }
}

size_t round_up(size_t x) { return (x + 255) & ~(size_t)255; }

}  // namespace

}  // namespace gsofa

extern "C" int gsofa_counts_noalloc(int32_t n, int64_t nnz, const int32_t* d_rowptr, const int32_t* d_col,
                                    int64_t* d_row_fill, int64_t* d_row_te, int64_t* d_row_reach,
                                    uint64_t* d_totals, int32_t device, void* stream) {
    using namespace gsofa;
    if (n < 1 || nnz < 0 || !d_rowptr || (nnz > 0 && !d_col) || !d_row_fill || !d_row_te || !d_row_reach ||
        !d_totals || device < 0 || device >= kMaxDevices)
        return GSOFA_ERR_INVALID;
    if (n > kMaxN || nnz >= kMaxBlockNnz) return GSOFA_ERR_TOO_LARGE;
    if (hipSetDevice(device) != hipSuccess) {
        (void)hipGetLastError();
        return GSOFA_ERR_NO_DEVICE;
    }
    void* base = nullptr;
    uint32_t resident = 0;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (!g_ptr[device]) {  // set-up, once per device, with the device current
            void* p = nullptr;
            if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
                (void)hipGetLastError();
                return GSOFA_ERR_HIP;
            }
            uint32_t r = 0;
            if (resident_blocks(kSlots, device, &r) != GSOFA_OK) return GSOFA_ERR_HIP;
            g_ptr[device] = p;
            g_resident[device] = r;
        }
        base = g_ptr[device];
        resident = g_resident[device];
    }
    // carve the scratch: control block, work table, window, per-workgroup
    // offsets, then per-workgroup labels [W][8], masks [2][W] and queues [2][W]
    const uint32_t W = (uint32_t)n;
    const uint32_t nchunks = (W + kSlots - 1) / kSlots;
    size_t off = 0;
    const size_t o_ctrl = off;
    off += round_up(sizeof(Ctrl));
    const size_t o_chunks = off;
    off += round_up(sizeof(BlkChunk) * (size_t)nchunks);
    const size_t o_mat = off;
    off += round_up(2 * sizeof(uint32_t));
    const size_t o_bstate = off;
    const size_t per_wg = (size_t)W * (4 * kSlots + 8 + 32);
    const size_t fixed = off + round_up(2 * sizeof(uint32_t) * (size_t)std::max<uint32_t>(resident, 1));
    if (fixed + per_wg > (size_t)GSOFA_NOALLOC_SCRATCH_BYTES) return GSOFA_ERR_BUDGET;
    uint32_t nb = std::min<uint32_t>(std::max<uint32_t>(resident, 1), nchunks);
    nb = (uint32_t)std::min<size_t>(nb, ((size_t)GSOFA_NOALLOC_SCRATCH_BYTES - fixed) / per_wg);
    off += round_up(2 * sizeof(uint32_t) * (size_t)nb);
    const size_t o_lab = off;
    off += round_up(4 * (size_t)kSlots * W * nb);
    const size_t o_mask = off;
    off += round_up(8 * (size_t)W * nb);
    const size_t o_q = off;
    off += 32 * (size_t)W * nb;
    if (off > (size_t)GSOFA_NOALLOC_SCRATCH_BYTES) return GSOFA_ERR_BUDGET;
    char* s = static_cast<char*>(base);
    hipStream_t st = static_cast<hipStream_t>(stream);
    BlkParams bp{};
    bp.rp = d_rowptr;
    bp.ci = d_col;
    bp.n = n;
    bp.W = W;
    bp.chunks = reinterpret_cast<BlkChunk*>(s + o_chunks);
    bp.n_chunks = nchunks;
    bp.mat_base = reinterpret_cast<uint32_t*>(s + o_mat);
    bp.nblocks = nb;
    bp.lab = reinterpret_cast<uint32_t*>(s + o_lab);
    bp.mask = reinterpret_cast<uint32_t*>(s + o_mask);
    bp.qg = reinterpret_cast<uint4*>(s + o_q);
    bp.fills = nullptr;
    bp.fill_cap = 0;  // counts only
    bp.ctrl = reinterpret_cast<Ctrl*>(s + o_ctrl);
    bp.bstate = reinterpret_cast<uint32_t*>(s + o_bstate);
    bp.row_te = reinterpret_cast<unsigned long long*>(d_row_te);
    bp.row_reach = reinterpret_cast<unsigned long long*>(d_row_reach);
    bp.row_fills = reinterpret_cast<unsigned long long*>(d_row_fill);
    bp.loops = 0xFFFFFFFFu / W;
    // the array keeps the last call's contents: clear what must start fresh
    // (labels stale, masks empty, counters zero) with kernels
    bool ok = launch_fill(reinterpret_cast<uint32_t*>(bp.ctrl), sizeof(Ctrl) / 4, 0u, st) == hipSuccess;
    ok = ok && launch_fill(bp.lab, (uint64_t)kSlots * W * nb, 0xFFFFFFFFu, st) == hipSuccess;
    ok = ok && launch_fill(bp.mask, 2ull * W * nb, 0u, st) == hipSuccess;
    if (ok) {
        hipLaunchKernelGGL(k_na_chunks, dim3((nchunks + 255) / 256), dim3(256), 0, st,
                           const_cast<BlkChunk*>(bp.chunks), W, nchunks);
        hipLaunchKernelGGL(k_na_state, dim3((nb + 255) / 256), dim3(256), 0, st, bp.bstate, nb, W,
                           const_cast<uint32_t*>(bp.mat_base));
        ok = hipGetLastError() == hipSuccess;
    }
    ok = ok && launch_traverse_blk(bp, kSlots, 0, st) == hipSuccess;
    if (ok) {
        hipLaunchKernelGGL(k_na_totals, dim3(1), dim3(64), 0, st, bp.ctrl,
                           reinterpret_cast<unsigned long long*>(d_totals));
        ok = hipGetLastError() == hipSuccess;
    }
    if (!ok) {
        (void)hipGetLastError();
        return GSOFA_ERR_HIP;
    }
    return GSOFA_OK;
}
