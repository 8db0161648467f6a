// gsofa_struct.hip -- device kernels that turn the traversal's fill stream into
// sorted per-row L and U structure, plus the supernode pass and small helpers
// (scans, fills).
//
// The reference never materialises L/U (its markers are transient,
// with_space_optimization/symbfact.cuh:2126) and has only a stub for
// supernodes (with_space_optimization/with_supernodes/symbfact.cuh:1087-1101);
// the outputs here are defined by the fill-path sets themselves.
#include "gsofa_device.h"
#include "gsofa_internal.h"

namespace gsofa {

namespace {

constexpr int SB = 256;          // threads per block for the helpers
constexpr int SCAN_ITEMS = 8;    // items per thread in a scan block
constexpr int SCAN_TILE = SB * SCAN_ITEMS;
constexpr int SORT_LDS = 4096;   // longest segment sorted by LDS bitonic
constexpr int BM_WORDS = 8192;   // 32 KB LDS bitmap for long segments
constexpr uint32_t BM_BITS = BM_WORDS * 32u;
constexpr int WSORT = 1024;      // longest segment sorted by one wave
constexpr int LTILE = 16;        // rows scanned per step for longer segments
using dev::WAVE;

// Fill-stream tiles: FT threads take FK entries each (FTILE per tile) and fold
// them into an LDS table of FH slots keyed by (row, part), at most half full.
constexpr int FT = 512;
constexpr int FK = 4;
constexpr int FTILE = FT * FK;
constexpr int FH = 2 * FTILE;
constexpr int FH_LOG2 = 12;
constexpr int FGRID = 1024;      // persistent blocks (LDS allows 2-4 per CU)
static_assert(FH == 1 << FH_LOG2, "table slots are a power of two");
static_assert(FTILE <= 65536 && FH <= 65536, "slots and ranks pack into 16 bits each");

__global__ void k_fill_u32(uint32_t* p, uint64_t count, uint32_t v) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void k_add_i32(int32_t* p, uint64_t count, int32_t d) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] += d;
}


__device__ __forceinline__ bool row_processed(int32_t s, int32_t n_gpu, int32_t gpu_id) {
    return gpu_id < 0 || (s % n_gpu) == gpu_id;
}

// Per processed row: A's strictly-lower / strictly-upper entry counts, and
// minus their sum in row_fill (k_row_fill adds the final counts: the fills).
__global__ void k_count_rows(FinalParams f) {
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < f.n;
         s += (int64_t)gridDim.x * blockDim.x) {
        uint32_t l = 0, u = 0;
        const bool proc = row_processed((int32_t)s, f.n_gpu, f.gpu_id);
        if (proc) {
            for (int32_t k = f.rp[s]; k < f.rp[s + 1]; k++) {
                const int32_t j = f.ci[k];
                l += (j < s);
                u += (j > s);
            }
        }
        f.cntL[s] = l;
        f.cntU[s] = u;
        f.row_fill[s] = proc ? -(int64_t)(l + u) : -1;
    }
}

__global__ void k_row_fill(FinalParams f) {
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < f.n;
         s += (int64_t)gridDim.x * blockDim.x)
        if (row_processed((int32_t)s, f.n_gpu, f.gpu_id)) f.row_fill[s] += (int64_t)f.cntL[s] + f.cntU[s];
}

// Fill-stream entries to read: the host's count, or the traversal's append
// position left in device memory (capped at the stream's length).  Entries
// a workgroup reserved and left unused hold ~0 (row >= n) and are skipped.
__device__ __forceinline__ uint64_t fills_in_stream(const FinalParams& f) {
    if (!f.n_fills_dev) return f.n_fills;
    const uint64_t pos = *f.n_fills_dev;
    return pos < f.n_fills ? pos : f.n_fills;
}

// The LDS table of one stream tile: key (row << 1 | part) + 1, 0 when empty
// (part 0: L, column below the row; part 1: U), and the slots in use.
struct FTab {
    uint32_t key[FH];
    uint32_t cnt[FH];
    uint16_t used[FTILE];
    uint32_t nused[2];  // by tile parity: one is counted while the other is reset
};

__device__ __forceinline__ void ftab_init(FTab& T) {
    for (uint32_t i = threadIdx.x; i < (uint32_t)FH; i += blockDim.x) {
        T.key[i] = 0u;
        T.cnt[i] = 0u;
    }
    if (threadIdx.x < 2) T.nused[threadIdx.x] = 0u;
}

// The slot of key, inserted if absent.  A tile holds at most FTILE distinct
// keys and the table FH = 2 * FTILE slots, cleared after every tile, so the
// probe always ends.
__device__ __forceinline__ uint32_t ftab_slot(FTab& T, uint32_t key, uint32_t par) {
    uint32_t s = (key * 2654435761u) >> (32 - FH_LOG2);
    for (;;) {
        const uint32_t cur = T.key[s];
        if (cur == key) return s;
        if (cur == 0u) {
            const uint32_t prev = atomicCAS(&T.key[s], 0u, key);
            if (prev == 0u) {
                T.used[atomicAdd(&T.nused[par], 1u)] = (uint16_t)s;
                return s;
            }
            if (prev == key) return s;
        }
        s = (s + 1u) & (uint32_t)(FH - 1);
    }
}

// Key of a stream entry, or 0 for the ~0 entries a workgroup reserved and
// left unused (row >= n).
__device__ __forceinline__ uint32_t fill_key(uint64_t e, int32_t n) {
    const uint32_t row = (uint32_t)(e >> 32), col = (uint32_t)e;
    if (row >= (uint32_t)n) return 0u;
    return ((row << 1) | (col < row ? 0u : 1u)) + 1u;
}

// Per-row L and U fill counts.  Neighbouring stream entries share rows (one
// source's fills over a relax batch, one target's fills over its slots), so
// a tile's entries are counted per (row, part) in LDS first: one global
// atomic per distinct key per tile instead of one per entry.
__global__ __launch_bounds__(FT) void k_count_fills(FinalParams f) {
    __shared__ FTab T;
    const uint32_t t = threadIdx.x;
    ftab_init(T);
    __syncthreads();
    const uint64_t nf = fills_in_stream(f);
    uint32_t par = 0;
    for (uint64_t base = (uint64_t)blockIdx.x * FTILE; base < nf; base += (uint64_t)gridDim.x * FTILE) {
        uint64_t e[FK];
#pragma unroll
        for (int k = 0; k < FK; k++) {
            const uint64_t i = base + (uint64_t)k * FT + t;
            e[k] = i < nf ? f.fills[i] : ~0ull;
        }
#pragma unroll
        for (int k = 0; k < FK; k++) {
            const uint32_t key = fill_key(e[k], f.n);
            if (key) atomicAdd(&T.cnt[ftab_slot(T, key, par)], 1u);
        }
        // the other parity's count was last read before the previous barrier
        if (t == 0) T.nused[par ^ 1u] = 0u;
        __syncthreads();
        const uint32_t nu = T.nused[par];
        for (uint32_t u = t; u < nu; u += FT) {
            const uint32_t sl = T.used[u];
            const uint32_t k1 = T.key[sl] - 1u, row = k1 >> 1;
            atomicAdd((k1 & 1u) ? &f.cntU[row] : &f.cntL[row], T.cnt[sl]);
            T.key[sl] = 0u;
            T.cnt[sl] = 0u;
        }
        __syncthreads();
        par ^= 1u;
    }
}

// A's off-diagonal entries of each processed row go first (one thread per row,
// so no atomics); cursors start at zero.
__global__ void k_scatter_rows(FinalParams f, int32_t* tmpL, int32_t* tmpU) {
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < f.n;
         s += (int64_t)gridDim.x * blockDim.x) {
        uint32_t l = 0, u = 0;
        if (row_processed((int32_t)s, f.n_gpu, f.gpu_id)) {
            const int64_t pl = f.L_rowptr[s], pu = f.U_rowptr[s];
            for (int32_t k = f.rp[s]; k < f.rp[s + 1]; k++) {
                const int32_t j = f.ci[k];
                if (j < s) {
                    if (pl + l < f.L_cap) tmpL[pl + l] = j;
                    l++;
                } else if (j > s) {
                    if (pu + u < f.U_cap) tmpU[pu + u] = j;
                    u++;
                }
            }
        }
        f.cntL[s] = l;
        f.cntU[s] = u;
    }
}

// Fill entries after A's: a tile's entries take local ranks per (row, part)
// in LDS, one global atomic per distinct key reserves the run in the row, and
// each entry lands at the run's start plus its rank.  Order within a row is
// arbitrary (the segment sort follows).
__global__ __launch_bounds__(FT) void k_scatter_fills(FinalParams f, int32_t* tmpL, int32_t* tmpU) {
    __shared__ FTab T;
    __shared__ int64_t rbase[FH];  // start of a slot's run, written after the tile's first barrier
    const uint32_t t = threadIdx.x;
    ftab_init(T);
    __syncthreads();
    const uint64_t nf = fills_in_stream(f);
    uint32_t par = 0;
    for (uint64_t base = (uint64_t)blockIdx.x * FTILE; base < nf; base += (uint64_t)gridDim.x * FTILE) {
        uint64_t e[FK];
        uint32_t sr[FK];  // slot << 16 | rank, or ~0 for no entry
#pragma unroll
        for (int k = 0; k < FK; k++) {
            const uint64_t i = base + (uint64_t)k * FT + t;
            e[k] = i < nf ? f.fills[i] : ~0ull;
        }
#pragma unroll
        for (int k = 0; k < FK; k++) {
            const uint32_t key = fill_key(e[k], f.n);
            sr[k] = ~0u;
            if (key) {
                const uint32_t sl = ftab_slot(T, key, par);
                sr[k] = (sl << 16) | atomicAdd(&T.cnt[sl], 1u);
            }
        }
        if (t == 0) T.nused[par ^ 1u] = 0u;
        // also: every entry of the previous tile has read rbase
        __syncthreads();
        const uint32_t nu = T.nused[par];
        for (uint32_t u = t; u < nu; u += FT) {
            const uint32_t sl = T.used[u];
            const uint32_t k1 = T.key[sl] - 1u, row = k1 >> 1;
            const uint32_t c = T.cnt[sl];
            rbase[sl] = (k1 & 1u) ? f.U_rowptr[row] + (int64_t)atomicAdd(&f.cntU[row], c)
                                  : f.L_rowptr[row] + (int64_t)atomicAdd(&f.cntL[row], c);
            T.key[sl] = 0u;
            T.cnt[sl] = 0u;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < FK; k++) {
            if (sr[k] == ~0u) continue;
            const uint32_t row = (uint32_t)(e[k] >> 32), col = (uint32_t)e[k];
            const int64_t q = rbase[sr[k] >> 16] + (int64_t)(sr[k] & 0xffffu);
            if (col < row) {
                if (q < f.L_cap) tmpL[q] = (int32_t)col;
            } else {
                if (q < f.U_cap) tmpU[q] = (int32_t)col;
            }
        }
        par ^= 1u;
    }
}

__device__ __forceinline__ uint32_t block_excl_scan_u32(uint32_t x, uint32_t* sh, uint32_t* total) {
    // sh: SB words
    const int t = threadIdx.x;
    sh[t] = x;
    __syncthreads();
    for (int d = 1; d < SB; d <<= 1) {
        uint32_t y = t >= d ? sh[t - d] : 0u;
        __syncthreads();
        sh[t] += y;
        __syncthreads();
    }
    uint32_t incl = sh[t];
    *total = sh[SB - 1];
    __syncthreads();
    return incl - x;
}

// Segments of one to WSORT entries, one wave each, no workgroup barriers:
// up to 64 entries by a bitonic network over lanes (shuffles), longer ones by
// a bitonic network over the wave's own LDS buffer.  Values within a segment
// are distinct and below INT32_MAX, which pads.
__global__ __launch_bounds__(SB) void k_sort_waves(const int64_t* rowptr, const int32_t* src, int32_t* dst,
                                                   int32_t n, int64_t cap) {
    constexpr int WPB = SB / WAVE;
    __shared__ int32_t buf[WPB][WSORT];
    const int lane = (int)__lane_id();
    const int w = (int)__builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);
    int32_t* s = buf[w];
    for (int64_t r = (int64_t)blockIdx.x * WPB + w; r < n; r += (int64_t)gridDim.x * WPB) {
        const int64_t b = rowptr[r], e = rowptr[r + 1];
        const int64_t len = e - b;
        if (len <= 0 || b < 0 || e > cap || len > WSORT) continue;
        if (len <= WAVE) {
            int32_t x = lane < len ? src[b + lane] : INT32_MAX;
            for (int k = 2; k <= WAVE; k <<= 1) {
                for (int j = k >> 1; j > 0; j >>= 1) {
                    const int32_t y = __shfl_xor(x, j);
                    const bool keep_min = ((lane & j) == 0) == ((lane & k) == 0);
                    x = keep_min ? (x < y ? x : y) : (x > y ? x : y);
                }
            }
            if (lane < len) dst[b + lane] = x;
            continue;
        }
        int P = 2 * WAVE;
        while (P < len) P <<= 1;
        for (int i = lane; i < P; i += WAVE) s[i] = i < len ? src[b + i] : INT32_MAX;
        dev::wave_sync();
        for (int k = 2; k <= P; k <<= 1) {
            for (int j = k >> 1; j > 0; j >>= 1) {
                for (int q = lane; q < P / 2; q += WAVE) {
                    const int i = ((q & ~(j - 1)) << 1) | (q & (j - 1));  // pair (i, i + j)
                    const int32_t a = s[i], c = s[i + j];
                    if ((a > c) == ((i & k) == 0)) {
                        s[i] = c;
                        s[i + j] = a;
                    }
                }
                dev::wave_sync();
            }
        }
        for (int i = lane; i < len; i += WAVE) dst[b + i] = s[i];
        dev::wave_sync();  // read out before the next segment overwrites the buffer
    }
}

// One segment longer than WSORT by the whole block.  Its value range comes
// first: a range of at most BM_WINDOWS windows of BM_BITS values (or a segment
// longer than SORT_LDS) is sorted by setting bits of an LDS bitmap per window
// (only the words the range covers) and emitting them in order; otherwise an
// LDS bitonic network sorts it.  Called by every thread of the block.
constexpr int64_t BM_WINDOWS = 4;
__device__ void block_sort_segment(int64_t b, int64_t len, const int32_t* src, int32_t* dst, int32_t* s,
                                   uint32_t* bm, uint32_t* sh, int32_t* red) {
    const int t = threadIdx.x;
    if (t == 0) {
        red[0] = INT32_MAX;
        red[1] = INT32_MIN;
    }
    __syncthreads();
    int32_t lo = INT32_MAX, hi = INT32_MIN;
    for (int64_t i = t; i < len; i += SB) {
        const int32_t v = src[b + i];
        lo = v < lo ? v : lo;
        hi = v > hi ? v : hi;
    }
    atomicMin(&red[0], lo);
    atomicMax(&red[1], hi);
    __syncthreads();
    lo = red[0];
    hi = red[1];
    __syncthreads();  // red is read before the next segment resets it
    const int64_t range = (int64_t)hi - lo + 1;
    if (len <= SORT_LDS && range > BM_WINDOWS * (int64_t)BM_BITS) {
        int P = 2;
        while (P < len) P <<= 1;
        for (int i = t; i < P; i += SB) s[i] = i < len ? src[b + i] : INT32_MAX;
        __syncthreads();
        for (int k = 2; k <= P; k <<= 1) {
            for (int j = k >> 1; j > 0; j >>= 1) {
                for (int i = t; i < P; i += SB) {
                    const int ixj = i ^ j;
                    if (ixj > i) {
                        const int32_t a = s[i], c = s[ixj];
                        const bool up = (i & k) == 0;
                        if ((a > c) == up) {
                            s[i] = c;
                            s[ixj] = a;
                        }
                    }
                }
                __syncthreads();
            }
        }
        for (int i = t; i < len; i += SB) dst[b + i] = s[i];
        __syncthreads();
        return;
    }
    int64_t written = 0;
    for (int64_t wb = lo; wb <= hi; wb += BM_BITS) {
        const int64_t span = (int64_t)hi - wb + 1;
        const int nwords = span >= (int64_t)BM_BITS ? BM_WORDS : (int)((span + 31) >> 5);
        for (int i = t; i < nwords; i += SB) bm[i] = 0u;
        __syncthreads();
        for (int64_t i = t; i < len; i += SB) {
            const int64_t v = src[b + i];
            if (v >= wb && v < wb + (int64_t)BM_BITS) {
                const uint32_t o = (uint32_t)(v - wb);
                atomicOr(&bm[o >> 5], 1u << (o & 31u));
            }
        }
        __syncthreads();
        // thread t takes words [w0, w1), consecutive so that its bits follow
        // the bits of threads below it
        const int wpt = (nwords + SB - 1) / SB;
        const int w0 = t * wpt < nwords ? t * wpt : nwords;
        const int w1 = w0 + wpt < nwords ? w0 + wpt : nwords;
        uint32_t cnt = 0;
        for (int w = w0; w < w1; w++) cnt += (uint32_t)__popc(bm[w]);
        uint32_t total;
        const uint32_t pos = block_excl_scan_u32(cnt, sh, &total);
        int64_t o = b + written + pos;
        for (int w = w0; w < w1; w++) {
            uint32_t bits = bm[w];
            while (bits) {
                const uint32_t bp = (uint32_t)__ffs(bits) - 1u;
                bits &= bits - 1u;
                dst[o++] = (int32_t)(wb + (int64_t)w * 32 + bp);
            }
        }
        written += total;
        __syncthreads();
    }
}

// Segments longer than WSORT: the first wave marks them in tiles of LTILE
// rows (small, so that runs of long rows spread over blocks), the block sorts
// each.  Rows ending beyond cap (entries allocated in src and
// dst) are left alone here and in k_sort_waves, so a row-pointer total above
// the allocation cannot write out of bounds (the host reports that case as an
// error).
__global__ __launch_bounds__(SB) void k_sort_long(const int64_t* rowptr, const int32_t* src, int32_t* dst,
                                                  int32_t n, int64_t cap) {
    static_assert(SORT_LDS <= BM_WORDS, "the bitonic buffer shares the bitmap's words");
    __shared__ uint32_t bm[BM_WORDS];  // bitmap windows, or the bitonic buffer
    __shared__ uint32_t sh[SB];
    __shared__ int32_t red[2];
    __shared__ uint64_t longs;
    int32_t* s = reinterpret_cast<int32_t*>(bm);
    const int t = threadIdx.x;
    for (int64_t r0 = (int64_t)blockIdx.x * LTILE; r0 < n; r0 += (int64_t)gridDim.x * LTILE) {
        if (t < WAVE) {
            const int64_t r = r0 + t;
            bool lng = false;
            if (t < LTILE && r < n) {
                const int64_t b = rowptr[r], e = rowptr[r + 1];
                lng = e - b > WSORT && b >= 0 && e <= cap;
            }
            const uint64_t m = __ballot(lng);
            if (t == 0) longs = m;
        }
        __syncthreads();
        uint64_t m = longs;
        while (m) {
            const int64_t r = r0 + (__ffsll((unsigned long long)m) - 1);
            m &= m - 1ull;
            const int64_t b = rowptr[r];
            block_sort_segment(b, rowptr[r + 1] - b, src, dst, s, bm, sh, red);
        }
        __syncthreads();  // every thread has read longs
    }
}

// ---- scans -----------------------------------------------------------------
struct OpAdd {
    __device__ static int64_t apply(int64_t a, int64_t b) { return a + b; }
    __device__ static int64_t identity() { return 0; }
};
struct OpMax {
    __device__ static int64_t apply(int64_t a, int64_t b) { return a > b ? a : b; }
    __device__ static int64_t identity() { return INT64_MIN; }
};

template <class Op>
__global__ __launch_bounds__(SB) void k_scan_tiles(int64_t* a, int64_t len, int64_t* sums) {
    __shared__ int64_t sh[SB];
    const int t = threadIdx.x;
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)t * SCAN_ITEMS;
    int64_t v[SCAN_ITEMS];
    int64_t acc = Op::identity();
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; i++) {
        const int64_t idx = base + i;
        v[i] = idx < len ? a[idx] : Op::identity();
        acc = Op::apply(acc, v[i]);
        v[i] = acc;
    }
    sh[t] = acc;
    __syncthreads();
    for (int d = 1; d < SB; d <<= 1) {
        int64_t y = t >= d ? sh[t - d] : Op::identity();
        __syncthreads();
        sh[t] = Op::apply(sh[t], y);
        __syncthreads();
    }
    const int64_t prefix = t > 0 ? sh[t - 1] : Op::identity();
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; i++) {
        const int64_t idx = base + i;
        if (idx < len) a[idx] = Op::apply(prefix, v[i]);
    }
    if (t == SB - 1 && sums) sums[blockIdx.x] = sh[SB - 1];
}

template <class Op>
__global__ void k_scan_add(int64_t* a, int64_t len, const int64_t* sums) {
    const int64_t tile = blockIdx.x + 1;
    const int64_t add = sums[tile - 1];
    for (int64_t i = tile * SCAN_TILE + threadIdx.x; i < (tile + 1) * SCAN_TILE && i < len; i += SB)
        a[i] = Op::apply(add, a[i]);
}

template <class Op>
hipError_t inclusive_scan(int64_t* a, int64_t len, int64_t* tmp, hipStream_t s) {
    if (len <= 0) return hipSuccess;
    const int64_t tiles = (len + SCAN_TILE - 1) / SCAN_TILE;
    hipLaunchKernelGGL(k_scan_tiles<Op>, dim3((uint32_t)tiles), dim3(SB), 0, s, a, len,
                       tiles > 1 ? tmp : nullptr);
    if (tiles > 1) {
        hipError_t e = inclusive_scan<Op>(tmp, tiles, tmp + tiles, s);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_scan_add<Op>, dim3((uint32_t)(tiles - 1)), dim3(SB), 0, s, a, len, tmp);
    }
    return hipGetLastError();
}

__global__ void k_u32_to_i64_shift(const uint32_t* in, int64_t* out, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        out[i + 1] = in[i];
    if (blockIdx.x == 0 && threadIdx.x == 0) out[0] = 0;
}

// ---- supernodes --------------------------------------------------------------
// Row i joins row i-1's supernode iff (i-1) is the last entry of L_i (i.e. the
// traversal from i passes through i-1), |U_i| = |U_{i-1}| - 1, and the
// supernode holds fewer than max_size rows.  The cap splits a maximal run of
// joining rows into consecutive pieces of max_size rows.
__global__ void k_sn_flags(SuperParams p, int64_t* key) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < p.n;
         i += (int64_t)gridDim.x * blockDim.x) {
        bool join = false;
        if (i > 0) {
            bool pass;
            if (p.pass) {
                pass = p.pass[i] != 0u;
            } else {
                const int64_t lb = p.L_rowptr[i], le = p.L_rowptr[i + 1];
                pass = le > lb && p.L_col[le - 1] == (int32_t)(i - 1);
            }
            const int64_t ui = p.U_rowptr[i + 1] - p.U_rowptr[i];
            const int64_t up = p.U_rowptr[i] - p.U_rowptr[i - 1];
            join = pass && (ui == up - 1);
        }
        p.flag[i] = join ? 1u : 0u;
        key[i] = join ? 0 : i;  // run start via inclusive max-scan
    }
}

__global__ void k_sn_starts(SuperParams p, const int64_t* runstart, int64_t* cnt) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < p.n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const bool join = p.flag[i] != 0u;
        const int64_t pos = i - runstart[i];
        const bool start = !join || (p.max_size > 0 && (pos % p.max_size) == 0);
        p.flag[i] = start ? 1u : 0u;
        cnt[i + 1] = start ? 1 : 0;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) cnt[0] = 0;
}

__global__ void k_sn_write(SuperParams p, const int64_t* excl) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < p.n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t id = excl[i] + (int64_t)p.flag[i] - 1;
        p.supno[i] = (int32_t)id;
        if (p.flag[i]) p.super_start[excl[i]] = (int32_t)i;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const int32_t ns = (int32_t)excl[p.n];
        *p.n_super = ns;
        p.super_start[ns] = p.n;
    }
}

inline dim3 grid_for(int64_t work, int per = SB, int cap = 65536) {
    int64_t g = (work + per - 1) / per;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return dim3((uint32_t)g);
}

// Rows of one share: their L and U lengths and whether row i-1 is in L_i (the
// last, largest column of the sorted row), kept after the share's columns
// leave the device.
__global__ void k_spill_rows(const int64_t* Lrp, const int32_t* Lcol, const int64_t* Urp, int32_t n,
                             int32_t n_gpu, int32_t gpu_id, uint32_t* lcnt, uint32_t* ucnt, uint32_t* pass) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        if (!row_processed((int32_t)i, n_gpu, gpu_id)) continue;
        const int64_t lb = Lrp[i], le = Lrp[i + 1];
        lcnt[i] = (uint32_t)(le - lb);
        ucnt[i] = (uint32_t)(Urp[i + 1] - Urp[i]);
        pass[i] = (i > 0 && le > lb && Lcol[le - 1] == (int32_t)(i - 1)) ? 1u : 0u;
    }
}

// Column counts of a row-major structure: one increment per entry.
__global__ void k_col_count(const int32_t* col, int64_t nnz, uint32_t* cnt) {
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nnz;
         k += (int64_t)gridDim.x * blockDim.x)
        atomicAdd(&cnt[col[k]], 1u);
}

// One wave per row: entry (r, c) takes the next free place of column c.  The
// order within a column is arbitrary until the segment sort.
__global__ void k_col_scatter(const int64_t* rowptr, const int32_t* col, int32_t n, const int64_t* colptr,
                              uint32_t* cursor, int32_t* rows) {
    const int lane = (int)(threadIdx.x % 64u);
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x / 64u);
    for (int64_t r = (int64_t)blockIdx.x * (blockDim.x / 64u) + threadIdx.x / 64u; r < n; r += waves) {
        const int64_t e = rowptr[r + 1];
        for (int64_t k = rowptr[r] + lane; k < e; k += 64) {
            const int32_t c = col[k];
            rows[colptr[c] + atomicAdd(&cursor[c], 1u)] = (int32_t)r;
        }
    }
}

// Subscript count of supernode s: its first row, then that row's U columns.
__global__ void k_sub_count(const int32_t* sstart, int32_t ns, const int64_t* Urp, uint32_t* cnt) {
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < ns;
         s += (int64_t)gridDim.x * blockDim.x) {
        const int32_t f = sstart[s];
        cnt[s] = 1u + (uint32_t)(Urp[f + 1] - Urp[f]);
    }
}

// One wave per supernode: sub[xsub[s]..] = f, U columns of row f (sorted).
__global__ void k_sub_write(const int32_t* sstart, int32_t ns, const int64_t* Urp, const int32_t* Ucol,
                            const int64_t* xsub, int32_t* sub) {
    const int lane = (int)(threadIdx.x % 64u);
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x / 64u);
    for (int64_t s = (int64_t)blockIdx.x * (blockDim.x / 64u) + threadIdx.x / 64u; s < ns; s += waves) {
        const int32_t f = sstart[s];
        const int64_t o = xsub[s], b = Urp[f], len = Urp[f + 1] - b;
        if (lane == 0) sub[o] = f;
        for (int64_t k = lane; k < len; k += 64) sub[o + 1 + k] = Ucol[b + k];
    }
}

}  // namespace

hipError_t spill_rows(const int64_t* Lrp, const int32_t* Lcol, const int64_t* Urp, int32_t n, int32_t n_gpu,
                      int32_t gpu_id, uint32_t* lcnt, uint32_t* ucnt, uint32_t* pass, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_spill_rows, grid_for(n), dim3(SB), 0, s, Lrp, Lcol, Urp, n, n_gpu, gpu_id, lcnt, ucnt, pass);
    return hipGetLastError();
}

// Column-major copy of a row-major structure (n x n, nnz entries): colptr
// (n+1) and the row indices of each column in increasing order.  cnt (n) and
// tmp (nnz) are scratch; scan_tmp holds scan_tmp_words(n) words.
hipError_t transpose_structure(const int64_t* rowptr, const int32_t* col, int64_t nnz, int32_t n,
                               uint32_t* cnt, int64_t* colptr, int64_t* scan_tmp, int32_t* tmp,
                               int32_t* rows, hipStream_t s) {
    if (n <= 0) return hipMemsetAsync(colptr, 0, sizeof(int64_t), s);
    hipError_t e = launch_fill(cnt, (uint64_t)n, 0u, s);
    if (e != hipSuccess) return e;
    if (nnz > 0) hipLaunchKernelGGL(k_col_count, grid_for(nnz), dim3(SB), 0, s, col, nnz, cnt);
    e = exclusive_scan_u32_to_i64(cnt, colptr, n, scan_tmp, s);
    if (e != hipSuccess) return e;
    if (nnz == 0) return hipGetLastError();
    e = launch_fill(cnt, (uint64_t)n, 0u, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_col_scatter, grid_for(n, SB / 64), dim3(SB), 0, s, rowptr, col, n, colptr, cnt, tmp);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launch_sort_segments(colptr, tmp, rows, n, nnz, s);
}

// Per-supernode subscripts: counts into cnt (ns), then (after the caller's
// exclusive scan into xsub) the lists themselves.
hipError_t supernode_sub_count(const int32_t* sstart, int32_t ns, const int64_t* Urp, uint32_t* cnt,
                               hipStream_t s) {
    if (ns <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_sub_count, grid_for(ns), dim3(SB), 0, s, sstart, ns, Urp, cnt);
    return hipGetLastError();
}
hipError_t supernode_sub_write(const int32_t* sstart, int32_t ns, const int64_t* Urp, const int32_t* Ucol,
                               const int64_t* xsub, int32_t* sub, hipStream_t s) {
    if (ns <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_sub_write, grid_for(ns, SB / 64), dim3(SB), 0, s, sstart, ns, Urp, Ucol, xsub, sub);
    return hipGetLastError();
}

hipError_t launch_add(int32_t* p, uint64_t count, int32_t delta, hipStream_t s) {
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(k_add_i32, grid_for((int64_t)count), dim3(SB), 0, s, p, count, delta);
    return hipGetLastError();
}

hipError_t launch_fill(uint32_t* p, uint64_t count, uint32_t value, hipStream_t s) {
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_u32, grid_for((int64_t)count), dim3(SB), 0, s, p, count, value);
    return hipGetLastError();
}

size_t scan_tmp_words(int64_t n) {
    size_t words = 0;
    int64_t len = n + 1;
    while (len > 1) {
        len = (len + SCAN_TILE - 1) / SCAN_TILE;
        words += (size_t)len;
    }
    return words + 16;
}

hipError_t exclusive_scan_u32_to_i64(const uint32_t* in, int64_t* out, int64_t n, int64_t* tmp,
                                     hipStream_t s) {
    hipLaunchKernelGGL(k_u32_to_i64_shift, grid_for(n), dim3(SB), 0, s, in, out, n);
    return inclusive_scan<OpAdd>(out + 1, n, tmp, s);
}

hipError_t launch_count(const FinalParams& f, hipStream_t s) {
    hipLaunchKernelGGL(k_count_rows, grid_for(f.n), dim3(SB), 0, s, f);
    if (f.n_fills)
        hipLaunchKernelGGL(k_count_fills, grid_for((int64_t)f.n_fills, FTILE, FGRID), dim3(FT), 0, s, f);
    hipLaunchKernelGGL(k_row_fill, grid_for(f.n), dim3(SB), 0, s, f);
    return hipGetLastError();
}

// tmp columns live in L_col/U_col's twin buffers passed through FinalParams:
// the caller passes the final arrays in L_col/U_col and scratch via cntL-sized
// buffers; see gsofa_api.hip for the buffer rotation.
hipError_t launch_scatter(const FinalParams& f, hipStream_t s) {
    // f.L_col / f.U_col here are the *unsorted* staging arrays
    hipLaunchKernelGGL(k_scatter_rows, grid_for(f.n), dim3(SB), 0, s, f, f.L_col, f.U_col);
    if (f.n_fills)
        hipLaunchKernelGGL(k_scatter_fills, grid_for((int64_t)f.n_fills, FTILE, FGRID), dim3(FT), 0, s, f,
                           f.L_col, f.U_col);
    return hipGetLastError();
}

hipError_t launch_sort_segments(const int64_t* rowptr, const int32_t* src, int32_t* dst, int32_t n, int64_t cap,
                                hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_sort_waves, grid_for(n, SB / WAVE, 8192), dim3(SB), 0, s, rowptr, src, dst, n, cap);
    hipLaunchKernelGGL(k_sort_long, grid_for(n, LTILE, 8192), dim3(SB), 0, s, rowptr, src, dst, n, cap);
    return hipGetLastError();
}

hipError_t launch_supernodes(const SuperParams& p, hipStream_t s) {
    if (p.n <= 0) {
        hipLaunchKernelGGL(k_sn_write, dim3(1), dim3(SB), 0, s, p, p.tmp64);
        return hipGetLastError();
    }
    int64_t* key = p.tmp64;             // n
    int64_t* cnt = p.tmp64 + p.n;       // n+1
    int64_t* scratch = cnt + p.n + 1;   // scan temporaries
    hipLaunchKernelGGL(k_sn_flags, grid_for(p.n), dim3(SB), 0, s, p, key);
    hipError_t e = inclusive_scan<OpMax>(key, p.n, scratch, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_sn_starts, grid_for(p.n), dim3(SB), 0, s, p, key, cnt);
    e = inclusive_scan<OpAdd>(cnt + 1, p.n, scratch, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_sn_write, grid_for(p.n), dim3(SB), 0, s, p, cnt);
    return hipGetLastError();
}

}  // namespace gsofa
