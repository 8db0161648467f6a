// gsofa_block.hip -- block-resident multi-source traversal (default mode).
//
// Same per-source arithmetic as the reference's warp-centric level kernel
// (with_space_optimization/symbfact.cuh:883-1082) and seeding (:639-721), with
// a different execution model for CDNA4: each workgroup of a persistent,
// occupancy-bounded grid owns S source slots and runs a "mini-group" of S
// sources through all of its frontier levels by itself, synchronising only
// with __syncthreads; the next mini-group is taken from one global counter.
// There is no grid-wide barrier and no hot global counter on the level path.
//
// * Sources are taken in decreasing order so that the deepest traversals start
//   first and the tail of the run consists of short ones.  Marker stamps are
//   stamp_base + (n - s), so they still increase per slot (no marker reset,
//   with_space_optimization/symbfact.cuh:1006).
// * Max-id labels are per slot with the no-reset offset of
//   with_space_optimization/symbfact.cuh:2388-2407, kept per workgroup.
// * A frontier entry carries the label it was relaxed with, so popping needs no
//   fresh label read; when a (possibly stale, never too low) cached read of the
//   label is already lower, a better entry for the same vertex was enqueued by
//   whoever lowered it, and the entry is dropped.
// * Frontier queues are two-tier: the first ql_cap entries of each level live
//   in LDS, the rest in a per-workgroup global region; beyond that, entries are
//   parked in a per-slot overflow bitmap and re-enqueued when the queue drains.
// * Wave64 merge-path expansion: a wave takes 64 entries, scans their row
//   lengths and streams the concatenated CSR words across its lanes.
#include "gsofa_device.h"
#include "gsofa_internal.h"

namespace gsofa {

namespace {

using namespace dev;

constexpr int BB = 512;           // threads per workgroup
constexpr int BW = BB / WAVE;     // waves per workgroup
constexpr int FB = 128;           // staged fills per wave
constexpr int BIGCAP = 32;        // workgroup-expanded rows per level
constexpr uint32_t VBITS = 25;    // vertex / relative-label field width
constexpr uint32_t VMASK = (1u << VBITS) - 1u;

struct BigRowB {
    int32_t beg, end;
    uint32_t cost, slot;
};

struct BlkLds {
    uint32_t incl[BW][WAVE];
    int32_t beg[BW][WAVE];
    uint32_t cost[BW][WAVE];
    uint32_t slot[BW][WAVE];
    uint64_t fbuf[BW][FB];
    BigRowB big[BIGCAP];
    uint32_t src[kMaxSlots];
    uint32_t qcnt[2];
    uint32_t cursor;
    uint32_t nbig;
    uint32_t orph;
    uint32_t base;
};

struct Blk {
    const BlkParams* p;
    uint32_t* lab0;
    uint32_t* mark0;
    uint32_t* orph0;
    uint64_t* qg0;   // 2 x qg_cap
    uint64_t* ql;    // 2 x ql_cap (LDS)
    uint32_t qtot;
    uint32_t off, stale;
};

struct TallyB {
    unsigned long long fills, te, reach, checks;
};

// entry: hi = slot << 25 | (label - off), lo = vertex
__device__ __forceinline__ uint64_t make_entry(uint32_t slot, uint32_t lab_rel, uint32_t v) {
    return ((uint64_t)((slot << VBITS) | lab_rel) << 32) | v;
}

__device__ __forceinline__ uint64_t q_read(const Blk& B, uint32_t buf, uint32_t i) {
    const uint32_t qc = B.p->ql_cap;
    return i < qc ? B.ql[buf * qc + i] : B.qg0[(uint64_t)buf * B.p->qg_cap + (i - qc)];
}

__device__ __forceinline__ void park(const Blk& B, BlkLds& S, uint64_t e) {
    const uint32_t v = (uint32_t)e;
    const uint32_t slot = (uint32_t)(e >> 32) >> VBITS;
    const uint32_t bit = 1u << (v & 31u);
    uint32_t* w = B.orph0 + (uint64_t)slot * B.p->ow + (v >> 5);
    const uint32_t old = __hip_atomic_fetch_or(w, bit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (!(old & bit)) atomicAdd(&S.orph, 1u);
}

__device__ __forceinline__ void q_write(const Blk& B, BlkLds& S, uint32_t buf, uint32_t i, uint64_t e) {
    const uint32_t qc = B.p->ql_cap;
    if (i < qc) B.ql[buf * qc + i] = e;
    else if (i < B.qtot) B.qg0[(uint64_t)buf * B.p->qg_cap + (i - qc)] = e;
    else park(B, S, e);
}

__device__ void flush_fills_w(const BlkParams& p, uint64_t* fb, uint32_t cnt) {
    wave_sync();
    const uint32_t lane = __lane_id();
    uint64_t pos = 0;
    if (lane == 0) pos = aadd64(&p.ctrl->fill_pos, cnt);
    pos = bcast0_64(pos);
    for (uint32_t t = lane; t < cnt; t += WAVE) {
        const uint64_t at = pos + t;
        if (at < p.fill_cap) p.fills[at] = fb[t];
    }
    wave_sync();
}

// Wave-uniform: enqueue lanes with push, stage lanes with fill.
__device__ __forceinline__ void emit(const Blk& B, BlkLds& S, uint64_t* fb, uint32_t& fcount,
                                     uint32_t nbuf, bool push, uint64_t entry, bool fill,
                                     uint64_t fentry) {
    const uint64_t pm = __ballot(push);
    if (pm) {
        uint32_t pos = 0;
        if (__lane_id() == 0) pos = atomicAdd(&S.qcnt[nbuf], (uint32_t)__popcll(pm));
        pos = bcast0(pos);
        if (push) q_write(B, S, nbuf, pos + rank_in(pm), entry);
    }
    const uint64_t fm = __ballot(fill);
    if (fm) {
        const uint32_t c = (uint32_t)__popcll(fm);
        if (fcount + c > (uint32_t)FB) {
            flush_fills_w(*B.p, fb, fcount);
            fcount = 0;
        }
        if (fill) fb[fcount + rank_in(fm)] = fentry;
        fcount += c;
    }
}

// One edge check (with_space_optimization/symbfact.cuh:989-1066).
// Returns bit0 = enqueue m with label `cost`, bit1 = new fill.
__device__ __forceinline__ uint32_t relax_b(const Blk& B, uint32_t m, uint32_t slot, uint32_t cost,
                                            uint32_t src, TallyB& t) {
    const BlkParams& p = *B.p;
    const uint32_t n = (uint32_t)p.n;
    const uint32_t stamp = p.stamp_base + (n - src);
    uint32_t* mark = B.mark0 + (uint64_t)slot * n;
    if (m > src) {
        if (ld_cached(mark + m) != stamp && amax(mark + m, stamp) < stamp) return 2u;
        return 0u;
    }
    uint32_t* lp = B.lab0 + (uint64_t)slot * n + m;
    if (ld_cached(lp) <= cost) return 0u;
    const uint32_t old = amin(lp, cost);
    if (old <= cost) return 0u;
    if (old >= B.stale) {  // first time m is reached from this source
        t.reach++;
        t.te += (uint32_t)(p.rp[m + 1] - p.rp[m]);
    }
    uint32_t r = (m < src) ? 1u : 0u;
    if (m + B.off > cost && amax(mark + m, stamp) < stamp) r |= 2u;
    return r;
}

__device__ void expand_strided_b(const Blk& B, BlkLds& S, uint64_t* fb, uint32_t& fcount,
                                 uint32_t nbuf, TallyB& t, int32_t beg, int32_t end,
                                 uint32_t cost, uint32_t slot, uint32_t lane0, uint32_t stride) {
    const uint32_t src = S.src[slot];
    for (int64_t e0 = beg; e0 < end; e0 += stride) {
        const int64_t e = e0 + lane0;
        uint32_t r = 0, m = 0;
        if (e < end) {
            m = (uint32_t)B.p->ci[e];
            t.checks++;
            r = relax_b(B, m, slot, cost, src, t);
            if (r & 2u) t.fills++;
        }
        emit(B, S, fb, fcount, nbuf, r & 1u, make_entry(slot, cost - B.off, m), (r & 2u) != 0,
             ((uint64_t)src << 32) | m);
    }
}

__device__ void expand_level_b(const Blk& B, BlkLds& S, uint64_t* fb, uint32_t& fcount,
                               uint32_t cbuf, uint32_t size, TallyB& t) {
    const BlkParams& p = *B.p;
    const uint32_t lane = __lane_id();
    const uint32_t w = threadIdx.x / WAVE;
    const uint32_t nbuf = cbuf ^ 1u;
    const uint32_t n = (uint32_t)p.n;
    const uint32_t T = (uint32_t)p.block_threshold;
    for (;;) {
        uint32_t i0 = 0;
        if (lane == 0) i0 = atomicAdd(&S.cursor, (uint32_t)WAVE);
        i0 = bcast0(i0);
        if (i0 >= size) break;
        const uint32_t i = i0 + lane;
        uint32_t deg = 0, cost = 0, slot = 0;
        int32_t b = 0;
        if (i < size) {
            const uint64_t e = q_read(B, cbuf, i);
            const uint32_t v = (uint32_t)e;
            const uint32_t hi = (uint32_t)(e >> 32);
            slot = hi >> VBITS;
            const uint32_t lab = B.off + (hi & VMASK);
            // a lower label means a better entry for v is already queued
            if (ld_cached(B.lab0 + (uint64_t)slot * n + v) >= lab) {
                cost = (v + B.off) > lab ? (v + B.off) : lab;  // symbfact.cuh:924
                b = p.rp[v];
                const int32_t en = p.rp[v + 1];
                deg = (uint32_t)(en - b);
                if (deg >= T) {
                    const uint32_t k = atomicAdd(&S.nbig, 1u);
                    if (k < (uint32_t)BIGCAP) {
                        S.big[k] = BigRowB{b, en, cost, slot};
                        deg = 0;
                    }
                }
            }
        }
        const uint32_t incl = wave_incl_scan(deg);
        const uint32_t total = __shfl(incl, WAVE - 1, WAVE);
        S.incl[w][lane] = incl;
        S.beg[w][lane] = b - (int32_t)(incl - deg);
        S.cost[w][lane] = cost;
        S.slot[w][lane] = slot;
        wave_sync();
        for (uint32_t e0 = 0; e0 < total; e0 += WAVE) {
            const uint32_t e = e0 + lane;
            uint32_t r = 0, m = 0, sl = 0, src = 0, cst = 0;
            if (e < total) {
                uint32_t lo = 0, hi = WAVE - 1;
                while (lo < hi) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (S.incl[w][mid] > e) hi = mid;
                    else lo = mid + 1;
                }
                m = (uint32_t)p.ci[S.beg[w][lo] + (int32_t)e];
                sl = S.slot[w][lo];
                cst = S.cost[w][lo];
                src = S.src[sl];
                t.checks++;
                r = relax_b(B, m, sl, cst, src, t);
                if (r & 2u) t.fills++;
            }
            emit(B, S, fb, fcount, nbuf, r & 1u, make_entry(sl, cst - B.off, m), (r & 2u) != 0,
                 ((uint64_t)src << 32) | m);
        }
        wave_sync();
    }
}

// Move parked entries from the overflow bitmap into queue `buf` while it has
// room (entries beyond the room stay parked).
__device__ void rescan_b(const Blk& B, BlkLds& S, uint32_t buf, uint32_t nsrc) {
    const BlkParams& p = *B.p;
    const uint64_t words = (uint64_t)nsrc * p.ow;
    uint32_t taken = 0;
    for (uint64_t wd = threadIdx.x; wd < words; wd += BB) {
        uint32_t bits = B.orph0[wd];
        if (!bits) continue;
        const uint32_t c = (uint32_t)__popc(bits);
        const uint32_t pos = atomicAdd(&S.qcnt[buf], c);
        const uint32_t fit = pos >= B.qtot ? 0u : ((B.qtot - pos) < c ? (B.qtot - pos) : c);
        const uint32_t slot = (uint32_t)(wd / p.ow);
        const uint32_t vb = (uint32_t)(wd % p.ow) * 32u;
        for (uint32_t k = 0; k < fit; k++) {
            const uint32_t bp = (uint32_t)__ffs(bits) - 1u;
            bits &= bits - 1u;
            const uint32_t v = vb + bp;
            // re-enqueue with the vertex's current label
            const uint32_t lab = ld_fresh(B.lab0 + (uint64_t)slot * p.n + v);
            q_write(B, S, buf, pos + k, make_entry(slot, lab - B.off, v));
        }
        B.orph0[wd] = bits;
        taken += fit;
    }
    if (taken) atomicSub(&S.orph, taken);
}

__global__ __launch_bounds__(BB) void k_traverse_blk(BlkParams p) {
    __shared__ BlkLds S;
    extern __shared__ uint64_t qlds[];
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = __lane_id();
    const uint32_t w = tid / WAVE;
    const uint32_t n = (uint32_t)p.n;
    const uint32_t N = (uint32_t)p.n_gpu;
    const uint32_t Sl = p.slots;
    const uint64_t cnt = ((uint64_t)n - (uint64_t)p.gpu_id + N - 1) / N;
    const uint32_t loops = 0xFFFFFFFFu / n;
    Blk B;
    B.p = &p;
    B.lab0 = p.lab + (uint64_t)blockIdx.x * Sl * n;
    B.mark0 = p.mark + (uint64_t)blockIdx.x * Sl * n;
    B.orph0 = p.orphan + (uint64_t)blockIdx.x * Sl * p.ow;
    B.qg0 = p.qg + (uint64_t)blockIdx.x * 2 * p.qg_cap;
    B.ql = qlds;
    B.qtot = p.ql_cap + p.qg_cap;
    uint32_t off = p.bstate[2 * blockIdx.x];
    uint32_t since = p.bstate[2 * blockIdx.x + 1];
    TallyB t{0, 0, 0, 0};
    unsigned long long levels = 0, rescans = 0;
    uint64_t* fb = S.fbuf[w];
    uint32_t fcount = 0;
    if (tid == 0) {
        S.qcnt[0] = S.qcnt[1] = 0;
        S.cursor = 0;
        S.nbig = 0;
        S.orph = 0;
    }
    for (;;) {
        if (tid == 0) S.base = atomicAdd(&p.ctrl->src_next, Sl);
        __syncthreads();
        const uint32_t base = S.base;
        if ((uint64_t)base >= cnt) break;
        const uint32_t nsrc = (uint32_t)((cnt - base) < Sl ? (cnt - base) : Sl);
        if (tid < nsrc) S.src[tid] = (uint32_t)(p.gpu_id + (cnt - 1 - (base + tid)) * N);
        if (since >= loops) {  // offsets exhausted: re-initialise this block's labels
            for (uint64_t i = tid; i < (uint64_t)Sl * n; i += BB) B.lab0[i] = 0xFFFFFFFFu;
            off = 0xFFFFFFFFu - n;
            since = 0;
        }
        __syncthreads();
        B.off = off;
        B.stale = off + n;
        // seeding (with_space_optimization/symbfact.cuh:639-721), one wave per slot
        for (uint32_t j = w; j < nsrc; j += BW) {
            const uint32_t src = S.src[j];
            const uint32_t stamp = p.stamp_base + (n - src);
            uint32_t* lab = B.lab0 + (uint64_t)j * n;
            uint32_t* mark = B.mark0 + (uint64_t)j * n;
            const int32_t b = p.rp[src], e = p.rp[src + 1];
            if (lane == 0) {
                lab[src] = off;
                t.te += (uint32_t)(e - b);
                t.checks += (uint32_t)(e - b);
            }
            for (int32_t t0 = b; t0 < e; t0 += WAVE) {
                const int32_t i = t0 + (int32_t)lane;
                bool push = false;
                uint32_t m = 0;
                if (i < e) {
                    m = (uint32_t)p.ci[i];
                    mark[m] = stamp;
                    if (m < src) {
                        const uint32_t old = lab[m];
                        if (old >= B.stale) {
                            t.reach++;
                            t.te += (uint32_t)(p.rp[m + 1] - p.rp[m]);
                        }
                        lab[m] = off;
                        push = true;
                    }
                }
                emit(B, S, fb, fcount, 0, push, make_entry(j, 0, m), false, 0);
            }
        }
        __syncthreads();
        uint32_t cur = 0;
        for (;;) {
            // decision inputs: read by every thread, then a barrier before any
            // wave can change them (parking, refills)
            const uint32_t appended = S.qcnt[cur];
            const uint32_t orph = S.orph;
            __syncthreads();
            const uint32_t size = appended < B.qtot ? appended : B.qtot;
            if (size < B.qtot && orph != 0) {
                // drain-and-resume: top the queue up from the overflow bitmap
                rescan_b(B, S, cur, nsrc);
                rescans++;
                __syncthreads();
                continue;
            }
            if (size == 0) break;
            expand_level_b(B, S, fb, fcount, cur, size, t);
            __syncthreads();
            const uint32_t nbig = S.nbig < (uint32_t)BIGCAP ? S.nbig : (uint32_t)BIGCAP;
            for (uint32_t r = 0; r < nbig; r++) {
                const BigRowB br = S.big[r];
                expand_strided_b(B, S, fb, fcount, cur ^ 1u, t, br.beg, br.end, br.cost, br.slot,
                                 tid, BB);
            }
            __syncthreads();
            if (tid == 0) {
                S.qcnt[cur] = 0;
                S.cursor = 0;
                S.nbig = 0;
            }
            __syncthreads();
            cur ^= 1u;
            levels++;
        }
        off -= n;
        since++;
    }
    if (fcount) flush_fills_w(p, fb, fcount);
    const unsigned long long f = wave_sum64(t.fills), te = wave_sum64(t.te),
                             rc = wave_sum64(t.reach), ch = wave_sum64(t.checks);
    if (lane == 0) {
        if (f) aadd64(&p.ctrl->fill_count, f);
        if (te) aadd64(&p.ctrl->te, te);
        if (rc) aadd64(&p.ctrl->reach, rc);
        if (ch) aadd64(&p.ctrl->edge_checks, ch);
    }
    if (tid == 0) {
        aadd64(&p.ctrl->levels, levels);
        aadd64(&p.ctrl->rescans, rescans);
        p.bstate[2 * blockIdx.x] = off;
        p.bstate[2 * blockIdx.x + 1] = since;
    }
}

}  // namespace

int blk_block_size() { return BB; }
const void* blk_kernel_ptr() { return reinterpret_cast<const void*>(&k_traverse_blk); }

hipError_t launch_traverse_blk(const BlkParams& p, size_t lds_bytes, hipStream_t s) {
    hipLaunchKernelGGL(k_traverse_blk, dim3(p.nblocks), dim3(BB), lds_bytes, s, p);
    return hipGetLastError();
}

}  // namespace gsofa
