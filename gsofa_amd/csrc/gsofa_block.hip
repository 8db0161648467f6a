// gsofa_block.hip -- block-resident multi-source traversal (default mode).
//
// Same per-source arithmetic as the reference's warp-centric level kernel
// (with_space_optimization/symbfact.cuh:883-1082) and seeding (:639-721), with
// an execution model built for CDNA4:
//
// * Each workgroup of a persistent, occupancy-bounded grid owns 32 source
//   slots and runs a "mini-group" of 32 consecutive sources through all of
//   their frontier levels by itself, synchronising only with __syncthreads.
//   The next mini-group comes from one global counter.  There is no grid-wide
//   barrier and no hot global counter on the level path.
// * Labels and markers are stored slot-interleaved, [vertex][32 slots]: the
//   32 sources' max-id labels of one vertex share one 128-byte line.  Nearby
//   sources reach mostly the same vertices in the same levels, so the frontier
//   is kept per vertex with a 32-bit mask of the slots that must expand it
//   (deduplicated: a vertex enters the next queue when its mask leaves zero).
// * A half-wave (32 lanes, lane = slot) expands one frontier vertex for all of
//   its slots: the row is read once, and each CSR word costs one coalesced
//   line access for the 32 labels.  Edge checks stay per source: lane j relaxes
//   label(j, m) with its own cost max(v, label(j, v)) exactly as the reference
//   does for source j.
// * Sources are taken in decreasing order so the deepest traversals start
//   first; marker stamps are stamp_base + (n - s), still increasing per slot,
//   so markers are never reset (symbfact.cuh:1006).  Labels use the per-group
//   decreasing offset of symbfact.cuh:2388-2407, kept per workgroup.
// * The frontier queue is two-tier: the first ql_cap entries of a level in
//   LDS, the rest in a per-workgroup global region of W entries; with the
//   per-vertex deduplication it can never overflow.
// * Work is a table of chunks (up to 32 sources of one matrix).  A batch of
//   matrices is one block-diagonal matrix in which no traversal can leave its
//   own matrix, so per-workgroup arrays are indexed relative to the chunk's
//   matrix and sized by the largest matrix W, and labels hold
//   offset + (vertex - matrix base).
#include "gsofa_device.h"
#include "gsofa_internal.h"

// Scope of the label/marker filter loads.  Agent scope reads the L2 (where the
// atomics land) rather than a possibly stale L1 line.
#ifndef GSOFA_FILTER_SCOPE
#define GSOFA_FILTER_SCOPE __HIP_MEMORY_SCOPE_AGENT
#endif

namespace gsofa {

namespace {

using namespace dev;

constexpr int BB = 512;          // threads per workgroup
constexpr int BW = BB / WAVE;    // waves per workgroup
constexpr int SL = kBlockSlots;  // slots per workgroup (lanes of a half-wave)
constexpr int FB = 128;          // staged fills per wave
#ifndef GSOFA_KB
#define GSOFA_KB 8
#endif
constexpr int KB = GSOFA_KB;     // edges probed together per half-wave

struct BlkLds {
    uint64_t fbuf[BW][FB];
    unsigned long long s_te[SL];     // per-slot TE of the current mini-group
    unsigned long long s_reach[SL];  // per-slot |R_<(s)|
    uint32_t src[SL];
    uint32_t qcnt[2];
    uint32_t cursor;
    uint32_t chunk;
};

struct TallyB {
    unsigned long long fills, te, reach, checks;
};

struct Blk {
    const BlkParams* p;
    uint32_t* lab0;   // [W][SL]
    uint32_t* mark0;  // [W][SL]
    uint32_t* mask0;  // [2][W]
    uint32_t* qg0;    // [2][W]
    uint32_t* ql;     // [2][ql_cap] (LDS)
    uint32_t W;
};

__device__ __forceinline__ uint32_t ld_filter(uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, GSOFA_FILTER_SCOPE);
}

__device__ __forceinline__ uint32_t q_read(const Blk& B, uint32_t buf, uint32_t i) {
    const uint32_t qc = B.p->ql_cap;
    return i < qc ? B.ql[buf * qc + i] : B.qg0[(uint64_t)buf * B.W + (i - qc)];
}

__device__ __forceinline__ void q_write(const Blk& B, uint32_t buf, uint32_t i, uint32_t v) {
    const uint32_t qc = B.p->ql_cap;
    if (i < qc) B.ql[buf * qc + i] = v;
    else B.qg0[(uint64_t)buf * B.W + (i - qc)] = v;
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

__device__ __forceinline__ void stage_fill(const BlkParams& p, uint64_t* fb, uint32_t& fcount,
                                           bool fill, uint64_t entry) {
    const uint64_t fm = __ballot(fill);
    if (fm) {
        const uint32_t c = (uint32_t)__popcll(fm);
        if (fcount + c > (uint32_t)FB) {
            flush_fills_w(p, fb, fcount);
            fcount = 0;
        }
        if (fill) fb[fcount + rank_in(fm)] = entry;
        fcount += c;
    }
}

__global__ __launch_bounds__(BB) void k_traverse_blk(BlkParams p) {
    __shared__ BlkLds S;
    extern __shared__ uint32_t qlds[];
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = __lane_id();
    const uint32_t w = tid / WAVE;
    const uint32_t half = lane >> 5;  // which half-wave
    const uint32_t j = lane & 31u;    // slot of this lane
    const bool leader = j == 0;
    const uint32_t n = (uint32_t)p.n;
    const uint32_t W = p.W;
    const uint32_t loops = 0xFFFFFFFFu / W;
    Blk B;
    B.p = &p;
    B.W = W;
    B.lab0 = p.lab + (uint64_t)blockIdx.x * SL * W;
    B.mark0 = p.mark + (uint64_t)blockIdx.x * SL * W;
    B.mask0 = p.mask + (uint64_t)blockIdx.x * 2 * W;
    B.qg0 = p.qg + (uint64_t)blockIdx.x * 2 * W;
    B.ql = qlds;
    uint32_t off = p.bstate[2 * blockIdx.x];
    uint32_t since = p.bstate[2 * blockIdx.x + 1];
    TallyB t{0, 0, 0, 0};
    unsigned long long levels = 0, tot_te = 0, tot_reach = 0;
    uint64_t* fb = S.fbuf[w];
    uint32_t fcount = 0;
    if (tid == 0) {
        S.qcnt[0] = S.qcnt[1] = 0;
        S.cursor = 0;
    }
    for (;;) {
        if (tid == 0) S.chunk = atomicAdd(&p.ctrl->src_next, 1u);
        __syncthreads();
        const uint32_t ck = S.chunk;
        if (ck >= p.n_chunks) break;
        const BlkChunk C = p.chunks[ck];
        const uint32_t nsrc = C.count;
        const uint32_t wb = p.mat_base[C.mat];  // window base: vertex wb is local 0
        if (tid < (uint32_t)SL) {
            S.src[tid] = tid < nsrc ? C.first - tid * C.stride : 0u;
            S.s_te[tid] = 0;
            S.s_reach[tid] = 0;
        }
        if (since >= loops) {  // offsets exhausted: re-initialise this block's labels
            for (uint64_t i = tid; i < (uint64_t)SL * W; i += BB) B.lab0[i] = 0xFFFFFFFFu;
            off = 0xFFFFFFFFu - W;
            since = 0;
        }
        __syncthreads();
        const uint32_t stale = off + W;
        // this lane's source (local id); lanes of slots >= nsrc never get a mask bit
        const uint32_t my_srcg = S.src[j];
        const uint32_t my_src = my_srcg - wb;
        const uint32_t my_stamp = p.stamp_base + (n - my_srcg);

        // ---- seeding (symbfact.cuh:639-721): wave w seeds slots w, w+BW, ...
        for (uint32_t k = w; k < nsrc; k += BW) {
            const uint32_t srcg = S.src[k];
            const uint32_t src = srcg - wb;
            const uint32_t stamp = p.stamp_base + (n - srcg);
            const int32_t b = p.rp[srcg], e = p.rp[srcg + 1];
            if (lane == 0) {
                B.lab0[(uint64_t)src * SL + k] = off;
                atomicAdd(&S.s_te[k], (unsigned long long)(e - b));
                t.checks += (uint32_t)(e - b);
            }
            for (int32_t i = b + (int32_t)lane; i < e; i += WAVE) {
                const uint32_t mg = (uint32_t)p.ci[i];
                const uint32_t m = mg - wb;
                B.mark0[(uint64_t)m * SL + k] = stamp;
                if (m < src) {
                    const uint32_t old = B.lab0[(uint64_t)m * SL + k];
                    if (old >= stale) {
                        atomicAdd(&S.s_reach[k], 1ull);
                        atomicAdd(&S.s_te[k], (unsigned long long)(p.rp[mg + 1] - p.rp[mg]));
                    }
                    B.lab0[(uint64_t)m * SL + k] = off;
                    const uint32_t o = __hip_atomic_fetch_or(B.mask0 + m, 1u << k, __ATOMIC_RELAXED,
                                                             __HIP_MEMORY_SCOPE_AGENT);
                    if (o == 0u) q_write(B, 0, atomicAdd(&S.qcnt[0], 1u), m);
                }
            }
        }
        __syncthreads();

        // ---- level loop: relaxation (symbfact.cuh:883-1082), lane = slot
        uint32_t cur = 0;
        for (;;) {
            const uint32_t size = S.qcnt[cur];
            __syncthreads();  // every thread has the size before it can change
            if (size == 0) break;
            const uint32_t nbuf = cur ^ 1u;
            uint32_t* maskc = B.mask0 + (uint64_t)cur * W;
            uint32_t* maskn = B.mask0 + (uint64_t)nbuf * W;
            // Software pipeline: the queue entry, row bounds, slot mask and
            // label of the next vertex pair are fetched while the current
            // pair's edges are probed.
            uint32_t i0 = 0;
            if (lane == 0) i0 = atomicAdd(&S.cursor, 2u);
            i0 = bcast0(i0);
            bool nhave = false;
            uint32_t nv = 0, nM = 0, nlv = 0;
            int32_t nrb = 0, nre = 0;
            auto prefetch = [&](uint32_t base_i) {
                const uint32_t qi = base_i + half;
                nhave = base_i < size && qi < size;
                nv = 0;
                nrb = nre = 0;
                nM = 0;
                nlv = 0;
                if (nhave) {
                    nv = q_read(B, cur, qi);
                    nrb = p.rp[wb + nv];
                    nre = p.rp[wb + nv + 1];
                    nlv = ld_agent(B.lab0 + (uint64_t)nv * SL + j);
                    if (leader)
                        nM = __hip_atomic_exchange(maskc + nv, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
            };
            if (i0 < size) prefetch(i0);
            while (i0 < size) {
                const bool have = nhave;
                const uint32_t v = nv;
                const int32_t rb = nrb, re = nre;
                const uint32_t Mv = nM, lv = nlv;
                if (lane == 0) i0 = atomicAdd(&S.cursor, 2u);
                i0 = bcast0(i0);
                if (i0 < size) prefetch(i0);
                const uint32_t M = __shfl(Mv, 0, 32);
                const bool active = have && ((M >> j) & 1u);
                const uint32_t cost = (v + off) > lv ? (v + off) : lv;  // symbfact.cuh:924
                const uint32_t deg = have ? (uint32_t)(re - rb) : 0u;
                const uint32_t d0 = __shfl(deg, 0, WAVE), d1 = __shfl(deg, 32, WAVE);
                const uint32_t dmax = d0 > d1 ? d0 : d1;
                for (uint32_t c0 = 0; c0 < dmax; c0 += 32) {
                    // each half loads up to 32 CSR words of its row at once
                    uint32_t cm = 0;
                    if (c0 + j < deg) cm = (uint32_t)p.ci[rb + (int32_t)(c0 + j)] - wb;
                    const uint32_t lim = dmax - c0 < 32u ? dmax - c0 : 32u;
                    for (uint32_t k0 = 0; k0 < lim; k0 += KB) {
                        // KB edges at a time: all label/marker probes in flight
                        // together, then the atomics, then the bookkeeping
                        uint32_t mm[KB], val[KB], old[KB];
                        uint32_t* adr[KB];
                        bool go[KB], up[KB], need[KB];
#pragma unroll
                        for (int u = 0; u < KB; u++) {
                            mm[u] = __shfl(cm, (int)((k0 + u) & 31u), 32);
                            go[u] = active && (k0 + u) < lim && (c0 + k0 + u) < deg;
                            up[u] = mm[u] > my_src;
                            adr[u] = (up[u] ? B.mark0 : B.lab0) + (uint64_t)mm[u] * SL + j;
                        }
#pragma unroll
                        for (int u = 0; u < KB; u++) val[u] = go[u] ? ld_filter(adr[u]) : 0u;
#pragma unroll
                        for (int u = 0; u < KB; u++)
                            need[u] = go[u] && (up[u] ? val[u] != my_stamp : val[u] > cost);
#pragma unroll
                        for (int u = 0; u < KB; u++)
                            old[u] = need[u] ? (up[u] ? amax(adr[u], my_stamp) : amin(adr[u], cost)) : 0u;
                        uint32_t r[KB];  // bit0: relaxed below the source, bit1: new fill
#pragma unroll
                        for (int u = 0; u < KB; u++) {
                            r[u] = 0;
                            if (go[u]) t.checks++;
                            if (!need[u]) continue;
                            if (up[u]) {
                                r[u] = old[u] < my_stamp ? 2u : 0u;
                            } else if (old[u] > cost) {
                                const uint32_t m = mm[u];
                                if (old[u] >= stale) {  // first reach of m from this source
                                    t.reach++;
                                    t.te += (uint32_t)(p.rp[wb + m + 1] - p.rp[wb + m]);
                                }
                                r[u] = m < my_src ? 1u : 0u;
                                if (m + off > cost &&
                                    amax(B.mark0 + (uint64_t)m * SL + j, my_stamp) < my_stamp)
                                    r[u] |= 2u;
                            }
                            if (r[u] & 2u) t.fills++;
                        }
                        // slots that must expand m next level: one atomic per
                        // half and edge, all issued before any result is used
                        uint32_t bits[KB], prev[KB];
#pragma unroll
                        for (int u = 0; u < KB; u++)
                            bits[u] = (uint32_t)(__ballot(r[u] & 1u) >> (32u * half));
#pragma unroll
                        for (int u = 0; u < KB; u++)
                            prev[u] = (leader && bits[u])
                                          ? __hip_atomic_fetch_or(maskn + mm[u], bits[u], __ATOMIC_RELAXED,
                                                                  __HIP_MEMORY_SCOPE_AGENT)
                                          : 1u;
#pragma unroll
                        for (int u = 0; u < KB; u++)
                            if (leader && bits[u] && prev[u] == 0u)
                                q_write(B, nbuf, atomicAdd(&S.qcnt[nbuf], 1u), mm[u]);
#pragma unroll
                        for (int u = 0; u < KB; u++)
                            stage_fill(p, fb, fcount, (r[u] & 2u) != 0,
                                       ((uint64_t)my_srcg << 32) | (mm[u] + wb));
                    }
                }
            }
            __syncthreads();
            if (tid == 0) {
                S.qcnt[cur] = 0;
                S.cursor = 0;
            }
            __syncthreads();
            cur = nbuf;
            levels++;
        }
        // per-row TE and |R_<(s)|: lane j's tallies belong to slot j
        if (t.te) atomicAdd(&S.s_te[j], t.te);
        if (t.reach) atomicAdd(&S.s_reach[j], t.reach);
        t.te = t.reach = 0;
        __syncthreads();
        if (tid < nsrc) {
            p.row_te[S.src[tid]] = S.s_te[tid];
            p.row_reach[S.src[tid]] = S.s_reach[tid];
            tot_te += S.s_te[tid];
            tot_reach += S.s_reach[tid];
        }
        off -= W;
        since++;
    }
    if (fcount) flush_fills_w(p, fb, fcount);
    const unsigned long long f = wave_sum64(t.fills), te = wave_sum64(tot_te),
                             rc = wave_sum64(tot_reach), ch = wave_sum64(t.checks);
    if (lane == 0) {
        if (f) aadd64(&p.ctrl->fill_count, f);
        if (te) aadd64(&p.ctrl->te, te);
        if (rc) aadd64(&p.ctrl->reach, rc);
        if (ch) aadd64(&p.ctrl->edge_checks, ch);
    }
    if (tid == 0) {
        aadd64(&p.ctrl->levels, levels);
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
