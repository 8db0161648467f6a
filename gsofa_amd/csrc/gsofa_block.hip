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
// * Labels are stored slot-interleaved, [vertex][32 slots]: the
//   32 sources' max-id labels of one vertex share one 128-byte line.  Nearby
//   sources reach mostly the same vertices in the same levels, so the frontier
//   is kept per vertex with a 32-bit mask of the slots that must expand it
//   (deduplicated: a vertex enters the next queue when its mask leaves zero).
// * A level runs in rounds of two phases, and no atomic touches global memory
//   on the level path (on this part a global atomic executes at the memory
//   side, not in the L2):
//     A. combine: each half-wave pops a frontier vertex v, computes lane j's
//        relaxation cost max(v, label_j(v)) for the slots j of v's mask
//        (symbfact.cuh:924), looks each neighbour m up in an LDS hash table
//        of the round's distinct targets, and folds the cost into that
//        target's 32 per-slot cells with an LDS atomic min;
//     B. relax: every table entry (a distinct target m) is relaxed once per
//        slot against its label exactly as the reference relaxes
//        one edge (symbfact.cuh:989-1066), with the round's minimum cost.
//        Targets are unique within a round, so plain loads and stores suffice.
//   Folding several edges into one relaxation with their minimum cost gives
//   the same labels, fills and first reaches as applying them one by one: the
//   label ends at the minimum either way, and a fill is detected iff some
//   cost is below the target (so iff the minimum is).
//   A frontier vertex's own label may drop during a later round of the same
//   level; that drop re-enqueues the vertex, so expanding it with the older
//   cost is correct (label-correcting).
// * Fills are read off the labels, so there is no marker array.  The
//   reference stamps a marker per (source, vertex) to count each fill once
//   (symbfact.cuh:992-1009, 1046-1063).  Here labels only decrease during a
//   source's traversal and use the per-group decreasing offset of
//   symbfact.cuh:2388-2407 (kept per workgroup; older groups' labels compare
//   as stale), so:
//     - a vertex m below the source is a new L entry exactly when its label
//       first drops below m + off (the reference's "m + off > cost" on the
//       first stamp); neighbours of the source start at off and never count;
//     - a vertex m above the source is a new U entry exactly when it is first
//       touched: its label cell (unused as a label for this source) goes
//       from stale to a current value; neighbours start touched.
// * Sources are taken in decreasing order so the deepest traversals start
//   first.
// * Work is a table of chunks (up to 32 sources of one matrix).  A batch of
//   matrices is one block-diagonal matrix in which no traversal can leave its
//   own matrix, so per-workgroup arrays are indexed relative to the chunk's
//   matrix and sized by the largest matrix W, and labels hold
//   offset + (vertex - matrix base).
#include "gsofa_device.h"
#include "gsofa_internal.h"

namespace gsofa {

namespace {

using namespace dev;

constexpr int BB = 512;              // threads per workgroup
constexpr int BW = BB / WAVE;        // waves per workgroup
constexpr int NH = 2 * BW;           // half-waves
constexpr int SL = kBlockSlots;      // slots per workgroup (lanes of a half-wave)
constexpr int T = kBlkTable;         // distinct targets per round
constexpr int HB = 10;               // log2 of the hash slots
constexpr int H = 1 << HB;           // hash slots (load factor <= T / H)
constexpr int FB = WAVE;             // staged fills per wave (one ballot's worth)
constexpr uint32_t INF = 0xFFFFFFFFu;
constexpr uint32_t NONE = 0xFFFFFFFFu;   // no table entry (edge not in this group)
constexpr uint32_t FULL = 0xFFFFFFFEu;   // table full: retry the edge next round
constexpr uint32_t HOLE = 0xFFFFFFFFu;   // ent_m of an index that lost its insert
#ifndef GSOFA_KR
#define GSOFA_KR 4
#endif
#ifndef GSOFA_KV
#define GSOFA_KV 4
#endif
constexpr int KR = GSOFA_KR;         // table entries relaxed together per half-wave
constexpr int KV = GSOFA_KV;         // frontier vertices fetched together per half-wave
static_assert(T <= H / 2, "hash load factor");

struct BlkLds {
    uint32_t cell[T][SL];            // min cost per (round target, slot); INF = none
    unsigned long long hkey[H];      // ((m + 1) << 32) | table index; 0 = empty
    uint32_t ent_m[T];               // table index -> target (window-local) or HOLE
    uint32_t eidx[NH][32];           // table index of each lane's edge
    uint32_t fbuf[BW][FB];           // staged fills, see flush_fills_w
    unsigned long long s_te[SL];     // per-slot TE of the current mini-group
    unsigned long long s_reach[SL];  // per-slot |R_<(s)|
    uint32_t src[SL];
    uint32_t qcnt[2];
    uint32_t cursor;
    uint32_t tcnt;
    uint32_t held;
    uint32_t chunk;
};

struct Blk {
    const BlkParams* p;
    uint32_t* lab0;    // [W][SL]
    uint32_t* mask0;   // [2][W]
    uint32_t* q0;      // [2][W]
    uint32_t W;
    uint32_t wb;       // window base (global id of local vertex 0)
    uint32_t off, stale;
};

// Element at a 32-bit byte offset from a workgroup-uniform base: the offset
// stays one VGPR and the load/store uses the scalar-base addressing form.
// Every per-workgroup array here is far below 4 GiB.
template <class T>
__device__ __forceinline__ T& at(T* base, uint32_t byte_off) {
    return *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + byte_off);
}
template <class T>
__device__ __forceinline__ const T& at(const T* base, uint32_t byte_off) {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + byte_off);
}

struct TallyB {
    unsigned long long fills, te, reach, checks;
};

// Staged fills are (slot << 25) | global target (n < 2^25); the flush turns
// them into (source << 32) | target with the slot's source.
__device__ void flush_fills_w(const BlkParams& p, const BlkLds& S, const uint32_t* fb, uint32_t cnt) {
    wave_sync();
    const uint32_t lane = __lane_id();
    uint64_t pos = 0;
    if (lane == 0) pos = aadd64(&p.ctrl->fill_pos, cnt);
    pos = bcast0_64(pos);
    for (uint32_t t = lane; t < cnt; t += WAVE) {
        const uint64_t at = pos + t;
        const uint32_t f = fb[t];
        if (at < p.fill_cap) p.fills[at] = ((uint64_t)S.src[f >> 25] << 32) | (f & 0x1FFFFFFu);
    }
    wave_sync();
}

// Wave-uniform call: every lane of the wave must be active.  One call stages
// at most 64 fills, so a 64-entry buffer never overflows.
__device__ __forceinline__ void stage_fill(const BlkParams& p, const BlkLds& S, uint32_t* fb,
                                           uint32_t& fcount, bool fill, uint32_t entry) {
    const uint64_t fm = __ballot(fill);
    if (fm) {
        const uint32_t c = (uint32_t)__popcll(fm);
        if (fcount + c > (uint32_t)FB) {
            flush_fills_w(p, S, fb, fcount);
            fcount = 0;
        }
        if (fill) fb[fcount + rank_in(fm)] = entry;
        fcount += c;
    }
}

// Table index of target m in this round, inserting it if new; FULL when the
// table has no room left.  Linear probing over H slots; an index allocated for
// an insert that loses its race to an equal key becomes a HOLE entry.
__device__ __forceinline__ uint32_t tab_find(BlkLds& S, uint32_t m) {
    const unsigned long long key = (unsigned long long)(m + 1u) << 32;
    uint32_t s = (m * 2654435761u) >> (32 - HB);
    uint32_t mine = NONE, res = FULL;
    for (int probe = 0; probe < H; probe++) {
        unsigned long long cur = S.hkey[s];
        if (cur == 0ull) {
            if (mine == NONE) {
                mine = atomicAdd(&S.tcnt, 1u);
                if (mine >= (uint32_t)T) {
                    mine = NONE;
                    break;
                }
            }
            const unsigned long long prev = atomicCAS(&S.hkey[s], 0ull, key | mine);
            if (prev == 0ull) {
                S.ent_m[mine] = m;
                res = mine;
                mine = NONE;
                break;
            }
            cur = prev;
        }
        if ((cur & 0xFFFFFFFF00000000ull) == key) {
            res = (uint32_t)cur;
            break;
        }
        s = (s + 1u) & (uint32_t)(H - 1);
    }
    if (mine != NONE) S.ent_m[mine] = HOLE;
    return res;
}

// Frontier vertices held by a half-wave across rounds: up to KV vertices are
// fetched together so their dependent loads overlap.  All lanes of the half
// hold the same vertices; lane j holds slot j's cost of each.
struct HalfVtx {
    bool held;
    uint32_t cost[KV];  // lane j: max(v + off, label_j(v)), INF if slot j is not in v's mask
    uint32_t pend[KV];  // edges of the current 32-edge group not yet folded
    int32_t pos[KV];    // CSR index of the current group's first edge
    int32_t end[KV];    // CSR end of the row
};

__device__ __forceinline__ uint32_t group_mask(int32_t left) {
    return left >= 32 ? 0xFFFFFFFFu : (left <= 0 ? 0u : ((1u << left) - 1u));
}

// Phase A: pop frontier vertices and fold their edges' costs into the table
// until the queue is exhausted or the table is full.  Control flow diverges
// between the two halves of a wave; nothing here needs the whole wave.
__device__ void combine_phase(const Blk& B, BlkLds& S, HalfVtx& hv, uint32_t cur, uint32_t size,
                              TallyB& t) {
    const BlkParams& p = *B.p;
    const uint32_t lane = __lane_id();
    const uint32_t half = lane >> 5;
    const uint32_t j = lane & 31u;
    const uint32_t h = (threadIdx.x / WAVE) * 2 + half;
    uint32_t* maskc = B.mask0 + (uint64_t)cur * B.W;
    const uint32_t* qc = B.q0 + (uint64_t)cur * B.W;
    for (;;) {
        if (!hv.held) {
            if (S.tcnt >= (uint32_t)T) return;  // table full: take no new vertex
            uint32_t i0 = 0;
            if (j == 0) i0 = atomicAdd(&S.cursor, (uint32_t)KV);
            i0 = __shfl(i0, (int)(lane & 32u), WAVE);
            if (i0 >= size) return;
            uint32_t vv[KV], M[KV], lv[KV];
            int32_t r0[KV], r1[KV];
#pragma unroll
            for (int u = 0; u < KV; u++) vv[u] = i0 + u < size ? at(qc, (i0 + u) << 2) : NONE;
#pragma unroll
            for (int u = 0; u < KV; u++) {  // independent loads, issued together
                const bool ok = vv[u] != NONE;
                M[u] = ok ? at(maskc, vv[u] << 2) : 0u;
                r0[u] = ok ? at(p.rp, (B.wb + vv[u]) << 2) : 0;
                r1[u] = ok ? at(p.rp, (B.wb + vv[u] + 1) << 2) : 0;
                lv[u] = ok ? at(B.lab0, (vv[u] * SL + j) << 2) : 0u;
            }
#pragma unroll
            for (int u = 0; u < KV; u++) {
                // consumed; this buffer is refilled two levels later
                if (j == 0 && vv[u] != NONE) at(maskc, vv[u] << 2) = 0u;
                const uint32_t vo = vv[u] + B.off;
                hv.cost[u] = ((M[u] >> j) & 1u) ? (vo > lv[u] ? vo : lv[u]) : INF;  // symbfact.cuh:924
                hv.pos[u] = r0[u];
                hv.end[u] = r1[u];
                hv.pend[u] = group_mask(r1[u] - r0[u]);
                if (j == 0) t.checks += (unsigned long long)(uint32_t)(r1[u] - r0[u]) * (uint32_t)__popc(M[u]);
            }
            hv.held = true;
        }
        // one group of up to 32 edges of every held vertex
        uint32_t mv[KV];
#pragma unroll
        for (int u = 0; u < KV; u++)
            mv[u] = ((hv.pend[u] >> j) & 1u) ? (uint32_t)at(p.ci, (uint32_t)(hv.pos[u] + (int32_t)j) << 2) - B.wb : 0u;
        bool full = false;
#pragma unroll
        for (int u = 0; u < KV; u++) {
            const bool mine = (hv.pend[u] >> j) & 1u;
            uint32_t idx = NONE;
            if (mine) idx = tab_find(S, mv[u]);
            S.eidx[h][j] = idx;
            const uint32_t okm = (uint32_t)(__ballot(mine && idx != FULL) >> (32u * half));
            full = full || (okm != hv.pend[u]);
            wave_sync();
            if (hv.cost[u] != INF) {
                for (uint32_t b = okm; b; b &= b - 1u) {
                    const uint32_t id = S.eidx[h][__builtin_ctz(b)];
                    atomicMin(&S.cell[id][j], hv.cost[u]);
                }
            }
            wave_sync();
            hv.pend[u] &= ~okm;
        }
        if (full) return;  // table full: the rest waits for the next round
        bool any = false;
#pragma unroll
        for (int u = 0; u < KV; u++) {
            if (hv.pos[u] + 32 < hv.end[u]) {  // next group of a long row
                hv.pos[u] += 32;
                hv.pend[u] = group_mask(hv.end[u] - hv.pos[u]);
            } else {
                hv.pos[u] = hv.end[u];
            }
            any = any || hv.pend[u] != 0u;
        }
        if (!any) hv.held = false;
    }
}

// Phase B: relax every table entry once per slot.  Wave-uniform trip count.
// The degree of a target below the source is loaded with its label, whether
// or not the relaxation turns out to be a first reach, so each batch of KR
// entries costs one memory round trip.
__device__ void relax_phase(const Blk& B, BlkLds& S, uint32_t nbuf, uint32_t my_src,
                            uint32_t* fb, uint32_t& fcount, TallyB& t) {
    const BlkParams& p = *B.p;
    const uint32_t lane = __lane_id();
    const uint32_t half = lane >> 5;
    const uint32_t j = lane & 31u;
    const bool leader = j == 0;
    const uint32_t w = threadIdx.x / WAVE;
    uint32_t* maskn = B.mask0 + (uint64_t)nbuf * B.W;
    uint32_t* qn = B.q0 + (uint64_t)nbuf * B.W;
    const uint32_t tc = S.tcnt;
    const uint32_t ntab = tc < (uint32_t)T ? tc : (uint32_t)T;
    const uint32_t off = B.off;
    for (uint32_t base = w * 2 * KR; base < ntab; base += BW * 2 * KR) {
        uint32_t mm[KR], c[KR], lm[KR], nmk[KR];
        int32_t d0[KR], d1[KR];
        bool ok[KR], act[KR], up[KR];
#pragma unroll
        for (int u = 0; u < KR; u++) {
            const uint32_t e = base + half * KR + u;
            ok[u] = e < ntab;
            mm[u] = ok[u] ? S.ent_m[e] : HOLE;
            ok[u] = ok[u] && mm[u] != HOLE;
            c[u] = INF;
            if (ok[u]) {
                c[u] = S.cell[e][j];
                S.cell[e][j] = INF;  // ready for the next round
            }
            act[u] = c[u] != INF;
            up[u] = mm[u] > my_src;
        }
        // the one round trip: label, degree, next-level slot mask
        uint32_t o[KR];
#pragma unroll
        for (int u = 0; u < KR; u++) {
            const bool lo = act[u] && !up[u];
            o[u] = (mm[u] * SL + j) << 2;
            const uint32_t ro = (B.wb + mm[u]) << 2;
            lm[u] = act[u] ? at(B.lab0, o[u]) : 0u;
            d0[u] = lo ? at(p.rp, ro) : 0;
            d1[u] = lo ? at(p.rp, ro + 4u) : 0;
            nmk[u] = (ok[u] && leader) ? at(maskn, mm[u] << 2) : 0u;
        }
        bool push[KR], fill[KR];
#pragma unroll
        for (int u = 0; u < KR; u++) {
            push[u] = fill[u] = false;
            if (act[u]) {
                if (up[u]) {  // m above the source: U entry on first touch (symbfact.cuh:992-1009)
                    if (lm[u] >= B.stale) {
                        at(B.lab0, o[u]) = c[u];
                        fill[u] = true;
                    }
                } else if (lm[u] > c[u]) {  // relaxation (symbfact.cuh:1018-1063)
                    at(B.lab0, o[u]) = c[u];
                    push[u] = mm[u] < my_src;
                    if (lm[u] >= B.stale) {  // first reach: |R_<(s)| and TE_alg
                        t.reach++;
                        t.te += (uint32_t)(d1[u] - d0[u]);
                    }
                    const uint32_t mo = mm[u] + off;
                    fill[u] = lm[u] >= mo && c[u] < mo;  // label crosses below m: new L entry
                }
            }
        }
        // next-level slot masks: the half leader writes maskn[m] (m is unique
        // in the round, so no other lane touches it)
        uint32_t bits[KR];
#pragma unroll
        for (int u = 0; u < KR; u++) bits[u] = (uint32_t)(__ballot(push[u]) >> (32u * half));
        if (leader) {
#pragma unroll
            for (int u = 0; u < KR; u++) {
                if (bits[u] && (nmk[u] | bits[u]) != nmk[u]) {
                    at(maskn, mm[u] << 2) = nmk[u] | bits[u];
                    if (nmk[u] == 0u) qn[atomicAdd(&S.qcnt[nbuf], 1u)] = mm[u];
                }
            }
        }
#pragma unroll
        for (int u = 0; u < KR; u++) {
            if (fill[u]) t.fills++;
            stage_fill(p, S, fb, fcount, fill[u], (j << 25) | (mm[u] + B.wb));
        }
    }
}

__global__ __launch_bounds__(BB) void k_traverse_blk(BlkParams p) {
    __shared__ BlkLds S;
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = __lane_id();
    const uint32_t w = tid / WAVE;
    const uint32_t j = lane & 31u;
    const uint32_t W = p.W;
    const uint32_t loops = 0xFFFFFFFFu / W;
    Blk B;
    B.p = &p;
    B.W = W;
    B.lab0 = p.lab + (uint64_t)blockIdx.x * SL * W;
    B.mask0 = p.mask + (uint64_t)blockIdx.x * 2 * W;
    B.q0 = p.qg + (uint64_t)blockIdx.x * 2 * W;
    uint32_t off = p.bstate[2 * blockIdx.x];
    uint32_t since = p.bstate[2 * blockIdx.x + 1];
    TallyB t{0, 0, 0, 0};
    unsigned long long levels = 0, tot_te = 0, tot_reach = 0;
    uint32_t* fb = S.fbuf[w];
    uint32_t fcount = 0;
    for (uint32_t i = tid; i < (uint32_t)(T * SL); i += BB) (&S.cell[0][0])[i] = INF;
    for (uint32_t i = tid; i < (uint32_t)H; i += BB) S.hkey[i] = 0ull;
    if (tid == 0) {
        S.qcnt[0] = S.qcnt[1] = 0;
        S.cursor = S.tcnt = S.held = 0;
    }
    for (;;) {
        if (tid == 0) S.chunk = atomicAdd(&p.ctrl->src_next, 1u);
        __syncthreads();
        const uint32_t ck = S.chunk;
        if (ck >= p.n_chunks) break;
        const BlkChunk C = p.chunks[ck];
        const uint32_t nsrc = C.count;
        B.wb = p.mat_base[C.mat];
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
        B.off = off;
        B.stale = off + W;
        // this lane's source (window-local); lanes of slots >= nsrc never
        // receive a mask bit
        const uint32_t my_src = S.src[j] - B.wb;

        // ---- seeding (symbfact.cuh:639-721) by wave 0 alone, slot after slot,
        // so the mask read-modify-writes of different slots cannot race.
        // Every neighbour's label becomes off: below the source that is the
        // reference's seed label, above it the "touched" state; either way
        // neighbours never count as fills (the reference's pre-stamp).
        if (w == 0) {
            for (uint32_t k = 0; k < nsrc; k++) {
                const uint32_t srcg = S.src[k];
                const uint32_t src = srcg - B.wb;
                const int32_t b = p.rp[srcg], e = p.rp[srcg + 1];
                unsigned long long te = lane == 0 ? (unsigned long long)(e - b) : 0ull, rc = 0;
                if (lane == 0) {
                    at(B.lab0, (src * SL + k) << 2) = off;
                    t.checks += (uint32_t)(e - b);
                }
                for (int32_t i = b + (int32_t)lane; i < e; i += WAVE) {
                    const uint32_t mg = (uint32_t)p.ci[i];
                    const uint32_t m = mg - B.wb;
                    const uint32_t o = (m * SL + k) << 2;
                    const uint32_t old = at(B.lab0, o);
                    at(B.lab0, o) = off;
                    if (m < src) {
                        if (old >= B.stale) {
                            rc++;
                            te += (uint32_t)(p.rp[mg + 1] - p.rp[mg]);
                        }
                        const uint32_t mo = B.mask0[m];  // level-0 buffer
                        B.mask0[m] = mo | (1u << k);
                        if (mo == 0u) B.q0[atomicAdd(&S.qcnt[0], 1u)] = m;
                    }
                }
                te = wave_sum64(te);
                rc = wave_sum64(rc);
                if (lane == 0) {
                    S.s_te[k] += te;
                    S.s_reach[k] += rc;
                }
            }
        }
        __syncthreads();

        // ---- level loop
        uint32_t cur = 0;
        for (;;) {
            const uint32_t size = S.qcnt[cur];
            __syncthreads();  // every thread has the size before it can change
            if (size == 0) break;
            const uint32_t nbuf = cur ^ 1u;
            HalfVtx hv;
            hv.held = false;
            for (;;) {  // rounds
                combine_phase(B, S, hv, cur, size, t);
                if (hv.held && j == 0) atomicOr(&S.held, 1u);
                __syncthreads();
                // final for this round: the cursor moves only in phase A
                const bool more = S.held != 0u || S.cursor < size;
                relax_phase(B, S, nbuf, my_src, fb, fcount, t);
                __syncthreads();
                for (uint32_t i = tid; i < (uint32_t)H; i += BB) S.hkey[i] = 0ull;
                if (tid == 0) {
                    S.tcnt = 0;
                    S.held = 0;
                }
                __syncthreads();
                if (!more) break;
            }
            if (tid == 0) {
                S.qcnt[cur] = 0;
                S.cursor = 0;
            }
            __syncthreads();
            cur = nbuf;
            levels++;
        }
        // staged fills name their slot: flush before the slots change
        if (fcount) flush_fills_w(p, S, fb, fcount);
        fcount = 0;
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
