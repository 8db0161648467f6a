// gsofa_reach.hip -- two-pass reachability traversal (GSOFA_MODE_REACH).
//
// The fill-path theorem splits row s of struct(L+U) into two reachability
// questions that need no labels:
//   U_s = {m > s : m is reached from s through vertices below s}, and with it
//         R_<(s) (the vertices below s so reached) and TE_alg;
//   L_s = {m < s : s is reached from m through vertices below m in A^T}, since
//         a path s -> ... -> m of A whose intermediates lie below m is a path
//         m -> ... -> s of A^T whose intermediates lie below its source m.
// So one kernel answers both: over A it yields U rows, |R| and TE exactly as
// the reference's traversal does for U (with_space_optimization/
// symbfact.cuh:883-1082, the "m > src" branch at :992-1009); over A^T it
// yields L columns, written to the fill stream as (row, column) pairs so the
// structure pass builds the same sorted rows.  Each (vertex, source) pair is
// expanded at most once, where the reference's max-id labels re-expand a
// vertex each time its label drops (2-10 times per pair on 3D
// nested-dissection grids).
//
// Execution model (CDNA4, persistent grid of co-resident workgroups: one of
// 1024 threads per CU with the largest target table, or two of 512 threads
// with half the table each, so one runs while the other waits at a barrier):
// * A workgroup owns MB source slots (32, 64, 128 or 256: the bits of a mask
//   of one, two or four 64-bit words) and runs a mini-group of up to MB
//   sources of one matrix, in one of the two passes, through all of its
//   levels alone.  The next mini-group comes from one global counter over a
//   work table that interleaves the passes' chunks, deepest first, so both
//   passes share one launch and its tail.
// * Per vertex, one word holds the MB slots' visited bits with the
//   mini-group's stamp beside it; an older stamp reads as "nothing visited",
//   so the array is never cleared between mini-groups (the reference's
//   source-stamped markers, symbfact.cuh:2126, generalised to a bit per slot).
// * A level runs in rounds of two phases separated by workgroup barriers:
//     A. combine: lane groups of G lanes pop frontier vertices (all lanes of
//        a group load the same queue entry and mask, so nothing is
//        broadcast between lanes); a step covers G*E edges of the vertex,
//        lane j taking edges j, j+G, ... with all E loads issued first; each
//        target is found or inserted in an LDS hash table of the round's
//        distinct targets and the vertex's slot mask ORed into the target's
//        mask (ds_or);
//     B. relax: one lane per table entry: new = mask & ~visited; the new bits
//        of slots whose source lies above the target push it to the next
//        frontier (its R_<(s) membership), the others are U entries (fills
//        from level 1 on: level 0 expands the sources themselves, whose
//        neighbours are A's own entries).
//   Targets are unique within a round, so the visited word, the next-level
//   mask and the queue append need no global atomics.  Fills go to the
//   fill stream through blocks that a workgroup reserves with one global
//   atomic and hands out to its waves from LDS (the reserved entries it
//   leaves unused hold ~0, which the structure pass skips), so waves of the
//   whole grid do not queue on one counter.  The round's counters
//   alternate by parity and the level's by level parity, so a round costs two
//   barriers and a level no more.
// * Sources of a mini-group form an arithmetic sequence first, first-stride,
//   ... (the cyclic source map of symbfact.cuh:1084-1099 within a share), so
//   "source j lies above m" is the prefix of slots j < ceil((first-m)/stride).
#include <algorithm>

#include "gsofa_device.h"
#include "gsofa_internal.h"

namespace gsofa {

namespace {

using namespace dev;

// Workgroup shapes (threads, hash slots of the target table; a round takes
// at most half the slots in new targets, so the load stays below 1/2 +
// threads/slots):
//   wide: one 1024-thread workgroup per CU and the largest table that fits
//         the CU's 160 KB of LDS at 64 slots: fewest rounds per level;
//   pair: two 512-thread workgroups per CU with half the table each: one
//         workgroup runs while the other waits at a barrier.
constexpr int NT_WIDE = 1024, NH_WIDE = 11264;
constexpr int NT_PAIR = 512, NH_PAIR = 5632;
constexpr uint32_t FULLI = 0xFFFFFFFFu;
constexpr uint32_t PASS_T = 0x80000000u;  // chunk stride bit: the A^T pass
// Lane groups: 256 per workgroup, each taking 12 edges of its vertex per
// step: 4 lanes x 3 edges at 1024 threads, 2 lanes x 6 edges at 512 (256
// groups measured faster than 128 or 512 at either size).  GSOFA_RG and
// GSOFA_RE, when defined at build time, set one shape for both.
#ifndef GSOFA_RG
#define GSOFA_RG 0
#endif
#ifndef GSOFA_RE
#define GSOFA_RE 0
#endif
template <int NT>
constexpr int lanes_per_vertex() { return GSOFA_RG > 0 ? GSOFA_RG : NT / 256; }
template <int NT>
constexpr int edges_per_lane() { return GSOFA_RE > 0 ? GSOFA_RE : 12 / lanes_per_vertex<NT>(); }

template <class M>
struct MaskOps;
// Bit helpers shared by the scalar mask types.
template <class M, int B>
struct ScalarBits {
    static __device__ __forceinline__ M bit(uint32_t j) { return (M)1 << j; }
    static __device__ __forceinline__ M prefix(uint32_t c) {  // the low c bits
        return c >= (uint32_t)B ? ~(M)0 : (((M)1 << c) - (M)1);
    }
    static __device__ __forceinline__ M low_clear(M b) { return b & (b - (M)1); }
    template <class F>  // f(j) for every set bit j, in increasing order
    static __device__ __forceinline__ void for_each_bit(M m, F f) {
        if constexpr (B == 128) {
            for (uint64_t b = (uint64_t)m; b; b &= b - 1ull) f((uint32_t)__builtin_ctzll(b));
            for (uint64_t b = (uint64_t)(m >> 64); b; b &= b - 1ull) f(64u + (uint32_t)__builtin_ctzll(b));
        } else {
            for (M b = m; b; b &= b - (M)1) f((uint32_t)(B == 64 ? __builtin_ctzll(b) : __builtin_ctz(b)));
        }
    }
};
template <>
struct MaskOps<uint32_t> : ScalarBits<uint32_t, 32> {
    using V = uint2;  // visited word: (mask, stamp)
    static constexpr int BITS = 32;
    static __device__ __forceinline__ uint32_t get(const V& e, uint32_t st) { return e.y == st ? e.x : 0u; }
    static __device__ __forceinline__ V make(uint32_t m, uint32_t st) { return make_uint2(m, st); }
    static __device__ __forceinline__ uint32_t popc(uint32_t m) { return (uint32_t)__popc(m); }
    static __device__ __forceinline__ uint32_t ctz(uint32_t m) { return (uint32_t)__builtin_ctz(m); }
};
template <>
struct MaskOps<uint64_t> : ScalarBits<uint64_t, 64> {
    using V = uint4;  // visited word: (mask lo, mask hi, stamp, -)
    static constexpr int BITS = 64;
    static __device__ __forceinline__ uint64_t get(const V& e, uint32_t st) {
        return e.z == st ? (((uint64_t)e.y << 32) | e.x) : 0ull;
    }
    static __device__ __forceinline__ V make(uint64_t m, uint32_t st) {
        return make_uint4((uint32_t)m, (uint32_t)(m >> 32), st, 0u);
    }
    static __device__ __forceinline__ uint32_t popc(uint64_t m) { return (uint32_t)__popcll(m); }
    static __device__ __forceinline__ uint32_t ctz(uint64_t m) { return (uint32_t)__builtin_ctzll(m); }
};
typedef unsigned __int128 u128;
struct __align__(16) V128 {  // visited word of 128 slots: mask halves, stamp (32 B)
    uint64_t lo, hi;
    uint32_t stamp, pad[3];
};
template <>
struct MaskOps<u128> : ScalarBits<u128, 128> {
    using V = V128;
    static constexpr int BITS = 128;
    static __device__ __forceinline__ u128 get(const V& e, uint32_t st) {
        return e.stamp == st ? (((u128)e.hi << 64) | e.lo) : (u128)0;
    }
    static __device__ __forceinline__ V make(u128 m, uint32_t st) {
        V v;
        v.lo = (uint64_t)m;
        v.hi = (uint64_t)(m >> 64);
        v.stamp = st;
        v.pad[0] = v.pad[1] = v.pad[2] = 0u;
        return v;
    }
    static __device__ __forceinline__ uint32_t popc(u128 m) {
        return (uint32_t)(__popcll((uint64_t)m) + __popcll((uint64_t)(m >> 64)));
    }
    static __device__ __forceinline__ uint32_t ctz(u128 m) {
        const uint64_t lo = (uint64_t)m;
        return lo ? (uint32_t)__builtin_ctzll(lo) : 64u + (uint32_t)__builtin_ctzll((uint64_t)(m >> 64));
    }
};

// 256 slots: four 64-bit words.
struct M256 {
    uint64_t w[4];
    M256() = default;
    __device__ __forceinline__ constexpr M256(int z) : w{(uint64_t)z, 0ull, 0ull, 0ull} {}
    __device__ __forceinline__ explicit operator bool() const { return (w[0] | w[1] | w[2] | w[3]) != 0ull; }
};
__device__ __forceinline__ M256 operator&(const M256& a, const M256& b) {
    M256 r;
    for (int k = 0; k < 4; k++) r.w[k] = a.w[k] & b.w[k];
    return r;
}
__device__ __forceinline__ M256 operator|(const M256& a, const M256& b) {
    M256 r;
    for (int k = 0; k < 4; k++) r.w[k] = a.w[k] | b.w[k];
    return r;
}
__device__ __forceinline__ M256 operator~(const M256& a) {
    M256 r;
    for (int k = 0; k < 4; k++) r.w[k] = ~a.w[k];
    return r;
}
__device__ __forceinline__ bool operator==(const M256& a, const M256& b) {
    return ((a.w[0] ^ b.w[0]) | (a.w[1] ^ b.w[1]) | (a.w[2] ^ b.w[2]) | (a.w[3] ^ b.w[3])) == 0ull;
}
struct __align__(16) V256 {  // visited word of 256 slots: mask words, stamp (48 B)
    uint64_t w[4];
    uint32_t stamp, pad[3];
};
template <>
struct MaskOps<M256> {
    using V = V256;
    static constexpr int BITS = 256;
    static __device__ __forceinline__ M256 get(const V& e, uint32_t st) {
        M256 r(0);
        if (e.stamp == st)
            for (int k = 0; k < 4; k++) r.w[k] = e.w[k];
        return r;
    }
    static __device__ __forceinline__ V make(const M256& m, uint32_t st) {
        V v;
        for (int k = 0; k < 4; k++) v.w[k] = m.w[k];
        v.stamp = st;
        v.pad[0] = v.pad[1] = v.pad[2] = 0u;
        return v;
    }
    static __device__ __forceinline__ uint32_t popc(const M256& m) {
        return (uint32_t)(__popcll(m.w[0]) + __popcll(m.w[1]) + __popcll(m.w[2]) + __popcll(m.w[3]));
    }
    static __device__ __forceinline__ uint32_t ctz(const M256& m) {
        for (int k = 0; k < 3; k++)
            if (m.w[k]) return 64u * (uint32_t)k + (uint32_t)__builtin_ctzll(m.w[k]);
        return 192u + (uint32_t)__builtin_ctzll(m.w[3]);
    }
    static __device__ __forceinline__ M256 bit(uint32_t j) {
        M256 r(0);
        r.w[j >> 6] = 1ull << (j & 63u);
        return r;
    }
    static __device__ __forceinline__ M256 prefix(uint32_t c) {
        M256 r;
        for (int k = 0; k < 4; k++) {
            const uint32_t lo = 64u * (uint32_t)k;
            r.w[k] = c >= lo + 64u ? ~0ull : c <= lo ? 0ull : ((1ull << (c - lo)) - 1ull);
        }
        return r;
    }
    template <class F>
    static __device__ __forceinline__ void for_each_bit(const M256& m, F f) {
        for (int k = 0; k < 4; k++)
            for (uint64_t b = m.w[k]; b; b &= b - 1ull) f(64u * (uint32_t)k + (uint32_t)__builtin_ctzll(b));
    }
    static __device__ __forceinline__ M256 low_clear(M256 b) {
        for (int k = 0; k < 4; k++)
            if (b.w[k]) {
                b.w[k] &= b.w[k] - 1ull;
                break;
            }
        return b;
    }
};

// OR into an LDS mask (128 and 256 slots: per 64-bit word that has bits).
__device__ __forceinline__ void mask_or(uint32_t* p, uint32_t v) { atomicOr(p, v); }
__device__ __forceinline__ void mask_or(uint64_t* p, uint64_t v) {
    atomicOr(reinterpret_cast<unsigned long long*>(p), (unsigned long long)v);
}
__device__ __forceinline__ void mask_or(u128* p, u128 v) {
    unsigned long long* q = reinterpret_cast<unsigned long long*>(p);
    if ((uint64_t)v) atomicOr(q, (unsigned long long)(uint64_t)v);
    if ((uint64_t)(v >> 64)) atomicOr(q + 1, (unsigned long long)(uint64_t)(v >> 64));
}
__device__ __forceinline__ void mask_or(M256* p, const M256& v) {
    unsigned long long* q = reinterpret_cast<unsigned long long*>(p);
    for (int k = 0; k < 4; k++)
        if (v.w[k]) atomicOr(q + k, (unsigned long long)v.w[k]);
}

// One- and two-word masks (32, 64 slots) merge bits that reach a vertex still
// queued for the current level into that level's mask, so it expands them in
// the same pop: on 3D nested-dissection grids, where a vertex gets bits from
// different sources at neighbouring levels, that halves the traversal.  The
// two masks of a vertex (this level's, the next's) then sit side by side, so
// relax reads both with one access.  Wider masks, whose sources' searches
// mostly advance together (banded matrices), keep the two levels' masks in
// separate arrays and do not merge: the extra reads cost more there.
template <class M>
constexpr bool merges() { return sizeof(M) <= 8; }

// Hash slots of the target table for a shape and mask width: the largest
// table that fits the CU's LDS (4 B key, the mask, 2 B entry per slot).
template <class M>
constexpr int tab_slots(int threads) {
    return sizeof(M) == 32   ? (threads == NT_WIDE ? 4224 : 2048)
           : sizeof(M) == 16 ? (threads == NT_WIDE ? 7168 : 3584)
                             : (threads == NT_WIDE ? NH_WIDE : NH_PAIR);
}

template <class M, int H>
struct RLds {
    uint32_t hkey[H];                    // target + 1 (window-local); 0 = empty
    M hmask[H];                          // OR of the masks folded into the target this round
    uint16_t ent[H];                     // occupied hash slots in insertion order
    unsigned long long s_fill[MaskOps<M>::BITS];  // counts only: U fills per slot (A pass)
    unsigned long long acc[8];           // workgroup totals: fills, checks, -, expansions,
                                         // TE and |R| of the A pass, of the A^T pass
    unsigned long long cte, creach;      // the current mini-group's TE and |R|
    uint32_t qcnt[2];                    // queue sizes, by buffer
    uint32_t cursor[2];                  // pop cursors, by level parity
    uint32_t tcnt[2];                    // inserts of this round, by round parity
    uint32_t held[2];                    // some group holds unfolded edges, by round parity
    uint32_t chunk;
    uint32_t fcur;                       // entries handed out of the current fill block
    unsigned long long fbase;            // the block's first entry in the fill stream
};

// Hash slot of window-local target m, inserting it if new; FULLI when it is
// new and the round already holds T targets.  Linear probing.
template <class M, int H>
__device__ __forceinline__ uint32_t rtab_find(RLds<M, H>& S, uint32_t* tcnt, uint32_t m) {
    constexpr int T = H / 2;
    const uint32_t key = m + 1u;
    uint32_t s = (uint32_t)(((uint64_t)(m * 2654435761u) * (uint32_t)H) >> 32);
    {
        const uint32_t cur = S.hkey[s];  // usual case: present at its first slot
        if (cur == key) return s;
    }
    for (int probe = 0; probe < H; probe++) {
        const uint32_t cur = S.hkey[s];
        if (cur == key) return s;
        if (cur == 0u) {
            if (*tcnt >= (uint32_t)T) return FULLI;
            const uint32_t prev = atomicCAS(&S.hkey[s], 0u, key);
            if (prev == 0u) {
                S.ent[atomicAdd(tcnt, 1u)] = (uint16_t)s;
                return s;
            }
            if (prev == key) return s;
        }
        s = s + 1u == (uint32_t)H ? 0u : s + 1u;
    }
    return FULLI;
}

// Slots j < count whose source first - j*stride lies above global vertex mg.
template <class M>
__device__ __forceinline__ M above_mask(uint32_t mg, uint32_t first, uint32_t stride, uint32_t count) {
    if (mg >= first) return (M)0;
    const uint32_t d = first - mg;
    uint32_t c = stride == 1u ? d : (d + stride - 1u) / stride;
    c = c < count ? c : count;
    return MaskOps<M>::prefix(c);
}

// The calling lane group's bits of a wave ballot (bit k: lane k of the group).
template <int G>
__device__ __forceinline__ uint32_t grp_ballot(bool x, uint32_t grp) {
    return (uint32_t)(__ballot(x) >> (G * grp)) & ((1u << G) - 1u);
}

// The frontier vertex a lane group holds across rounds (group-uniform: every
// lane of the group loads the same queue entry and mask, so no broadcast is
// needed): its slot mask, edge cursor, row end and, per lane, the edges of
// its current step not yet folded.
template <class M>
struct RGrp {
    bool held;           // a vertex is in progress
    bool done;           // the queue is exhausted for this level
    uint32_t pos, end, pend;
    M vm;
};

struct RTally {
    unsigned long long fills, checks, expansions, te, reach;
};

// The mini-group being run: its chunk, matrix window and pass.
struct RCur {
    const int32_t* rp;
    const int32_t* ci;
    uint32_t first, count, stride, mat, wb;
    bool tpass;
};

// Phase A: lane groups of G lanes pop frontier vertices one at a time; a
// step of a group covers the next G*E edges of its vertex, lane j taking
// edges j, j+G, ..., j+(E-1)G of the step (all E loads issued before the
// first fold).  Each target is found or inserted in the table and the
// vertex's mask folded into it, until the queue is exhausted or the table is
// full.  Wave-uniform control flow; the groups hold independent vertices.
template <int H, int G, int E, class M>
__device__ void rcombine(RLds<M, H>& S, RGrp<M>& hv, const RCur& c, M* maskc, const uint4* qc, uint32_t size,
                         uint32_t* cursor, uint32_t rb, RTally& t) {
    using O = MaskOps<M>;
    constexpr int T = H / 2;
    const uint32_t lane = __lane_id();
    const uint32_t grp = lane / G;
    const uint32_t j = lane % G;
    constexpr uint32_t ALL = (1u << E) - 1u;  // a lane's edges of one step
    for (;;) {
        const bool want = !hv.held && !hv.done && S.tcnt[rb] < (uint32_t)T;
        const uint64_t wl = __ballot(want && j == 0u);
        uint32_t base = 0;
        if (lane == 0 && wl) base = atomicAdd(cursor, (uint32_t)__popcll(wl));
        base = bcast0(base);
        uint32_t m[E];     // targets of the current step
        bool fresh = false;  // m holds the first step of a vertex just popped
        if (want) {
            const uint32_t idx = base + (uint32_t)__popcll(wl & ((1ull << (G * grp)) - 1ull));
            if (idx >= size) {
                hv.done = true;
            } else {
                const uint4 e = qc[idx];
                // the mask and the first step's edges load together
                const M mk = maskc[(merges<M>() ? 2u : 1u) * e.x];
#pragma unroll
                for (int k = 0; k < E; k++) {
                    const uint32_t ek = e.y + j + (uint32_t)(G * k);
                    m[k] = ek < e.z ? (uint32_t)c.ci[ek] - c.wb : FULLI;
                }
                fresh = true;
                hv.pos = e.y;
                hv.end = e.z;
                hv.pend = ALL;
                hv.vm = mk;
                hv.held = e.y < e.z;
                if (j == 0u) {
                    maskc[(merges<M>() ? 2u : 1u) * e.x] = (M)0;  // consumed; refilled two levels later
                    t.checks += (unsigned long long)(e.z - e.y) * O::popc(mk);
                    t.expansions++;
                }
            }
        }
        if (!__ballot(hv.held)) {
            // no group holds edges: pop again while the queue and the table allow
            if (__ballot(!hv.done && S.tcnt[rb] < (uint32_t)T)) continue;
            return;
        }
        bool failed = false;
        if (hv.held) {
            if (!fresh) {
#pragma unroll
                for (int k = 0; k < E; k++) {
                    const uint32_t e = hv.pos + j + (uint32_t)(G * k);
                    m[k] = FULLI;
                    if (((hv.pend >> k) & 1u) && e < hv.end) m[k] = (uint32_t)c.ci[e] - c.wb;
                }
            }
            uint32_t left = 0;
#pragma unroll
            for (int k = 0; k < E; k++) {
                if (m[k] != FULLI) {
                    const uint32_t hs = rtab_find<M, H>(S, &S.tcnt[rb], m[k]);
                    if (hs == FULLI) left |= 1u << k;
                    else mask_or(&S.hmask[hs], hv.vm);
                }
            }
            hv.pend = left;  // the unplaced edges stay for the next round
            failed = left != 0u;
        }
        const uint32_t fm = grp_ballot<G>(failed, grp);
        if (hv.held && !fm) {
            hv.pos += (uint32_t)(G * E);
            hv.pend = ALL;
            hv.held = hv.pos < hv.end;
        }
        if (__ballot(fm != 0u)) return;  // the table is full for everyone
    }
}

template <int RB, int H, int G, int E, class M>
__global__ __launch_bounds__(RB) void k_reach_blk(RchParams p) {
    static_assert(H / 2 + RB <= H, "insert overshoot stays within the entry list");
    static_assert(H <= 65536, "entries are 16-bit slot numbers");
    static_assert(WAVE % G == 0 && G < 32 && E >= 1 && E < 32, "lane groups tile a wave; a lane's edges fit its pending bits");
    using O = MaskOps<M>;
    using VW = typename O::V;
    __shared__ RLds<M, H> S;
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = __lane_id();
    const uint32_t w = tid / WAVE;
    const uint32_t W = p.W;
    VW* vis = reinterpret_cast<VW*>(p.vis) + (uint64_t)blockIdx.x * W;
    M* nmask0 = reinterpret_cast<M*>(p.nmask) + (uint64_t)blockIdx.x * 2 * W;
    uint4* q0 = p.q + (uint64_t)blockIdx.x * 2 * W;
    uint32_t epoch = p.bstate[blockIdx.x];
    RTally t{0, 0, 0, 0, 0};
    unsigned long long levels = 0;
    for (uint32_t i = tid; i < (uint32_t)H; i += RB) {
        S.hkey[i] = 0u;
        S.hmask[i] = (M)0;
    }
    if (tid < 8) S.acc[tid] = 0;
    const uint32_t fblk = p.fill_blk, fmax = p.fill_blk >> 4;  // block; largest batch served from it
    if (tid == 0) {
        S.qcnt[0] = S.qcnt[1] = 0;
        S.tcnt[0] = S.tcnt[1] = S.held[0] = S.held[1] = 0;
        S.fcur = fblk;  // no block yet
        S.fbase = 0;
    }
    uint32_t rb = 0;  // round parity
    for (;;) {
        if (tid == 0) {
            S.chunk = atomicAdd(&p.ctrl->src_next, 1u);
            S.cte = S.creach = 0;
            S.cursor[0] = S.cursor[1] = 0;
        }
        if (tid < (uint32_t)O::BITS) S.s_fill[tid] = 0;
        __syncthreads();
        const uint32_t ck = S.chunk;
        if (ck >= p.n_chunks) break;
        const BlkChunk C = p.chunks[ck];
        RCur c;
        c.tpass = (C.stride & PASS_T) != 0u;
        c.rp = c.tpass ? p.rpT : p.rp;
        c.ci = c.tpass ? p.ciT : p.ci;
        c.first = C.first;
        c.count = C.count;
        c.stride = C.stride & ~PASS_T;
        c.mat = C.mat;
        c.wb = p.mat_base[C.mat];
        if (epoch >= p.loops) {  // stamps exhausted: clear this workgroup's visited words
            for (uint32_t i = tid; i < W; i += RB) vis[i] = O::make((M)0, 0u);
            epoch = 0;
            __syncthreads();  // cleared before the sources are stamped
        }
        epoch++;
        const uint32_t stamp = epoch;
        // level 0: the sources themselves, each visited by and queued for its own slot
        if (tid < c.count) {
            const uint32_t sg = c.first - tid * c.stride;
            const uint32_t s = sg - c.wb;
            const M bit = O::bit(tid);
            vis[s] = O::make(bit, stamp);
            nmask0[(merges<M>() ? 2u : 1u) * s] = bit;  // level 0 uses the first mask
            const uint32_t r0 = (uint32_t)c.rp[sg], r1 = (uint32_t)c.rp[sg + 1];
            q0[tid] = make_uint4(s, r0, r1, 0u);
            t.te += r1 - r0;
        }
        if (tid == 0) S.qcnt[0] = c.count;
        __syncthreads();

        uint32_t cur = 0;
        for (uint32_t lev = 0;; lev++) {
            // final since the last barrier; reset during this level's relax
            const uint32_t size = S.qcnt[cur];
            if (size == 0) break;
            const uint32_t nxt = cur ^ 1u;
            // this level's and the next level's masks: side by side per vertex
            // (merging widths, index 2v) or two arrays (index v)
            constexpr uint32_t MS = merges<M>() ? 2u : 1u;
            M* maskc = nmask0 + (merges<M>() ? (uint64_t)cur : (uint64_t)cur * W);
            M* maskn = nmask0 + (merges<M>() ? (uint64_t)nxt : (uint64_t)nxt * W);
            const uint4* qc = q0 + (uint64_t)cur * W;
            uint4* qn = q0 + (uint64_t)nxt * W;
            uint32_t* cursor = &S.cursor[lev & 1u];
            RGrp<M> hv;
            hv.held = false;
            hv.done = false;
            hv.pend = 0u;
            hv.pos = hv.end = 0u;
            hv.vm = (M)0;
            for (;;) {  // rounds, two barriers each
                // a fresh fill block once the current one may not hold a
                // batch; relax reads fbase and fcur only after the next barrier
                if (tid == 0 && p.fill_cap != 0ull && S.fcur + fmax > fblk) {
                    S.fbase = aadd64(&p.ctrl->fill_pos, fblk);
                    S.fcur = 0;
                }
                rcombine<H, G, E, M>(S, hv, c, maskc, qc, size, cursor, rb, t);
                if (hv.held && (lane % G) == 0u) atomicOr(&S.held[rb], 1u);
                __syncthreads();
                // final for this round: the cursor moves only in phase A
                const bool more = S.held[rb] != 0u || *cursor < size;
                // ---- B. relax: one lane per table entry
                const uint32_t tc = S.tcnt[rb];
                const uint32_t ntab = tc < (uint32_t)H ? tc : (uint32_t)H;
                for (uint32_t b0 = w * WAVE; b0 < ntab; b0 += RB) {
                    const uint32_t i = b0 + lane;
                    const bool ok = i < ntab;
                    uint32_t m = 0;
                    M mk = (M)0;
                    if (ok) {
                        const uint32_t hs = S.ent[i];
                        m = S.hkey[hs] - 1u;
                        mk = S.hmask[hs];
                        S.hkey[hs] = 0u;
                        S.hmask[hs] = (M)0;
                    }
                    const uint32_t mg = m + c.wb;
                    M up = (M)0;
                    uint32_t r0 = 0, r1 = 0;
                    bool enq = false;
                    if (ok) {
                        const VW vw = vis[m];
                        const M mc = merges<M>() ? maskc[MS * m] : (M)0;
                        const M mn = maskn[MS * m];
                        r0 = (uint32_t)c.rp[mg];
                        r1 = (uint32_t)c.rp[mg + 1];
                        const M vv = O::get(vw, stamp);
                        const M nw = mk & ~vv;
                        if (nw) {
                            vis[m] = O::make(vv | nw, stamp);
                            const M bl = nw & above_mask<M>(mg, c.first, c.stride, c.count);
                            up = lev ? (nw & ~bl) : (M)0;  // level 0: A's own entries
                            if (bl) {
                                if (merges<M>() && lev != 0u && mc) {
                                    // still queued for this level and not yet
                                    // popped: it expands these bits too, one
                                    // expansion fewer (not at level 0, whose
                                    // edges are A's own entries of the sources)
                                    maskc[MS * m] = mc | bl;
                                } else {
                                    maskn[MS * m] = mn | bl;
                                    enq = mn == (M)0;
                                }
                                const uint32_t nb = O::popc(bl);
                                t.reach += nb;
                                t.te += (unsigned long long)nb * (r1 - r0);
                            }
                        }
                    }
                    const uint64_t em = __ballot(enq);
                    if (em) {
                        uint32_t qb = 0;
                        if (lane == 0) qb = atomicAdd(&S.qcnt[nxt], (uint32_t)__popcll(em));
                        qb = bcast0(qb);
                        if (enq) qn[qb + rank_in(em)] = make_uint4(m, r0, r1, 0u);
                    }
                    const uint32_t nf = O::popc(up);
                    t.fills += nf;
                    if (__ballot(nf != 0u)) {
                        if (p.fill_cap != 0ull) {
                            // (row << 32) | column, one reservation per wave:
                            // from the workgroup's block, or (a batch larger
                            // than a sixteenth of a block, or one that no
                            // longer fits it) straight from the stream
                            const uint32_t incl = wave_incl_scan(nf);
                            const uint32_t total = __shfl(incl, WAVE - 1, WAVE);
                            uint64_t fp = 0;
                            if (lane == 0) {
                                const uint32_t old = total <= fmax ? atomicAdd(&S.fcur, total) : FULLI;
                                fp = (old != FULLI && old + total <= fblk) ? S.fbase + old
                                                                           : aadd64(&p.ctrl->fill_pos, total);
                            }
                            fp = bcast0_64(fp) + (incl - nf);
                            O::for_each_bit(up, [&](uint32_t j) {
                                const uint32_t src = c.first - j * c.stride;
                                const uint64_t pr = c.tpass ? (((uint64_t)mg << 32) | src)
                                                            : (((uint64_t)src << 32) | mg);
                                if (fp < p.fill_cap) p.fills[fp] = pr;
                                fp++;
                            });
                        } else if (nf) {  // counts only: fills per row
                            if (c.tpass) {
                                aadd64(&p.row_fills[mg], nf);
                            } else {
                                O::for_each_bit(up, [&](uint32_t j) { atomicAdd(&S.s_fill[j], 1ull); });
                            }
                        }
                    }
                }
                // counters of the next round, the consumed queue and the next
                // level's cursor, cleared while others still relax: each was
                // last read before this round's first barrier
                if (tid == 0) {
                    S.tcnt[rb ^ 1u] = 0;
                    S.held[rb ^ 1u] = 0;
                    S.qcnt[cur] = 0;
                    S.cursor[(lev + 1u) & 1u] = 0;
                }
                __syncthreads();
                rb ^= 1u;
                if (!more) break;
            }
            cur = nxt;
            levels++;
        }
        // the mini-group's TE and |R| (per matrix for the A pass); counts-only
        // fills per source row
        const unsigned long long te = wave_sum64(t.te), rc = wave_sum64(t.reach);
        if (lane == 0 && (te | rc)) {
            atomicAdd(&S.cte, te);
            atomicAdd(&S.creach, rc);
        }
        t.te = t.reach = 0;
        __syncthreads();
        if (tid == 0) {
            atomicAdd(&S.acc[c.tpass ? 6 : 4], S.cte);
            atomicAdd(&S.acc[c.tpass ? 7 : 5], S.creach);
            if (!c.tpass && p.mat_te) {
                aadd64(&p.mat_te[c.mat], S.cte);
                aadd64(&p.mat_reach[c.mat], S.creach);
            }
        }
        if (p.fill_cap == 0ull && !c.tpass && tid < c.count && S.s_fill[tid])
            aadd64(&p.row_fills[c.first - tid * c.stride], S.s_fill[tid]);
        __syncthreads();
    }
    if (t.fills) atomicAdd(&S.acc[0], t.fills);
    if (t.checks) atomicAdd(&S.acc[1], t.checks);
    if (t.expansions) atomicAdd(&S.acc[3], t.expansions);
    __syncthreads();
    if (tid == 0) {
        Ctrl* ct = p.ctrl;
        if (S.acc[0]) aadd64(&ct->fill_count, S.acc[0]);
        if (S.acc[1]) aadd64(&ct->edge_checks, S.acc[1]);
        if (S.acc[3]) aadd64(&ct->expansions, S.acc[3]);
        if (S.acc[4]) aadd64(&ct->te, S.acc[4]);
        if (S.acc[5]) aadd64(&ct->reach, S.acc[5]);
        if (S.acc[6]) aadd64(&ct->te_t, S.acc[6]);
        if (S.acc[7]) aadd64(&ct->reach_t, S.acc[7]);
        aadd64(&ct->levels, levels);
        p.bstate[blockIdx.x] = epoch;
    }
}

}  // namespace

bool rch_bits_valid(int bits) { return bits == 32 || bits == 64 || bits == 128 || bits == 256; }
bool rch_threads_valid(int threads) { return threads == NT_WIDE || threads == NT_PAIR; }

namespace {
template <int NT, class M>
const void* kptr() {
    return reinterpret_cast<const void*>(
        &k_reach_blk<NT, tab_slots<M>(NT), lanes_per_vertex<NT>(), edges_per_lane<NT>(), M>);
}
template <int NT, class M>
void klaunch(const RchParams& p, hipStream_t s) {
    hipLaunchKernelGGL((k_reach_blk<NT, tab_slots<M>(NT), lanes_per_vertex<NT>(), edges_per_lane<NT>(), M>),
                       dim3(p.nblocks), dim3(NT), 0, s, p);
}
template <int NT>
const void* kptr_bits(int bits) {
    return bits == 256   ? kptr<NT, M256>()
           : bits == 128 ? kptr<NT, u128>()
           : bits == 64  ? kptr<NT, uint64_t>()
                         : kptr<NT, uint32_t>();
}
template <int NT>
void klaunch_bits(const RchParams& p, int bits, hipStream_t s) {
    if (bits == 256) klaunch<NT, M256>(p, s);
    else if (bits == 128) klaunch<NT, u128>(p, s);
    else if (bits == 64) klaunch<NT, uint64_t>(p, s);
    else klaunch<NT, uint32_t>(p, s);
}
}  // namespace

const void* rch_kernel_ptr(int bits, int threads) {
    return threads == NT_PAIR ? kptr_bits<NT_PAIR>(bits) : kptr_bits<NT_WIDE>(bits);
}
uint32_t rch_pass_bit() { return PASS_T; }
// Per workgroup: visited words (8, 16 or 32 B), two next-level masks, two
// 16-byte frontier queues, per vertex of the window.
int64_t reach_bytes(uint32_t nb, uint32_t W, int bits) {
    const int64_t vw = rch_visit_bytes(bits), mw = bits / 8;
    return (int64_t)nb * ((vw + 2 * mw + 32) * (int64_t)W + 8);
}

// Every block a workgroup gives up held more than 15/16 of a block of fills:
// a batch is served from a block only if it is at most a sixteenth of it,
// and the block is given up once the rest may not hold such a batch.  With
// its last block, a workgroup's fill-stream span is below 16/15 of its fills
// plus one block; over the grid, below 16/15 need + nblocks * block.  The
// block size keeps nblocks * block within need/8 + 2^20, which
// rch_fill_room covers (the 64-entry floor keeps that for nblocks <= 16384).
uint32_t rch_fill_block(uint64_t need, uint32_t nblocks) {
    const uint64_t nb = std::max<uint32_t>(nblocks, 1u);
    uint32_t blk = 1u << 16;
    while (blk > 64u && nb * blk > need / 8 + (1ull << 20)) blk >>= 1;
    return blk;
}
int64_t rch_visit_bytes(int bits) { return bits == 256 ? 48 : bits == 128 ? 32 : bits == 64 ? 16 : 8; }
uint64_t rch_fill_room(uint64_t need) { return need + need / 15 + need / 8 + (1ull << 20) + 1; }

hipError_t launch_reach(const RchParams& p, int bits, int threads, hipStream_t s) {
    if (!rch_bits_valid(bits) || !rch_threads_valid(threads)) return hipErrorInvalidValue;
    if (threads == NT_PAIR) klaunch_bits<NT_PAIR>(p, bits, s);
    else klaunch_bits<NT_WIDE>(p, bits, s);
    return hipGetLastError();
}

}  // namespace gsofa
