// gsofa_block.hip -- block-resident multi-source traversal (default mode).
//
// Same per-source arithmetic as the reference's warp-centric level kernel
// (with_space_optimization/symbfact.cuh:883-1082) and seeding (:639-721), with
// an execution model built for CDNA4:
//
// * Each workgroup (16 waves, one per CU) of a persistent grid owns SL source
//   slots (32, 16 or 8; one kernel instance each) and runs a "mini-group" of
//   SL consecutive sources through all of their frontier levels by itself,
//   synchronising only with __syncthreads.  The next mini-group comes from
//   one global counter.  There is no grid-wide barrier and no hot global
//   counter on the level path.
// * Labels are stored slot-interleaved, [vertex][SL slots]: the SL sources'
//   max-id labels of one vertex share one line (128 bytes at 32 slots).
//   Nearby sources reach mostly the same vertices in the same levels, so the
//   frontier is kept per vertex with a mask of the slots that must expand it
//   (deduplicated: a vertex enters the next queue when its mask leaves zero).
// * Work is spread over lane groups of SL lanes (two per wave at 32 slots,
//   eight at 8): lane j of a group always serves slot j.
// * A level runs in rounds of two phases.  The only global atomic on the
//   level path is one fill-stream reservation per wave and relax batch that
//   found fills (on this part a global atomic executes at the memory side,
//   not in the L2); everything else is LDS atomics and plain stores:
//     A. combine: each lane group pops frontier vertices v, computes lane
//        j's relaxation cost max(v, label_j(v)) for the slots j of v's mask
//        (symbfact.cuh:924), looks each neighbour m up in an LDS hash table
//        of the round's distinct targets, and folds the cost into that
//        target's SL per-slot cells with an LDS atomic min;
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
// * Work is a table of chunks (up to SL sources of one matrix).  A batch of
//   matrices is one block-diagonal matrix in which no traversal can leave its
//   own matrix, so per-workgroup arrays are indexed relative to the chunk's
//   matrix and sized by the largest matrix W, and labels hold
//   offset + (vertex - matrix base).
#include "gsofa_device.h"
#include "gsofa_internal.h"

namespace gsofa {

namespace {

using namespace dev;

#ifndef GSOFA_BB
#define GSOFA_BB 1024
#endif
#ifndef GSOFA_WPE
#define GSOFA_WPE 4
#endif
constexpr int BB = GSOFA_BB;         // threads per workgroup
constexpr int BW = BB / WAVE;        // waves per workgroup
constexpr uint32_t INF = 0xFFFFFFFFu;
constexpr uint32_t NONE = 0xFFFFFFFFu;   // no table entry (edge not in this group)
constexpr uint32_t FULL = 0xFFFFFFFEu;   // table full: retry the edge next round
constexpr uint32_t HOLE = 0xFFFFFFFFu;   // ent_m of an index that lost its insert

// Shape of one instantiation, by slots per workgroup SL (= lanes per lane
// group).  One workgroup of 16 waves holds a CU and nearly all of its 160 KB
// of LDS, most of it the per-round target table: the more distinct targets a
// round takes, the fewer rounds (and barriers) a level needs, and that is
// worth more than a second, smaller workgroup per CU.  Fewer slots leave
// room for more targets.
//   T  distinct targets per round, HB log2 of the hash slots (load <= 1/2),
//   KR table entries relaxed together per lane group (<= SL),
//   KV frontier vertices fetched together per lane group (<= SL).
template <int SL> struct Cfg;
template <> struct Cfg<32> { static constexpr int T = 1024, HB = 11, KR = 16, KV = 4; };
template <> struct Cfg<16> { static constexpr int T = 1792, HB = 12, KR = 16, KV = 4; };
template <> struct Cfg<8> { static constexpr int T = 2048, HB = 12, KR = 8, KV = 4; };

template <int SL>
struct Shape {
    static constexpr int T = Cfg<SL>::T;
    static constexpr int HB = Cfg<SL>::HB;
    static constexpr int H = 1 << HB;            // hash slots
    static constexpr int KR = Cfg<SL>::KR;
    static constexpr int KV = Cfg<SL>::KV;
    static constexpr int GPW = WAVE / SL;        // lane groups per wave
    static constexpr int NG = BW * GPW;          // lane groups per workgroup
    static constexpr uint32_t GM = SL == 32 ? 0xFFFFFFFFu : ((1u << SL) - 1u);  // a group's lanes
    static_assert(SL == 8 || SL == 16 || SL == 32, "slots per workgroup");
    static_assert(KR <= SL && KV <= SL, "one entry / vertex per lane of a group");
    static_assert(T <= H / 2, "hash load factor");
    static_assert(T % (GPW * KR) == 0, "relax batches tile the table");
    static_assert((T + 1) * SL * 4 + H * 8 + T * 4 + NG * SL * 4 + 1024 <= 160 * 1024, "LDS of one CU");
};

template <int SL>
struct BlkLds {
    using Z = Shape<SL>;
    uint32_t cell[Z::T + 1][SL];         // min cost per (round target, slot); INF = none.  Row T
                                         // is a sink for the padding of edge quads, never read
    unsigned long long hkey[Z::H];       // ((m + 1) << 32) | table index; 0 = empty
    uint32_t ent_m[Z::T];                // table index -> target (window-local) or HOLE
    uint4 eidx[Z::NG][SL / 4];           // table index of each lane's edge (SL per group)
    unsigned long long s_te[SL];         // per-slot TE of the current mini-group
    unsigned long long s_reach[SL];      // per-slot |R_<(s)|
    unsigned long long s_fill[SL];       // per-slot fills
    unsigned long long acc[4];           // workgroup totals: fills, TE, |R|, edge checks
    uint32_t src[SL];
    uint32_t qcnt[2];
    uint32_t cursor;
    uint32_t tcnt[2];                    // table entries in use, by round parity
    uint32_t held[2];                    // some group holds unfolded edges, by round parity
    uint32_t chunk;
};

struct Blk {
    const BlkParams* p;
    uint32_t* lab0;    // [W][SL]
    uint32_t* mask0;   // [2][W]
    uint4* q0;         // [2][W] frontier entries (v, row begin, row end, -)
    uint32_t W;
    uint32_t wb;       // window base (global id of local vertex 0)
    uint32_t off, stale;
    uint32_t rb;       // round parity: which tcnt / held belong to this round
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

// Per-lane tallies of the current mini-group (lane j of a group: slot j's te /
// reach / fills; one source's counts fit 32 bits since nnz < 2^31).
struct TallyB {
    uint32_t fills, te, reach;
    unsigned long long checks;
};

// Table index of target m in this round, inserting it if new; FULL when the
// table has no room left.  Linear probing over H slots; an index allocated for
// an insert that loses its race to an equal key becomes a HOLE entry.
template <int SL>
__device__ __forceinline__ uint32_t tab_find(BlkLds<SL>& S, uint32_t* tcnt, uint32_t m) {
    using Z = Shape<SL>;
    const unsigned long long key = (unsigned long long)(m + 1u) << 32;
    uint32_t s = (m * 2654435761u) >> (32 - Z::HB);
    {
        const unsigned long long cur = S.hkey[s];  // usual case: present at its first slot
        if ((cur & 0xFFFFFFFF00000000ull) == key) return (uint32_t)cur;
    }
    uint32_t mine = NONE, res = FULL;
    for (int probe = 0; probe < Z::H; probe++) {
        unsigned long long cur = S.hkey[s];
        if (cur == 0ull) {
            if (mine == NONE) {
                mine = atomicAdd(tcnt, 1u);
                if (mine >= (uint32_t)Z::T) {
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
        s = (s + 1u) & (uint32_t)(Z::H - 1);
    }
    if (mine != NONE) S.ent_m[mine] = HOLE;
    return res;
}

// Lane u's value of x within the caller's lane group of SL lanes: a swizzle
// broadcast inside each group of 32 lanes whose and-mask keeps the group's
// lane bits above log2(SL) and whose or-mask is u.  u must fold to a constant
// (the swizzle pattern is an immediate); the source lane must be active.
template <int SL, int U>
__device__ __forceinline__ uint32_t bcast_grp(uint32_t x) {
    return (uint32_t)__builtin_amdgcn_ds_swizzle((int)x, (0x1F & ~(SL - 1)) | (U << 5));
}
template <int SL, int U = 0>
__device__ __forceinline__ uint32_t from_lane(uint32_t x, uint32_t u) {
    if constexpr (U == SL) {
        return 0u;
    } else {
        if (u == (uint32_t)U) return bcast_grp<SL, U>(x);
        return from_lane<SL, U + 1>(x, u);
    }
}

// Edges 0 .. top-1 cover every group's placed edges: the OR of all groups'
// ballot bits, as a wave-uniform count.
template <int SL>
__device__ __forceinline__ uint32_t wave_top(uint64_t m) {
    if constexpr (SL <= 16) m |= m >> 32;
    if constexpr (SL <= 8) m |= m >> 16;
    m |= m >> SL;
    const uint32_t g = (uint32_t)m & Shape<SL>::GM;
    return g ? 32u - (uint32_t)__builtin_clz(g) : 0u;
}

// The calling lane group's bits of a wave ballot (bit k: lane k of the group).
template <int SL>
__device__ __forceinline__ uint32_t grp_ballot(bool x, uint32_t grp) {
    return (uint32_t)(__ballot(x) >> (SL * grp)) & Shape<SL>::GM;
}

// Frontier vertices held by a lane group across rounds.  Lane k of the group
// keeps vertex k's CSR cursor (pos, end) and its pending-edge mask for the
// current SL-edge group; lane j keeps slot j's cost of every vertex and the
// j-th edge target of its current group.
template <int KV>
struct GrpVtx {
    bool held;           // group-uniform: a batch is in progress
    bool done;           // group-uniform: the queue is exhausted for this level
    int32_t pos, end;    // lane k: vertex k's current group start, row end
    uint32_t pend;       // lane k: vertex k's edges of the group not yet folded
    uint32_t cost[KV];   // lane j: max(v_u + off, label_j(v_u)), INF if j not in v_u's mask
    uint32_t civ[KV];    // lane j: target of edge j of v_u's current group (window-local)
};

template <int SL>
__device__ __forceinline__ uint32_t group_mask(int32_t left) {
    return left >= SL ? Shape<SL>::GM : (left <= 0 ? 0u : ((1u << left) - 1u));
}

// Phase A: pop frontier vertices (KV per lane group) and fold their edges'
// costs into the table until the queue is exhausted or the table is full.
// Wave-uniform control flow; the lane groups carry independent batches.
template <int SL>
__device__ void combine_phase(const Blk& B, BlkLds<SL>& S, GrpVtx<Shape<SL>::KV>& hv, uint32_t cur,
                              uint32_t size, TallyB& t) {
    using Z = Shape<SL>;
    constexpr int KV = Z::KV;
    const BlkParams& p = *B.p;
    const uint32_t lane = __lane_id();
    const uint32_t grp = lane / SL;
    const uint32_t j = lane % SL;
    const uint32_t h = (threadIdx.x / WAVE) * Z::GPW + grp;
    uint32_t* maskc = B.mask0 + (uint64_t)cur * B.W;
    const uint4* qc = B.q0 + (uint64_t)cur * B.W;
    if (hv.held) {  // resuming after a full table: reload the pending groups' targets
#pragma unroll
        for (int u = 0; u < KV; u++) {
            const uint32_t pu = from_lane<SL>(hv.pend, u);
            const int32_t pos = (int32_t)from_lane<SL>((uint32_t)hv.pos, u);
            hv.civ[u] = ((pu >> j) & 1u) ? at(p.ci, (uint32_t)(pos + (int32_t)j) << 2) - B.wb : 0u;
        }
    }
    for (;;) {
        // -- take a new batch where a group has none
        const bool want = !hv.held && !hv.done && S.tcnt[B.rb] < (uint32_t)Z::T;
        uint32_t i0 = 0;
        if (j == 0 && want) i0 = atomicAdd(&S.cursor, (uint32_t)KV);
        i0 = from_lane<SL>(i0, 0);
        if (want && i0 >= size) hv.done = true;
        const bool fresh = want && i0 < size;
        if (__ballot(fresh)) {
            // lane k: queue entry (v, row begin, row end) and slot mask of vertex k
            const bool mine = fresh && j < (uint32_t)KV && i0 + j < size;
            uint4 q = make_uint4(NONE, 0u, 0u, 0u);
            if (mine) q = at(qc, (i0 + j) << 4);
            uint32_t M = 0;
            if (mine) M = at(maskc, q.x << 2);
            if (fresh) {
                hv.pos = (int32_t)q.y;
                hv.end = (int32_t)q.z;
                hv.pend = group_mask<SL>((int32_t)(q.z - q.y));
            }
            // every vertex's label line and first edge group, issued together
            uint32_t lv[KV];
#pragma unroll
            for (int u = 0; u < KV; u++) {
                const uint32_t vu = from_lane<SL>(q.x, u);
                const uint32_t pu = from_lane<SL>(hv.pend, u);
                const int32_t pos = (int32_t)from_lane<SL>((uint32_t)hv.pos, u);
                lv[u] = 0u;
                if (fresh && vu != NONE) lv[u] = at(B.lab0, (vu * SL + j) << 2);
                if (fresh) hv.civ[u] = ((pu >> j) & 1u) ? at(p.ci, (uint32_t)(pos + (int32_t)j) << 2) - B.wb : 0u;
            }
            if (mine) {
                at(maskc, q.x << 2) = 0u;  // consumed; this buffer is refilled two levels later
                t.checks += (unsigned long long)(q.z - q.y) * (uint32_t)__popc(M);
            }
#pragma unroll
            for (int u = 0; u < KV; u++) {
                const uint32_t vu = from_lane<SL>(q.x, u);
                const uint32_t Mu = from_lane<SL>(M, u);
                const uint32_t vo = vu + B.off;
                if (fresh)  // symbfact.cuh:924
                    hv.cost[u] = ((Mu >> j) & 1u) && vu != NONE ? (vo > lv[u] ? vo : lv[u]) : INF;
            }
            if (fresh) hv.held = true;
        }
        if (!__ballot(hv.held)) return;  // nothing left for any group
        // -- fold the held vertices' edges, vertex by vertex.  Bit k of each
        // group's share of longm: vertex k's row has more edge groups after
        // the current one (rare), so only those read back pos and end.
        const uint64_t longm = __ballot(hv.held && j < (uint32_t)KV && hv.pos + SL < hv.end);
        bool full = false;
#pragma unroll
        for (int u = 0; u < KV; u++) {
            uint32_t pu = hv.held ? from_lane<SL>(hv.pend, u) : 0u;
            while (pu) {  // group-uniform
                const bool mine = (pu >> j) & 1u;
                uint32_t idx = NONE;
                if (mine) idx = tab_find<SL>(S, &S.tcnt[B.rb], hv.civ[u]);
                const bool placed = mine && idx != FULL;
                // unplaced edges (past the row end, or the table is full) fold
                // into the sink row, so every group folds whole quads
                reinterpret_cast<uint32_t*>(S.eidx[h])[j] = placed ? idx : (uint32_t)Z::T;
                const uint64_t wm = __ballot(placed);
                const uint32_t okm = (uint32_t)(wm >> (SL * grp)) & Z::GM;
                const uint32_t nq = (wave_top<SL>(wm) + 3u) >> 2;  // wave-uniform
                wave_sync();
                if (hv.cost[u] != INF) {
                    const uint32_t cu = hv.cost[u];
                    const uint4* ep = S.eidx[h];
                    for (uint32_t q = 0; q < nq; q++) {
                        const uint4 id = ep[q];
                        atomicMin(&S.cell[id.x][j], cu);
                        atomicMin(&S.cell[id.y][j], cu);
                        atomicMin(&S.cell[id.z][j], cu);
                        atomicMin(&S.cell[id.w][j], cu);
                    }
                }
                wave_sync();
                pu &= ~okm;
                if (pu) {  // table full: keep the rest of this group
                    full = true;
                    break;
                }
                // next group of a long row (rare): load its targets now
                if ((longm >> (SL * grp + u)) & 1u) {  // group-uniform
                    int32_t pos = (int32_t)from_lane<SL>((uint32_t)hv.pos, u);
                    const int32_t end = (int32_t)from_lane<SL>((uint32_t)hv.end, u);
                    if (pos + SL < end) {
                        pos += SL;
                        pu = group_mask<SL>(end - pos);
                        if (j == (uint32_t)u) hv.pos = pos;
                        hv.civ[u] = ((pu >> j) & 1u) ? at(p.ci, (uint32_t)(pos + (int32_t)j) << 2) - B.wb : 0u;
                    }
                }
            }
            if (hv.held && j == (uint32_t)u) hv.pend = pu;
            if (__ballot(full)) return;  // the table is full for everyone
        }
        hv.held = false;  // every vertex of the batch is folded
    }
}

// Phase B: relax every table entry once per slot.  Each lane group takes KR
// entries: lane k loads entry k's CSR range and next-level mask, then the
// group issues all KR label-line loads before using any, so a batch costs
// one memory round trip.  Wave-uniform control flow.
template <int SL>
__device__ void relax_phase(const Blk& B, BlkLds<SL>& S, uint32_t nbuf, uint32_t my_src, uint32_t my_srcg,
                            TallyB& t) {
    using Z = Shape<SL>;
    constexpr int KR = Z::KR;
    const BlkParams& p = *B.p;
    const uint32_t lane = __lane_id();
    const uint32_t grp = lane / SL;
    const uint32_t j = lane % SL;
    const uint32_t w = threadIdx.x / WAVE;
    uint32_t* maskn = B.mask0 + (uint64_t)nbuf * B.W;
    uint4* qn = B.q0 + (uint64_t)nbuf * B.W;
    const uint32_t tc = S.tcnt[B.rb];
    const uint32_t ntab = tc < (uint32_t)Z::T ? tc : (uint32_t)Z::T;
    const uint32_t off = B.off;
    for (uint32_t base = w * (Z::GPW * KR); base < ntab; base += BW * (Z::GPW * KR)) {
        const uint32_t eb = base + grp * KR;
        // lane k: entry eb + k
        const uint32_t ek = eb + j;
        // lanes past KR or past the table read a valid cell and drop it
        const uint32_t mr = S.ent_m[ek < (uint32_t)Z::T ? ek : (uint32_t)Z::T - 1u];
        const uint32_t mk = (j < (uint32_t)KR && ek < ntab) ? mr : HOLE;
        const bool okk = mk != HOLE;
        int32_t d0 = 0, d1 = 0;
        uint32_t nm = 0;
        if (okk) {
            const uint32_t ro = (B.wb + mk) << 2;
            d0 = at(p.rp, ro);
            d1 = at(p.rp, ro + 4u);
            nm = at(maskn, mk << 2);
        }
        const uint32_t dd = (uint32_t)(d1 - d0);
        // lane j: slot j's label of every entry (a hole reads vertex 0's
        // line: its cells are INF, so the value is never used)
        uint32_t lm[KR];
#pragma unroll
        for (int u = 0; u < KR; u++) {
            const uint32_t mu = from_lane<SL>(mk, u);
            lm[u] = at(B.lab0, ((mu == HOLE ? 0u : mu) * SL + j) << 2);
        }
        uint32_t mybits = 0;    // lane k: slots that push entry k's target
        uint32_t fillbits = 0;  // lane j: bit u = slot j gains entry u's target
#pragma unroll
        for (int u = 0; u < KR; u++) {
            // broadcasts while every lane is active
            const uint32_t mu = from_lane<SL>(mk, u);
            const uint32_t du = from_lane<SL>(dd, u);
            // hole and unused cells are INF: no entry check needed
            const uint32_t c = S.cell[eb + u][j];
            S.cell[eb + u][j] = INF;  // ready for the next round
            bool push = false, fill = false;
            if (c != INF) {
                const uint32_t l = lm[u];
                const bool up = mu > my_src;  // above the source (symbfact.cuh:992-1009)
                // U entry: first touch (the stale label, l >= stale); below: the
                // relaxation of symbfact.cuh:1018-1063 (l > c).  One compare:
                // stale >= W >= 1.
                if (l > (up ? B.stale - 1u : c)) {
                    at(B.lab0, (mu * SL + j) << 2) = c;
                    const uint32_t mo = mu + off;
                    fill = up || (l >= mo && c < mo);  // below: the label crosses below m
                    push = !up;  // the source itself never decreases (its label is off)
                    const bool fr = push && l >= B.stale;  // first reach: |R_<(s)| and TE_alg
                    t.reach += fr ? 1u : 0u;
                    t.te += fr ? du : 0u;
                }
            }
            const uint32_t pb = grp_ballot<SL>(push, grp);
            if (j == (uint32_t)u) mybits = pb;
            if (fill) fillbits |= 1u << u;
        }
        // next-level slot masks and queue entries, one lane per entry (each
        // target is unique in the round, so no other lane touches its mask)
        const bool upd = okk && mybits && (nm | mybits) != nm;
        if (upd) at(maskn, mk << 2) = nm | mybits;
        const bool enq = upd && nm == 0u;
        const uint64_t em = __ballot(enq);
        if (em) {
            uint32_t qb = 0;
            if (lane == 0) qb = atomicAdd(&S.qcnt[nbuf], (uint32_t)__popcll(em));
            qb = bcast0(qb);
            if (enq) qn[qb + rank_in(em)] = make_uint4(mk, (uint32_t)d0, (uint32_t)d1, 0u);
        }
        // fills of the batch, (source << 32) | target: one reservation per wave
        // (none when only counts are wanted: fill_cap = 0)
        const uint32_t nf = (uint32_t)__popc(fillbits);
        t.fills += nf;
        if (p.fill_cap != 0ull && __ballot(nf != 0u)) {
            const uint32_t incl = wave_incl_scan(nf);
            const uint32_t total = __shfl(incl, WAVE - 1, WAVE);
            uint64_t pos = 0;
            if (lane == 0) pos = aadd64(&p.ctrl->fill_pos, total);
            pos = bcast0_64(pos) + (incl - nf);
            for (uint32_t b = fillbits; b; b &= b - 1u, pos++) {
                const uint32_t mu = S.ent_m[eb + (uint32_t)__builtin_ctz(b)];  // per-lane index
                if (pos < p.fill_cap) p.fills[pos] = ((uint64_t)my_srcg << 32) | (mu + B.wb);
            }
        }
    }
}

template <int SL>
__global__ __launch_bounds__(BB) __attribute__((amdgpu_waves_per_eu(GSOFA_WPE, GSOFA_WPE))) void k_traverse_blk(BlkParams p) {
    using Z = Shape<SL>;
    __shared__ BlkLds<SL> S;
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = __lane_id();
    const uint32_t w = tid / WAVE;
    const uint32_t j = lane % SL;
    const uint32_t W = p.W;
    const uint32_t loops = p.loops;  // offsets before the labels must be reset (<= 2^32 / W)
    Blk B;
    B.p = &p;
    B.W = W;
    B.lab0 = p.lab + (uint64_t)blockIdx.x * SL * W;
    B.mask0 = p.mask + (uint64_t)blockIdx.x * 2 * W;
    B.q0 = p.qg + (uint64_t)blockIdx.x * 2 * W;
    B.rb = 0;
    uint32_t off = p.bstate[2 * blockIdx.x];
    uint32_t since = p.bstate[2 * blockIdx.x + 1];
    TallyB t{0, 0, 0, 0};
    unsigned long long levels = 0;
    for (uint32_t i = tid; i < (uint32_t)((Z::T + 1) * SL); i += BB) (&S.cell[0][0])[i] = INF;
    for (uint32_t i = tid; i < (uint32_t)Z::H; i += BB) S.hkey[i] = 0ull;
    if (tid < 4) S.acc[tid] = 0;
    if (tid == 0) {
        S.qcnt[0] = S.qcnt[1] = 0;
        S.cursor = 0;
        S.tcnt[0] = S.tcnt[1] = S.held[0] = S.held[1] = 0;
    }
    for (;;) {
        if (tid == 0) S.chunk = atomicAdd(&p.ctrl->src_next, 1u);
        __syncthreads();
        const uint32_t ck = S.chunk;
        if (ck >= p.n_chunks) break;
        const BlkChunk C = p.chunks[ck];
        const uint32_t nsrc = C.count;  // <= SL (the host's chunking)
        B.wb = p.mat_base[C.mat];
        if (tid < (uint32_t)SL) {
            S.src[tid] = tid < nsrc ? C.first - tid * C.stride : 0u;
            S.s_te[tid] = 0;
            S.s_reach[tid] = 0;
            S.s_fill[tid] = 0;
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
        const uint32_t my_srcg = S.src[j];
        const uint32_t my_src = my_srcg - B.wb;

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
                        const int32_t r0 = p.rp[mg], r1 = p.rp[mg + 1];
                        if (old >= B.stale) {
                            rc++;
                            te += (uint32_t)(r1 - r0);
                        }
                        const uint32_t mo = B.mask0[m];  // level-0 buffer
                        B.mask0[m] = mo | (1u << k);
                        if (mo == 0u)
                            B.q0[atomicAdd(&S.qcnt[0], 1u)] = make_uint4(m, (uint32_t)r0, (uint32_t)r1, 0u);
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
            GrpVtx<Z::KV> hv;
            hv.held = false;
            hv.done = false;
            hv.pend = 0u;
            hv.pos = hv.end = 0;
            for (;;) {  // rounds, two barriers each
                combine_phase<SL>(B, S, hv, cur, size, t);
                if (hv.held && j == 0) atomicOr(&S.held[B.rb], 1u);
                __syncthreads();
                // final for this round: the cursor moves only in phase A
                const bool more = S.held[B.rb] != 0u || S.cursor < size;
                relax_phase<SL>(B, S, nbuf, my_src, my_srcg, t);
                // the next round's table, prepared while others still relax:
                // relax never reads the keys, and the other parity's counter
                // and flag were last read before the previous barrier
                for (uint32_t i = tid; i < (uint32_t)Z::H; i += BB) S.hkey[i] = 0ull;
                if (tid == 0) {
                    S.tcnt[B.rb ^ 1u] = 0;
                    S.held[B.rb ^ 1u] = 0;
                }
                __syncthreads();
                B.rb ^= 1u;
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
        // per-row TE, |R_<(s)| and fills: lane j's tallies belong to slot j
        if (t.te) atomicAdd(&S.s_te[j], (unsigned long long)t.te);
        if (t.reach) atomicAdd(&S.s_reach[j], (unsigned long long)t.reach);
        if (t.fills) atomicAdd(&S.s_fill[j], (unsigned long long)t.fills);
        t.te = t.reach = t.fills = 0;
        __syncthreads();
        if (tid < nsrc) {
            p.row_te[S.src[tid]] = S.s_te[tid];
            p.row_reach[S.src[tid]] = S.s_reach[tid];
            p.row_fills[S.src[tid]] = S.s_fill[tid];
            atomicAdd(&S.acc[0], S.s_fill[tid]);
            atomicAdd(&S.acc[1], S.s_te[tid]);
            atomicAdd(&S.acc[2], S.s_reach[tid]);
        }
        off -= W;
        since++;
    }
    if (t.checks) atomicAdd(&S.acc[3], t.checks);
    __syncthreads();
    if (tid == 0) {
        if (S.acc[0]) aadd64(&p.ctrl->fill_count, S.acc[0]);
        if (S.acc[1]) aadd64(&p.ctrl->te, S.acc[1]);
        if (S.acc[2]) aadd64(&p.ctrl->reach, S.acc[2]);
        if (S.acc[3]) aadd64(&p.ctrl->edge_checks, S.acc[3]);
        aadd64(&p.ctrl->levels, levels);
        p.bstate[2 * blockIdx.x] = off;
        p.bstate[2 * blockIdx.x + 1] = since;
    }
}

}  // namespace

bool blk_slots_valid(int slots) { return slots == 32 || slots == 16 || slots == 8; }
int blk_block_size() { return BB; }
const void* blk_kernel_ptr(int slots) {
    switch (slots) {
        case 16: return reinterpret_cast<const void*>(&k_traverse_blk<16>);
        case 8: return reinterpret_cast<const void*>(&k_traverse_blk<8>);
        default: return reinterpret_cast<const void*>(&k_traverse_blk<32>);
    }
}

hipError_t launch_traverse_blk(const BlkParams& p, int slots, size_t lds_bytes, hipStream_t s) {
    switch (slots) {
        case 32: hipLaunchKernelGGL(k_traverse_blk<32>, dim3(p.nblocks), dim3(BB), lds_bytes, s, p); break;
        case 16: hipLaunchKernelGGL(k_traverse_blk<16>, dim3(p.nblocks), dim3(BB), lds_bytes, s, p); break;
        case 8: hipLaunchKernelGGL(k_traverse_blk<8>, dim3(p.nblocks), dim3(BB), lds_bytes, s, p); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace gsofa
