// gsofa_kernels.hip -- CDNA4 (gfx950) kernels of the fill-path symbolic
// factorization: the persistent multi-source traversal, the L/U structure
// build and the supernode pass.
//
// Traversal semantics follow the reference's per-source max-id labels:
//   seeding     with_space_optimization/symbfact.cuh:639-721
//   relaxation  with_space_optimization/symbfact.cuh:883-1082
//   packed max-id offsets  symbfact.cuh:661,695,922 ; no-reset offset 2388-2407
//   persistent group loop  without_space_optimization/symbfact.cuh:706-962
// The execution model is new: one cooperative, occupancy-bounded grid runs
// every group and level of a share; levels are separated by an atomic
// generation barrier (replacing the lock-array barrier of
// without_space_optimization/symbfact.cuh:107-167); each wave64 expands 64
// frontier pairs at once through a wave-level merge of their row lengths, so
// lanes read consecutive CSR words; pushes and fills are compacted with
// ballot + mbcnt into per-wave LDS staging buffers and flushed with one global
// atomic per buffer; rows at least block_threshold long are expanded by the
// whole workgroup at the end of the level.
#include "gsofa_device.h"
#include "gsofa_internal.h"

namespace gsofa {

namespace {

using namespace dev;

constexpr int BLOCK = 256;
constexpr int WAVES = BLOCK / WAVE;
constexpr int PBUF = 256;   // staged frontier pairs per wave
constexpr int FBUF = 128;   // staged fills per wave
constexpr int BIGCAP = 64;  // block-expanded rows per workgroup per level

struct WaveLds {
    uint64_t pbuf[PBUF];
    uint64_t fbuf[FBUF];
    uint32_t incl[WAVE];
    int32_t beg[WAVE];
    uint32_t cost[WAVE];
    uint32_t slot[WAVE];
};

struct BigRow {
    int32_t beg, end;
    uint32_t cost, slot;
};

struct Snap {
    unsigned long long q, orph;
};

struct BlockLds {
    WaveLds w[WAVES];
    BigRow big[BIGCAP];
    uint32_t nbig;
    uint32_t flag;
    Snap snap;
};

// Atomic generation barrier over a co-resident (cooperatively launched) grid.
// The last block to arrive resets the arrival count, snapshots the decision
// inputs of the next phase (*snap_q and the orphan count) into the slot of this
// generation, and bumps the generation; the others poll the generation with
// acquire loads.  Every block then reads the same snapshot, so control flow
// stays uniform even though the counters change again during the next phase.
// A wall-clock bound turns a lost block into an abort flag instead of a hang.
// Returns false on abort.
__device__ bool grid_sync(Ctrl* c, uint32_t nblocks, uint64_t spin_ticks, uint32_t* s_flag,
                          unsigned long long* snap_q, Snap* s_snap) {
    __syncthreads();  // orders this block's writes before thread 0's release
    if (threadIdx.x == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        const uint32_t gen = ld_agent(&c->bar_gen);
        const uint32_t arrived =
            __hip_atomic_fetch_add(&c->bar_count, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (arrived == nblocks - 1) {
            st_agent(&c->bar_count, 0u);
            st_agent(&c->snap_q[gen & 1u], snap_q ? ld_agent(snap_q) : 0ull);
            st_agent(&c->snap_orph[gen & 1u], ld_agent(&c->orphans));
            __hip_atomic_store(&c->bar_gen, gen + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            const uint64_t t0 = wall_clock64();
            uint32_t polls = 0;
            while (__hip_atomic_load(&c->bar_gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == gen) {
                __builtin_amdgcn_s_sleep(2);
                if ((++polls & 255u) == 0) {
                    if (ld_agent(&c->abort) || wall_clock64() - t0 > spin_ticks) {
                        st_agent(&c->abort, 1u);
                        break;
                    }
                }
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        s_snap->q = ld_agent(&c->snap_q[gen & 1u]);
        s_snap->orph = ld_agent(&c->snap_orph[gen & 1u]);
        *s_flag = ld_agent(&c->abort);
    }
    __syncthreads();
    return *s_flag == 0;
}

// Per-group constants shared by every phase.
struct Group {
    uint32_t first;  // source of slot 0
    uint32_t nsrc;   // valid slots
    uint32_t off;    // max-id offset of this group
    uint32_t stale;  // labels >= stale belong to earlier groups
};

// Packed max-id layout: slot k holds labels of vertices 0..src_k back to back,
// base_k = sum_{j<k} (src_j + 1) with src_j = first + j*N
// (with_space_optimization/symbfact.cuh:661).
__device__ __forceinline__ uint64_t slot_base(uint64_t k, uint64_t first, uint64_t N) {
    return k * (first + 1) + N * (k * (k - 1) / 2);
}

struct Level {
    uint64_t* qn;                     // next queue
    unsigned long long* nextcnt;
    uint32_t qcap;
};

__device__ __forceinline__ void park_orphan(const TravParams& p, uint64_t pr) {
    uint32_t k = (uint32_t)(pr >> 32), v = (uint32_t)pr;
    uint32_t bit = 1u << (v & 31u);
    uint32_t old = __hip_atomic_fetch_or(&p.orphan[(uint64_t)k * p.orphan_words + (v >> 5)], bit,
                                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (!(old & bit)) aadd64(&p.ctrl->orphans, 1ull);
}

__device__ void flush_pushes(const TravParams& p, WaveLds& W, uint32_t cnt, const Level& lv) {
    wave_sync();
    const uint32_t lane = __lane_id();
    uint64_t pos = 0;
    if (lane == 0) pos = aadd64(lv.nextcnt, cnt);
    pos = bcast0_64(pos);
    for (uint32_t t = lane; t < cnt; t += WAVE) {
        uint64_t pr = W.pbuf[t];
        uint64_t at = pos + t;
        if (at < lv.qcap) lv.qn[at] = pr;
        else park_orphan(p, pr);
    }
    wave_sync();
}

__device__ void flush_fills(const TravParams& p, WaveLds& W, uint32_t cnt) {
    wave_sync();
    const uint32_t lane = __lane_id();
    uint64_t pos = 0;
    if (lane == 0) pos = aadd64(&p.ctrl->fill_pos, cnt);
    pos = bcast0_64(pos);
    for (uint32_t t = lane; t < cnt; t += WAVE) {
        uint64_t at = pos + t;
        if (at < p.fill_cap) p.fills[at] = W.fbuf[t];
    }
    wave_sync();
}

// Wave-uniform staging of one pass of per-lane results.
__device__ __forceinline__ void stage(const TravParams& p, WaveLds& W, uint32_t& pcount,
                                      uint32_t& fcount, bool push, uint64_t pair, bool fill,
                                      uint64_t fentry, const Level& lv) {
    uint64_t pm = __ballot(push);
    if (pm) {
        uint32_t c = (uint32_t)__popcll(pm);
        if (pcount + c > (uint32_t)PBUF) {
            flush_pushes(p, W, pcount, lv);
            pcount = 0;
        }
        if (push) W.pbuf[pcount + rank_in(pm)] = pair;
        pcount += c;
    }
    uint64_t fm = __ballot(fill);
    if (fm) {
        uint32_t c = (uint32_t)__popcll(fm);
        if (fcount + c > (uint32_t)FBUF) {
            flush_fills(p, W, fcount);
            fcount = 0;
        }
        if (fill) W.fbuf[fcount + rank_in(fm)] = fentry;
        fcount += c;
    }
}

struct Tally {
    unsigned long long fills, te, reach, checks;
};

// One edge check of the relaxation (with_space_optimization/symbfact.cuh:989-1066).
// Returns bit0 = enqueue m, bit1 = m is a new fill of the source.
__device__ __forceinline__ uint32_t relax(const TravParams& p, uint32_t m, uint32_t k,
                                          uint32_t cost, uint32_t src, uint64_t base,
                                          const Group& g, Tally& t) {
    const uint32_t stamp = p.stamp_base + src + 1u;
    if (m > src) {
        uint32_t* mp = p.mark + (uint64_t)k * (uint32_t)p.n + m;
        if (ld_cached(mp) != stamp && amax(mp, stamp) < stamp) return 2u;
        return 0u;
    }
    uint32_t* lp = p.lab + base + m;
    if (ld_cached(lp) <= cost) return 0u;
    uint32_t old = amin(lp, cost);
    if (old <= cost) return 0u;
    if (old >= g.stale) {  // first time m is reached from this source
        t.reach++;
        t.te += (uint32_t)(p.rp[m + 1] - p.rp[m]);
    }
    uint32_t r = (m < src) ? 1u : 0u;
    if (m + g.off > cost) {
        uint32_t* mp = p.mark + (uint64_t)k * (uint32_t)p.n + m;
        if (amax(mp, stamp) < stamp) r |= 2u;
    }
    return r;
}

__device__ void seed_group(const TravParams& p, const Group& g, WaveLds& W, uint32_t& pcount,
                           uint32_t& fcount, const Level& lv, Tally& t) {
    const uint32_t N = (uint32_t)p.n_gpu;
    for (uint32_t k = blockIdx.x; k < g.nsrc; k += p.nblocks) {
        const uint32_t src = g.first + k * N;
        const uint32_t stamp = p.stamp_base + src + 1u;
        const uint64_t base = slot_base(k, g.first, N);
        const int32_t b = p.rp[src], e = p.rp[src + 1];
        if (threadIdx.x == 0) {
            p.lab[base + src] = g.off;  // with_space_optimization/symbfact.cuh:662
            t.te += (uint32_t)(e - b);
            t.checks += (uint32_t)(e - b);
        }
        for (int32_t t0 = b; t0 < e; t0 += BLOCK) {
            const int32_t i = t0 + (int32_t)threadIdx.x;
            bool push = false;
            uint32_t m = 0;
            if (i < e) {
                m = (uint32_t)p.ci[i];
                p.mark[(uint64_t)k * (uint32_t)p.n + m] = stamp;  // symbfact.cuh:699
                if (m < src) {                                    // symbfact.cuh:700-714
                    uint32_t old = p.lab[base + m];
                    if (old >= g.stale) {
                        t.reach++;
                        t.te += (uint32_t)(p.rp[m + 1] - p.rp[m]);
                    }
                    p.lab[base + m] = g.off;
                    push = true;
                }
            }
            stage(p, W, pcount, fcount, push, ((uint64_t)k << 32) | m, false, 0, lv);
        }
    }
}

// Expand one row segment [beg + lo, beg + hi) for one frontier pair with all
// lanes of the calling group (wave or block) striding it.
__device__ void expand_strided(const TravParams& p, const Group& g, WaveLds& W, uint32_t& pcount,
                               uint32_t& fcount, const Level& lv, Tally& t, int32_t beg,
                               int32_t end, uint32_t cost, uint32_t k, uint32_t lane0,
                               uint32_t stride) {
    const uint32_t N = (uint32_t)p.n_gpu;
    const uint32_t src = g.first + k * N;
    const uint64_t base = slot_base(k, g.first, N);
    for (int64_t e0 = beg; e0 < end; e0 += stride) {
        int64_t e = e0 + lane0;
        uint32_t r = 0, m = 0;
        if (e < end) {
            m = (uint32_t)p.ci[e];
            t.checks++;
            r = relax(p, m, k, cost, src, base, g, t);
            if (r & 2u) t.fills++;
        }
        stage(p, W, pcount, fcount, r & 1u, ((uint64_t)k << 32) | m, (r & 2u) != 0,
              ((uint64_t)src << 32) | m, lv);
    }
}

__device__ void expand_level(const TravParams& p, const Group& g, BlockLds& S, WaveLds& W,
                             uint32_t& pcount, uint32_t& fcount, const Level& lv,
                             const uint64_t* qc, uint32_t size, uint32_t* cursor, Tally& t) {
    const uint32_t lane = __lane_id();
    const uint32_t N = (uint32_t)p.n_gpu;
    const uint32_t T = (uint32_t)p.block_threshold;
    for (;;) {
        uint32_t i0 = 0;
        if (lane == 0) i0 = __hip_atomic_fetch_add(cursor, (uint32_t)WAVE, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_AGENT);
        i0 = bcast0(i0);
        if (i0 >= size) break;
        const uint32_t i = i0 + lane;
        uint32_t deg = 0, cost = 0, k = 0;
        int32_t b = 0;
        if (i < size) {
            const uint64_t pr = qc[i];
            const uint32_t v = (uint32_t)pr;
            k = (uint32_t)(pr >> 32);
            const uint64_t base = slot_base(k, g.first, N);
            const uint32_t lv_ = ld_agent(p.lab + base + v);
            cost = (v + g.off) > lv_ ? (v + g.off) : lv_;  // symbfact.cuh:924
            b = p.rp[v];
            const int32_t e = p.rp[v + 1];
            deg = (uint32_t)(e - b);
            if (deg >= T) {  // hand the row to the whole workgroup
                uint32_t slot = atomicAdd(&S.nbig, 1u);
                if (slot < (uint32_t)BIGCAP) {
                    S.big[slot] = BigRow{b, e, cost, k};
                    deg = 0;
                }
            }
        }
        // rows that stayed with the wave: wave-level merge of their lengths
        uint32_t incl = wave_incl_scan(deg);
        const uint32_t total = __shfl(incl, WAVE - 1, WAVE);
        W.incl[lane] = incl;
        W.beg[lane] = b - (int32_t)(incl - deg);
        W.cost[lane] = cost;
        W.slot[lane] = k;
        wave_sync();
        for (uint32_t e0 = 0; e0 < total; e0 += WAVE) {
            const uint32_t e = e0 + lane;
            uint32_t r = 0, m = 0, kk = 0, src = 0;
            if (e < total) {
                uint32_t lo = 0, hi = WAVE - 1;
                while (lo < hi) {  // first lane whose inclusive sum exceeds e
                    uint32_t mid = (lo + hi) >> 1;
                    if (W.incl[mid] > e) hi = mid;
                    else lo = mid + 1;
                }
                m = (uint32_t)p.ci[W.beg[lo] + (int32_t)e];
                kk = W.slot[lo];
                src = g.first + kk * N;
                t.checks++;
                r = relax(p, m, kk, W.cost[lo], src, slot_base(kk, g.first, N), g, t);
                if (r & 2u) t.fills++;
            }
            stage(p, W, pcount, fcount, r & 1u, ((uint64_t)kk << 32) | m, (r & 2u) != 0,
                  ((uint64_t)src << 32) | m, lv);
        }
        wave_sync();
    }
}

__device__ void rescan_orphans(const TravParams& p, const Group& g, uint64_t* qc,
                               unsigned long long* cnt) {
    const uint64_t words = (uint64_t)g.nsrc * p.orphan_words;
    const uint64_t nth = (uint64_t)p.nblocks * BLOCK;
    unsigned long long taken = 0;
    for (uint64_t w = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; w < words; w += nth) {
        uint32_t bits = ld_agent(p.orphan + w);
        if (!bits) continue;
        const uint32_t c = (uint32_t)__popc(bits);
        const uint64_t pos = aadd64(cnt, c);
        const uint32_t fit = pos >= p.qcap ? 0u : (uint32_t)((p.qcap - pos) < c ? (p.qcap - pos) : c);
        const uint64_t k = w / p.orphan_words;
        const uint32_t vb = (uint32_t)(w % p.orphan_words) * 32u;
        for (uint32_t t = 0; t < fit; t++) {
            uint32_t bpos = (uint32_t)__ffs(bits) - 1u;
            bits &= bits - 1u;
            qc[pos + t] = (k << 32) | (vb + bpos);
        }
        st_agent(p.orphan + w, bits);
        taken += fit;
    }
    if (taken) aadd64(&p.ctrl->orphans, (unsigned long long)(-(long long)taken));
}

__global__ __launch_bounds__(BLOCK) void k_traverse(TravParams p) {
    __shared__ BlockLds S;
    const uint32_t wid = threadIdx.x / WAVE;
    const uint32_t lane = __lane_id();
    WaveLds& W = S.w[wid];
    Ctrl* c = p.ctrl;
    Tally t{0, 0, 0, 0};
    unsigned long long levels = 0, rescans = 0;
    const uint32_t n = (uint32_t)p.n;
    const uint32_t N = (uint32_t)p.n_gpu;
    const uint32_t loops = p.loops;  // groups before the labels must be reset (<= 2^32 / n)
    const uint64_t share_cnt = ((uint64_t)n - (uint64_t)p.gpu_id + N - 1) / N;
    uint32_t off = p.off_start;
    uint32_t since = p.groups_since_init;
    uint64_t hw = p.lab_hw;
    const uint64_t gtid = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
    const uint64_t nth = (uint64_t)p.nblocks * BLOCK;
    const uint64_t spin_ticks = p.spin_ticks;
    if (threadIdx.x == 0) S.nbig = 0;

    for (int32_t gi = 0; gi < p.n_groups; gi++) {
        const uint64_t k0 = (uint64_t)(p.first_group + gi) * (uint64_t)p.chunk;
        Group g;
        g.nsrc = (uint32_t)((share_cnt - k0) < (uint64_t)p.chunk ? (share_cnt - k0) : (uint64_t)p.chunk);
        g.first = (uint32_t)(p.gpu_id + k0 * N);
        const uint64_t size_g = slot_base(g.nsrc, g.first, N);
        uint64_t init_lo = hw;
        if (since >= loops) {  // offsets exhausted: full re-initialisation
            off = 0xFFFFFFFFu - n;
            since = 0;
            init_lo = 0;
            hw = 0;
        }
        g.off = off;
        g.stale = off + n;
        for (uint64_t i = init_lo + gtid; i < size_g; i += nth) p.lab[i] = 0xFFFFFFFFu;
        if (size_g > hw) hw = size_g;
        if (gtid == 0) {
            for (int j = 0; j < 3; j++) {
                st_agent(&c->qcnt[j], 0ull);
                st_agent(&c->cursor[j], 0u);
            }
        }
        if (!grid_sync(c, p.nblocks, spin_ticks, &S.flag, nullptr, &S.snap)) return;

        uint32_t pcount = 0, fcount = 0;
        Level lv0{p.qbuf0, &c->qcnt[0], p.qcap};
        seed_group(p, g, W, pcount, fcount, lv0, t);
        if (pcount) flush_pushes(p, W, pcount, lv0);
        if (fcount) flush_fills(p, W, fcount);
        pcount = fcount = 0;
        if (!grid_sync(c, p.nblocks, spin_ticks, &S.flag, &c->qcnt[0], &S.snap)) return;

        // Level loop.  Level L reads queue L&1 with counter qcnt[L%3] and
        // cursor[L%3], and appends to queue (L+1)&1 with counter qcnt[(L+1)%3].
        // During level L block 0 clears the counters level L+1 will append
        // to / fetch from; the barrier that ends level L publishes them.
        // All decisions use the snapshot taken at the preceding barrier.
        uint32_t L = 0;
        for (;;) {
            uint64_t* qc = (L & 1u) ? p.qbuf1 : p.qbuf0;
            const unsigned long long appended = S.snap.q;
            const unsigned long long orph = S.snap.orph;
            const uint32_t size = appended < p.qcap ? (uint32_t)appended : p.qcap;
            if (size < p.qcap && orph != 0ull) {
                // drain-and-resume: top this level's queue up from the
                // overflow bitmap, then run the level as usual
                rescan_orphans(p, g, qc, &c->qcnt[L % 3u]);
                rescans++;
                if (!grid_sync(c, p.nblocks, spin_ticks, &S.flag, &c->qcnt[L % 3u], &S.snap)) return;
                continue;
            }
            if (size == 0) break;
            if (gtid == 0) {
                st_agent(&c->qcnt[(L + 2u) % 3u], 0ull);
                st_agent(&c->cursor[(L + 1u) % 3u], 0u);
            }
            Level lv{(L & 1u) ? p.qbuf0 : p.qbuf1, &c->qcnt[(L + 1u) % 3u], p.qcap};
            expand_level(p, g, S, W, pcount, fcount, lv, qc, size, &c->cursor[L % 3u], t);
            // block-centric expansion of the long rows this block collected
            __syncthreads();
            const uint32_t nbig = S.nbig < (uint32_t)BIGCAP ? S.nbig : (uint32_t)BIGCAP;
            for (uint32_t r = 0; r < nbig; r++) {
                const BigRow br = S.big[r];
                expand_strided(p, g, W, pcount, fcount, lv, t, br.beg, br.end, br.cost, br.slot,
                               threadIdx.x, BLOCK);
            }
            __syncthreads();
            if (threadIdx.x == 0) S.nbig = 0;
            if (pcount) flush_pushes(p, W, pcount, lv);
            if (fcount) flush_fills(p, W, fcount);
            pcount = fcount = 0;
            if (!grid_sync(c, p.nblocks, spin_ticks, &S.flag, &c->qcnt[(L + 1u) % 3u], &S.snap)) return;
            L++;
            levels++;
        }
        off -= n;
        since++;
    }
    // statistics: one atomic per wave
    unsigned long long f = wave_sum64(t.fills), te = wave_sum64(t.te), rc = wave_sum64(t.reach),
                       ch = wave_sum64(t.checks);
    if (lane == 0) {
        if (f) aadd64(&c->fill_count, f);
        if (te) aadd64(&c->te, te);
        if (rc) aadd64(&c->reach, rc);
        if (ch) aadd64(&c->edge_checks, ch);
    }
    if (gtid == 0) {
        aadd64(&c->levels, levels);
        aadd64(&c->rescans, rescans);
    }
}

}  // namespace

hipError_t launch_traverse(const TravParams& p, hipStream_t s, int cooperative) {
    TravParams copy = p;
    void* args[] = {(void*)&copy};
    if (cooperative)
        return hipLaunchCooperativeKernel(reinterpret_cast<const void*>(&k_traverse), dim3(p.nblocks),
                                          dim3(BLOCK), args, 0, s);
    hipLaunchKernelGGL(k_traverse, dim3(p.nblocks), dim3(BLOCK), 0, s, copy);
    return hipGetLastError();
}

int traverse_block_size() { return BLOCK; }
const void* traverse_kernel_ptr() { return reinterpret_cast<const void*>(&k_traverse); }
size_t traverse_lds_bytes() { return sizeof(BlockLds); }

}  // namespace gsofa
