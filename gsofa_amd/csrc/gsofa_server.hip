// gsofa_server.hip -- request server with continuous batching (host code).
//
// No counterpart in the reference, whose driver runs one matrix per process
// (with_space_optimization/symbfact.cuh:1553).  Requests arrive one at a time
// from any host thread.  Submission does the host work in the caller's thread:
// the request is validated, its rows (and the rows of its transpose) packed,
// and its output bounded (gsofa_fill_bound's rule), so a malformed request
// fails to its caller alone and workers only concatenate.
//
// max_inflight worker threads each own an arena: a reach-mode plan with its
// own HIP stream whose device buffers (CSR of A and A^T, per-workgroup
// visited words, masks and queues, fill stream, L/U columns, per-row arrays)
// grow to the largest batch the worker has run and then stay.  A free worker
// admits the waiting requests, oldest first, while the batch fits its share
// of the device budget, loads them into its arena (one upload), runs them as
// one persistent launch plus the structure and supernode passes, and copies
// each request's rows straight from the device into its result.  A worker in
// steady state makes no device allocation, so batches of different workers
// overlap on the device: requests that arrive while others run start as soon
// as a worker (slot) and its share of memory are free.
//
// Admission: each worker owns budget / max_inflight bytes.  A batch's bytes
// are its fixed arrays (CSR of A and A^T, per-row arrays, and the fill stream
// and output columns sized from the requests' bounds; reach_fixed_bytes) plus
// its traversal workgroups, which take what the fixed part leaves of the share
// (reach_grid).  A request is added while at least one workgroup still fits.
// One that cannot fit an otherwise empty share runs alone: its worker waits
// until no batch is in flight, no other worker starts one meanwhile, the
// arenas are released, and a plan of its own gets the whole budget: in core
// (reach mode) when the budget holds it, else in spilled shares (block mode,
// GSOFA_FLAG_SPILL, powers of two up to kMaxSpill: the fewest shares that
// leave room for 7/8 of the traversal workgroups the most shares would).
// Only a request with room for no workgroup fails, with GSOFA_ERR_BUDGET.
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "gsofa.h"
#include "gsofa_internal.h"

namespace {

using Clock = std::chrono::steady_clock;

double ms_between(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

struct Request {
    int64_t ticket = 0;
    gsofa::PackedMatrix pm;  // packed rows of A and A^T, output bounds
    Clock::time_point t_submit, t_start, t_done;
    bool done = false;
    int status = GSOFA_PENDING;
    gsofa_result res{};
    gsofa_request_info info{};
    gsofa_matrix view() const {  // the packed rows as a matrix (plans made outside an arena)
        return {pm.n, pm.nnz, pm.rp.data(), pm.rp.data() + 1, pm.ci.data()};
    }
};

}  // namespace

struct gsofa_server {
    gsofa_server_options opt{};
    int64_t share = 0;      // device bytes per worker
    uint32_t resident = 1;  // co-resident traversal workgroups of a batch plan
    std::mutex mu;
    std::condition_variable cv_work;  // waiting requests or shutdown
    std::condition_variable cv_done;  // a request became final
    std::condition_variable cv_idle;  // no batch in flight
    int32_t running = 0;              // batches in flight
    bool exclusive = false;           // a request is waiting for or using the whole budget
    std::deque<Request*> waiting;
    std::unordered_map<int64_t, Request*> live;  // every request not yet taken
    int64_t next_ticket = 1;
    int32_t next_batch = 1;
    bool stop = false;
    std::vector<gsofa_plan*> arenas;  // one per worker, created on its first batch
    std::vector<std::thread> workers;
};

namespace {

constexpr int32_t kMaxSpill = 256;  // most spilled shares tried for one request
constexpr int kReachBits = 64;      // sources per workgroup of an in-core batch (reach mode)

// Bytes the batch `reqs` needs apart from its traversal workgroups.  In core
// (shares = 1) a batch runs in reach mode; with `shares` spilled shares (one
// request, block mode) the device holds one share's output: the cyclic
// shares' bounds are close to equal, and a quarter more covers the largest.
int64_t fixed_bytes(const std::vector<Request*>& reqs, int32_t shares = 1) {
    int64_t n = 0, nnz = 0;
    uint64_t nchunks = 0, fb = 0, lu = 0;
    const uint64_t per = shares > 1 ? (uint64_t)gsofa::kBatchSlots : (uint64_t)kReachBits;
    for (const Request* r : reqs) {
        n += r->pm.n;
        nnz += r->pm.nnz;
        fb += r->pm.need_fill;
        lu += r->pm.need_L + r->pm.need_U;
        nchunks += 2 * (((uint64_t)r->pm.n + per - 1) / per);
    }
    if (shares > 1) {
        fb = fb / (uint64_t)shares + fb / (4 * (uint64_t)shares) + 1;
        lu = lu / (uint64_t)shares + lu / (4 * (uint64_t)shares) + 1;
        return gsofa::block_fixed_bytes(n, nnz, (int32_t)reqs.size(), nchunks, std::max<uint64_t>(fb, 1),
                                        (int64_t)lu);
    }
    return gsofa::reach_fixed_bytes(n, nnz, (int32_t)reqs.size(), nchunks, fb, lu);
}

uint32_t largest(const std::vector<Request*>& reqs) {
    uint32_t W = 1;
    for (const Request* r : reqs) W = std::max<uint32_t>(W, (uint32_t)r->pm.n);
    return W;
}

// Traversal bytes of one workgroup: reach mode in core, block mode spilled.
int64_t per_workgroup(const std::vector<Request*>& reqs, int32_t shares = 1) {
    return shares > 1 ? gsofa::block_bytes(1, largest(reqs), gsofa::kBatchSlots)
                      : gsofa::reach_bytes(1, largest(reqs), kReachBits);
}

// Traversal bytes left for the batch within `budget` (< 0: does not fit).
int64_t trav_room(int64_t budget, const std::vector<Request*>& reqs, int32_t shares = 1) {
    return budget - fixed_bytes(reqs, shares) - per_workgroup(reqs, shares);
}

void finish(gsofa_server* S, Request* r, int status) {
    r->status = status;
    r->done = true;
    r->t_done = Clock::now();
    r->info.latency_ms = ms_between(r->t_submit, r->t_done);
    r->info.queued_ms = ms_between(r->t_submit, r->t_start);
}

gsofa_options batch_options(const gsofa_server* S) {
    gsofa_options o;
    gsofa_options_default(&o);
    o.device = S->opt.device;
    o.max_supernode_size = S->opt.max_supernode_size;
    o.flags = S->opt.flags;
    return o;
}

// Publishes the batch's results (st[i] the status of request i).
void publish(gsofa_server* S, std::vector<Request*>& batch, int32_t batch_id, std::vector<gsofa_result>& res,
             const std::vector<int>& st, double batch_ms, int32_t spill) {
    std::lock_guard<std::mutex> lk(S->mu);
    for (size_t i = 0; i < batch.size(); i++) {
        Request* r = batch[i];
        if (st[i] == GSOFA_OK) r->res = res[i];
        r->info.batch_id = batch_id;
        r->info.batch_size = (int32_t)batch.size();
        r->info.batch_ms = batch_ms;
        r->info.spill_shares = spill;
        // the packed input is no longer needed
        std::vector<int32_t>().swap(r->pm.rp);
        std::vector<int32_t>().swap(r->pm.ci);
        std::vector<int32_t>().swap(r->pm.rpT);
        std::vector<int32_t>().swap(r->pm.ciT);
        finish(S, r, st[i]);
    }
    S->cv_done.notify_all();
}

// In core, in the worker's arena: one upload, one execute, per-request copies.
void run_arena(gsofa_server* S, int32_t worker, std::vector<Request*>& batch, int32_t batch_id) {
    std::vector<int> st(batch.size(), GSOFA_OK);
    std::vector<gsofa_result> res(batch.size());
    double batch_ms = 0;
    gsofa_plan*& A = S->arenas[(size_t)worker];
    int rc = GSOFA_OK;
    if (!A) {
        const gsofa_options o = batch_options(S);
        rc = gsofa::arena_create(&o, &A);
    }
    std::vector<const gsofa::PackedMatrix*> mats(batch.size());
    for (size_t i = 0; i < batch.size(); i++) mats[i] = &batch[i]->pm;
    if (rc == GSOFA_OK) rc = gsofa::arena_load(A, mats.data(), (int32_t)mats.size(), S->share);
    if (rc == GSOFA_OK) rc = gsofa_plan_execute(A);
    if (rc == GSOFA_OK) {
        for (size_t i = 0; i < batch.size(); i++) {
            st[i] = gsofa::arena_fetch(A, (int32_t)i, &res[i]);
            if (st[i] == GSOFA_OK) batch_ms = res[i].ms;
        }
    } else {
        std::fill(st.begin(), st.end(), rc);
    }
    publish(S, batch, batch_id, res, st, batch_ms, 1);
}

// Alone on the device with `budget` bytes: a plan of its own, in core (reach
// mode) or in `spill` spilled shares (block mode).
void run_alone(gsofa_server* S, Request* r, int32_t batch_id, int64_t budget, int32_t spill) {
    std::vector<Request*> one{r};
    gsofa_options o = batch_options(S);
    if (spill > 1) {
        o.mode = GSOFA_MODE_BLOCK;
        o.block_slots = gsofa::kBatchSlots;
        o.flags |= GSOFA_FLAG_SPILL;
        o.n_virtual_gpu = spill;
    } else {
        o.mode = GSOFA_MODE_REACH;
        o.block_slots = kReachBits;
    }
    // the traversal arrays get what the fixed part leaves of the budget
    o.mem_budget_bytes = std::max<int64_t>(budget - fixed_bytes(one, spill), 1);
    const gsofa_matrix M = r->view();
    gsofa_plan* P = nullptr;
    int rc = gsofa_plan_create(&M, &o, &P);
    if (rc == GSOFA_OK) rc = gsofa_plan_execute(P);
    std::vector<int> st(1, rc);
    std::vector<gsofa_result> res(1);
    double batch_ms = 0;
    if (rc == GSOFA_OK) {
        st[0] = gsofa_plan_fetch(P, &res[0]);
        if (st[0] == GSOFA_OK) batch_ms = res[0].ms;
    }
    if (P) gsofa_plan_destroy(P);
    publish(S, one, batch_id, res, st, batch_ms, spill);
}

void worker_main(gsofa_server* S, int32_t worker) {
    if (hipSetDevice(S->opt.device) != hipSuccess) (void)hipGetLastError();
    std::unique_lock<std::mutex> lk(S->mu);
    for (;;) {
        S->cv_work.wait(lk, [&] { return (S->stop || !S->waiting.empty()) && !S->exclusive; });
        if (S->waiting.empty()) break;  // stopping and drained
        std::vector<Request*> batch;
        int64_t sum_n = 0, sum_nnz = 0;
        while (!S->waiting.empty() && (int32_t)batch.size() < S->opt.max_batch) {
            Request* r = S->waiting.front();
            // a batch is one block-diagonal matrix: it keeps the index limits
            // every request met alone
            if (sum_n + r->pm.n > gsofa::kMaxN || sum_nnz + r->pm.nnz >= (int64_t)INT32_MAX) break;
            batch.push_back(r);
            if (trav_room(S->share, batch) < 0) {
                batch.pop_back();
                break;
            }
            sum_n += r->pm.n;
            sum_nnz += r->pm.nnz;
            S->waiting.pop_front();
        }
        int64_t budget = S->share;
        int32_t spill = 1;
        bool alone = false;
        if (batch.empty()) {  // the oldest request cannot fit a share on its own
            Request* r = S->waiting.front();
            S->waiting.pop_front();
            const std::vector<Request*> one{r};
            const int32_t most = std::min<int32_t>(kMaxSpill, std::max<int32_t>(r->pm.n, 1));
            auto wgs = [&](int32_t sh) {  // traversal workgroups that fit with sh shares
                const int64_t per_wg = per_workgroup(one, sh);
                const int64_t room = S->opt.mem_budget_bytes - fixed_bytes(one, sh);
                return room < per_wg ? (int64_t)0 : std::min<int64_t>(room / per_wg, (int64_t)S->resident);
            };
            // in core (reach mode) when the whole budget holds it with at
            // least one workgroup; otherwise the fewest spilled shares
            // (block mode) that keep 7/8 of the workgroups the most shares would
            int32_t top = 1;
            while (top * 2 <= most) top *= 2;
            if (wgs(1) == 0 && top > 1 && r->pm.nnz < gsofa::kMaxBlockNnz) {
                spill = 2;
                while (spill < top && 8 * wgs(spill) < 7 * wgs(top)) spill *= 2;
            }
            if (wgs(spill) == 0) {
                r->t_start = Clock::now();
                finish(S, r, GSOFA_ERR_BUDGET);
                S->cv_done.notify_all();
                continue;
            }
            // it fits the whole budget (in core or spilled): run it alone once
            // the device drains, with the workers' arenas released
            S->exclusive = true;
            S->cv_idle.wait(lk, [&] { return S->running == 0; });
            for (gsofa_plan*& a : S->arenas) {
                if (a) gsofa_plan_destroy(a);
                a = nullptr;
            }
            batch.push_back(r);
            budget = S->opt.mem_budget_bytes;
            alone = true;
        }
        const int32_t id = S->next_batch++;
        const Clock::time_point t0 = Clock::now();
        for (Request* r : batch) r->t_start = t0;
        S->running++;
        if (!alone && !S->waiting.empty()) S->cv_work.notify_one();  // another worker may take the rest
        lk.unlock();
        if (alone) run_alone(S, batch[0], id, budget, spill);
        else run_arena(S, worker, batch, id);
        lk.lock();
        S->running--;
        if (alone) {
            S->exclusive = false;
            S->cv_work.notify_all();
        }
        if (S->running == 0) S->cv_idle.notify_all();
    }
}

int take(gsofa_server* S, Request* r, gsofa_result* out, gsofa_request_info* info) {
    // caller holds S->mu; r is final.  Other waiters on the same ticket wake
    // and find it gone.
    const int st = r->status;
    if (info) *info = r->info;
    if (st == GSOFA_OK && out) *out = r->res;
    else if (st == GSOFA_OK) gsofa_result_free(&r->res);
    S->live.erase(r->ticket);
    delete r;
    S->cv_done.notify_all();
    return st;
}

}  // namespace

extern "C" {

void gsofa_server_options_default(gsofa_server_options* o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->device = 0;
    o->mem_budget_bytes = 0;
    o->max_inflight = 2;
    o->max_batch = 256;
    o->max_supernode_size = 0;
    o->flags = 0;
}

int gsofa_server_create(const gsofa_server_options* opt, gsofa_server** out) {
    if (!out) return GSOFA_ERR_INVALID;
    *out = nullptr;
    gsofa_server_options o;
    if (opt) o = *opt;
    else gsofa_server_options_default(&o);
    // spilled output is chosen per request by the server itself (a request
    // that does not fit runs alone, spilled if it must); the flag is refused
    if (o.max_inflight < 1 || o.max_inflight > 16 || o.max_batch < 1 || (o.flags & GSOFA_FLAG_SPILL))
        return GSOFA_ERR_INVALID;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        (void)hipGetLastError();
        return GSOFA_ERR_NO_DEVICE;
    }
    if (o.device < 0 || o.device >= ndev) return GSOFA_ERR_INVALID;
    if (o.mem_budget_bytes <= 0) {
        size_t freeb = 0, totalb = 0;
        if (hipSetDevice(o.device) != hipSuccess || hipMemGetInfo(&freeb, &totalb) != hipSuccess) {
            (void)hipGetLastError();
            return GSOFA_ERR_HIP;
        }
        o.mem_budget_bytes = (int64_t)((double)freeb * 0.80);
    }
    uint32_t resident = 0;
    if (hipSetDevice(o.device) != hipSuccess || gsofa::reach_resident(o.device, 1024, &resident) != GSOFA_OK) {
        (void)hipGetLastError();
        return GSOFA_ERR_HIP;
    }
    gsofa_server* S = new gsofa_server();
    S->opt = o;
    S->resident = std::max<uint32_t>(resident, 1);
    S->share = o.mem_budget_bytes / o.max_inflight;
    S->arenas.assign((size_t)o.max_inflight, nullptr);
    for (int32_t i = 0; i < o.max_inflight; i++) S->workers.emplace_back(worker_main, S, i);
    *out = S;
    return GSOFA_OK;
}

int gsofa_server_submit(gsofa_server* S, const gsofa_matrix* A, int64_t* ticket) {
    if (!S || !A || !ticket) return GSOFA_ERR_INVALID;
    if (A->n < 1) return GSOFA_ERR_INVALID;
    Request* r = new Request();
    // the checks a plan would make, answered to this caller alone (a
    // malformed request never reaches a batch shared with others), and the
    // packing and output bounds, done here in the submitting thread
    const int rc = gsofa::pack_matrix(A, &r->pm);
    if (rc != GSOFA_OK) {
        delete r;
        return rc;
    }
    r->t_submit = Clock::now();
    std::lock_guard<std::mutex> lk(S->mu);
    if (S->stop) {
        delete r;
        return GSOFA_ERR_STATE;
    }
    r->ticket = S->next_ticket++;
    S->live[r->ticket] = r;
    S->waiting.push_back(r);
    *ticket = r->ticket;
    S->cv_work.notify_one();
    return GSOFA_OK;
}

int gsofa_server_poll(gsofa_server* S, int64_t ticket, gsofa_result* out, gsofa_request_info* info) {
    if (!S) return GSOFA_ERR_INVALID;
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->live.find(ticket);
    if (it == S->live.end()) return GSOFA_ERR_INVALID;
    if (!it->second->done) return GSOFA_PENDING;
    return take(S, it->second, out, info);
}

int gsofa_server_wait(gsofa_server* S, int64_t ticket, gsofa_result* out, gsofa_request_info* info) {
    if (!S) return GSOFA_ERR_INVALID;
    std::unique_lock<std::mutex> lk(S->mu);
    // the ticket is looked up afresh on every wake-up: another thread may take
    // it (poll or a second wait) meanwhile, and then this call fails cleanly
    auto it = S->live.find(ticket);
    S->cv_done.wait(lk, [&] {
        it = S->live.find(ticket);
        return it == S->live.end() || it->second->done;
    });
    if (it == S->live.end()) return GSOFA_ERR_INVALID;
    return take(S, it->second, out, info);
}

void gsofa_server_destroy(gsofa_server* S) {
    if (!S) return;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
    }
    S->cv_work.notify_all();
    for (std::thread& t : S->workers) t.join();
    for (gsofa_plan* a : S->arenas)
        if (a) gsofa_plan_destroy(a);
    for (auto& kv : S->live) {
        if (kv.second->done && kv.second->status == GSOFA_OK) gsofa_result_free(&kv.second->res);
        delete kv.second;
    }
    delete S;
}

}  // extern "C"
