// bpmax_server.cpp -- continuous-batching request server over the engine.
//
// One device arena of `memory_budget` bytes is carved into per-batch extents (first fit,
// coalescing free list).  Client threads submit pairs into a FIFO queue; one server
// thread
//   * retires finished batches (an event per batch, polled), publishing each score and
//     its submit-to-completion latency, and frees the batch's extent (or keeps it until
//     every request of the batch is released, when tables are kept);
//   * admits the queue head into a new batch whenever a batch slot is idle, while the
//     batch fits the largest free extent and max_batch, and launches it asynchronously
//     on the slot's own engine context (two streams each), so up to `slots` batches run
//     on the device at once.
// Admission is strictly FIFO: a large pair at the head waits for memory rather than
// being overtaken, so no request starves.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "bpmax.h"
#include "bpmax_engine_impl.h"

namespace bpx {

using Clock = std::chrono::steady_clock;

enum ReqState { RQ_QUEUED, RQ_RUNNING, RQ_DONE, RQ_FAILED };

struct Request {
    std::vector<int> s1, s2;
    int M = 0, N = 0;
    size_t bytes = 0;
    Clock::time_point t_submit, t_done;
    int state = RQ_QUEUED;
    float score = 0.f;
    long long batch = -1;
    PairDesc desc{}; // layout inside its batch's extent (kept tables)
    std::string err;
};

struct Batch {
    size_t off = 0, bytes = 0;
    int live = 0; // requests not yet released (kept tables)
};

struct Slot {
    bpmax_ctx *ctx = nullptr;
    hipEvent_t done = nullptr;
    float *h_scores = nullptr; // pinned
    std::vector<long long> ids;
    long long batch = -1;
    bool busy = false;
};

struct Server {
    bpmax_ctx *owner = nullptr;
    bpmax_server_config cfg{};
    float intra[25], inter[25];
    void *arena = nullptr;
    std::map<size_t, size_t> free_ext; // offset -> length
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<long long> queue;
    std::unordered_map<long long, Request> reqs;
    std::unordered_map<long long, Batch> batches;
    std::vector<Slot> slots;
    std::thread th;
    bool stopping = false;
    long long next_id = 1, next_batch = 1;
    bpmax_server_stats st{};
    long long inflight_pairs = 0;
};

static size_t largest_free(const Server &S)
{
    size_t m = 0;
    for (auto &e : S.free_ext) m = std::max(m, e.second);
    return m;
}

static bool ext_alloc(Server &S, size_t bytes, size_t *off)
{
    for (auto it = S.free_ext.begin(); it != S.free_ext.end(); ++it) {
        if (it->second < bytes) continue;
        *off = it->first;
        const size_t rest = it->second - bytes;
        const size_t o = it->first + bytes;
        S.free_ext.erase(it);
        if (rest) S.free_ext[o] = rest;
        S.st.bytes_in_use += bytes;
        S.st.peak_bytes_in_use = std::max(S.st.peak_bytes_in_use, S.st.bytes_in_use);
        return true;
    }
    return false;
}

static void ext_free(Server &S, size_t off, size_t bytes)
{
    S.st.bytes_in_use -= bytes;
    auto it = S.free_ext.emplace(off, bytes).first;
    auto next = std::next(it);
    if (next != S.free_ext.end() && it->first + it->second == next->first) {
        it->second += next->second;
        S.free_ext.erase(next);
    }
    if (it != S.free_ext.begin()) {
        auto prev = std::prev(it);
        if (prev->first + prev->second == it->first) {
            prev->second += it->second;
            S.free_ext.erase(it);
        }
    }
}

static size_t round_up(size_t x) { return (x + 4095) / 4096 * 4096; }

// Retire finished batches.  Called with the lock held.
static bool retire(Server &S)
{
    bool any = false;
    for (Slot &sl : S.slots) {
        if (!sl.busy) continue;
        const hipError_t q = hipEventQuery(sl.done);
        if (q == hipErrorNotReady) continue;
        const auto now = Clock::now();
        Batch &b = S.batches[sl.batch];
        for (size_t k = 0; k < sl.ids.size(); ++k) {
            Request &r = S.reqs[sl.ids[k]];
            r.t_done = now;
            if (q == hipSuccess) {
                r.state = RQ_DONE;
                r.score = sl.h_scores[k];
                ++S.st.completed;
            } else {
                r.state = RQ_FAILED;
                r.err = std::string("batch failed: ") + hipGetErrorString(q);
                ++S.st.failed;
            }
        }
        S.inflight_pairs -= (long long)sl.ids.size();
        if (!S.cfg.keep_tables || q != hipSuccess) {
            ext_free(S, b.off, b.bytes);
            S.batches.erase(sl.batch);
            for (long long id : sl.ids) S.reqs[id].batch = -1;
        } else {
            b.live = (int)sl.ids.size();
        }
        sl.ids.clear();
        sl.batch = -1;
        sl.busy = false;
        any = true;
    }
    if (any) S.cv_done.notify_all();
    return any;
}

// Form and launch one batch on an idle slot.  Called with the lock held; releases it
// around the launch.
static bool admit(Server &S, std::unique_lock<std::mutex> &lk)
{
    Slot *sl = nullptr;
    for (Slot &s : S.slots)
        if (!s.busy) {
            sl = &s;
            break;
        }
    if (!sl || S.queue.empty()) return false;
    const size_t room = largest_free(S);
    std::vector<long long> ids;
    size_t bytes = 0;
    for (long long id : S.queue) {
        const Request &r = S.reqs[id];
        if ((int)ids.size() >= S.cfg.max_batch || round_up(bytes + r.bytes) > room) break;
        ids.push_back(id);
        bytes += r.bytes;
    }
    if (ids.empty()) return false;
    bytes = round_up(bytes);
    size_t off = 0;
    if (!ext_alloc(S, bytes, &off)) return false;
    S.queue.erase(S.queue.begin(), S.queue.begin() + ids.size());
    const long long bid = S.next_batch++;
    S.batches[bid] = Batch{off, bytes, 0};
    std::vector<const int *> p1, p2;
    std::vector<int> Ms, Ns;
    for (long long id : ids) {
        Request &r = S.reqs[id];
        r.state = RQ_RUNNING;
        r.batch = bid;
        p1.push_back(r.s1.data());
        p2.push_back(r.s2.data());
        Ms.push_back(r.M);
        Ns.push_back(r.N);
    }
    sl->busy = true;
    sl->ids = ids;
    sl->batch = bid;
    ++S.st.batches;
    S.st.max_batch_seen = std::max<long long>(S.st.max_batch_seen, (long long)ids.size());
    S.inflight_pairs += (long long)ids.size();
    S.st.max_inflight_pairs = std::max(S.st.max_inflight_pairs, S.inflight_pairs);

    // launch outside the lock (schedule building and submission take host time); the
    // request vectors are not touched by other threads while the request is running
    lk.unlock();
    bpmax_ctx *c = sl->ctx;
    c->arena_ext = (char *)S.arena + off;
    c->arena_ext_bytes = bytes;
    int rc = bpx_run_pairs(c, (int)ids.size(), p1.data(), Ms.data(), p2.data(), Ns.data(), S.intra, S.inter, false);
    if (rc == BPMAX_OK) rc = bpx_gather_scores(c, sl->h_scores, false);
    if (rc == BPMAX_OK && hipEventRecord(sl->done, c->stream) != hipSuccess) rc = BPMAX_EDEVICE;
    std::vector<PairDesc> descs = c->pairs_host;
    const std::string err = c->err;
    lk.lock();
    if (rc != BPMAX_OK) {
        // nothing was left running for this batch that can be waited on: fail it now
        (void)hipStreamSynchronize(c->stream);
        (void)hipStreamSynchronize(c->stream2);
        for (long long id : ids) {
            Request &r = S.reqs[id];
            r.state = RQ_FAILED;
            r.err = err;
            r.t_done = Clock::now();
            r.batch = -1;
            ++S.st.failed;
        }
        ext_free(S, off, bytes);
        S.batches.erase(bid);
        S.inflight_pairs -= (long long)ids.size();
        sl->ids.clear();
        sl->batch = -1;
        sl->busy = false;
        S.cv_done.notify_all();
        return true;
    }
    for (size_t k = 0; k < ids.size(); ++k) S.reqs[ids[k]].desc = descs[k];
    return true;
}

static void server_loop(Server *S)
{
    std::unique_lock<std::mutex> lk(S->mu);
    for (;;) {
        bool progress = retire(*S);
        while (admit(*S, lk)) progress = true;
        bool busy = false;
        for (auto &sl : S->slots) busy = busy || sl.busy;
        if (S->stopping && S->queue.empty() && !busy) break;
        if (!progress) S->cv_work.wait_for(lk, std::chrono::microseconds(busy ? 50 : 2000));
    }
}

static void server_free(Server *S)
{
    for (Slot &sl : S->slots) {
        if (sl.ctx) (void)bpmax_ctx_destroy(sl.ctx);
        if (sl.done) (void)hipEventDestroy(sl.done);
        if (sl.h_scores) (void)hipHostFree(sl.h_scores);
    }
    if (S->arena) (void)hipFree(S->arena);
    delete S;
}

} // namespace bpx

using namespace bpx;

void bpx_server_shutdown(bpmax_ctx *ctx)
{
    if (ctx && ctx->server) (void)bpmax_server_stop(ctx);
}

extern "C" int bpmax_server_start(bpmax_ctx *ctx, const bpmax_server_config *cfg_in)
{
    if (!ctx) return BPMAX_EINVAL;
    if (ctx->server) {
        ctx->err = "server already running on this context";
        return BPMAX_EINVAL;
    }
    bpmax_server_config cfg{};
    if (cfg_in) cfg = *cfg_in;
    if (cfg.max_batch <= 0) cfg.max_batch = 64;
    if (cfg.slots <= 0) cfg.slots = 3;
    if (cfg.queue_capacity <= 0) cfg.queue_capacity = 4096;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    if (cfg.memory_budget == 0) {
        size_t fr = 0, tot = 0;
        HIPCHK(ctx, hipMemGetInfo(&fr, &tot));
        cfg.memory_budget = fr / 10 * 8;
    }
    cfg.memory_budget = cfg.memory_budget / 4096 * 4096;
    Server *S = new Server();
    S->owner = ctx;
    S->cfg = cfg;
    float di[25], de[25];
    bpmax_engine_score_table(di, de);
    memcpy(S->intra, cfg.intra ? cfg.intra : di, sizeof S->intra);
    memcpy(S->inter, cfg.inter ? cfg.inter : (cfg.intra ? cfg.intra : de), sizeof S->inter);
    S->cfg.intra = S->intra;
    S->cfg.inter = S->inter;
    if (hipMalloc(&S->arena, cfg.memory_budget) != hipSuccess) {
        (void)hipGetLastError();
        S->arena = nullptr;
        ctx->err = "server arena: hipMalloc of " + std::to_string(cfg.memory_budget) + " bytes failed";
        server_free(S);
        return BPMAX_ENOMEM;
    }
    S->free_ext[0] = cfg.memory_budget;
    S->st.budget = cfg.memory_budget;
    S->slots.resize(cfg.slots);
    for (Slot &sl : S->slots) {
        int rc = bpmax_ctx_create(&sl.ctx, ctx->device);
        if (rc == BPMAX_OK && hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess) rc = BPMAX_EDEVICE;
        if (rc == BPMAX_OK && hipHostMalloc((void **)&sl.h_scores, sizeof(float) * cfg.max_batch) != hipSuccess)
            rc = BPMAX_ENOMEM;
        if (rc != BPMAX_OK) {
            ctx->err = "server slot setup failed";
            server_free(S);
            return rc;
        }
        sl.ctx->schedule = ctx->schedule == BPMAX_SCHED_TASKS ? BPMAX_SCHED_PIPE : ctx->schedule;
        sl.ctx->outer_split = ctx->outer_split;
        sl.ctx->outer_items = ctx->outer_items;
        sl.ctx->outer_kgroup = ctx->outer_kgroup;
        sl.ctx->closure_ahead = ctx->closure_ahead;
        sl.ctx->intra_wide = ctx->intra_wide;
        sl.ctx->graphs = false; // batches change shape every time
        sl.ctx->tuning = ctx->tuning; // tuned launch schedules of the owning context
    }
    ctx->server = S;
    S->th = std::thread(server_loop, S);
    return BPMAX_OK;
}

extern "C" int bpmax_server_submit(bpmax_ctx *ctx, const int *seq1, int M, const int *seq2, int N, long long *id)
{
    if (!ctx || !ctx->server || !seq1 || !seq2 || M < 1 || N < 1 || !id) return BPMAX_EINVAL;
    Server &S = *ctx->server;
    for (int i = 0; i < M; ++i)
        if (seq1[i] < 0 || seq1[i] > 4) return BPMAX_EINVAL;
    for (int i = 0; i < N; ++i)
        if (seq2[i] < 0 || seq2[i] > 4) return BPMAX_EINVAL;
    std::string why;
    if (!weights_exact(S.intra, S.inter, M, N, &why)) {
        ctx->err = why;
        return BPMAX_EWEIGHTS;
    }
    const size_t bytes = bpx_arena_bytes(1, &M, &N);
    std::lock_guard<std::mutex> g(S.mu);
    if (S.stopping) return BPMAX_EINVAL;
    if (round_up(bytes) > S.cfg.memory_budget) {
        ctx->err = "pair " + std::to_string(M) + "x" + std::to_string(N) + " needs " + std::to_string(bytes) +
                   " bytes, more than the server budget";
        return BPMAX_ETOOBIG;
    }
    if ((int)S.queue.size() >= S.cfg.queue_capacity) return BPMAX_EBUSY;
    const long long rid = S.next_id++;
    Request &r = S.reqs[rid];
    r.s1.assign(seq1, seq1 + M);
    r.s2.assign(seq2, seq2 + N);
    r.M = M;
    r.N = N;
    r.bytes = bytes;
    r.t_submit = Clock::now();
    S.queue.push_back(rid);
    ++S.st.submitted;
    *id = rid;
    S.cv_work.notify_one();
    return BPMAX_OK;
}

static int poll_locked(bpmax_ctx *ctx, Server &S, long long id, float *score, double *latency_ms)
{
    auto it = S.reqs.find(id);
    if (it == S.reqs.end()) return BPMAX_ENOTFOUND;
    const Request &r = it->second;
    if (r.state == RQ_FAILED) {
        ctx->err = r.err;
        return BPMAX_EDEVICE;
    }
    if (r.state != RQ_DONE) return BPMAX_ENOTREADY;
    if (score) *score = r.score;
    if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(r.t_done - r.t_submit).count();
    return BPMAX_OK;
}

extern "C" int bpmax_server_poll(bpmax_ctx *ctx, long long id, float *score, double *latency_ms)
{
    if (!ctx || !ctx->server) return BPMAX_EINVAL;
    std::lock_guard<std::mutex> g(ctx->server->mu);
    return poll_locked(ctx, *ctx->server, id, score, latency_ms);
}

extern "C" int bpmax_server_wait(bpmax_ctx *ctx, long long id, double timeout_ms, float *score, double *latency_ms)
{
    if (!ctx || !ctx->server) return BPMAX_EINVAL;
    Server &S = *ctx->server;
    std::unique_lock<std::mutex> lk(S.mu);
    const auto until = Clock::now() + std::chrono::microseconds((long long)(timeout_ms * 1000.0));
    for (;;) {
        const int rc = poll_locked(ctx, S, id, score, latency_ms);
        if (rc != BPMAX_ENOTREADY) return rc;
        if (timeout_ms < 0) S.cv_done.wait(lk);
        else if (S.cv_done.wait_until(lk, until) == std::cv_status::timeout)
            return poll_locked(ctx, S, id, score, latency_ms);
    }
}

extern "C" int bpmax_server_fetch(bpmax_ctx *ctx, long long id, bpmax_layout layout, float *F_out)
{
    if (!ctx || !ctx->server || !F_out) return BPMAX_EINVAL;
    if (layout != BPMAX_LAYOUT_PACKED_TRI_TRI && layout != BPMAX_LAYOUT_DENSE_4D) return BPMAX_EINVAL;
    Server &S = *ctx->server;
    PairDesc d;
    const float *base;
    {
        std::lock_guard<std::mutex> g(S.mu);
        auto it = S.reqs.find(id);
        if (it == S.reqs.end()) return BPMAX_ENOTFOUND;
        const Request &r = it->second;
        if (r.state == RQ_FAILED) return BPMAX_EDEVICE;
        if (r.state != RQ_DONE) return BPMAX_ENOTREADY;
        if (!S.cfg.keep_tables || r.batch < 0) {
            ctx->err = "tables are not kept (keep_tables = 0)";
            return BPMAX_EINVAL;
        }
        d = r.desc;
        base = (const float *)((const char *)S.arena + S.batches[r.batch].off);
    }
    // the extent stays allocated while the request is unreleased
    return bpx_copy_out(ctx, base, d, layout, F_out);
}

extern "C" int bpmax_server_release(bpmax_ctx *ctx, long long id)
{
    if (!ctx || !ctx->server) return BPMAX_EINVAL;
    Server &S = *ctx->server;
    std::lock_guard<std::mutex> g(S.mu);
    auto it = S.reqs.find(id);
    if (it == S.reqs.end()) return BPMAX_ENOTFOUND;
    if (it->second.state == RQ_QUEUED || it->second.state == RQ_RUNNING) return BPMAX_EBUSY;
    const long long bid = it->second.batch;
    S.reqs.erase(it);
    if (bid >= 0) {
        auto b = S.batches.find(bid);
        if (b != S.batches.end() && --b->second.live <= 0) {
            ext_free(S, b->second.off, b->second.bytes);
            S.batches.erase(b);
            S.cv_work.notify_one();
        }
    }
    return BPMAX_OK;
}

extern "C" int bpmax_server_get_stats(bpmax_ctx *ctx, bpmax_server_stats *stats)
{
    if (!ctx || !ctx->server || !stats) return BPMAX_EINVAL;
    std::lock_guard<std::mutex> g(ctx->server->mu);
    *stats = ctx->server->st;
    return BPMAX_OK;
}

extern "C" int bpmax_server_stop(bpmax_ctx *ctx)
{
    if (!ctx || !ctx->server) return BPMAX_EINVAL;
    Server *S = ctx->server;
    {
        std::lock_guard<std::mutex> g(S->mu);
        S->stopping = true;
        S->cv_work.notify_all();
    }
    if (S->th.joinable()) S->th.join();
    ctx->server = nullptr;
    server_free(S);
    return BPMAX_OK;
}
