// bpmax_server.cpp -- continuous-batching request server over the engine.
//
// The server advances every admitted pair by one outer diagonal per "tick".  A tick is one
// step of the outer wavefront for all resident pairs at once, each pair on its own diagonal,
// built with the same emitters as a batched solve (emit_s2, emit_tick) and enqueued on the
// server's two streams under the pipelined schedule's event rules: R0 of tick t waits for the
// closure of tick t-1 to reach the logical diagonals it reads, and the closure of tick t waits
// for the R0 ranges of tick t that hold its tiles.  Requests that arrive while others run join
// at the next tick that has room (prologue on the R0 stream: fill, S1, S2), and pairs leave
// after their last diagonal; the device always holds one running schedule that pairs enter
// and leave.  Admission is FIFO (a large pair at the head waits for room instead of being
// overtaken, so nothing starves) by three budgets: arena memory (the pair's tiled F layout,
// S1, S2 and codes), resident pairs, and an upper bound of the work items a tick can hold.
//
// Everything the loop uses on the device is allocated at start: the arena (per-pair extents,
// first fit with a coalescing free list) and `depth` ring slots, each a staging block
// (weights, descriptors, work items) with a pinned host copy, a codes staging buffer and
// score buffers.  The server thread never allocates or frees device memory, it drives exactly
// two streams, and every stream wait it issues names an event recorded earlier in its own
// issue order.
//
// Planning (admission, tick building and the ordered list of stream operations) is host-only
// code (ServeCore); the executor turns a plan's operations into HIP calls.  bpmax_server_check
// replays a request sequence through the same planner without a GPU and checks every
// dependency, event wait and extent reuse of the operations it produces.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <string>
#include <thread>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "bpmax.h"
#include "bpmax_engine_impl.h"
#include "bpmax_launch.h"

namespace bpx {

using Clock = std::chrono::steady_clock;

enum ReqState { RQ_QUEUED, RQ_RUNNING, RQ_DONE, RQ_FAILED };

struct Request {
    std::vector<int> s1, s2;
    int M = 0, N = 0;
    size_t bytes = 0;          // extent bytes (multiple of 4096)
    long long items = 0;       // work-item bound of one tick of this pair (item_bound)
    Clock::time_point t_submit, t_done;
    int state = RQ_QUEUED;
    float score = 0.f;
    std::string err;
    size_t off = 0;            // extent offset in the arena (bytes)
    bool resident = false;     // the extent is allocated
    PairDesc desc{};           // layout inside the arena (offsets from the arena base)
    int d1 = 0;                // outer diagonal of the pair's next tick
    bool half = false;         // R0 and the cross-tile steps on FP16 operands (exact for it)
};

// Operations of one tick, in issue order, on stream 0 (R0 / prologue) or 1 (closure).
enum OpKind { OP_UPLOAD, OP_CODES, OP_FILL, OP_S1, OP_LAUNCH, OP_RECORD, OP_WAIT, OP_SCORES, OP_DONE, OP_S2SHIFT };
struct Op {
    int kind, stream, a; // a: launch index, joiner index or a count
    long long ev;        // event key (OP_RECORD / OP_WAIT)
};

constexpr int EV_OUT = 0, EV_IN = 1;
// Event key: (tick, tile size, kind, index).  The executor maps it to the event pool of ring
// slot tick % depth; the key without the tick identifies the event within the slot.
inline long long ev_key(long long tick, int T, int kind, int idx)
{
    return (tick << 20) | ((long long)(T == 64) << 18) | ((long long)kind << 16) | idx;
}
inline long long ev_tick(long long key) { return key >> 20; }
inline int ev_local(long long key) { return (int)(key & ((1 << 20) - 1)); }

struct TickGroup {
    int T = 0, rs = 1, nbmax = 0, nblmax = 0;
    std::vector<int> out_end;    // storage diagonal -> last storage diagonal of its R0 range
    std::vector<int> rec_levels; // closure levels that record an event, ascending
};

struct TickPlan {
    long long tick = -1;
    std::vector<long long> ids;  // tick-local pair -> request id
    std::vector<PairDesc> P;     // tick-local descriptors (work items index these)
    std::vector<int> d1s;        // outer diagonal of each pair in this tick
    int njoin = 0;               // P[0..njoin) join with this tick
    std::vector<int> fin;        // tick-local indices of the pairs whose last tick this is
    Schedule S;
    std::vector<TickGroup> groups;
    std::vector<Op> ops;
    std::vector<int> codes;      // codes of the joiners (seq1 then reversed, padded seq2)
    std::vector<size_t> codes_at; // per joiner: first int of its codes in `codes`
};

static long long c3ll(long long n) { return n * (n - 1) * (n - 2) / 6; }
static size_t al64(size_t x) { return (x + 63) / 64 * 64; }
static size_t round4k(size_t x) { return (x + 4095) / 4096 * 4096; }

// FP16 tables of a pair whose R0 runs on FP16 operands, in floats: the mirror of F and the
// shifted copies of S2 (PairDesc::Fh, S2r, S2c).
static size_t half_floats(int M, int N)
{
    const int T = choose_tile(N), nb = (N + T - 1) / T;
    const size_t ts = (size_t)tri_ll(nb) * T * T;
    return al64((size_t)tri_ll(M) * ts / 2) + 2 * al64(ts / 2);
}

// Floats of one pair's extent: F, S2, S1 (as bpx_run_pairs lays them out), the FP16 tables
// when `half`, and the codes (seq1, then reversed seq2 padded to whole tiles) as ints.
static size_t extent_floats(int M, int N, bool half)
{
    const int T = choose_tile(N), nb = (N + T - 1) / T;
    return pair_device_floats(M, N, T) + (half ? half_floats(M, N) : 0) + al64((size_t)M) + al64((size_t)nb * T);
}

// Descriptor of a pair whose extent starts at float offset e of the arena.
static PairDesc layout_extent(int M, int N, long long e, bool half)
{
    PairDesc d{};
    d.M = M;
    d.N = N;
    d.T = choose_tile(N);
    d.nb = (N + d.T - 1) / d.T;
    d.nbl = d.nb * (d.T / inner_tile(d.T));
    d.nblr = (N + inner_tile(d.T) - 1) / inner_tile(d.T);
    d.tstride = tri_ll(d.nb) * d.T * d.T;
    d.F = e;
    d.S2 = d.F + (long long)al64((size_t)(tri_ll(M) * d.tstride));
    d.S1 = d.S2 + (long long)al64((size_t)d.tstride);
    d.a = d.S1 + (long long)al64((size_t)M * M);
    if (half) {
        d.Fh = d.a;
        d.S2r = d.Fh + (long long)al64((size_t)(tri_ll(M) * d.tstride / 2));
        d.S2c = d.S2r + (long long)al64((size_t)(d.tstride / 2));
        d.a = d.S2c + (long long)al64((size_t)(d.tstride / 2));
    }
    d.b = d.a + (long long)al64((size_t)M);
    return d;
}

// Upper bound of the work items one pair contributes to any tick (its own items; the
// per-launch terms are in tick_item_const).  R0: every (triangle, storage tile, K) of a
// range at most once per k1 chunk beyond the first, whose total is bounded by the launch's
// item target; closure: every logical tile of every triangle, plus its cross-tile chunks
// (closure ahead: two ranges of D steps in chunks of >= 8; otherwise 2D steps in chunks of
// >= 2); S2 of a joiner likewise.
static long long item_bound(int M, int N, bool ahead)
{
    const int T = choose_tile(N), nb = (N + T - 1) / T, nbl = nb * (T / inner_tile(T));
    long long cross = 0, s2 = tri_ll(nbl);
    for (int D = 1; D < nbl; ++D) {
        cross += ahead ? (long long)(nbl - D - 1) * 2 * ((D + 7) / 8) : (long long)(nbl - D) * D;
        s2 += (long long)(nbl - D) * D;
    }
    return (long long)M * (c3ll(nb + 2) + tri_ll(nbl) + cross) + s2;
}

// Items a tick may hold beyond the pairs' own: per R0 launch (two tile-size groups, `split`
// launches each) the launch's item target and the split of the last round of every XCD.
static long long tick_item_const(const Tuning &e, int nbmax)
{
    const long long split = e.outer_split > 0 ? std::min(e.outer_split, std::max(1, nbmax)) : std::max(1, nbmax);
    return 2 * split * ((long long)std::max(64, e.outer_items) + 8 * 128);
}

struct ServeCore {
    size_t budget = 0;
    int max_active = 256, depth = 3;
    long long cap_items = 0;
    size_t cap_codes = 0;        // ints of joiner codes one tick can stage
    bool keep_tables = false;
    bool r0_half = true;         // FP16 operands for the pairs they are exact for
    Tuning eff{};
    float w[50];

    std::map<size_t, size_t> free_ext; // offset -> length
    std::deque<long long> queue;
    std::unordered_map<long long, Request> reqs;
    std::vector<long long> active;     // resident pairs still running, in join order
    long long items_sum = 0;           // item bounds of `active`
    long long next_id = 1, next_tick = 0;
    long long last_tick = -1;
    std::vector<TickGroup> last_groups;
    bpmax_server_stats st{};

    bool ext_alloc(size_t bytes, size_t *off)
    {
        for (auto it = free_ext.begin(); it != free_ext.end(); ++it) {
            if (it->second < bytes) continue;
            *off = it->first;
            const size_t rest = it->second - bytes, o = it->first + bytes;
            free_ext.erase(it);
            if (rest) free_ext[o] = rest;
            st.bytes_in_use += bytes;
            st.peak_bytes_in_use = std::max(st.peak_bytes_in_use, st.bytes_in_use);
            return true;
        }
        return false;
    }
    void ext_free(size_t off, size_t bytes)
    {
        st.bytes_in_use -= bytes;
        auto it = free_ext.emplace(off, bytes).first;
        auto next = std::next(it);
        if (next != free_ext.end() && it->first + it->second == next->first) {
            it->second += next->second;
            free_ext.erase(next);
        }
        if (it != free_ext.begin()) {
            auto prev = std::prev(it);
            if (prev->first + prev->second == it->first) {
                prev->second += it->second;
                free_ext.erase(it);
            }
        }
    }
    size_t codes_ints(int M, int N) const
    {
        const int T = choose_tile(N), nb = (N + T - 1) / T;
        return al64((size_t)M) + al64((size_t)nb * T);
    }

    // Checks that a pair can ever be admitted; fills its footprint.
    int prepare(Request &r, std::string *why) const
    {
        // FP16 tables where exact, unless they would make the pair too big for the budget
        r.half = r0_half && choose_tile(r.N) == 64 && weights_half_exact(w, w + 25, r.M, r.N, r.s1.data(), r.s2.data()) &&
                 round4k(extent_floats(r.M, r.N, true) * sizeof(float)) <= budget;
        r.bytes = round4k(extent_floats(r.M, r.N, r.half) * sizeof(float));
        const int T = choose_tile(r.N), nb = (r.N + T - 1) / T;
        r.items = item_bound(r.M, r.N, eff.closure_ahead != 0);
        if (r.bytes > budget) {
            *why = "pair " + std::to_string(r.M) + "x" + std::to_string(r.N) + " needs " + std::to_string(r.bytes) +
                   " bytes, more than the server budget";
            return BPMAX_ETOOBIG;
        }
        if (r.items + tick_item_const(eff, nb) > cap_items || codes_ints(r.M, r.N) > cap_codes) {
            *why = "pair " + std::to_string(r.M) + "x" + std::to_string(r.N) + " exceeds the per-tick staging capacity";
            return BPMAX_ETOOBIG;
        }
        return BPMAX_OK;
    }

    long long submit(Request &&r)
    {
        const long long id = next_id++;
        reqs[id] = std::move(r);
        queue.push_back(id);
        ++st.submitted;
        return id;
    }

    bool idle() const { return queue.empty() && active.empty(); }

    // Admission and the next tick's plan.  False when no pair is resident or admissible.
    bool plan(TickPlan &tp)
    {
        std::vector<long long> join;
        size_t codes_used = 0;
        int nbmax = 0;
        for (long long id : active) nbmax = std::max(nbmax, reqs[id].desc.nb);
        while (!queue.empty() && (int)(active.size() + join.size()) < max_active) {
            Request &r = reqs[queue.front()];
            const int T = choose_tile(r.N), nb = (r.N + T - 1) / T;
            const size_t ci = codes_ints(r.M, r.N);
            if (items_sum + r.items + tick_item_const(eff, std::max(nbmax, nb)) > cap_items) break;
            if (codes_used + ci > cap_codes) break;
            size_t off = 0;
            if (!ext_alloc(r.bytes, &off)) break;
            r.off = off;
            r.resident = true;
            r.state = RQ_RUNNING;
            r.d1 = 0;
            r.desc = layout_extent(r.M, r.N, (long long)(off / sizeof(float)), r.half);
            items_sum += r.items;
            codes_used += ci;
            nbmax = std::max(nbmax, nb);
            join.push_back(queue.front());
            queue.pop_front();
        }
        if (active.empty() && join.empty()) return false;

        tp = TickPlan{};
        tp.tick = next_tick++;
        tp.njoin = (int)join.size();
        tp.ids = join;
        tp.ids.insert(tp.ids.end(), active.begin(), active.end());
        for (long long id : tp.ids) {
            const Request &r = reqs[id];
            tp.P.push_back(r.desc);
            tp.d1s.push_back(r.d1);
        }
        for (int j = 0; j < tp.njoin; ++j) { // codes, with offsets relative to the arena base
            const Request &r = reqs[tp.ids[j]];
            tp.codes_at.push_back(tp.codes.size());
            const size_t n = codes_ints(r.M, r.N), na = al64((size_t)r.M);
            tp.codes.resize(tp.codes.size() + n, 4);
            int *c = tp.codes.data() + tp.codes_at[j];
            std::copy(r.s1.begin(), r.s1.end(), c);
            for (int x = 0; x < r.N; ++x) c[na + x] = r.s2[r.N - 1 - x];
        }
        for (int k = 0; k < (int)tp.P.size(); ++k)
            if (tp.d1s[k] == tp.P[k].M - 1) tp.fin.push_back(k);
        build(tp);

        // advance the pairs; finishing pairs leave the resident set (their extents stay
        // allocated until their tick retires)
        std::vector<long long> still;
        for (size_t k = 0; k < tp.ids.size(); ++k) {
            Request &r = reqs[tp.ids[k]];
            if (++r.d1 < r.M) still.push_back(tp.ids[k]);
            else items_sum -= r.items;
        }
        active = std::move(still);
        ++st.batches;
        st.max_batch_seen = std::max<long long>(st.max_batch_seen, (long long)tp.ids.size());
        st.max_inflight_pairs = std::max<long long>(st.max_inflight_pairs, (long long)tp.ids.size());
        last_tick = tp.tick;
        last_groups = tp.groups;
        return true;
    }

    // Launch list and operations of a tick (see the file comment for the event rules).
    void build(TickPlan &tp)
    {
        const std::vector<PairDesc> &P = tp.P;
        for (const PairDesc &d : P) {
            bool have = false;
            for (auto &g : tp.groups) have = have || g.T == d.T;
            if (!have) {
                TickGroup g;
                g.T = d.T;
                g.rs = d.nbl / d.nb;
                tp.groups.push_back(g);
            }
        }
        std::sort(tp.groups.begin(), tp.groups.end(), [](const TickGroup &a, const TickGroup &b) { return a.T < b.T; });
        std::vector<char> sel(P.size(), 0);
        for (int j = 0; j < tp.njoin; ++j) sel[j] = 1;
        for (int g = 0; g < (int)tp.groups.size(); ++g) {
            TickGroup &G = tp.groups[g];
            int s2nbl = 0;
            for (size_t k = 0; k < P.size(); ++k)
                if (P[k].T == G.T) {
                    G.nbmax = std::max(G.nbmax, P[k].nb);
                    G.nblmax = std::max(G.nblmax, P[k].nblr);
                    if (sel[k]) s2nbl = std::max(s2nbl, P[k].nblr);
                }
            if (s2nbl) emit_s2(P, sel, g, G.T, s2nbl, tp.S);
        }
        for (int g = 0; g < (int)tp.groups.size(); ++g) {
            TickGroup &G = tp.groups[g];
            emit_tick(P, tp.d1s, g, G.T, G.nbmax, G.nblmax, (int)tp.tick, tp.S, G.out_end, eff.outer_split,
                      eff.outer_items, eff.outer_kgroup, eff.closure_ahead != 0);
            // R0 items of pairs with FP16 tables first (in order), so that a launch runs them
            // on FP16 operands and the rest on FP32 ones
            for (Launch &L : tp.S.launches) {
                if (L.group != g || L.d1 < 0 || L.kind != BPMAX_K_OUTER || L.half_items) continue;
                auto b = tp.S.items.begin() + L.first, e = b + L.count;
                L.half_items = (int)(std::stable_partition(b, e, [&](const WorkItem &w) { return P[w.pair].Fh >= 0; }) - b);
            }
            std::set<int> rec;
            for (const Launch &L : tp.S.launches)
                if (L.group == g && L.d1 >= 0 && L.kind == BPMAX_K_OUTER)
                    rec.insert(std::min(G.rs * L.diag + G.rs - 1, G.nblmax - 1));
            rec.insert(G.nblmax - 1);
            G.rec_levels.assign(rec.begin(), rec.end());
        }

        std::vector<Op> &ops = tp.ops;
        ops.push_back(Op{OP_UPLOAD, 0, 0, 0});
        for (int j = 0; j < tp.njoin; ++j) {
            ops.push_back(Op{OP_CODES, 0, j, 0});
            ops.push_back(Op{OP_FILL, 0, j, 0});
        }
        if (tp.njoin) ops.push_back(Op{OP_S1, 0, tp.njoin, 0});
        const int nl = (int)tp.S.launches.size();
        for (int k = 0; k < nl; ++k)
            if (tp.S.launches[k].d1 < 0) ops.push_back(Op{OP_LAUNCH, 0, k, 0});
        // the shifted FP16 copies of the joiners' S2 (cross-tile operands of their closure)
        bool shift = false;
        for (int j = 0; j < tp.njoin; ++j) shift = shift || P[j].S2r >= 0;
        if (shift) ops.push_back(Op{OP_S2SHIFT, 0, tp.njoin, 0});
        // R0 ranges, each after the closure of the previous tick reached what it reads
        for (int g = 0; g < (int)tp.groups.size(); ++g) {
            const TickGroup &G = tp.groups[g];
            const TickGroup *prev = nullptr;
            if (last_tick == tp.tick - 1)
                for (const TickGroup &q : last_groups)
                    if (q.T == G.T) prev = &q;
            int waited = -1;
            for (int k = 0; k < nl; ++k) {
                const Launch &L = tp.S.launches[k];
                if (L.group != g || L.d1 < 0 || L.kind != BPMAX_K_OUTER) continue;
                if (prev) {
                    const int need = std::min(G.rs * L.diag + G.rs - 1, prev->nblmax - 1);
                    const int lvl = *std::lower_bound(prev->rec_levels.begin(), prev->rec_levels.end(), need);
                    if (lvl > waited) {
                        ops.push_back(Op{OP_WAIT, 0, 0, ev_key(tp.tick - 1, G.T, EV_IN, lvl)});
                        waited = lvl;
                    }
                }
                ops.push_back(Op{OP_LAUNCH, 0, k, 0});
                ops.push_back(Op{OP_RECORD, 0, 0, ev_key(tp.tick, G.T, EV_OUT, L.diag)});
            }
        }
        // closure levels, interleaved across groups by level, each after the R0 range that
        // holds its tiles
        std::vector<int> waited_out(tp.groups.size(), -1);
        std::vector<std::tuple<int, int, int>> order; // (level, group, launch)
        for (int k = 0; k < nl; ++k) {
            const Launch &L = tp.S.launches[k];
            if (L.d1 >= 0 && L.kind != BPMAX_K_OUTER) order.emplace_back(L.diag, L.group, k);
        }
        std::stable_sort(order.begin(), order.end(),
                         [](const std::tuple<int, int, int> &a, const std::tuple<int, int, int> &b) {
                             return std::get<0>(a) != std::get<0>(b) ? std::get<0>(a) < std::get<0>(b)
                                                                     : std::get<1>(a) < std::get<1>(b);
                         });
        for (auto &o : order) {
            const int k = std::get<2>(o);
            const Launch &L = tp.S.launches[k];
            const TickGroup &G = tp.groups[L.group];
            const int sd = std::min((L.diag + (L.ahead ? 1 : 0) + G.rs - 1) / G.rs, G.nbmax - 1);
            const int need = G.out_end[sd];
            if (need > waited_out[L.group]) {
                ops.push_back(Op{OP_WAIT, 1, 0, ev_key(tp.tick, G.T, EV_OUT, need)});
                waited_out[L.group] = need;
            }
            ops.push_back(Op{OP_LAUNCH, 1, k, 0});
            if (L.kind == BPMAX_K_INNER_INTRA &&
                std::binary_search(G.rec_levels.begin(), G.rec_levels.end(), L.diag))
                ops.push_back(Op{OP_RECORD, 1, 0, ev_key(tp.tick, G.T, EV_IN, L.diag)});
        }
        if (!tp.fin.empty()) ops.push_back(Op{OP_SCORES, 1, (int)tp.fin.size(), 0});
        ops.push_back(Op{OP_DONE, 1, 0, 0});
    }

    // A tick finished (ok) or failed: publish the finishing pairs' scores and free their
    // extents (kept ones stay until released).
    void retire(const TickPlan &tp, bool ok, const float *scores, const std::string &err, Clock::time_point now)
    {
        for (size_t k = 0; k < tp.fin.size(); ++k) {
            Request &r = reqs[tp.ids[tp.fin[k]]];
            r.t_done = now;
            if (ok) {
                r.state = RQ_DONE;
                r.score = scores ? scores[k] : 0.f;
                ++st.completed;
            } else {
                r.state = RQ_FAILED;
                r.err = err;
                ++st.failed;
            }
            if (!keep_tables || !ok) {
                ext_free(r.off, r.bytes);
                r.resident = false;
            }
        }
    }

    // Fails every request that is queued or resident (device error).
    void fail_all(const std::string &err, Clock::time_point now)
    {
        for (auto &kv : reqs) {
            Request &r = kv.second;
            if (r.state != RQ_QUEUED && r.state != RQ_RUNNING) continue;
            r.state = RQ_FAILED;
            r.err = err;
            r.t_done = now;
            ++st.failed;
        }
        queue.clear();
        active.clear();
        items_sum = 0;
    }

    // Stopping with nothing resident: fails the queued requests that can never be admitted
    // because kept tables hold the memory they need.
    void fail_unadmittable(Clock::time_point now)
    {
        while (!queue.empty()) {
            Request &r = reqs[queue.front()];
            r.state = RQ_FAILED;
            r.err = "not admitted: the arena is held by kept tables that were not released";
            r.t_done = now;
            ++st.failed;
            queue.pop_front();
        }
    }
};

// ------------------------------------------------------------------------- executor
struct Ring {
    void *d_block = nullptr;     // device staging block
    char *h_block = nullptr;     // pinned host copy
    int *h_codes = nullptr;      // pinned joiner codes
    float *d_scores = nullptr, *h_scores = nullptr;
    hipEvent_t done = nullptr;
    std::map<int, hipEvent_t> ev; // event pool of the slot, by local key
};

struct Server {
    bpmax_ctx *owner = nullptr;
    bpmax_ctx *eng = nullptr;    // the server's two streams and launch settings
    bpmax_server_config cfg{};
    ServeCore core;
    void *arena = nullptr;
    size_t block_cap = 0;
    std::vector<Ring> ring;
    std::deque<std::unique_ptr<TickPlan>> inflight;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::thread th;
    bool stopping = false, broken = false;
};

static size_t pad256(size_t x) { return (x + 255) / 256 * 256; }

// Byte offsets of a tick's staging block: weights | P | P of joiners | P of finishing pairs | items.
struct BlockLayout {
    size_t p, pj, pf, items, end;
};
static BlockLayout block_layout(size_t np, size_t nj, size_t nf, size_t nitems)
{
    BlockLayout b;
    b.p = 256;
    b.pj = b.p + pad256(np * sizeof(PairDesc));
    b.pf = b.pj + pad256(nj * sizeof(PairDesc));
    b.items = b.pf + pad256(nf * sizeof(PairDesc));
    b.end = b.items + nitems * sizeof(WorkItem);
    return b;
}

static int exec_tick(Server &S, const TickPlan &tp, std::string *err)
{
    bpmax_ctx *c = S.eng;
    Ring &R = S.ring[(size_t)(tp.tick % (long long)S.ring.size())];
    const BlockLayout bl = block_layout(tp.P.size(), tp.njoin, tp.fin.size(), tp.S.items.size());
    if (bl.end > S.block_cap || tp.codes.size() > S.core.cap_codes || tp.fin.size() > (size_t)S.core.max_active) {
        *err = "tick exceeds the staging capacity";
        return BPMAX_ENOMEM;
    }
    char *h = R.h_block;
    memcpy(h, S.core.w, sizeof S.core.w);
    memcpy(h + bl.p, tp.P.data(), tp.P.size() * sizeof(PairDesc));
    memcpy(h + bl.pj, tp.P.data(), tp.njoin * sizeof(PairDesc));
    for (size_t k = 0; k < tp.fin.size(); ++k) memcpy(h + bl.pf + k * sizeof(PairDesc), &tp.P[tp.fin[k]], sizeof(PairDesc));
    memcpy(h + bl.items, tp.S.items.data(), tp.S.items.size() * sizeof(WorkItem));
    if (!tp.codes.empty()) memcpy(R.h_codes, tp.codes.data(), tp.codes.size() * sizeof(int));
    char *d = (char *)R.d_block;
    const float *dw = (const float *)d;
    const PairDesc *dP = (const PairDesc *)(d + bl.p), *dPj = (const PairDesc *)(d + bl.pj),
                   *dPf = (const PairDesc *)(d + bl.pf);
    const WorkItem *dI = (const WorkItem *)(d + bl.items);
    float *arena = (float *)S.arena;
    const int *dcodes = (const int *)S.arena;
    hipStream_t st[2] = {c->stream, c->stream2};
    auto event = [&](long long key) -> hipEvent_t * {
        Ring &Rk = S.ring[(size_t)(ev_tick(key) % (long long)S.ring.size())];
        auto it = Rk.ev.find(ev_local(key));
        if (it != Rk.ev.end()) return &it->second;
        hipEvent_t e = nullptr;
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return nullptr;
        return &(Rk.ev[ev_local(key)] = e);
    };
#define SRVCHK(expr)                                                                                              \
    do {                                                                                                          \
        hipError_t _e = (expr);                                                                                   \
        if (_e != hipSuccess) {                                                                                   \
            *err = std::string(#expr) + ": " + hipGetErrorString(_e);                                            \
            return BPMAX_EDEVICE;                                                                                 \
        }                                                                                                         \
    } while (0)
    for (const Op &o : tp.ops) {
        hipStream_t s = st[o.stream];
        switch (o.kind) {
        case OP_UPLOAD: SRVCHK(hipMemcpyAsync(d, h, bl.end, hipMemcpyHostToDevice, s)); break;
        case OP_CODES: {
            const PairDesc &p = tp.P[o.a];
            const size_t n = (o.a + 1 < tp.njoin ? tp.codes_at[o.a + 1] : tp.codes.size()) - tp.codes_at[o.a];
            SRVCHK(hipMemcpyAsync((int *)S.arena + p.a, R.h_codes + tp.codes_at[o.a], n * sizeof(int),
                                  hipMemcpyHostToDevice, s));
            break;
        }
        case OP_FILL: { // -inf over F, S2, S1 and (FP16 -inf) over the FP16 tables
            const PairDesc &p = tp.P[o.a];
            if (p.Fh >= 0) SRVCHK(launch_fill2(s, arena + p.F, p.Fh - p.F, -INFINITY, p.a - p.Fh, 0xFC00FC00u));
            else SRVCHK(launch_fill(s, arena + p.F, p.a - p.F, -INFINITY));
            break;
        }
        case OP_S2SHIFT: SRVCHK(launch_s2_shift(s, dPj, o.a, arena)); break;
        case OP_S1: SRVCHK(launch_s1(s, dPj, o.a, arena, dcodes, dw)); break;
        case OP_LAUNCH: {
            const int rc = bpx_run_launch(c, s, tp.S.launches[o.a], dP, dI, arena, dcodes, dw);
            if (rc != BPMAX_OK) {
                *err = c->err;
                return rc;
            }
            break;
        }
        case OP_RECORD: {
            hipEvent_t *e = event(o.ev);
            if (!e) {
                *err = "hipEventCreate failed";
                return BPMAX_EDEVICE;
            }
            SRVCHK(hipEventRecord(*e, s));
            break;
        }
        case OP_WAIT: {
            hipEvent_t *e = event(o.ev);
            if (!e) {
                *err = "hipEventCreate failed";
                return BPMAX_EDEVICE;
            }
            SRVCHK(hipStreamWaitEvent(s, *e, 0));
            break;
        }
        case OP_SCORES:
            SRVCHK(launch_scores(s, dPf, o.a, arena, R.d_scores));
            SRVCHK(hipMemcpyAsync(R.h_scores, R.d_scores, o.a * sizeof(float), hipMemcpyDeviceToHost, s));
            break;
        case OP_DONE: SRVCHK(hipEventRecord(R.done, s)); break;
        }
    }
#undef SRVCHK
    return BPMAX_OK;
}

static void server_loop(Server *S)
{
    std::unique_lock<std::mutex> lk(S->mu);
    ServeCore &C = S->core;
    for (;;) {
        bool progress = false;
        // retire finished ticks in order
        while (!S->inflight.empty()) {
            const TickPlan &tp = *S->inflight.front();
            Ring &R = S->ring[(size_t)(tp.tick % (long long)S->ring.size())];
            const hipError_t q = hipEventQuery(R.done);
            if (q == hipErrorNotReady) break;
            const auto now = Clock::now();
            if (q != hipSuccess) {
                const std::string e = std::string("tick failed: ") + hipGetErrorString(q);
                C.retire(tp, false, nullptr, e, now);
                for (auto &t : S->inflight)
                    if (t.get() != &tp) C.retire(*t, false, nullptr, e, now);
                S->inflight.clear();
                C.fail_all(e, now);
                S->broken = true;
            } else {
                C.retire(tp, true, R.h_scores, "", now);
                S->inflight.pop_front();
            }
            progress = true;
        }
        if (progress) S->cv_done.notify_all();
        if (!S->broken && (int)S->inflight.size() < C.depth) {
            auto tp = std::make_unique<TickPlan>();
            if (C.plan(*tp)) {
                lk.unlock();
                std::string err;
                const int rc = exec_tick(*S, *tp, &err);
                lk.lock();
                if (rc != BPMAX_OK) {
                    (void)hipStreamSynchronize(S->eng->stream);
                    (void)hipStreamSynchronize(S->eng->stream2);
                    const auto now = Clock::now();
                    for (auto &t : S->inflight) C.retire(*t, false, nullptr, err, now);
                    S->inflight.clear();
                    C.retire(*tp, false, nullptr, err, now);
                    C.fail_all(err, now);
                    S->broken = true;
                    S->cv_done.notify_all();
                } else {
                    S->inflight.push_back(std::move(tp));
                }
                continue;
            }
        }
        if (S->stopping && S->inflight.empty() && C.active.empty()) {
            if (C.queue.empty()) break;
            // nothing runs and the queue head cannot be admitted: kept tables hold the arena
            C.fail_unadmittable(Clock::now());
            S->cv_done.notify_all();
            continue;
        }
        if (!progress) S->cv_work.wait_for(lk, std::chrono::microseconds(S->inflight.empty() ? 2000 : 50));
    }
}

static void server_free(Server *S)
{
    if (S->eng) {
        (void)hipStreamSynchronize(S->eng->stream);
        (void)hipStreamSynchronize(S->eng->stream2);
    }
    for (Ring &R : S->ring) {
        if (R.d_block) (void)hipFree(R.d_block);
        if (R.h_block) (void)hipHostFree(R.h_block);
        if (R.h_codes) (void)hipHostFree(R.h_codes);
        if (R.d_scores) (void)hipFree(R.d_scores);
        if (R.h_scores) (void)hipHostFree(R.h_scores);
        if (R.done) (void)hipEventDestroy(R.done);
        for (auto &kv : R.ev) (void)hipEventDestroy(kv.second);
    }
    if (S->arena) (void)hipFree(S->arena);
    if (S->eng) (void)bpmax_ctx_destroy(S->eng);
    delete S;
}

// Planner settings shared by the server and its check.
static void core_setup(ServeCore &C, size_t budget, int max_active, int depth, const Tuning &eff, bool keep,
                       const float *intra, const float *inter)
{
    C.budget = budget;
    C.max_active = max_active;
    C.depth = depth;
    C.eff = eff;
    C.keep_tables = keep;
    C.cap_items = 4ll << 20;
    C.cap_codes = 4u << 20;
    C.free_ext.clear();
    if (budget) C.free_ext[0] = budget;
    C.st.budget = budget;
    memcpy(C.w, intra, 25 * sizeof(float));
    memcpy(C.w + 25, inter, 25 * sizeof(float));
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
    if (cfg.max_batch <= 0) cfg.max_batch = 256;
    if (cfg.slots <= 0) cfg.slots = 3;
    if (cfg.queue_capacity <= 0) cfg.queue_capacity = 4096;
    if (cfg.slots > 8) {
        ctx->err = "slots (ticks in flight) must be <= 8";
        return BPMAX_EINVAL;
    }
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
    const float *intra = cfg.intra ? cfg.intra : di;
    const float *inter = cfg.inter ? cfg.inter : (cfg.intra ? cfg.intra : de);
    const Tuning eff{BPMAX_SCHED_PIPE, ctx->outer_split, ctx->outer_items, 0, ctx->intra_wide, ctx->outer_kgroup,
                     ctx->closure_ahead};
    core_setup(S->core, cfg.memory_budget, cfg.max_batch, cfg.slots, eff, cfg.keep_tables != 0, intra, inter);
    S->core.r0_half = ctx->r0_half; // (bpmax_ctx_set_r0_half, BPMAX_R0_HALF)
    S->cfg.intra = S->core.w;
    S->cfg.inter = S->core.w + 25;
    int rc = bpmax_ctx_create(&S->eng, ctx->device);
    if (rc != BPMAX_OK) {
        ctx->err = "server streams: context creation failed";
        server_free(S);
        return rc;
    }
    S->eng->eff_intra_wide = ctx->intra_wide;
    S->eng->intra_blocked = ctx->intra_blocked;
    auto fail = [&](const std::string &m, int code) {
        (void)hipGetLastError();
        ctx->err = m;
        server_free(S);
        return code;
    };
    if (hipMalloc(&S->arena, cfg.memory_budget) != hipSuccess)
        return fail("server arena: hipMalloc of " + std::to_string(cfg.memory_budget) + " bytes failed", BPMAX_ENOMEM);
    const size_t np = (size_t)cfg.max_batch;
    S->block_cap = block_layout(np, np, np, (size_t)S->core.cap_items).end;
    S->ring.resize((size_t)cfg.slots);
    for (Ring &R : S->ring) {
        if (hipMalloc(&R.d_block, S->block_cap) != hipSuccess ||
            hipHostMalloc((void **)&R.h_block, S->block_cap) != hipSuccess ||
            hipHostMalloc((void **)&R.h_codes, S->core.cap_codes * sizeof(int)) != hipSuccess ||
            hipMalloc((void **)&R.d_scores, np * sizeof(float)) != hipSuccess ||
            hipHostMalloc((void **)&R.h_scores, np * sizeof(float)) != hipSuccess)
            return fail("server staging buffers: allocation failed", BPMAX_ENOMEM);
        if (hipEventCreateWithFlags(&R.done, hipEventDisableTiming) != hipSuccess)
            return fail("server events: creation failed", BPMAX_EDEVICE);
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
    if (!weights_exact(S.core.w, S.core.w + 25, M, N, &why)) {
        ctx->err = why;
        return BPMAX_EWEIGHTS;
    }
    Request r;
    r.s1.assign(seq1, seq1 + M);
    r.s2.assign(seq2, seq2 + N);
    r.M = M;
    r.N = N;
    std::lock_guard<std::mutex> g(S.mu);
    if (S.stopping || S.broken) return BPMAX_EINVAL;
    const int rc = S.core.prepare(r, &why);
    if (rc != BPMAX_OK) {
        ctx->err = why;
        return rc;
    }
    if ((int)S.core.queue.size() >= S.cfg.queue_capacity) return BPMAX_EBUSY;
    r.t_submit = Clock::now();
    *id = S.core.submit(std::move(r));
    S.cv_work.notify_one();
    return BPMAX_OK;
}

static int poll_locked(bpmax_ctx *ctx, Server &S, long long id, float *score, double *latency_ms)
{
    auto it = S.core.reqs.find(id);
    if (it == S.core.reqs.end()) return BPMAX_ENOTFOUND;
    const Request &r = it->second;
    if (r.state == RQ_FAILED) {
        ctx->err = r.err;
        return r.err.rfind("not admitted", 0) == 0 ? BPMAX_ENOMEM : BPMAX_EDEVICE;
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
    {
        std::lock_guard<std::mutex> g(S.mu);
        auto it = S.core.reqs.find(id);
        if (it == S.core.reqs.end()) return BPMAX_ENOTFOUND;
        const Request &r = it->second;
        if (r.state == RQ_FAILED) return BPMAX_EDEVICE;
        if (r.state != RQ_DONE) return BPMAX_ENOTREADY;
        if (!S.core.keep_tables || !r.resident) {
            ctx->err = "tables are not kept (keep_tables = 0)";
            return BPMAX_EINVAL;
        }
        d = r.desc;
    }
    // the extent stays allocated while the request is unreleased; the copy runs on the
    // owner context's own stream
    return bpx_copy_out(ctx, (const float *)S.arena, d, layout, F_out);
}

extern "C" int bpmax_server_release(bpmax_ctx *ctx, long long id)
{
    if (!ctx || !ctx->server) return BPMAX_EINVAL;
    Server &S = *ctx->server;
    std::lock_guard<std::mutex> g(S.mu);
    auto it = S.core.reqs.find(id);
    if (it == S.core.reqs.end()) return BPMAX_ENOTFOUND;
    if (it->second.state == RQ_QUEUED || it->second.state == RQ_RUNNING) return BPMAX_EBUSY;
    if (it->second.resident) S.core.ext_free(it->second.off, it->second.bytes);
    S.core.reqs.erase(it);
    S.cv_work.notify_one();
    return BPMAX_OK;
}

extern "C" int bpmax_server_get_stats(bpmax_ctx *ctx, bpmax_server_stats *stats)
{
    if (!ctx || !ctx->server || !stats) return BPMAX_EINVAL;
    std::lock_guard<std::mutex> g(ctx->server->mu);
    *stats = ctx->server->core.st;
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

// ------------------------------------------------------------------- host-only check
// Replays requests through the planner (no GPU): request k is submitted once the server has
// planned arrive[k] ticks (or at once when nothing else is left), at most `depth` ticks are in
// flight, and the oldest retires when no further tick can be planned or `depth` are in
// flight.  The operations of all ticks then run through a model of two FIFO streams with
// events, and every work item must start after all the work it reads; every wait must name
// an event recorded before it and not yet re-recorded by a later tick sharing its ring slot;
// extents of resident pairs must never overlap, and every operation on a pair must finish
// before its extent is released.
extern "C" int bpmax_server_check(int nreq, const int *Ms, const int *Ns, const int *arrive, size_t memory_budget,
                                  int max_active, int depth, const bpmax_tuning *t, int skip, long long *checked,
                                  char *msg, int msglen)
{
    auto fail = [&](const std::string &m) {
        if (msg && msglen > 0) snprintf(msg, (size_t)msglen, "%s", m.c_str());
        return BPMAX_EINVAL;
    };
    if (nreq < 1 || !Ms || !Ns || !arrive || !t || max_active < 1 || depth < 1 || depth > 8 || memory_budget < 4096 ||
        t->outer_items < 64 || t->outer_split < 0 || t->outer_kgroup < 0 || t->outer_kgroup > 255)
        return fail("bad arguments");
    ServeCore C;
    float di[25], de[25];
    bpmax_engine_score_table(di, de);
    const Tuning eff{BPMAX_SCHED_PIPE, t->outer_split, t->outer_items, 0, t->intra_wide, t->outer_kgroup,
                     t->closure_ahead ? 1 : 0};
    core_setup(C, memory_budget / 4096 * 4096, max_active, depth, eff, false, di, de);
    std::vector<long long> rid(nreq, -1);
    std::vector<std::unique_ptr<TickPlan>> plans;
    std::deque<TickPlan *> fl;
    // sequence of (kind 0: tick planned, 1: tick retired) for the extent checks
    std::vector<std::pair<int, TickPlan *>> seq;
    int k = 0;
    for (;;) {
        while (k < nreq && arrive[k] <= C.next_tick) {
            Request r;
            r.M = Ms[k];
            r.N = Ns[k];
            if (r.M < 1 || r.N < 1) return fail("bad shape");
            // all G: the largest best pair weight per base under the default table, so the
            // FP16 choice follows the length-only bound
            r.s1.assign(r.M, 2);
            r.s2.assign(r.N, 2);
            std::string why;
            if (C.prepare(r, &why) != BPMAX_OK) return fail(why);
            rid[k++] = C.submit(std::move(r));
        }
        if ((int)fl.size() < depth) {
            auto tp = std::make_unique<TickPlan>();
            if (C.plan(*tp)) {
                fl.push_back(tp.get());
                seq.emplace_back(0, tp.get());
                plans.push_back(std::move(tp));
                continue;
            }
        }
        if (!fl.empty()) {
            std::vector<float> sc(fl.front()->fin.size(), 0.f);
            C.retire(*fl.front(), true, sc.data(), "", Clock::now());
            seq.emplace_back(1, fl.front());
            fl.pop_front();
            continue;
        }
        if (k < nreq) { // idle server: the next request arrives now
            const int a = arrive[k];
            for (int q = k; q < nreq && arrive[q] == a; ++q) {
                Request r;
                r.M = Ms[q];
                r.N = Ns[q];
                r.s1.assign(r.M, 2);
                r.s2.assign(r.N, 2);
                std::string why;
                if (C.prepare(r, &why) != BPMAX_OK) return fail(why);
                rid[q] = C.submit(std::move(r));
                k = q + 1;
            }
            continue;
        }
        break;
    }
    for (int q = 0; q < nreq; ++q)
        if (C.reqs[rid[q]].state != RQ_DONE) return fail("request " + std::to_string(q) + " never finished");

    // --- stream model: launch-like operations get global ids; frontier per stream
    struct GOp {
        int tick, stream, kind, a;
        long long ev;
    };
    std::vector<GOp> ops;
    for (auto &tp : plans)
        for (const Op &o : tp->ops) ops.push_back(GOp{(int)tp->tick, o.stream, o.kind, o.a, o.ev});
    {
        int nw = 0;
        for (auto &o : ops)
            if (o.kind == OP_WAIT && nw++ == skip) o.kind = -1; // planted error: the wait is dropped
    }
    const int n = (int)ops.size();
    std::vector<int> fO(n, -1), fI(n, -1); // op i starts after every op <= fO[i] on stream 0, <= fI[i] on 1
    int cur[2][2] = {{-1, -1}, {-1, -1}};  // per stream: frontier (stream 0, stream 1)
    int last[2] = {-1, -1};
    std::map<long long, std::pair<int, int>> evf; // event key -> frontier at its record
    std::map<int, long long> slot_tick;           // (slot, local key) -> tick of the latest record
    for (int i = 0; i < n; ++i) {
        GOp &o = ops[i];
        const int s = o.stream;
        if (o.kind == OP_WAIT) {
            auto it = evf.find(o.ev);
            if (it == evf.end()) return fail("a wait names an event not recorded before it");
            const int sk = (int)(ev_tick(o.ev) % depth) << 20 | ev_local(o.ev);
            if (slot_tick[sk] != ev_tick(o.ev)) return fail("a wait names an event its ring slot already re-recorded");
            cur[s][0] = std::max(cur[s][0], it->second.first);
            cur[s][1] = std::max(cur[s][1], it->second.second);
            continue;
        }
        if (o.kind == -1) continue;
        if (o.kind == OP_RECORD) {
            int f[2] = {cur[s][0], cur[s][1]};
            f[s] = std::max(f[s], last[s]);
            evf[o.ev] = {f[0], f[1]};
            slot_tick[(int)(ev_tick(o.ev) % depth) << 20 | ev_local(o.ev)] = ev_tick(o.ev);
            continue;
        }
        cur[s][s] = std::max(cur[s][s], last[s]);
        fO[i] = cur[s][0];
        fI[i] = cur[s][1];
        last[s] = i;
    }
    auto before = [&](int a, int b) { return a >= 0 && (ops[a].stream == 0 ? a <= fO[b] : a <= fI[b]); };

    // --- per-request bookkeeping: op index of each launch, keyed by request id
    std::vector<int> first_op(plans.size());
    {
        int i = 0;
        for (size_t p = 0; p < plans.size(); ++p) {
            first_op[p] = i;
            i += (int)plans[p]->ops.size();
        }
    }
    using Key = std::tuple<long long, int, int, int, int>;
    std::map<Key, int> closed, r0_last, ext_at;
    std::map<long long, int> fill_op, s1_op, s2_last, done_op, shift_op;
    std::map<long long, std::vector<int>> touch; // ops that run items of a request
    long long nchk = 0;
    for (size_t p = 0; p < plans.size(); ++p) {
        const TickPlan &tp = *plans[p];
        for (size_t q = 0; q < tp.ops.size(); ++q) {
            const int i = first_op[p] + (int)q;
            const Op &o = tp.ops[q];
            if (o.kind == OP_FILL) fill_op[tp.ids[o.a]] = i;
            if (o.kind == OP_CODES) touch[tp.ids[o.a]].push_back(i);
            if (o.kind == OP_S1)
                for (int j = 0; j < o.a; ++j) s1_op[tp.ids[j]] = i;
            if (o.kind == OP_S2SHIFT)
                for (int j = 0; j < o.a; ++j)
                    if (tp.P[j].S2r >= 0) shift_op[tp.ids[j]] = i;
            if (o.kind == OP_DONE)
                for (int f : tp.fin) done_op[tp.ids[f]] = i;
            if (o.kind != OP_LAUNCH) continue;
            const Launch &L = tp.S.launches[o.a];
            for (int w = L.first; w < L.first + L.count; ++w) {
                const WorkItem &it = tp.S.items[w];
                const long long id = tp.ids[it.pair];
                touch[id].push_back(i);
                const Key key{id, it.i1, it.j1, it.I, it.J};
                if (L.d1 < 0) s2_last[id] = i;
                else if (L.kind == BPMAX_K_OUTER) r0_last[key] = i;
                else if (L.kind == BPMAX_K_INNER_INTRA && wi_type(it.flags) != TASK_EXT) {
                    if (closed.count(key)) return fail("a tile is closed twice");
                    closed[key] = i;
                } else ext_at[key] = std::max(ext_at.count(key) ? ext_at[key] : -1, i);
            }
        }
    }
    for (size_t p = 0; p < plans.size(); ++p) {
        const TickPlan &tp = *plans[p];
        for (size_t q = 0; q < tp.ops.size(); ++q) {
            const int i = first_op[p] + (int)q;
            const Op &o = tp.ops[q];
            if (o.kind == OP_S1 || o.kind == OP_FILL || o.kind == OP_CODES || o.kind == OP_S2SHIFT) {
                if (!before(first_op[p], i) && q > 0) return fail("prologue before the tick's upload");
                continue;
            }
            if (o.kind == OP_SCORES) {
                for (int f : tp.fin) {
                    const PairDesc &d = tp.P[f];
                    const auto c = closed.find(Key{tp.ids[f], 0, d.M - 1, 0, d.nblr - 1});
                    if (c == closed.end() || !before(c->second, i)) return fail("score read before its cell is final");
                    ++nchk;
                }
                continue;
            }
            if (o.kind != OP_LAUNCH) continue;
            if (!before(first_op[p], i)) return fail("a launch can start before its tick's staging block is uploaded");
            const Launch &L = tp.S.launches[o.a];
            char where[200];
            for (int w = L.first; w < L.first + L.count; ++w) {
                const WorkItem &it = tp.S.items[w];
                const long long id = tp.ids[it.pair];
                const PairDesc &d = tp.P[it.pair];
                const int rs = d.nbl / d.nb;
                snprintf(where, sizeof where, " (tick %lld, request %lld, triangle (%d,%d), tile (%d,%d))", tp.tick, id,
                         it.i1, it.j1, it.I, it.J);
                if (!before(fill_op.at(id), i)) return fail(std::string("work before the pair's fill") + where);
                if (L.d1 < 0) continue; // S2: in order on one stream after the fill
                if (L.kind == BPMAX_K_OUTER) {
                    if (!before(s2_last.at(id), i) || !before(s1_op.at(id), i))
                        return fail(std::string("R0 before S1/S2") + where);
                    auto src_closed = [&](int a, int b) {
                        const auto c = closed.find(Key{id, a, b, it.I * rs, std::min((it.J + 1) * rs, d.nblr) - 1});
                        return c != closed.end() && before(c->second, i);
                    };
                    for (int k1 = it.i1; k1 < it.j1; ++k1) {
                        if (!src_closed(it.i1, k1) || !src_closed(k1 + 1, it.j1))
                            return fail(std::string("R0 before the closure it reads") + where);
                        nchk += 2;
                    }
                    if ((it.flags & WI_EPILOGUE) && it.j1 - it.i1 >= 4 && !src_closed(it.i1 + 1, it.j1 - 1))
                        return fail(std::string("P1 before its source is closed") + where);
                    continue;
                }
                const auto r0 = r0_last.find(Key{id, it.i1, it.j1, it.I / rs, it.J / rs});
                if (r0 == r0_last.end() || !before(r0->second, i)) return fail(std::string("closure before R0") + where);
                if (d.S2r >= 0) { // FP16 cross-tile operands: S2's shifted copies
                    const auto sh = shift_op.find(id);
                    if (sh == shift_op.end() || !before(s2_last.at(id), sh->second) || !before(sh->second, i))
                        return fail(std::string("closure before the shifted copies of S2") + where);
                }
                const bool cross = L.kind == BPMAX_K_INNER_EXT || wi_type(it.flags) == TASK_EXT;
                const int D = it.J - it.I;
                auto tile_closed = [&](int X, int Y) {
                    const auto c = closed.find(Key{id, it.i1, it.j1, X, Y});
                    return c != closed.end() && before(c->second, i);
                };
                if (cross) {
                    for (int s = it.step_begin; s < it.step_end; ++s) {
                        if (!(s < D ? tile_closed(it.I, it.I + s) : tile_closed(it.I + 1 + s - D, it.J)))
                            return fail(std::string("cross-tile step before its operand") + where);
                        ++nchk;
                    }
                    continue;
                }
                for (int Y = it.I; Y < it.J; ++Y)
                    if (!tile_closed(it.I, Y)) return fail(std::string("closure before the tile left of it") + where);
                for (int X = it.I + 1; X <= it.J; ++X)
                    if (!tile_closed(X, it.J)) return fail(std::string("closure before the tile below it") + where);
                if (L.kind == BPMAX_K_INNER_INTRA && !L.ahead) {
                    const auto e = ext_at.find(Key{id, it.i1, it.j1, it.I, it.J});
                    if (e != ext_at.end() && !before(e->second, i))
                        return fail(std::string("in-tile before cross-tile") + where);
                }
                nchk += 1 + 2 * D;
            }
        }
    }
    // every tile of every pair closed; every operation on a pair done before its release
    for (int q = 0; q < nreq; ++q) {
        const Request &r = C.reqs[rid[q]];
        const int nbl = r.desc.nblr;
        for (int i1 = 0; i1 < r.M; ++i1)
            for (int j1 = i1; j1 < r.M; ++j1)
                for (int X = 0; X < nbl; ++X)
                    for (int Y = X; Y < nbl; ++Y)
                        if (!closed.count(Key{rid[q], i1, j1, X, Y})) return fail("a tile is never closed");
        const int dn = done_op.at(rid[q]);
        for (int i : touch[rid[q]])
            if (!before(i, dn) && i != dn) return fail("an operation on a pair can run after its release");
    }
    // extents: resident pairs never overlap (replayed in plan/retire order); a tick reuses a
    // ring slot only after the tick that used it before retired
    {
        std::map<long long, std::pair<size_t, size_t>> live;
        std::set<long long> retired;
        for (auto &e : seq) {
            const TickPlan &tp = *e.second;
            if (e.first == 1) retired.insert(tp.tick);
            if (e.first == 0 && tp.tick >= depth && !retired.count(tp.tick - depth))
                return fail("a tick reuses a ring slot whose previous tick has not retired");
            if (e.first == 0) {
                for (int j = 0; j < tp.njoin; ++j) {
                    const Request &r = C.reqs[tp.ids[j]];
                    for (auto &kv : live) {
                        const size_t a = kv.second.first, b = a + kv.second.second;
                        if (r.off < b && a < r.off + r.bytes) return fail("extents of two resident pairs overlap");
                    }
                    if (r.off + r.bytes > C.budget) return fail("an extent leaves the arena");
                    live[tp.ids[j]] = {r.off, r.bytes};
                }
            } else {
                for (int f : tp.fin) live.erase(tp.ids[f]);
            }
        }
    }
    if (checked) *checked = nchk;
    return BPMAX_OK;
}
