// bpmax_engine_impl.h -- private definition of bpmax_ctx and engine helpers.
#pragma once
#include <hip/hip_runtime.h>
#include <functional>
#include <map>
#include <string>
#include <vector>

#include "bpmax.h"
#include "bpmax_internal.h"

namespace bpx {

struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
};

struct Launch {
    int kind;            // bpmax_kernel_class
    int T;
    bool single_strand;  // S2 phases
    int first, count;    // work-item range (count 0: nothing to launch, events only)
    double updates;      // algorithmic cell updates (outer phases)
    int group;           // T group (pairs sharing a storage tile size)
    int d1;              // outer diagonal (-1: S2 phases)
    int diag;            // storage block diagonal (outer) or logical block diagonal (inner)
    bool ahead = false;  // in-tile launch that also carries diagonal diag+1's early
                         // cross-tile steps (k_inner_closure_blk)
    bool bulk = false;   // R0 launch of the k1 terms that read only the triangles two or more
                         // outer diagonals below (runs one outer diagonal ahead)
    int half_items = 0;  // R0 launch of the serving loop: its first half_items items belong to
                         // pairs with FP16 tables and run on FP16 operands
};

// launch-schedule settings of one problem class (bpmax_tuning)
struct Tuning {
    int schedule, outer_split, outer_items, graphs, intra_wide, outer_kgroup, closure_ahead;
    int tile = 0; // storage tile: 0 = choose_tile, 32 = 32 for every pair
};

struct GroupInfo {
    int T, rs;           // storage tile, logical tiles per storage tile side
    int maxM, maxnb, maxnbl;
    std::vector<int> out_end; // storage diagonal -> last diagonal of the R0 launch holding it
};

struct Schedule {
    std::vector<WorkItem> items;
    std::vector<Launch> launches;
    std::vector<GroupInfo> groups;
};

struct KStat {
    long long launches = 0;
    double ms = 0.0;
    double updates = 0.0;
};

struct PendingTiming {
    int kclass;
    double updates;
    hipEvent_t e0, e1;
};

struct Server; // continuous-batching server (bpmax_server.cpp)

int choose_tile(int N);
int pick_tile(int N, int tile);
size_t pair_device_floats(int M, int N, int T);
bool weights_exact(const float *intra, const float *inter, int M, int N, std::string *why);
bool weights_half_exact(const float *intra, const float *inter, int M, int N, const int *s1 = nullptr,
                        const int *s2 = nullptr);
// r0_ahead: the bulk of each outer diagonal's R0 (k1 terms reading only triangles two or
// more diagonals below) is emitted one diagonal early, for the R0 stream, and the two edge
// k1 terms with the closure (see enqueue_pipe); otherwise R0 of d1 is emitted whole.
void build_schedule(const std::vector<PairDesc> &P, Schedule &S, int outer_split, int outer_items, int outer_kgroup,
                    bool closure_ahead, bool r0_ahead = false);
// S2 of the pairs with sel[p] set, group g (tile size T), logical diagonals < nblmax.
void emit_s2(const std::vector<PairDesc> &P, const std::vector<char> &sel, int g, int T, int nblmax, Schedule &S);
// One step of the outer wavefront for group g: pair p on outer diagonal d1s[p] (< 0: not in
// the tick); out_end receives the last storage diagonal of the R0 range holding each Ds.
void emit_tick(const std::vector<PairDesc> &P, const std::vector<int> &d1s, int g, int T, int nbmax, int nblmax,
               int tick, Schedule &S, std::vector<int> &out_end, int outer_split, int outer_items,
               int outer_kgroup, bool closure_ahead);
long long tuning_class(int M, int N, int npairs);
unsigned long long splitmix64_host(unsigned long long x);
void build_tasks(const std::vector<PairDesc> &P, int T, int chunk, std::vector<WorkItem> &tasks,
                 std::vector<int> &flags);

} // namespace bpx

struct bpmax_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    bpx::DevBuf arena, codes, meta, out, scores, r0meta, flags;
    bpx::DevBuf r0half;                  // bpmax_tri_to_half's inexact flag
    std::vector<int> flags_host, flags_staging;
    int schedule = BPMAX_SCHED_PIPE;
    int task_chunk = 16;
    unsigned long long task_polls = 0;   // last task-scheduler run: waiting polls (sum over tasks
    long long task_count = 0;            // of the largest lane count) and tasks run
    std::vector<bpx::PairDesc> pairs_host;
    bool result_valid = false;           // pairs_host's F tables are resident in the arena
    bpx::Schedule sched;
    std::vector<long long> sched_key;   // (schedule, chunk, shapes) the schedule was built for
    void *items_on_device = nullptr;    // meta buffer that already holds sched.items
    std::vector<char> staging;
    std::vector<int> codes_host;
    void *arena_ext = nullptr;           // device arena slice owned by someone else
    size_t arena_ext_bytes = 0;
    // no-allocation mode (bpmax_engine_solve_noalloc): a run's arena, codes and meta are
    // carved from this device region, its scores land in carve_scores
    char *carve = nullptr;
    size_t carve_bytes = 0;
    float *carve_scores = nullptr;
    char *meta_dev = nullptr;            // meta block (weights | descs | items) of the last run
    // pipelined schedule: second stream, dependency events, cached graph
    hipStream_t stream2 = nullptr;       // inner closure
    int outer_split = 1;                 // R0 launches per outer diagonal (per part), pipelined schedule
                                         // (0: one per storage diagonal)
    int outer_items = 4096;              // work items an R0 launch aims for (split-K granularity)
    int outer_kgroup = 1;                // K values per R0 work item (0: every K of the tile)
    bool r0_ahead = true;                // pipelined schedule: bulk R0 one outer diagonal ahead
    bool r0_half = true;                 // R0 on FP16 operands where that is exact (setting)
    bool run_half = false;               // the current run's R0 launches use the FP16 mirrors
    long long fill_half = 0;             // words of FP16 mirror after the run's FP32 tables
    int intra_wide = 256;                // in-tile launches of at most this many tiles use the
                                         // 1024-lane (latency) form, larger ones the blocked
                                         // (or, with intra_blocked = 0, the 512-lane) form
    int eff_intra_wide = 256;            // the current run's threshold
    int closure_ahead = 1;               // inner closure in one launch per diagonal (bpmax_tuning)
    int tile = 0;                        // storage tile setting (bpmax_tuning)
    int intra_blocked = 1;               // F in-tile launches that are not wide use the blocked
                                         // 256-lane form (0: the 512-lane push form)
    std::map<long long, bpx::Tuning> tuning; // tuned settings per problem class
    std::vector<hipEvent_t> sync_ev;
    bool graphs = false;
    hipGraphExec_t graph = nullptr;
    std::vector<long long> graph_key;
    bool timing = false;
    unsigned timing_mask = ~0u;          // kernel classes timed while timing is on
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    std::vector<bpx::PendingTiming> pending;
    bpx::KStat stats[BPMAX_K_NCLASSES];
    std::vector<void *> user_allocs;
    void *pin[2] = {nullptr, nullptr};   // pinned staging of large device-to-host copies
    hipEvent_t pin_ev[2] = {nullptr, nullptr};
    bpx::Server *server = nullptr;
};

// Return BPMAX_EDEVICE from the enclosing function, with the message in ctx->err, when a
// HIP call fails.
#define HIPCHK(ctx, expr)                                                                   \
    do {                                                                                    \
        hipError_t _e = (expr);                                                             \
        if (_e != hipSuccess) {                                                             \
            (ctx)->err = std::string(#expr) + ": " + hipGetErrorString(_e);                 \
            return BPMAX_EDEVICE;                                                           \
        }                                                                                   \
    } while (0)

int bpx_grow(bpmax_ctx *ctx, bpx::DevBuf &b, size_t bytes);
// One launch of a schedule on stream st (nothing for an empty position); dP / dI / dcodes /
// dw: device descriptors, work items, codes base (PairDesc.a/.b are int offsets from it)
// and the 50 weights.
int bpx_run_launch(bpmax_ctx *ctx, hipStream_t st, const bpx::Launch &L, const bpx::PairDesc *dP,
                   const bpx::WorkItem *dI, float *arena, const int *dcodes, const float *dw);
// Lays out and solves the pairs; with sync = false only enqueues (the context's streams
// hold the work; the host-side copies live in the context until its next run).
int bpx_run_pairs(bpmax_ctx *ctx, int npairs, const int *const *seq1s, const int *Ms, const int *const *seq2s,
                  const int *Ns, const float *intra, const float *inter, bool sync);
// Scores of the last run into `scores` (pinned host memory when sync = false).
int bpx_gather_scores(bpmax_ctx *ctx, float *scores, bool sync);
float *bpx_arena(bpmax_ctx *ctx);
// F of pair P (resident in `arena`) in `layout` to host F_out, through ctx's stream.
int bpx_copy_out(bpmax_ctx *ctx, const float *arena, const bpx::PairDesc &P, bpmax_layout layout, float *F_out);
size_t bpx_arena_bytes(int npairs, const int *Ms, const int *Ns);
int bpx_collect_timing(bpmax_ctx *ctx);
// Synchronous device-to-host copy of `bytes` into pageable memory, on ctx's stream.
int bpx_d2h(bpmax_ctx *ctx, void *dst, const void *src, size_t bytes);
void bpx_server_shutdown(bpmax_ctx *ctx);
