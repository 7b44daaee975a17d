// bpmax_engine_impl.h -- private definition of bpmax_ctx and engine helpers.
#pragma once
#include <hip/hip_runtime.h>
#include <functional>
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
size_t pair_device_floats(int M, int N, int T);
bool weights_exact(const float *intra, const float *inter, int M, int N, std::string *why);
void build_schedule(const std::vector<PairDesc> &P, Schedule &S, int outer_split);
void build_tasks(const std::vector<PairDesc> &P, int T, int chunk, std::vector<WorkItem> &tasks,
                 std::vector<int> &flags);

} // namespace bpx

struct bpmax_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    bpx::DevBuf arena, codes, meta, out, scores, r0meta, flags;
    std::vector<int> flags_host, flags_staging;
    int schedule = BPMAX_SCHED_PIPE;
    int task_chunk = 16;
    std::vector<bpx::PairDesc> pairs_host;
    bpx::Schedule sched;
    std::vector<long long> sched_key;   // (schedule, chunk, shapes) the schedule was built for
    void *items_on_device = nullptr;    // meta buffer that already holds sched.items
    std::vector<char> staging;
    // pipelined schedule: second stream, dependency events, cached graph
    hipStream_t stream2 = nullptr;       // inner closure
    int outer_split = 2;                 // R0 launches per outer diagonal, pipelined schedule
                                         // (0: one per storage diagonal)
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
    bpx::Server *server = nullptr;
};

int bpx_grow(bpmax_ctx *ctx, bpx::DevBuf &b, size_t bytes);
int bpx_run_pairs(bpmax_ctx *ctx, int npairs, const int *const *seq1s, const int *Ms, const int *const *seq2s,
                  const int *Ns, const float *intra, const float *inter);
int bpx_gather_scores(bpmax_ctx *ctx, float *scores);
int bpx_collect_timing(bpmax_ctx *ctx);
void bpx_server_shutdown(bpmax_ctx *ctx);
